"""GPU band solves (forward, diagonal, backward substitution) against the oracle's band LDL^T solve.

The shapes cover the kernel's cases: order below one 64-row block and band narrower than the
block, orders that are not multiples of 64, the widest bands of the batch workload (bw 470), and
a band wider than 512, which takes the multi-workgroup solve.  FP64 factors must reproduce the oracle's
solution to rounding; FP32 factors (FP64 K) must reach it through FP64 refinement.
"""
import numpy as np
import pytest

from gsstructuralanalysis_amd import geometry as g
from oracle import ok
from oracle.ok import Oracle

pytestmark = pytest.mark.gpu

CASES = [(2, 3, None), (3, 5, None), (3, 16, None), (3, 21, 7), (3, 48, 5), (3, 56, 4)]


@pytest.mark.parametrize("p,n,ny", CASES)
@pytest.mark.parametrize("prec", [0, 1])
def test_solve_matches_oracle(p, n, ny, prec, ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = g.roof_problem(p=p, n=n, ny=ny)
    o = Oracle(pt)
    print(f"[case] solve roof p={p} n={n} ny={ny} nfree={o.nfree} bw={o.bw} prec={prec}", flush=True)
    sp = ShellProblem(pt)
    u = np.random.default_rng(11).standard_normal(o.nfree) * 1e-3
    K = sp.tangent(u, precision=prec)
    K.factor()
    fac = ok.BandLDLT(o.tangent_band(u), o.nfree, o.bw)
    rng = np.random.default_rng(12)
    for _ in range(2):  # two solves with one factor
        b = rng.standard_normal(o.nfree)
        xr = fac.solve(b)
        if prec == 0:
            x = K.solve(b)
            # forward error: rounding of another summation order (the multi-workgroup solve
            # above bw 383 updates column-wise), grown by cond(K); backward error: rounding
            Kb = o.tangent_band(u)
            r = b - ok.band_symv(Kb, o.nfree, o.bw, x)
            assert np.linalg.norm(r) <= 1e-14 * np.abs(Kb).max() * np.sqrt(o.bw) * np.linalg.norm(x)
            assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)
        else:
            x0 = K.solve(b, max_refine=0)
            x = K.solve(b, max_refine=20, tol=1e-14)
            e0 = np.linalg.norm(x0 - xr) / np.linalg.norm(xr)
            e = np.linalg.norm(x - xr) / np.linalg.norm(xr)
            assert e0 < 0.5  # an FP32 factor's solution (cond(K) * 6e-8)
            assert e <= 1e-10 and e < e0
