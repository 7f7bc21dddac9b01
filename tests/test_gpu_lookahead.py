"""Lookahead schedules of the multi-workgroup factor (the next panel factored while the rest of the
trailing update runs; KSH_LOOKAHEAD=1, 2: 8 waves x 1 block, 8 x 2 per workgroup)
against the plain schedule (panel, then the whole update, per step, KSH_LOOKAHEAD=0).  Every
32 x 32 block takes the same arithmetic on the same inputs in all of them, so pivots, solutions
and paths must be bit-identical; the oracle comparisons of the factor itself are in
test_gpu_tiled.py."""
import numpy as np
import pytest

from gsstructuralanalysis_amd import _lib as L
from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu

KW = dict(method=0, max_iter=20, tol=1e-6, tol_f=1e-8, tol_u=1e-6, scaling=0.0, bifurcation_method=0)

# (p, n, ny): band widths from one trailing block column to the general (wide) panel
SHAPES = [(2, 3, None), (3, 21, 7), (3, 40, None), (3, 140, 3)]


@pytest.mark.parametrize("p,n,ny", SHAPES)
@pytest.mark.parametrize("prec", [L.FP64, L.FP32_IR])
def test_lookahead_factor_equals_plain(p, n, ny, prec, ksh, monkeypatch):
    from gsstructuralanalysis_amd.problem import ShellProblem
    monkeypatch.setenv("KSH_BAND_PATH", "tiled")
    sp = ShellProblem(g.roof_problem(p=p, n=n, ny=ny))
    print(f"[case] lookahead roof p={p} n={n} ny={ny} nfree={sp.nfree} bw={sp.bw} precision={prec}", flush=True)
    rng = np.random.default_rng(31)
    u = rng.standard_normal(sp.nfree) * 1e-3
    b = rng.standard_normal(sp.nfree)
    out = []
    for la in ("0", "1", "2", "2"):  # the second run of schedule 2 replays a captured graph
        monkeypatch.setenv("KSH_LOOKAHEAD", la)
        K = sp.tangent(u, precision=prec)
        K.set_inertia_guard(0.0)
        K.factor()
        out.append((K.vector_d(), K.inertia()[0], K.solve(b, max_refine=3 if prec != L.FP64 else 0, tol=1e-12)))
    for d, neg, x in out[1:]:
        assert np.array_equal(d, out[0][0])
        assert neg == out[0][1]
        assert np.array_equal(x, out[0][2])


def test_lookahead_paths_equal_plain(ksh, monkeypatch):
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch
    sps = [ShellProblem(g.imperfect_roof_instance(i, sizes=(8, 16, 24))) for i in range(6)]
    print("[case] lookahead: 6 imperfect roofs, 10 steps, batched", flush=True)
    monkeypatch.setenv("KSH_BAND_PATH", "tiled")
    monkeypatch.setenv("KSH_LOOKAHEAD", "0")
    r0, _ = run_alm_batch(sps, 10, 6e-3, **KW)
    for la in ("1", "2"):
        monkeypatch.setenv("KSH_LOOKAHEAD", la)
        r1, _ = run_alm_batch(sps, 10, 6e-3, **KW)
        for a, b in zip(r0, r1):
            assert np.array_equal(a["lambda"], b["lambda"])
            assert np.array_equal(a["negatives"], b["negatives"])
            assert np.array_equal(a["iterations"], b["iterations"])
