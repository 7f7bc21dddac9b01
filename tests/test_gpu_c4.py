"""Eigen-analyses at the full size of the large plate: degree 3, 128 x 128 elements (49,411 free
dofs, band width 1166 on the free dofs), C0 plate with a 0.05 sin(pi x) sin(pi y) initial deflection, density 1e3,
10 lowest pairs (solvers/modal_shell_XML.cpp, solvers/buckling_shell_XML.cpp;
src/gsEigenSolvers/gsModalSolver.h, gsBucklingSolver.h).

Checks:
  * the 10 eigenvalues against ARPACK (scipy eigsh, shift-invert with the oracle's band LDL^T) on
    the oracle's K, M, K_NL at the same size, committed in tests/golden/c4_eigen.json
    (tests/golden/make_c4_eigen.py): relative error <= 1e-8 each, FP64 and FP32-factor forms;
  * Sturm counts: the number of negative pivots of K - sigma M (multi-workgroup FP64 factor) at
    sigma between consecutive computed eigenvalues is the number of eigenvalues below sigma
    (Sylvester), so none is missed and their order is right;
  * eigenpair residuals ||A x - lambda B x|| / ||A x||;
  * the FP32-factor form (FP64 refinement) against the FP64 form, within 1e-8.
The same solver is pinned against dense FP64 eigenvalues of the oracle's matrices at small sizes
in test_gpu_eigen.py.
"""
import json
import os
import time

import numpy as np
import pytest

from gsstructuralanalysis_amd import _lib as L
from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with open(os.path.join(HERE, "golden", "c4_eigen.json")) as f:
        return json.load(f)


def _check_golden(name, got, ref):
    got, ref = np.sort(np.asarray(got)[:len(ref)]), np.sort(np.asarray(ref))
    err = np.abs(got - ref) / np.abs(ref)
    print(f"[c4] {name}: max relative deviation from ARPACK {err.max():.2e}", flush=True)
    assert err.max() <= 1e-8, (got, ref)


def _z0(x, y):
    return 0.05 * np.sin(np.pi * x) * np.sin(np.pi * y)


@pytest.fixture(scope="module")
def plate(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = g.clamped_plate_problem(p=3, n=128, z0=_z0, rho=1e3)
    sp = ShellProblem(pt)
    print(f"[case] large plate p=3 n=128 nfree={sp.nfree} bw={sp.bw}", flush=True)
    assert sp.bw == 1166 and sp.nfree == 49411  # 1190 before the clamped edges are eliminated
    return sp


def _kernel_times(lib):
    import ctypes as C
    out = []
    for nm in ("band_ldlt_tiled", "band_solve_tiled", "band_symv", "band_ldlt"):
        cnt, ms = C.c_int64(), C.c_double()
        lib.ksh_profile_query(nm.encode(), C.byref(cnt), C.byref(ms))
        if cnt.value:
            out.append(f"{nm} {cnt.value} x {ms.value / cnt.value:.2f} ms")
    return ", ".join(out)


def test_large_plate_modal(plate):
    sp = plate
    n = sp.nfree
    lib = L.load()
    for prec in (L.FP64, L.FP32_IR):  # per-kernel times of each form
        lib.ksh_profile_reset()
        lib.ksh_profile_enable(1)
        t0 = time.perf_counter()
        r = sp.run_modal(10, precision=prec, tol=1e-10)
        dt = time.perf_counter() - t0
        lib.ksh_profile_enable(0)
        print(f"[c4] modal precision {prec}: {dt:.2f} s, {r['solves']} solves; {_kernel_times(lib)}", flush=True)
        if prec == L.FP32_IR:
            r32 = r
    t0 = time.perf_counter()
    r64 = sp.run_modal(12, tol=1e-10, keep_vectors=True)  # 12: the Sturm cuts need the gap above the tenth
    t64 = time.perf_counter() - t0
    t32 = dt
    print(f"[c4] modal FP64 {t64:.2f} s ({r64['solves']} solves, {r64['restarts']} restarts), "
          f"FP32 factor {t32:.2f} s ({r32['solves']} solves); omega^2 = {r64['values']}", flush=True)
    assert r64["converged"] == 12 and r32["converged"] == 10
    lam = r64["values"]
    assert np.all(np.diff(lam) >= 0) and lam[0] > 0
    assert np.abs(r32["values"] - lam[:10]).max() <= 1e-8 * np.abs(lam[:10]).max()
    fx = _golden()
    _check_golden("modal FP64", lam[:10], fx["modal_omega2"])
    _check_golden("modal FP32 factor", r32["values"], fx["modal_omega2"])
    z = np.zeros(n)
    K = sp.tangent(z)
    M = sp.mass()
    for lv, x in zip(lam, r64["vectors"]):
        kx = K.matvec(x)
        assert np.linalg.norm(kx - lv * M.matvec(x)) <= 1e-6 * np.linalg.norm(kx)
    # Sturm counts between distinct computed eigenvalues (double pairs of the symmetric plate
    # have no cut between them), below the first, and past the tenth
    cuts = [(0, 0.5 * lam[0])]
    cuts += [(k + 1, 0.5 * (lam[k] + lam[k + 1])) for k in range(11) if lam[k + 1] - lam[k] > 1e-6 * lam[k + 1]]
    assert max(c for c, _ in cuts) >= 10
    for count, sigma in cuts:
        A = sp.tangent(z)
        A.axpby(1.0, -sigma, M)
        A.factor()
        neg, _ = A.inertia()
        assert neg == count, (sigma, neg, count)


def test_large_plate_buckling(plate):
    sp = plate
    n = sp.nfree
    t0 = time.perf_counter()
    r64 = sp.run_buckling(10, tol=1e-10, keep_vectors=True)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    r32 = sp.run_buckling(10, precision=L.FP32_IR, tol=1e-10)
    t32 = time.perf_counter() - t0
    print(f"[c4] buckling FP64 {t64:.2f} s ({r64['solves']} solves), FP32 factor {t32:.2f} s; "
          f"lambda = {r64['values']}", flush=True)
    assert r64["converged"] == 10 and r32["converged"] == 10
    lam = r64["values"]
    assert np.abs(r32["values"] - lam).max() <= 1e-8 * np.abs(lam).max()
    fx = _golden()
    _check_golden("buckling FP64", lam, fx["buckling_mu"])
    _check_golden("buckling FP32 factor", r32["values"], fx["buckling_mu"])
    z = np.zeros(n)
    KL = sp.tangent(z)
    KL.factor()
    u = KL.solve(sp.force())
    B = sp.tangent(u)
    B.axpby(1.0, -1.0, KL)  # K_NL - K_L
    for lv, x in zip(lam, r64["vectors"]):
        ax = KL.matvec(x)
        assert np.linalg.norm(ax - lv * B.matvec(x)) <= 1e-6 * np.linalg.norm(ax)
