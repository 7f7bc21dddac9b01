"""GPU banded LDL^T parity: solves, D vector, inertia, refinement, mat-vec (through the C-ABI).

Checks: D against the oracle's band LDL^T (same ordering, so the pivots agree to rounding),
solution against scipy's dense solve, negative-pivot count against the number of negative
eigenvalues from scipy.linalg.eigvalsh (Sylvester's law of inertia: exact).
"""
import numpy as np
import pytest
import scipy.linalg as sl

from gsstructuralanalysis_amd import geometry as g
from oracle import ok
from oracle.ok import Oracle, band_to_dense

pytestmark = pytest.mark.gpu


def _problem(p=3, n=5, ny=None):
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = g.roof_problem(p=p, n=n, ny=ny)
    print(f"[case] roof p={p} n={n} ny={ny}", flush=True)
    return pt, ShellProblem(pt), Oracle(pt)


@pytest.mark.parametrize("p,n,ny", [(3, 5, None), (2, 6, 2), (3, 2, 9), (3, 9, None)])
def test_factor_solve_and_D(p, n, ny, ksh):
    pt, sp, o = _problem(p, n, ny)
    u = np.random.default_rng(2).standard_normal(o.nfree) * 1e-3
    K = sp.tangent(u)
    K.factor()
    ab = o.tangent_band(u)
    fac = ok.BandLDLT(ab, o.nfree, o.bw)
    d = K.vector_d()
    assert np.abs(d - fac.vectorD()).max() <= 1e-10 * np.abs(fac.vectorD()).max()
    b = np.random.default_rng(3).standard_normal(o.nfree)
    x = K.solve(b)
    A = band_to_dense(ab, o.nfree, o.bw)
    xr = sl.solve(A, b, assume_a="sym")
    assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b) * np.linalg.cond(A) ** 0.5


@pytest.mark.parametrize("k", [0, 1, 7, 40])
def test_inertia_matches_eigenvalues(k, ksh):
    """K - sigma M with sigma between eigenvalues k-1 and k: negative pivots = negative eigenvalues."""
    pt, sp, o = _problem(3, 5)
    from gsstructuralanalysis_amd.problem import BandMatrix
    K = sp.tangent(np.zeros(o.nfree))
    M = sp.mass()
    Kd = band_to_dense(K.download(), o.nfree, o.bw)
    Md = band_to_dense(M.download(), o.nfree, o.bw)
    ev = sl.eigh(Kd, Md, eigvals_only=True)
    sigma = 0.5 * ev[0] if k == 0 else 0.5 * (ev[k - 1] + ev[k])
    A = BandMatrix(sp)
    A.upload(K.download())
    A.axpby(1.0, -sigma, M)
    neg, dmin = A.inertia()
    ref = int(np.sum(np.linalg.eigvalsh(Kd - sigma * Md) < 0))
    assert neg == ref
    assert (dmin < 0) == (ref > 0)


@pytest.mark.parametrize("mode", [1, 2])
def test_fp32_factor_with_refinement(mode, ksh):
    """mode 1: FP64 K, FP32 LDL^T, FP64 refinement residuals -> the FP64 solution of K64;
    mode 2: K rounded to FP32 -> the FP64-accurate solution of K32."""
    pt, sp, o = _problem(3, 6)
    u = np.random.default_rng(4).standard_normal(o.nfree) * 1e-3
    K32 = sp.tangent(u, precision=mode)
    K32.factor()
    b = np.random.default_rng(5).standard_normal(o.nfree)
    A32 = band_to_dense(K32.download(), o.nfree, o.bw)  # the stored K values
    if mode == 1:
        assert np.abs(K32.download() - o.tangent_band(u)).max() <= 1e-11 * np.abs(o.tangent_band(u)).max()
    xr = sl.solve(A32, b, assume_a="sym")
    x0 = K32.solve(b, max_refine=0)
    x = K32.solve(b, max_refine=20, tol=1e-13)
    e0 = np.linalg.norm(x0 - xr) / np.linalg.norm(xr)
    e = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    assert e <= 1e-9 and e < e0
    neg32, _ = K32.inertia()
    K64 = sp.tangent(u)
    neg64, _ = K64.inertia()
    assert neg32 == neg64


def test_matvec_and_axpby(ksh):
    pt, sp, o = _problem(3, 4)
    u = np.random.default_rng(6).standard_normal(o.nfree) * 1e-3
    K = sp.tangent(u)
    x = np.random.default_rng(7).standard_normal(o.nfree)
    A = band_to_dense(o.tangent_band(u), o.nfree, o.bw)
    assert np.linalg.norm(K.matvec(x) - A @ x) <= 1e-12 * np.linalg.norm(A @ x)
    M = sp.mass()
    K.axpby(2.0, -3.0, M)
    ref = 2.0 * o.tangent_band(u) - 3.0 * o.mass_band()
    assert np.abs(K.download() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_zero_pivot_reports_solver_error(ksh):
    from gsstructuralanalysis_amd import _lib
    from gsstructuralanalysis_amd.problem import BandMatrix
    pt, sp, o = _problem(2, 3)
    A = BandMatrix(sp)
    ab = o.tangent_band(np.zeros(o.nfree)).reshape(o.nfree, o.bw + 1)
    ab[0, :] = 0.0  # zero first column -> zero pivot
    A.upload(ab.ravel())
    with pytest.raises(_lib.KshError) as e:
        A.factor()
    assert e.value.code == _lib.SOLVER_ERROR
