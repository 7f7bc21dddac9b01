"""GPU parity of the assembly kernels against the FP64 oracle (through the C-ABI).

Tolerances (stated by the FP64 plain-form bar): K and R entries within 1e-11 of max|.|
(norm-wise), plus 1e-9 relative on every entry above 1e-6 max|.|.  The FP32 tangent
(FP32 element GEMMs on MFMA, FP64 accumulation in the band gather) is held to 2e-6 of max|K|.
"""
import numpy as np
import pytest

from gsstructuralanalysis_amd import geometry as g
from oracle.ok import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "roof_p3_n4": lambda: g.roof_problem(p=3, n=4),
    "roof_p3_n6x3": lambda: g.roof_problem(p=3, n=6, ny=3),
    "plate_p2_n5": lambda: g.clamped_plate_problem(p=2, n=5),
    "roof_p2_n3": lambda: g.roof_problem(p=2, n=3),
    "imperfect_p3": lambda: g.imperfect_roof_instance(3, sizes=(5,)),
}


def _pair(name):
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = CASES[name]()
    pt.fsurf = (0.1, -0.2, -3.0)
    print(f"[case] {name} p={pt.p} nx={pt.nx} ny={pt.ny}", flush=True)
    return pt, ShellProblem(pt), Oracle(pt)


def _close(a, b, tol_norm=1e-11, tol_rel=1e-9):
    scale = np.abs(b).max()
    err = np.abs(a - b)
    assert err.max() <= tol_norm * scale, (err.max() / scale)
    big = np.abs(b) > 1e-6 * scale
    assert (err[big] / np.abs(b[big])).max() <= tol_rel


@pytest.mark.parametrize("name", list(CASES))
def test_force_and_bandwidth(name, ksh):
    pt, sp, o = _pair(name)
    assert sp.nfree == o.nfree and sp.bw == o.bw
    _close(sp.force(), o.force(), 1e-13, 1e-12)


@pytest.mark.parametrize("name", list(CASES))
def test_residual_parity(name, ksh):
    pt, sp, o = _pair(name)
    rng = np.random.default_rng(5)
    for scale in (0.0, 1e-3, 3e-2):
        u = rng.standard_normal(o.nfree) * scale
        _close(sp.residual(u, 0.7, conv=1), o.residual_alm(u, 0.7))
        _close(sp.residual(u, 1.3, conv=0), o.residual_static(u, 1.3))


@pytest.mark.parametrize("name", list(CASES))
def test_tangent_parity_fp64(name, ksh):
    pt, sp, o = _pair(name)
    rng = np.random.default_rng(7)
    for scale in (0.0, 1e-3, 3e-2):
        u = rng.standard_normal(o.nfree) * scale
        K = sp.tangent(u).download()
        _close(K, o.tangent_band(u))


@pytest.mark.parametrize("name", ["roof_p3_n4", "plate_p2_n5"])
def test_tangent_fp32(name, ksh):
    pt, sp, o = _pair(name)
    u = np.random.default_rng(8).standard_normal(o.nfree) * 1e-2
    K32 = sp.tangent(u, precision=2).download()
    ref = o.tangent_band(u)
    assert np.abs(K32 - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["roof_p3_n4", "plate_p2_n5"])
def test_mass_parity(name, ksh):
    pt, sp, o = _pair(name)
    _close(sp.mass().download(), o.mass_band())


def test_assembly_is_deterministic(ksh):
    pt, sp, o = _pair("roof_p3_n6x3")
    u = np.random.default_rng(9).standard_normal(o.nfree) * 1e-2
    K1 = sp.tangent(u).download()
    K2 = sp.tangent(u).download()
    R1 = sp.residual(u, 0.3)
    R2 = sp.residual(u, 0.3)
    assert np.array_equal(K1, K2) and np.array_equal(R1, R2)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_reassembly_over_written_band(prec, ksh):
    """A band assembled afresh, re-assembled in place, and re-assembled after upload / axpby left
    values in slots no element reaches holds the same bits each time; those slots are exactly 0,
    as in the oracle's band."""
    pt, sp, o = _pair("roof_p3_n6x3")
    rng = np.random.default_rng(10)
    u0, u1 = (rng.standard_normal(o.nfree) * 1e-2 for _ in range(2))
    ref = sp.tangent(u1, precision=prec).download()
    unreached = o.tangent_band(u1) == 0.0
    assert unreached.mean() > 0.3 and np.all(ref[unreached] == 0.0)
    K = sp.tangent(u0, precision=prec)
    assert np.array_equal(sp.tangent(u1, out=K).download(), ref)
    K.upload(np.ones(K.n * (K.bw + 1)))
    assert np.array_equal(sp.tangent(u1, out=K).download(), ref)
    K.axpby(1.0, 2.0, sp.tangent(u0, precision=prec))
    assert np.array_equal(sp.tangent(u1, out=K).download(), ref)
    assert np.array_equal(sp.tangent(u1, out=K).download(), ref)
