"""Follower pressure on the GPU (ksh_problem_set_pressure) against the FP64 oracle: the force vector
with P(0), both residual conventions with the follower term P(u) - P(0), the tangent's lower band
with -dP/du (FP64 storage, and FP32 storage within FP32 rounding), a load-stepped Newton path and an
arc-length path.  Bars as for the dead-load operators: entries within 1e-11 of max|.| (FP32 tangent:
2e-6), paths within 1e-7 with equal iteration counts and pivot counts.
"""
import numpy as np
import pytest

from gsstructuralanalysis_amd import _lib as L
from gsstructuralanalysis_amd import geometry as g
from oracle.ok import Oracle
from oracle.solvers import run_crisfield, run_static_newton

pytestmark = pytest.mark.gpu

CASES = {
    "roof_p3_n4": lambda: g.roof_problem(p=3, n=4),
    "plate_p2_n5": lambda: g.clamped_plate_problem(p=2, n=5),
    "imperfect_p3": lambda: g.imperfect_roof_instance(3, sizes=(5,)),
}


def _pair(name, pressure=-25.0):
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = CASES[name]()
    pt.fsurf = (0.1, -0.2, -3.0)
    pt.pressure = pressure
    print(f"[case] pressure {pressure} {name} p={pt.p} nx={pt.nx} ny={pt.ny}", flush=True)
    return pt, ShellProblem(pt), Oracle(pt)


def _close(a, b, tol_norm=1e-11):
    scale = np.abs(b).max()
    assert np.abs(a - b).max() <= tol_norm * scale, np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", list(CASES))
def test_pressure_force_and_residuals(name, ksh):
    pt, sp, o = _pair(name)
    _close(sp.force(), o.force(), 1e-13)
    rng = np.random.default_rng(5)
    for scale in (0.0, 1e-3, 3e-2):
        u = rng.standard_normal(o.nfree) * scale
        _close(sp.residual(u, 0.7, conv=L.RES_ALM), o.residual_alm(u, 0.7))
        _close(sp.residual(u, 1.3, conv=L.RES_STATIC), o.residual_static(u, 1.3))


@pytest.mark.parametrize("name", list(CASES))
def test_pressure_tangent(name, ksh):
    pt, sp, o = _pair(name)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(o.nfree) * 2e-2
    ref = o.tangent_band(u)
    _close(sp.tangent(u).download(), ref)
    _close(sp.tangent(u, precision=L.FP32).download(), ref, 2e-6)
    # removing the pressure restores the dead-load operators
    sp.set_pressure(0.0)
    pt.pressure = 0.0
    o0 = Oracle(pt)
    _close(sp.tangent(u).download(), o0.tangent_band(u))
    _close(sp.residual(u, 0.7, conv=L.RES_ALM), o0.residual_alm(u, 0.7))


def test_static_newton_under_follower_pressure(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = g.clamped_plate_problem(p=2, n=8, q=0.0)
    pt.pressure = -1e2
    print("[case] static newton clamped plate p=2 8x8, follower pressure", flush=True)
    ref = run_static_newton(pt, 5, 0.2, max_iter=25, tol=1e-10)
    got = ShellProblem(pt).run_static(5, 0.2, max_iter=25, tol=1e-10, keep_U=True)
    assert got["steps_done"] == ref["steps_done"] == 5
    assert np.abs(got["U"] - ref["U"]).max() <= 1e-7 * np.abs(ref["U"]).max()
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(got["negatives"], ref["negatives"])


def test_arc_length_under_follower_pressure(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = g.roof_problem(p=3, n=6, P=0.0)
    pt.pressure = -2.0
    print("[case] crisfield roof p=3 6x6, follower pressure", flush=True)
    kw = dict(max_iter=20, tol_f=1e-8, tol_u=1e-6, scaling=0.0)
    ref = run_crisfield(pt, 12, 2e-3, **kw)
    got = ShellProblem(pt).run_alm(12, 2e-3, method=0, tol=1e-6, keep_U=True, **kw)
    assert got["steps_done"] == ref["steps_done"] == 12
    assert np.abs(got["lambda"] - ref["lambda"]).max() <= 1e-7 * np.abs(ref["lambda"]).max()
    assert np.abs(got["U"] - ref["U"]).max() <= 1e-7 * np.abs(ref["U"]).max()
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(got["negatives"], ref["negatives"])
    assert (got["residue_f"] < 1e-8).all()
