"""The oracle's singular-point routines (gsALMBase.hpp:422-511, 631-856) pinned on two-dof systems
whose singular points are known in closed form: a symmetric (pitchfork) bifurcation and the
limit point of the shallow truss.  The GPU engine is compared with these routines in
test_gpu_singular.py."""
import numpy as np
import pytest

from oracle.ok import BandLDLT
from oracle.solvers import run_crisfield
from test_oracle_pins import TrussOps


class PitchforkOps:
    """Energy k1 u1^2 / 2 + k2 u2^2 / 2 + a u1 u2^2 + b u2^4 / 4, load F = (P, 0).
    Primary path u2 = 0, u1 = lam P / k1; it loses stability where k2 + 2 a u1 = 0, i.e. at
    u1c = -k2 / (2 a), lam_c = k1 u1c / P, a symmetric bifurcation with the critical mode (0, 1);
    the secondary branch is k2 + 2 a u1 + b u2^2 = 0, lam P = k1 u1 + a u2^2 (stable for
    2 b > 4 a^2).  Same interface as OracleOps (band storage, bw = 1)."""

    def __init__(self, k1=1.0, k2=0.5, a=-0.5, b=1.0, P=1.0):
        self.k1, self.k2, self.a, self.b = k1, k2, a, b
        self.n, self.bw = 2, 1
        self.F = np.array([P, 0.0])
        self.n_jac = self.n_res = self.n_fact = self.n_solve = 0

    @property
    def u1c(self):
        return -self.k2 / (2 * self.a)

    @property
    def lam_c(self):
        return self.k1 * self.u1c / self.F[0]

    def fint(self, u):
        u1, u2 = u
        return np.array([self.k1 * u1 + self.a * u2 * u2, self.k2 * u2 + 2 * self.a * u1 * u2 + self.b * u2 ** 3])

    def jacobian(self, u):
        self.n_jac += 1
        u1, u2 = u
        return np.array([self.k1, 2 * self.a * u2, self.k2 + 2 * self.a * u1 + 3 * self.b * u2 * u2, 0.0])

    def residual_alm(self, u, lam):
        self.n_res += 1
        return self.fint(u) - lam * self.F

    def factor(self, ab):
        self.n_fact += 1
        return BandLDLT(ab, self.n, self.bw)

    def solve(self, fac, b):
        self.n_solve += 1
        return fac.solve(b)


OPTS = dict(max_iter=20, tol_f=1e-10, tol_u=1e-10, scaling=0.0)


def _first(sing, kind):
    idx = np.flatnonzero(sing == kind)
    assert idx.size >= 1, sing
    return int(idx[0])


def test_bifurcation_point_and_branch_switch():
    ops = PitchforkOps()
    r = run_crisfield(None, 30, 0.04, ops=ops, singular_point=2, perturbation=1e3, **OPTS)
    assert r["steps_done"] == 30
    k = _first(r["singular"], 2)
    assert (r["singular"] != 0).sum() == 1
    # every step before it on the primary path, below the critical load
    assert np.all(r["U"][:k, 1] == 0) and np.all(r["lambda"][:k] < ops.lam_c)
    # the computed singular point is the bifurcation point, moved by 1 / tau along the mode (0, +-1)
    u = r["U"][k]
    assert abs(u[0] - ops.u1c) <= 1e-10 and abs(abs(u[1]) - 1e-3) <= 1e-10
    assert abs(r["lambda"][k] - ops.lam_c) <= 1e-10
    # then the path follows the secondary branch (stable: no negative pivot) away from the primary one
    for u, lam in zip(r["U"][k + 1:], r["lambda"][k + 1:]):
        assert abs(u[1]) > 2e-3
        assert abs(ops.k2 + 2 * ops.a * u[0] + ops.b * u[1] ** 2) <= 1e-9
        assert abs(ops.k1 * u[0] + ops.a * u[1] ** 2 - lam * ops.F[0]) <= 1e-9
    assert np.all(r["negatives"][k + 1:] == 0)
    assert r["lambda"][-1] > ops.lam_c + 1e-3  # stable-symmetric: the load rises on the branch


def test_bifurcation_point_by_bisection_then_extended_system():
    ops = PitchforkOps()
    r = run_crisfield(None, 17, 0.03, ops=ops, singular_point=2, sp_tol_b=1e-2, **OPTS)
    assert r["steps_done"] == 17
    k = _first(r["singular"], 2)
    assert k == 16
    u = r["U"][k]
    assert abs(u[0] - ops.u1c) <= 1e-10 and abs(abs(u[1]) - 1e-3) <= 1e-10
    assert abs(r["lambda"][k] - ops.lam_c) <= 1e-10
    # the pivot count is that of the bisection's last point, just past the critical load (the
    # stability bookkeeping then follows the bisection's steps, so later steps are not asserted)
    assert r["negatives"][k] == 1


def test_singular_point_without_switch_stops_on_the_bifurcation_point():
    ops = PitchforkOps()
    ref = run_crisfield(None, 30, 0.04, ops=ops, singular_point=2, **OPTS)
    k = _first(ref["singular"], 2)
    r = run_crisfield(None, k + 1, 0.04, ops=PitchforkOps(), singular_point=1, **OPTS)
    assert r["steps_done"] == k + 1 and r["singular"][k] == 2
    assert abs(r["U"][k, 0] - ops.u1c) <= 1e-10 and r["U"][k, 1] == 0.0
    assert abs(r["lambda"][k] - ops.lam_c) <= 1e-10


def test_limit_point_is_not_a_bifurcation():
    """At the truss's limit point the critical mode is not orthogonal to the load: the test says
    'limit point', the driver's computeSingularPoint sets the solution back to the previous point
    (gsALMBase.hpp:425-428), and the path goes on through the limit point."""
    ops = TrussOps()
    r = run_crisfield(None, 60, 0.01, ops=ops, singular_point=1, max_iter=20, tol_f=1e-10, tol_u=1e-10)
    plain = run_crisfield(None, 60, 0.01, ops=TrussOps(), max_iter=20, tol_f=1e-10, tol_u=1e-10)
    assert r["steps_done"] == 60
    k = _first(r["singular"], 1)
    assert (r["singular"] == 2).sum() == 0
    assert np.array_equal(r["U"][k], r["U"][k - 1]) and r["lambda"][k] == r["lambda"][k - 1]
    # the steps before it are those of the plain path; past it the pivot count stays 1 up to the
    # re-stiffening, as on the plain path
    assert np.array_equal(r["lambda"][:k], plain["lambda"][:k])
    assert r["negatives"][k + 1] == 1 and r["U"][:, 0].max() > ops.h * (1 - 1 / np.sqrt(3))


@pytest.mark.parametrize("tau", [1e2, 1e4])
def test_perturbation_size(tau):
    ops = PitchforkOps()
    r = run_crisfield(None, 16, 0.04, ops=ops, singular_point=2, perturbation=tau, **OPTS)
    k = _first(r["singular"], 2)
    assert abs(abs(r["U"][k, 1]) - 1.0 / tau) <= 1e-10
