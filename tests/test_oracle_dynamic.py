"""The oracle's dynamic relaxation (gsStaticDR.hpp) and nonlinear Newmark
(gsDynamicNewmark.hpp:69-142) pinned by what the mathematics fixes: the average-acceleration
rule on a linear oscillator is the trapezoidal rule, whose discrete solution is U0 cos(n theta)
with theta = 2 atan(omega dt / 2), and whose energy changes by exactly -dt Vbar^T C Vbar per step;
dynamic relaxation stops at the static equilibrium K^-1 F (linear) or at Newton's solution
(the shell).  The GPU engines are compared with these routines in test_gpu_dynamic.py."""
import numpy as np
import pytest

from gsstructuralanalysis_amd import geometry as g
from oracle.ok import BandLDLT, band_symv, dense_to_band
from oracle.solvers import run_dynamic_relaxation, run_newmark, run_static_newton


class LinearOps:
    """A linear spring-mass system in band storage: R(u) = lam F - K u, constant K and M."""

    def __init__(self, K, M, F):
        self.n = K.shape[0]
        self.bw = 1
        self.Kb, self.Mb = dense_to_band(K, 1), dense_to_band(M, 1)
        self.K, self.M, self.F = K, M, np.asarray(F, dtype=np.float64)
        self.n_jac = self.n_res = 0

    def jacobian(self, u):
        self.n_jac += 1
        return self.Kb.copy()

    def residual_static(self, u, lam):
        self.n_res += 1
        return lam * self.F - self.K @ u

    def mass(self):
        return self.Mb.copy()

    def factor(self, ab):
        return BandLDLT(ab, self.n, self.bw)

    def solve(self, fac, b):
        return fac.solve(b)


def _diag_ops(F=(0.0, 0.0)):
    return LinearOps(np.diag([4.0, 9.0]), np.diag([1.0, 2.0]), F)


def test_newmark_average_acceleration_is_the_trapezoidal_rule():
    ops = _diag_ops()
    U0 = np.array([1.0, 0.5])
    A0 = -np.linalg.solve(ops.M, ops.K @ U0)
    dt, n = 0.1, 200
    r = run_newmark(None, n, dt, ops=ops, U0=U0, A0=A0, tol_f=1e-12, tol_u=1e-12)
    assert r["steps_done"] == n
    omega = np.sqrt(np.diag(ops.K) / np.diag(ops.M))
    theta = 2 * np.arctan(omega * dt / 2)
    k = np.arange(1, n + 1)[:, None]
    exact = U0[None, :] * np.cos(k * theta[None, :])
    assert np.abs(r["U"] - exact).max() <= 1e-12
    # a slip in the predictor weights (e.g. beta for 0.5 - beta) breaks this: the period of the
    # discrete solution differs from the exact one by the known elongation only
    assert np.abs(r["U"][:, 0] - U0[0] * np.cos(k[:, 0] * omega[0] * dt)).max() > 1e-3


def test_newmark_energy_balance_with_mass_damping():
    K = np.array([[6.0, -2.0], [-2.0, 4.0]])
    M = np.array([[2.0, 0.5], [0.5, 1.0]])
    ops = LinearOps(K, M, (0.0, 0.0))
    U0 = np.array([1.0, -0.3])
    c = 0.4
    r = run_newmark(None, 80, 0.05, ops=ops, U0=U0, A0=-np.linalg.solve(M, K @ U0), damping_mass=c, tol_f=1e-13,
                    tol_u=1e-13)
    U = np.vstack([U0, r["U"]])
    V = np.vstack([np.zeros(2), r["V"]])
    E = 0.5 * np.einsum("ki,ij,kj->k", V, M, V) + 0.5 * np.einsum("ki,ij,kj->k", U, K, U)
    Vb = 0.5 * (V[1:] + V[:-1])
    loss = 0.05 * c * np.einsum("ki,ij,kj->k", Vb, M, Vb)
    assert np.abs(np.diff(E) + loss).max() <= 1e-12 * E[0]
    assert E[-1] < 0.5 * E[0]


def test_dynamic_relaxation_reaches_the_linear_equilibrium():
    F = np.array([1.0, -2.0])
    ops = _diag_ops(F)
    r = run_dynamic_relaxation(None, ops=ops, mass=np.diag(ops.M), max_it=2000, tol=1e-10, tol_e=1e-12)
    assert r["converged"]
    assert np.abs(r["U"] - np.linalg.solve(ops.K, F)).max() <= 1e-8


def dr_mass_coefficient(K_band, n, bw, mass, safety=1.5):
    """The fictitious-mass factor that keeps the explicit dt = 1 steps stable: alpha >= lambda_max(K,
    M) / 4, bounded by Gershgorin row sums of M^-1 K (the reference takes alpha as an input)."""
    rows = band_symv(np.abs(K_band), n, bw, np.ones(n))
    return safety * float(np.max(rows / mass)) / 4.0


def test_dynamic_relaxation_on_the_plate_matches_newton():
    """Clamped plate under a load that bends it well past its thickness: dynamic relaxation with
    kinetic damping (energy peaks reset the velocities) stops at Newton's equilibrium."""
    from oracle.ok import Oracle
    from oracle.solvers import OracleOps, lumped_mass
    pt = g.clamped_plate_problem(p=2, n=4, q=50.0)
    newton = run_static_newton(pt, 5, 0.2, max_iter=25, tol=1e-10)
    assert newton["steps_done"] == 5
    ops = OracleOps(pt)
    m = lumped_mass(ops)
    alpha = dr_mass_coefficient(Oracle(pt).tangent_band(np.zeros(ops.n)), ops.n, ops.bw, m)
    r = run_dynamic_relaxation(pt, damping=0.0, alpha=alpha, max_it=20000, tol=1e-9, tol_e=1e-12)
    assert r["converged"], r["iterations"]
    assert r["peaks"] > 0
    ref = newton["U"][-1]
    print(r["iterations"], r["peaks"], np.abs(ref).max())
    assert np.abs(r["U"] - ref).max() <= 1e-6 * np.abs(ref).max()
    assert np.abs(ref).max() > pt.t
