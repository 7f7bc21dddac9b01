"""Follower pressure in the oracle: P(u) = p sum_g w_g N_A (a1 x a2) on the deformed surface and its
derivative (gsThinShellAssembler::setPressure, examples/gsThinShell_ArcLength.cpp:956-957; the arc-length
residual with it, :997-1005).  Pins that a slipped sign, a transposed cross product or a lost term fails:
  - P is quadratic in u, so central differences equal dP/du to rounding;
  - on a flat plate P(0) is the dead pressure of the same magnitude along +z;
  - the components of sum_A P_A(0) are p times the vector area of the surface, here in closed form;
  - P is unchanged by a rigid translation;
  - the residual's derivative equals the assembled Jacobian's lower band.
"""
import numpy as np
import pytest

from gsstructuralanalysis_amd import geometry as g
from oracle.ok import Oracle, band_to_dense


def _roof(n=4, pressure=-3.0):
    pt = g.roof_problem(p=3, n=n)
    pt.pressure = pressure
    return pt


def test_pressure_tangent_is_the_exact_derivative():
    o = Oracle(_roof())
    rng = np.random.default_rng(1)
    u = 1e-2 * rng.standard_normal(o.ndof)
    Kp = o.pressure_tangent_dense(u)
    h = 1e-3
    for j in rng.choice(o.ndof, 16, replace=False):
        e = np.zeros(o.ndof)
        e[j] = h
        fd = (o.pressure_force(u + e) - o.pressure_force(u - e)) / (2 * h)
        assert np.abs(fd - Kp[:, j]).max() <= 1e-11 * np.abs(Kp).max()
    assert np.abs(Kp - Kp.T).max() > 1e-3 * np.abs(Kp).max()  # a follower load's tangent is not symmetric


def test_flat_plate_pressure_at_rest_is_the_dead_pressure():
    pt = g.clamped_plate_problem(p=3, n=4, q=0.0)
    pt.pressure = 7.5
    P0 = Oracle(pt).pressure_force(np.zeros(pt.ndof))
    dead = Oracle(g.clamped_plate_problem(p=3, n=4, q=-7.5)).external_force()  # fsurf = (0, 0, +7.5)
    assert np.abs(P0 - dead).max() <= 1e-15 * np.abs(dead).max() + 1e-16


def test_total_pressure_force_is_p_times_the_vector_area():
    # the roof x = L xi, y = -2 s eta, z = 2 h eta (1 - eta) (reproduced exactly by the cubic
    # interpolation): a1 x a2 = (0, -2 L h (1 - 2 eta), -2 L s), whose integral is (0, 0, -2 L s)
    R, th, L, p = 2.5, 0.1, 0.5, 2.0
    s = R * np.sin(th)
    o = Oracle(_roof(n=6, pressure=p))
    tot = o.pressure_force(np.zeros(o.ndof)).reshape(-1, 3).sum(0)
    assert np.allclose(tot, [0.0, 0.0, -2.0 * L * s * p], rtol=0, atol=1e-13)


def test_pressure_is_translation_invariant():
    o = Oracle(_roof())
    rng = np.random.default_rng(2)
    u = 1e-2 * rng.standard_normal(o.ndof)
    shift = np.tile([0.3, -0.2, 0.7], o.ndof // 3)
    assert np.abs(o.pressure_force(u + shift) - o.pressure_force(u)).max() <= 1e-14


@pytest.mark.parametrize("lam", [0.0, 0.7])
def test_residual_derivative_is_the_jacobian_lower_band(lam):
    pt = g.roof_problem(p=3, n=3)
    pt.pressure = -40.0
    o = Oracle(pt)
    rng = np.random.default_rng(3)
    u = 1e-3 * rng.standard_normal(o.nfree)
    J = np.zeros((o.nfree, o.nfree))
    h = 1e-6
    for j in range(o.nfree):
        e = np.zeros(o.nfree)
        e[j] = h
        J[:, j] = (o.residual_alm(u + e, lam) - o.residual_alm(u - e, lam)) / (2 * h)
    K = band_to_dense(o.tangent_band(u), o.nfree, o.bw)
    lower = np.tril_indices(o.nfree)
    assert np.abs(J[lower] - K[lower]).max() <= 1e-6 * np.abs(K).max()
    # the residual at rest: the follower part cancels, r(0) = -lam F
    assert np.abs(o.residual_alm(np.zeros(o.nfree), lam) + lam * o.force()).max() <= 1e-14
