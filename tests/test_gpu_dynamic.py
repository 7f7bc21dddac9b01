"""Dynamic relaxation and nonlinear Newmark on the GPU against the oracle's restatements
(gsStaticDR.hpp:93-266; gsDynamicNewmark.hpp:69-142 under tutorials/nonlinear_shell_dynamic.cpp),
which test_oracle_dynamic.py pins in closed form.  The case is the small clamped plate
(degree 2, 4 x 4 elements, 36 free dofs): its fundamental period T = 0.01582 comes from the
oracle's K and M."""
import numpy as np
import pytest
import scipy.linalg as sl

from gsstructuralanalysis_amd import _lib as L
from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu

T1 = 0.0158188184560007  # fundamental period of the 4 x 4 plate (oracle K, M)
NEWMARK = dict(tol_f=1e-6, tol_u=1e-8)  # Newton updates run 2e-5, then 5e-12; the residual floor is ~2e-10


def _dr_setup(pt):
    from oracle.ok import Oracle, band_to_dense
    from oracle.solvers import OracleOps, lumped_mass
    from test_oracle_dynamic import dr_mass_coefficient
    o = Oracle(pt)
    ops = OracleOps(pt)
    m = lumped_mass(ops)
    Kb = o.tangent_band(np.zeros(o.nfree))
    alpha = dr_mass_coefficient(Kb, o.nfree, o.bw, m)
    lam1 = sl.eigh(band_to_dense(Kb, o.nfree, o.bw), np.diag(m), eigvals_only=True)[0]
    return alpha, lam1


@pytest.mark.parametrize("kind", ["kinetic", "viscous"])
def test_dynamic_relaxation_against_oracle(kind, ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    from oracle.solvers import run_dynamic_relaxation, run_static_newton
    q = 50.0 if kind == "kinetic" else 1.0
    pt = g.clamped_plate_problem(p=2, n=4, q=q)
    alpha, lam1 = _dr_setup(pt)
    c = 0.0 if kind == "kinetic" else 0.9 * 2 * np.sqrt(alpha * lam1)
    print(f"[case] dynamic relaxation, clamped plate p=2 4x4, q={q}, {kind} damping (c={c:.4g}), alpha={alpha:.4g}",
          flush=True)
    kw = dict(damping=c, alpha=alpha, max_iter=20000, tol=1e-9, tol_e=1e-12)
    ref = run_dynamic_relaxation(pt, damping=c, alpha=alpha, max_it=20000, tol=1e-9, tol_e=1e-12)
    got = ShellProblem(pt).run_dr(**kw)
    print(f"[dr {kind}] GPU {got['iterations']} iterations, {got['peaks']} peaks; oracle {ref['iterations']}, "
          f"{ref['peaks']}", flush=True)
    assert got["converged"] and ref["converged"]
    assert abs(got["iterations"] - ref["iterations"]) <= 0.02 * ref["iterations"]
    scale = np.abs(ref["U"]).max()
    assert np.abs(got["U"] - ref["U"]).max() <= 1e-7 * scale
    # Newton's relative residual floor at this small load is about 1e-10 (tol 1e-9 there)
    newton = run_static_newton(pt, 5, 0.2, max_iter=25, tol=1e-10 if q > 10 else 1e-9)
    assert newton["steps_done"] == 5
    assert np.abs(got["U"] - newton["U"][-1]).max() <= 1e-6 * scale
    if kind == "kinetic":
        assert got["peaks"] > 0 and abs(got["peaks"] - ref["peaks"]) <= 0.02 * ref["peaks"]
    n = len(got["ek_history"])
    assert np.allclose(got["ek_history"][:50], ref["ek"][:50], rtol=1e-9, atol=0)
    assert n == got["iterations"] + 1


def test_dynamic_relaxation_batch_equals_single_runs(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem, run_dr_batch
    pts = [g.clamped_plate_problem(p=2, n=4, q=q) for q in (10.0, 50.0)] + [g.clamped_plate_problem(p=2, n=6, q=20.0)]
    alphas = [_dr_setup(p)[0] for p in pts]
    alpha = max(alphas)
    sps = [ShellProblem(p) for p in pts]
    print("[case] dynamic relaxation batch of three plates, kinetic damping", flush=True)
    kw = dict(damping=0.0, alpha=alpha, max_iter=20000, tol=1e-9, tol_e=1e-12)
    batch, _ = run_dr_batch(sps, keep_U=True, **kw)
    for sp, b in zip(sps, batch):
        one = sp.run_dr(**kw)
        assert one["converged"] and b["converged"]
        assert one["iterations"] == b["iterations"] and one["peaks"] == b["peaks"]
        assert np.array_equal(one["U"], b["U"])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_newmark_step_load_against_oracle(prec, ksh):
    """A suddenly applied small load: the GPU trajectory equals the oracle's, and the centre
    overshoots to about twice its static deflection (undamped, average acceleration)."""
    from gsstructuralanalysis_amd.problem import ShellProblem
    from oracle.ok import Oracle, band_to_dense
    from oracle.solvers import run_newmark
    pt = g.clamped_plate_problem(p=2, n=4, q=1.0)
    print(f"[case] Newmark, clamped plate p=2 4x4, step load q=1, dt = T/20, 30 steps, {prec}", flush=True)
    ref = run_newmark(pt, 30, T1 / 20, **NEWMARK)
    extra = dict(precision=L.FP64) if prec == "fp64" else dict(precision=L.FP32_IR, max_refine=4, refine_tol=1e-12)
    got = ShellProblem(pt).run_newmark(30, dt=T1 / 20, **NEWMARK, **extra)
    assert got["steps_done"] == ref["steps_done"] == 30
    assert np.array_equal(got["iterations"], ref["iterations"])
    for key in ("U", "V", "A"):
        err = np.abs(got[key] - ref[key]).max() / np.abs(ref[key]).max()
        print(f"[newmark {prec}] {key}: max relative deviation {err:.2e}", flush=True)
        assert err <= 1e-8
    o = Oracle(pt)
    us = np.linalg.solve(band_to_dense(o.tangent_band(np.zeros(o.nfree)), o.nfree, o.bw), o.force())
    c = int(np.argmax(np.abs(us)))
    amp = np.abs(got["U"][:, c]).max() / abs(us[c])
    print(f"[newmark {prec}] dynamic amplification {amp:.4f}", flush=True)
    assert 1.9 <= amp <= 2.02


def test_newmark_nonlinear_against_oracle(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    from oracle.solvers import run_newmark
    pt = g.clamped_plate_problem(p=2, n=4, q=3e2)
    print("[case] Newmark, clamped plate p=2 4x4, q=300 (deflection several thicknesses), 20 steps, mass damping",
          flush=True)
    ref = run_newmark(pt, 20, T1 / 20, damping_mass=50.0, **NEWMARK)
    got = ShellProblem(pt).run_newmark(20, dt=T1 / 20, damping_mass=50.0, precision=L.FP64, **NEWMARK)
    assert got["steps_done"] == ref["steps_done"] == 20
    assert np.abs(got["iterations"] - ref["iterations"]).max() <= 1
    err = np.abs(got["U"] - ref["U"]).max() / np.abs(ref["U"]).max()
    print(f"[newmark nonlinear] U max relative deviation {err:.2e}; max |U| {np.abs(ref['U']).max():.3e}", flush=True)
    assert err <= 1e-7
    assert np.abs(ref["U"]).max() > 2 * pt.t


def test_newmark_batch_equals_single_runs(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem, run_newmark_batch
    pts = [g.clamped_plate_problem(p=2, n=4, q=q) for q in (1.0, 3e2)] + [g.clamped_plate_problem(p=2, n=6, q=50.0)]
    sps = [ShellProblem(p) for p in pts]
    print("[case] Newmark batch of three plates", flush=True)
    batch, _ = run_newmark_batch(sps, 12, keep=True, dt=T1 / 20, precision=L.FP64, **NEWMARK)
    for sp, b in zip(sps, batch):
        one = sp.run_newmark(12, dt=T1 / 20, precision=L.FP64, **NEWMARK)
        assert one["steps_done"] == b["steps_done"] == 12
        assert np.array_equal(one["iterations"], b["iterations"]) and np.array_equal(one["U"], b["U"])
