"""Singular points and branch switching on the GPU path engine against the oracle's restatement
(gsALMBase.hpp:422-856; drivers examples/gsThinShell_ArcLength.cpp:1107-1117 and
solvers/quasistatic_shell_XML.cpp:305-311).

The case is the reference example's clamped strip under compression (gsThinShell_ArcLength.cpp
test case 1): perfect and flat, its compressed state loses stability at a symmetric bifurcation
near the clamped Euler load 4 pi^2 EI / L^2 (EI = 1).  The oracle's routines are pinned in
closed form on two-dof systems in test_oracle_singular.py.  The extended iterations stop at
||K phi|| < 1e-6 here: their residuals run 3e-3, 5e-8, 9e-11 and then sit at the rounding floor
(~5e-11) just under the reference's 1e-10, too close to that tolerance for two implementations
to agree on the iteration count."""
import numpy as np
import pytest

from gsstructuralanalysis_amd import _lib as L
from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu

OPTS = dict(max_iter=20, tol_f=1e-8, tol_u=1e-6, scaling=-1.0, max_halvings=10)
SP = dict(sp_tol_e=1e-6)


def _euler_lambda(pt, b=0.1):
    EI = pt.E * b * pt.t ** 3 / 12
    return 4 * np.pi ** 2 * EI / (pt.E * b * pt.t * 1e-6 * b)


def _compare(got, ref, tag):
    k = ref["steps_done"]
    assert got["steps_done"] == k, (got["steps_done"], k)
    assert np.array_equal(got["singular"], ref["singular"]), (got["singular"], ref["singular"])
    assert np.array_equal(got["negatives"], ref["negatives"])
    for key in ("lambda", "unorm"):
        err = float(np.abs(got[key] - ref[key]).max() / np.abs(ref[key]).max())
        print(f"[{tag}] {key}: max relative deviation from the oracle {err:.2e}", flush=True)
        assert err <= 1e-7, key
    err = float(np.abs(got["U"] - ref["U"]).max() / np.abs(ref["U"]).max())
    print(f"[{tag}] U: max relative deviation {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_strip_bifurcation_and_branch_switch(prec, ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    from oracle.solvers import run_crisfield
    pt = g.compressed_strip_problem(nx=16)
    print(f"[case] clamped strip p=3 16x1 ({pt.nfree} dofs), singular_point=2, {prec}", flush=True)
    ref = run_crisfield(pt, 24, 2e-4, singular_point=2, **OPTS, **SP)
    extra = dict(precision=L.FP64) if prec == "fp64" else dict(precision=L.FP32_IR, max_refine=4, refine_tol=1e-12)
    got = ShellProblem(pt).run_alm(24, 2e-4, keep_U=True, singular_point=2, **OPTS, **SP, **extra)
    _compare(got, ref, f"strip {prec}")
    k = int(np.flatnonzero(got["singular"] == 2)[0])
    lam_e = _euler_lambda(pt)
    print(f"[strip] bifurcation at step {k}: lambda {got['lambda'][k]:.8f} = {got['lambda'][k] / lam_e:.6f} "
          f"x the Euler load", flush=True)
    assert abs(got["lambda"][k] / lam_e - 1) <= 1e-3
    # past it the strip deflects out of its plane on the stable branch, at an almost constant load
    w = np.array([np.abs(pt.to_full(u).reshape(-1, 3)[:, 2]).max() for u in got["U"]])
    assert np.all(w[:k] == 0) and np.all(np.diff(w[k:]) > 0) and w[-1] > 2 * w[k]
    assert np.all(np.abs(got["lambda"][k:] / got["lambda"][k] - 1) <= 1e-3)
    assert np.all(got["negatives"][k + 1:] == 0)


def test_strip_bisection_then_extended_system(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    from oracle.solvers import run_crisfield
    pt = g.compressed_strip_problem(nx=16)
    print("[case] clamped strip, bisection (TolB 1e-2) then extended iterations", flush=True)
    ref = run_crisfield(pt, 15, 2e-4, singular_point=2, sp_tol_b=1e-2, **OPTS, **SP)
    got = ShellProblem(pt).run_alm(15, 2e-4, keep_U=True, singular_point=2, sp_tol_b=1e-2, precision=L.FP64, **OPTS, **SP)
    _compare(got, ref, "strip bisection")
    assert (got["singular"] == 2).sum() == 1


def test_strip_singular_point_without_switch(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    from oracle.solvers import run_crisfield
    pt = g.compressed_strip_problem(nx=16)
    print("[case] clamped strip, singular_point=1 (no branch switch), up to the singular point", flush=True)
    ref = run_crisfield(pt, 14, 2e-4, singular_point=1, **OPTS, **SP)
    got = ShellProblem(pt).run_alm(14, 2e-4, keep_U=True, singular_point=1, precision=L.FP64, **OPTS, **SP)
    _compare(got, ref, "strip no switch")
    assert got["singular"][13] == 2
    assert np.abs(pt.to_full(got["U"][13]).reshape(-1, 3)[:, 2]).max() <= 1e-12  # still flat


def test_roof_limit_point_with_singular_point_handling(ksh):
    """The roof snaps through at a limit point: the test says so, the solution goes back to the
    previous point and the path goes on (gsALMBase.hpp:425-428)."""
    from gsstructuralanalysis_amd.problem import ShellProblem
    from oracle.solvers import run_crisfield
    pt = g.roof_problem(p=2, n=4)
    print(f"[case] roof p=2 4x4 ({pt.nfree} dofs), singular_point=1", flush=True)
    opts = dict(max_iter=20, tol_f=1e-8, tol_u=1e-6, scaling=0.0)
    ref = run_crisfield(pt, 14, 6e-3, singular_point=1, **opts)
    got = ShellProblem(pt).run_alm(14, 6e-3, keep_U=True, singular_point=1, precision=L.FP64, **opts)
    _compare(got, ref, "roof limit point")
    assert (got["singular"] == 1).sum() >= 1


def test_batch_with_singular_points_matches_single_runs(ksh):
    """Instances at a singular point advance through its stages while the others go on stepping:
    each path of a mixed batch equals its own single-problem run."""
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch
    pts = [g.compressed_strip_problem(nx=16), g.compressed_strip_problem(nx=24), g.roof_problem(p=3, n=8)]
    sps = [ShellProblem(p) for p in pts]
    print("[case] batch: two strips and a roof, singular_point=2", flush=True)
    ds = np.array([2e-4, 2e-4, 4e-3])
    batch, _ = run_alm_batch(sps, 20, ds, keep_U=True, singular_point=2, precision=L.FP64, **OPTS, **SP)
    for sp, d, b in zip(sps, ds, batch):
        one = sp.run_alm(20, float(d), keep_U=True, singular_point=2, precision=L.FP64, **OPTS, **SP)
        assert one["steps_done"] == b["steps_done"]
        assert np.array_equal(one["singular"], b["singular"]) and np.array_equal(one["negatives"], b["negatives"])
        assert np.array_equal(one["lambda"], b["lambda"]) and np.array_equal(one["U"], b["U"])
    assert (batch[0]["singular"] == 2).sum() == 1 and (batch[1]["singular"] == 2).sum() == 1
