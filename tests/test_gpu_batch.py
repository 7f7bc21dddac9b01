"""Batched engine: many uneven imperfect panels run the unchanged Crisfield loop with merged
device launches; every instance must reproduce its own single-problem run."""
import numpy as np
import pytest

from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu

KW = dict(method=0, max_iter=20, tol=1e-6, tol_f=1e-8, tol_u=1e-6, scaling=0.0, bifurcation_method=0)


def _instances(k, sizes=(4, 6, 8)):
    return [g.imperfect_roof_instance(i, sizes=sizes) for i in range(k)]


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single_runs(precision, ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch
    pts = _instances(6)
    print(f"[case] batch of {len(pts)} imperfect roofs, precision {precision}", flush=True)
    extra = dict(max_refine=6, refine_tol=1e-11) if precision == 1 else {}
    sps = [ShellProblem(pt) for pt in pts]
    batch, wall = run_alm_batch(sps, 15, 2e-3, precision=precision, keep_U=True, **KW, **extra)
    assert wall > 0
    for sp, b in zip(sps, batch):
        s = sp.run_alm(15, 2e-3, precision=precision, keep_U=True, **KW, **extra)
        assert b["steps_done"] == s["steps_done"] == 15
        assert np.array_equal(b["negatives"], s["negatives"])
        assert np.array_equal(b["iterations"], s["iterations"])
        tol = 1e-12 if precision == 0 else 1e-9
        assert np.abs(b["U"] - s["U"]).max() <= tol * np.abs(s["U"]).max()
        assert np.abs(b["lambda"] - s["lambda"]).max() <= tol * np.abs(s["lambda"]).max()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single_runs_desynchronised(precision, ksh):
    """Instances of different sizes take different iteration counts, so one engine round mixes
    request kinds (tangent, residual, factor, solve of different instances); each instance must
    still reproduce its single-problem path exactly (the C2 study's settings on smaller meshes)."""
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch
    pts = [g.imperfect_roof_instance(i, sizes=(6, 8, 10, 12)) for i in range(12)]
    print(f"[case] desynchronised batch of {len(pts)} imperfect roofs, precision {precision}", flush=True)
    extra = dict(max_refine=6, refine_tol=1e-11) if precision == 1 else {}
    sps = [ShellProblem(pt) for pt in pts]
    batch, _ = run_alm_batch(sps, 20, 6e-3, precision=precision, keep_U=True, **KW, **extra)
    its = set()
    for sp, b in zip(sps, batch):
        s = sp.run_alm(20, 6e-3, precision=precision, keep_U=True, **KW, **extra)
        assert b["steps_done"] == s["steps_done"] == 20
        assert np.array_equal(b["negatives"], s["negatives"])
        assert np.array_equal(b["iterations"], s["iterations"])
        tol = 1e-12 if precision == 0 else 1e-9
        assert np.abs(b["U"] - s["U"]).max() <= tol * np.abs(s["U"]).max()
        assert np.abs(b["lambda"] - s["lambda"]).max() <= tol * np.abs(s["lambda"]).max()
        its.add(int(np.sum(s["iterations"])))
    assert len(its) > 1  # the instances really are out of step


def test_batch_with_failures_and_uneven_work(ksh):
    """One instance cannot converge in 2 iterations and halves its arc length; the others proceed."""
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch
    pts = _instances(4, sizes=(4, 5))
    sps = [ShellProblem(pt) for pt in pts]
    kw = dict(KW)
    kw["max_iter"] = 4
    batch, _ = run_alm_batch(sps, 6, 2e-2, max_halvings=6, **kw)
    for sp, b in zip(sps, batch):
        s = sp.run_alm(6, 2e-2, max_halvings=6, **kw)
        assert b["steps_done"] == s["steps_done"]
        assert b["halvings"] == s["halvings"]
        assert np.allclose(b["lambda"], s["lambda"], rtol=1e-12, atol=0)


def _count(name):
    import ctypes as C
    from gsstructuralanalysis_amd import _lib as L
    cnt, ms = C.c_int64(), C.c_double()
    L.load().ksh_profile_query(name.encode(), C.byref(cnt), C.byref(ms))
    return cnt.value


def test_early_fp64_factors_change_nothing(ksh, monkeypatch):
    """The engine starts the FP64 factor of a matrix whose previous FP32 factor needed the FP64
    fallback on a second stream, before any solve asks for it.  With one refinement sweep and an
    acceptance of 1e-13 nearly every FP32 solve falls back, so most fallbacks take an early
    factor: the batch with early factors, the batch without them (KSH_EARLY64=0) and the single
    runs give the same path."""
    from gsstructuralanalysis_amd import _lib as L
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch
    pts = [g.imperfect_roof_instance(i, sizes=(6, 8, 10)) for i in range(6)]
    print(f"[case] early FP64 factors, batch of {len(pts)} imperfect roofs, forced fallbacks", flush=True)
    extra = dict(precision=1, max_refine=1, refine_tol=1e-12, refine_accept=1e-13)
    sps = [ShellProblem(pt) for pt in pts]
    lib = L.load()
    lib.ksh_profile_reset()
    lib.ksh_profile_enable(1)
    try:
        early, _ = run_alm_batch(sps, 12, 4e-3, keep_U=True, **KW, **extra)
    finally:
        lib.ksh_profile_enable(0)
    used, stalls = _count("fp64_refactor_early_used"), _count("fp64_refactor_stall")
    print(f"[case] fallbacks {stalls}, taken from early factors {used}", flush=True)
    assert used > stalls // 2
    monkeypatch.setenv("KSH_EARLY64", "0")
    plain, _ = run_alm_batch(sps, 12, 4e-3, keep_U=True, **KW, **extra)
    for sp, a, b in zip(sps, early, plain):
        s = sp.run_alm(12, 4e-3, keep_U=True, **KW, **extra)
        for r in (a, b):
            assert r["steps_done"] == s["steps_done"] == 12
            assert np.array_equal(r["negatives"], s["negatives"])
            assert np.array_equal(r["iterations"], s["iterations"])
            assert np.abs(r["U"] - s["U"]).max() <= 1e-9 * np.abs(s["U"]).max()
            assert np.abs(r["lambda"] - s["lambda"]).max() <= 1e-9 * np.abs(s["lambda"]).max()
        assert np.array_equal(a["U"], b["U"]) and np.array_equal(a["lambda"], b["lambda"])
