"""Learned arc-length predictor: features against a numpy projection, and what the trained network
gains over the tangent predictor on held-out imperfect roofs (corrector iterations, fallbacks)
while the Crisfield corrector keeps the path (load factors, negative-pivot counts)."""
import numpy as np
import pytest

from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu

KW = dict(method=0, max_iter=20, tol=1e-6, tol_f=1e-8, tol_u=1e-6, scaling=0.0, bifurcation_method=0)


def test_features_are_the_modal_projection(ksh):
    import ctypes as C
    from gsstructuralanalysis_amd import _lib as L
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = g.roof_problem(p=3, n=12)
    sp = ShellProblem(pt)
    mx = pt.nx + pt.p
    gx = g.greville(pt.p, pt.nx)
    gy = g.greville(pt.p, pt.ny)
    rng = np.random.default_rng(3)
    c = rng.standard_normal(64)
    X, Y = np.meshgrid(gx, gy)  # [j][i]
    Phi = np.stack([np.sin((a + 1) * np.pi * X) * np.sin((b + 1) * np.pi * Y)
                    for b in range(8) for a in range(8)], axis=-1).reshape(-1, 64)  # cp = i + mx j
    free = pt.free_map()
    U = np.zeros(sp.nfree)
    zfull = Phi @ c
    for cp in range(len(zfull)):
        f = free[3 * cp + 2]
        if f >= 0:
            U[f] = zfull[cp]
    out = np.zeros(66)
    L.check(L.load().ksh_predictor_features(sp.handle, U.ctypes.data_as(C.POINTER(C.c_double)), 0.25,
                                            out.ctypes.data_as(C.POINTER(C.c_double))))
    zfree = np.array([zfull[cp] if free[3 * cp + 2] >= 0 else 0.0 for cp in range(len(zfull))])
    ref = np.linalg.lstsq(Phi, zfree, rcond=None)[0]
    assert out[0] == 0.25 and abs(out[1] - np.linalg.norm(U)) <= 1e-12 * np.linalg.norm(U)
    assert np.abs(out[2:] - ref).max() <= 1e-6 * np.abs(ref).max()
    assert mx * mx >= 64


def test_learned_predictor_gains_on_held_out_roofs(ksh):
    from gsstructuralanalysis_amd.learned import LearnedPredictor, dataset
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch
    sizes, steps, ds = (8, 12, 16), 40, 6e-3
    train = [ShellProblem(g.imperfect_roof_instance(i, sizes=sizes)) for i in range(48)]
    tr, _ = run_alm_batch(train, steps, ds, features=True, **KW)
    X, Y = dataset([r["features"] for r in tr])
    print(f"[case] learned predictor: {len(X)} training samples", flush=True)
    net = LearnedPredictor(seed=1)
    losses = net.fit(X, Y, epochs=400, batch=256, lr=1e-3)
    assert losses[-1] < 0.2 * losses[0]

    held = [ShellProblem(g.imperfect_roof_instance(i, sizes=sizes, seed0=30261202)) for i in range(12)]
    base, _ = run_alm_batch(held, steps, ds, **KW)
    lrn, _ = run_alm_batch(held, steps, ds, learned=net, **KW)
    it_b = sum(int(r["iterations"].sum()) for r in base)
    it_l = sum(int(r["iterations"].sum()) for r in lrn)
    jac_b = sum(r["jacobians"] for r in base)
    jac_l = sum(r["jacobians"] for r in lrn)
    used = sum(r["learned_steps"] for r in lrn)
    fb = sum(r["fallbacks"] for r in lrn)
    print(f"[learned] corrector iterations tangent {it_b} learned {it_l}; jacobians {jac_b} -> {jac_l}; "
          f"learned steps {used}, fallbacks {fb}", flush=True)
    for a, b in zip(base, lrn):
        assert a["steps_done"] == b["steps_done"] == steps
        assert np.array_equal(a["negatives"], b["negatives"])
        assert np.abs(a["lambda"] - b["lambda"]).max() <= 1e-6 * np.abs(a["lambda"]).max()
    assert used > 0.5 * steps * len(held)
    assert fb <= 0.2 * used
    assert jac_l < jac_b


def test_network_of_another_shape_is_refused(ksh):
    """The predictor reads 4 x 66 features and writes 64 mode coefficients and a load factor per
    instance; a network of other widths passed through the C-ABI is refused before any step (the
    host buffers are sized for those widths)."""
    import ctypes as C
    from gsstructuralanalysis_amd import _lib as L
    from gsstructuralanalysis_amd.learned import WIDTHS
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch

    class _Net:  # what run_alm_batch takes from a predictor: its handle
        def __init__(self):
            self.handle = C.c_void_p()

    print(f"[case] learned predictor of widths other than {WIDTHS}", flush=True)
    lib = L.load()
    sps = [ShellProblem(g.imperfect_roof_instance(i, sizes=(4,))) for i in range(2)]
    for widths in [(WIDTHS[0] + 1, 32, 32, WIDTHS[3]), (WIDTHS[0], 32, 32, WIDTHS[3] - 1)]:
        net = _Net()
        L.check(lib.ksh_mlp_create((C.c_int * 4)(*widths), 1, C.byref(net.handle)))
        try:
            with pytest.raises(L.KshError, match="must map"):
                run_alm_batch(sps, 2, 6e-3, learned=net, **KW)
        finally:
            lib.ksh_mlp_destroy(net.handle)
