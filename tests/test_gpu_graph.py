"""Captured factor launch sequences (hipGraph replays, opt-in with KSH_GRAPHS=1) against the same
factorizations run eagerly (the default, KSH_GRAPHS=0): every replay, with the program's own eager work and synchronizations
between replays, must give bit-identical pivots, solutions and paths."""
import numpy as np
import pytest

from gsstructuralanalysis_amd import _lib as L
from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu

KW = dict(method=0, max_iter=20, tol=1e-6, tol_f=1e-8, tol_u=1e-6, scaling=0.0, bifurcation_method=0)


@pytest.mark.parametrize("prec", [L.FP64, L.FP32_IR])
def test_replayed_factor_equals_eager(prec, ksh, monkeypatch):
    from gsstructuralanalysis_amd.problem import ShellProblem
    sp = ShellProblem(g.roof_problem(p=3, n=12))
    print(f"[case] graph replays roof p=3 n=12 nfree={sp.nfree} precision={prec}", flush=True)
    rng = np.random.default_rng(21)
    us = [rng.standard_normal(sp.nfree) * 1e-3 for _ in range(6)]
    b = rng.standard_normal(sp.nfree)

    def run(graphs):
        monkeypatch.setenv("KSH_GRAPHS", "1" if graphs else "0")
        K = sp.tangent(us[0], precision=prec)
        out = []
        for u in us:  # same matrix object: the same launch sequence each time
            sp.tangent(u, out=K)
            K.factor()
            out.append((K.vector_d(), K.solve(b, max_refine=3 if prec != L.FP64 else 0, tol=1e-12)))
        return out

    eager = run(False)
    replayed = run(True)
    for (d0, x0), (d1, x1) in zip(eager, replayed):
        assert np.array_equal(d0, d1)
        assert np.array_equal(x0, x1)


def test_replayed_paths_equal_eager(ksh, monkeypatch):
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch
    sps = [ShellProblem(g.imperfect_roof_instance(i, sizes=(8, 12))) for i in range(6)]
    print("[case] graph replays: 6 imperfect roofs, 10 steps, batched", flush=True)
    monkeypatch.setenv("KSH_GRAPHS", "0")
    r0, _ = run_alm_batch(sps, 10, 6e-3, **KW)
    monkeypatch.setenv("KSH_GRAPHS", "1")
    r1, _ = run_alm_batch(sps, 10, 6e-3, **KW)
    s0 = sps[0].run_alm(10, 6e-3, **KW)
    monkeypatch.setenv("KSH_GRAPHS", "0")
    s1 = sps[0].run_alm(10, 6e-3, **KW)
    for a, b in zip(r0, r1):
        assert np.array_equal(a["lambda"], b["lambda"])
        assert np.array_equal(a["negatives"], b["negatives"])
        assert np.array_equal(a["iterations"], b["iterations"])
    assert np.array_equal(s0["lambda"], s1["lambda"]) and np.array_equal(s0["negatives"], s1["negatives"])
