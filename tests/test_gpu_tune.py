"""Launch-shape autotuner: every variant it tries (elements per workgroup of the tangent tiles,
threads per block of the band gather, schedule of the band factor) computes the same result bit
for bit (KSH_TUNE=v<k> forces variant k), and after a warm-up launch and one timed launch per
variant it settles on one."""
import ctypes as C

import numpy as np
import pytest

from gsstructuralanalysis_amd import _lib as L
from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu


def _choice(name):
    v, nv = C.c_int64(), C.c_int64()
    L.check(L.load().ksh_tune_choice(name.encode(), C.byref(v), C.byref(nv)))
    return v.value, nv.value


def test_tuned_variants_give_identical_tangents(ksh, monkeypatch):
    from gsstructuralanalysis_amd.problem import ShellProblem
    sp = ShellProblem(g.roof_problem(p=3, n=20))
    print(f"[case] autotuner roof p=3 n=20 nfree={sp.nfree}", flush=True)
    u = np.random.default_rng(4).standard_normal(sp.nfree) * 1e-3
    monkeypatch.setenv("KSH_TUNE", "0")
    ref = sp.tangent(u).download()
    for v in range(3):  # each variant, forced
        monkeypatch.setenv("KSH_TUNE", f"v{v}")
        assert np.array_equal(sp.tangent(u).download(), ref)
        assert _choice("tangent_tiles") == (v, 3) and _choice("band_gather") == (v, 3)
    monkeypatch.setenv("KSH_TUNE", "1")
    choices = []
    for _ in range(6):  # at most a warm-up and 3 measured launches (one per variant), then the chosen one
        assert np.array_equal(sp.tangent(u).download(), ref)
        choices.append(_choice("tangent_tiles")[0])
    assert all(0 <= c < 3 for c in choices) and choices[-1] == choices[-2]
    v, nv = _choice("band_gather")
    assert nv == 3 and 0 <= v < 3


def test_tuned_factor_schedules_give_identical_factors(ksh, monkeypatch):
    from gsstructuralanalysis_amd.problem import ShellProblem
    monkeypatch.setenv("KSH_BAND_PATH", "tiled")
    sp = ShellProblem(g.roof_problem(p=3, n=24))
    print(f"[case] autotuner factor roof p=3 n=24 nfree={sp.nfree} bw={sp.bw}", flush=True)
    u = np.random.default_rng(9).standard_normal(sp.nfree) * 1e-3
    K = sp.tangent(u)
    monkeypatch.setenv("KSH_TUNE", "0")
    K.factor()
    ref = K.vector_d()
    for v in range(3):
        monkeypatch.setenv("KSH_TUNE", f"v{v}")
        sp.tangent(u, out=K)
        K.factor()
        assert np.array_equal(K.vector_d(), ref)
        assert _choice("band_factor") == (v, 3)
    monkeypatch.setenv("KSH_TUNE", "1")
    choices = []
    for _ in range(6):
        sp.tangent(u, out=K)
        K.factor()
        assert np.array_equal(K.vector_d(), ref)
        choices.append(_choice("band_factor")[0])
    assert all(0 <= c < 3 for c in choices) and choices[-1] == choices[-2]
