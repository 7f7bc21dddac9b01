"""Launch-shape autotuner: every variant it tries (elements per workgroup of the tangent tiles,
threads per block of the band gather) computes the same K bit for bit (KSH_TUNE=v<k> forces
variant k), and after one launch per variant it settles on one."""
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
    for _ in range(5):  # at most 3 measured launches (one per variant), then the chosen one
        assert np.array_equal(sp.tangent(u).download(), ref)
        choices.append(_choice("tangent_tiles")[0])
    assert all(0 <= c < 3 for c in choices) and choices[-1] == choices[-2]
    v, nv = _choice("band_gather")
    assert nv == 3 and 0 <= v < 3
