"""The C-ABI library loads and exports every symbol include/ksh.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ksh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ksh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("ksh_problem_create", "ksh_assemble_residual", "ksh_assemble_tangent", "ksh_factor", "ksh_solve",
                 "ksh_inertia", "ksh_vector_d", "ksh_run_alm", "ksh_run_static"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from gsstructuralanalysis_amd import _lib
    path = _lib.LIB_PATH
    if not os.path.exists(path):
        pytest.fail("libksh.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's symbols
    assert sorted(_lib.SIGNATURES) == _declared()


def test_no_gpu_error_is_loud():
    """Creating a problem without a HIP device fails with a clear status (no CPU fallback)."""
    from gsstructuralanalysis_amd import _lib
    lib = _lib.load()
    n = ctypes.c_int(-1)
    rc = lib.ksh_device_count(ctypes.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    import numpy as np
    from gsstructuralanalysis_amd import geometry as g
    from gsstructuralanalysis_amd.problem import ShellProblem
    with pytest.raises(_lib.KshError):
        ShellProblem(g.roof_problem(p=3, n=2))
