"""The allocation-free residual entry (ksh_assemble_residual_dev_noalloc) against the ordinary one
(ksh_assemble_residual_dev): the same residual bit for bit on two calls in a row, the second on a
smaller problem (the shared array keeps the first call's contents), and a clear error for a problem
beyond the array.  Device inputs and outputs are the caller's memory, from the HIP runtime that
libksh.so itself links (loaded here through ctypes)."""
import numpy as np
import pytest

from gsstructuralanalysis_amd import _lib as L
from gsstructuralanalysis_amd import geometry as g

pytestmark = pytest.mark.gpu


class DevArray:
    """n doubles of device memory (hipMalloc of the HIP runtime libksh.so links)."""

    def __init__(self, n):
        import ctypes as C
        L.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = n
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(8 * n)) == 0

    def upload(self, a):
        import ctypes as C
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0  # H2D

    def download(self):
        import ctypes as C
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * self.n), 2) == 0  # D2H
        return a

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def _residuals(sp, u, lam, conv):
    lib = L.load()
    ud, r1, r2 = DevArray(sp.nfree), DevArray(sp.nfree), DevArray(sp.nfree)
    try:
        ud.upload(u)
        r1.upload(np.full(sp.nfree, np.nan))
        r2.upload(np.full(sp.nfree, np.nan))
        L.check(lib.ksh_assemble_residual_dev(sp.handle, ud.p, lam, conv, r1.p))
        L.check(lib.ksh_synchronize(sp.handle))
        L.check(lib.ksh_assemble_residual_dev_noalloc(sp.handle, ud.p, lam, conv, r2.p))
        L.check(lib.ksh_synchronize(sp.handle))
        return r1.download(), r2.download()
    finally:
        for a in (ud, r1, r2):
            a.free()


def test_noalloc_residual_equals_the_ordinary_entry(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    rng = np.random.default_rng(11)
    for n, conv, lam in ((16, L.RES_ALM, 0.7), (8, L.RES_STATIC, 0.3)):  # second call smaller
        pt = g.imperfect_roof_instance(5, sizes=(n,))
        sp = ShellProblem(pt)
        print(f"[case] no-alloc residual roof p=3 {n}x{n} ({sp.nfree} dofs) conv {conv}", flush=True)
        u = rng.standard_normal(sp.nfree) * 1e-3
        a, b = _residuals(sp, u, lam, conv)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b)
        assert np.array_equal(b, sp.residual(u, lam, conv))


def test_noalloc_residual_with_follower_pressure(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    rng = np.random.default_rng(12)
    pt = g.imperfect_roof_instance(5, sizes=(8,))
    pt.pressure = -3.0
    sp = ShellProblem(pt)
    print(f"[case] no-alloc residual roof p=3 8x8 with follower pressure ({sp.nfree} dofs)", flush=True)
    u = rng.standard_normal(sp.nfree) * 1e-3
    a, b = _residuals(sp, u, 0.4, L.RES_ALM)
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, b)


def test_noalloc_residual_rejects_a_problem_beyond_its_array(ksh):
    from gsstructuralanalysis_amd.problem import ShellProblem
    pt = g.plate(2, 200)  # 40,000 degree-2 elements: 52 MB of partial results
    sp = ShellProblem(pt)
    ud, rd = DevArray(sp.nfree), DevArray(sp.nfree)
    try:
        rc = L.load().ksh_assemble_residual_dev_noalloc(sp.handle, ud.p, 1.0, L.RES_ALM, rd.p)
        assert rc == L.OTHER_ERROR
        assert b"allocation-free entry holds" in L.load().ksh_last_error()
    finally:
        ud.free()
        rd.free()
