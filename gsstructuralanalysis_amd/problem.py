"""Python front-end of libksh.so: one shell problem on the GPU, its band matrices and solvers.

Mirrors the reference's operator surface (src/gsStructuralAnalysisTools/gsStructuralAnalysisTypes.h:58-93)
for scripting and tests:

* ``ShellProblem.residual(u, lam, conv)``  -- Residual_t / ALResidual_t bodies
* ``ShellProblem.tangent(u, precision)``  -- Jacobian_t body (device-resident BandMatrix)
* ``BandMatrix.factor / solve / vector_d / inertia`` -- gsSparseSolver<T>::SimplicialLDLT
* ``ShellProblem.run_alm / run_static``   -- gsALMCrisfield / gsALMRiks / gsStaticNewton drivers
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .geometry import Patch as GeoPatch


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def alm_options(nsteps, length, method=0, precision=L.FP64, max_iter=20, tol=1e-6, tol_u=1e-6, tol_f=1e-3,
                scaling=-1.0, angle_method=0, bifurcation_method=0, relaxation=1.0, quasi=0, quasi_iterations=-1,
                adaptive=0, adaptive_iterations=10, max_refine=0, refine_tol=1e-14, max_halvings=10, verbose=0,
                inertia_guard=0.0, learned=None, refine_accept=0.0, learned_mode=0, singular_point=0,
                sp_test_tol=1e-6, sp_test_it=5, sp_tol_e=1e-10, sp_tol_b=0.0, perturbation=1e3, sp_fail=1,
                length_after_sp=0.0):
    """ksh_alm_options with the reference's option names in snake case (gsALMBase.hpp:22-54);
    `learned` is a LearnedPredictor (Crisfield only); learned_mode 1 keeps the predictor's
    extrapolation and leaves out the network's correction (ablation); singular_point 1 / 2:
    computeSingularPoint on a change of stability, 2 followed by switchBranch."""
    return L.AlmOptions(method, precision, nsteps, length, max_iter, tol, tol_u, tol_f, scaling, angle_method,
                        bifurcation_method, relaxation, quasi, quasi_iterations, adaptive, adaptive_iterations,
                        max_refine, refine_tol, max_halvings, verbose, inertia_guard,
                        learned.handle if learned is not None else None, refine_accept, learned_mode, None,
                        singular_point, sp_test_tol, sp_test_it, sp_tol_e, sp_tol_b, perturbation, sp_fail,
                        length_after_sp)


def _alm_result(nsteps, nfree, keep_U, probe_dof, features=False):
    out = {k: np.zeros(nsteps) for k in ("lambda", "unorm", "indicator", "residue_f", "length", "w")}
    if features:
        out["features"] = np.zeros((nsteps, 66))
    out["negatives"] = np.zeros(nsteps, dtype=np.int64)
    out["iterations"] = np.zeros(nsteps, dtype=np.int32)
    out["singular"] = np.zeros(nsteps, dtype=np.int32)
    U = np.zeros(nsteps * nfree) if keep_U else None
    r = L.AlmResult()
    r.lambda_ = _dptr(out["lambda"])
    r.unorm = _dptr(out["unorm"])
    r.indicator = _dptr(out["indicator"])
    r.negatives = out["negatives"].ctypes.data_as(C.POINTER(C.c_int64))
    r.iterations = out["iterations"].ctypes.data_as(C.POINTER(C.c_int))
    r.residue_f = _dptr(out["residue_f"])
    r.length = _dptr(out["length"])
    r.w = _dptr(out["w"])
    r.probe_dof = probe_dof
    r.U = _dptr(U) if U is not None else None
    r.features = _dptr(out["features"]) if features else None
    r.singular = out["singular"].ctypes.data_as(C.POINTER(C.c_int))
    return r, out, U


def _alm_collect(r, out, U, nsteps, nfree):
    k = r.steps_done
    res = {key: val[:k] for key, val in out.items()}
    res.update(steps_done=k, halvings=r.halvings, final_status=r.final_status, wall_seconds=r.wall_seconds,
               phi=r.phi, jacobians=r.jacobians, residuals=r.residuals, solves=r.solves,
               learned_steps=r.learned_steps, fallbacks=r.fallbacks)
    if U is not None:
        res["U"] = U.reshape(nsteps, nfree)[:k]
    return res


def run_alm_batch(problems, nsteps, length, keep_U=False, features=False, **kw):
    """Arc-length paths of many problems at once (ksh_batch_run_alm): one host thread advances
    every path by one predictor or corrector iteration per round, in batched launches.
    Returns (list of per-problem result dicts, wall seconds of the whole batch)."""
    lengths = None
    if np.ndim(length) > 0:  # one arc length per problem
        lengths = np.ascontiguousarray(length, dtype=np.float64)
        if lengths.shape != (len(problems),):
            raise ValueError("one arc length per problem")
        length = float(lengths[0])
    o = alm_options(nsteps, length, **kw)
    if lengths is not None:
        o.lengths = _dptr(lengths)
    n = len(problems)
    handles = (C.c_void_p * n)(*[p.handle for p in problems])
    rs = (L.AlmResult * n)()
    keep = []
    for i, p in enumerate(problems):
        r, out, U = _alm_result(nsteps, p.nfree, keep_U, -1, features)
        rs[i] = r
        keep.append((out, U))
    wall = C.c_double()
    code = L.load().ksh_batch_run_alm(handles, n, C.byref(o), rs, C.byref(wall))
    if code not in (L.SUCCESS, L.NOT_CONVERGED):
        L.check(code)
    results = [_alm_collect(rs[i], keep[i][0], keep[i][1], nsteps, p.nfree) for i, p in enumerate(problems)]
    return results, wall.value


def eig_options(number=10, mode=L.EIG_SHIFT_INVERT, ncv_fac=3, shift=0.0, tol=1e-10, max_restarts=0,
                precision=L.FP64, max_refine=8, refine_tol=1e-13, seed=0):
    """ksh_eig_options with gsEigenProblemBase's option meanings (gsEigenProblemBase.h:40-89)."""
    return L.EigOptions(mode, number, ncv_fac, shift, tol, max_restarts, precision,
                        max_refine if precision != L.FP64 else 0, refine_tol, seed)


def _eig_collect(values, vectors, res):
    return {"values": values, "vectors": vectors, "converged": res.converged, "restarts": res.restarts,
            "solves": res.solves, "products": res.products, "wall_seconds": res.wall_seconds}


def eigen(A: "BandMatrix", B: "BandMatrix", number=10, keep_vectors=False, **kw):
    """Eigenpairs of A x = lambda B x nearest `shift` (ksh_eigen)."""
    o = eig_options(number, **kw)
    vals = np.zeros(number)
    vecs = np.zeros((number, A.n)) if keep_vectors else None
    res = L.EigResult()
    L.check(L.load().ksh_eigen(A.handle, B.handle, C.byref(o), _dptr(vals),
                               _dptr(vecs) if vecs is not None else None, C.byref(res)))
    return _eig_collect(vals, vecs, res)


class BandMatrix:
    """Device band matrix (lower band, ab[J*(bw+1)+(I-J)]) with its LDL^T factor."""

    def __init__(self, problem: "ShellProblem", precision: int = L.FP64):
        self.problem = problem
        self.handle = C.c_void_p()
        L.check(L.load().ksh_matrix_create(problem.handle, precision, C.byref(self.handle)))
        self.n, self.bw, self.precision = problem.nfree, problem.bw, precision

    def __del__(self):
        try:
            if self.handle:
                L.load().ksh_matrix_destroy(self.handle)
                self.handle = C.c_void_p()
        except Exception:
            pass

    def download(self) -> np.ndarray:
        ab = np.zeros(self.n * (self.bw + 1))
        L.check(L.load().ksh_matrix_download(self.handle, _dptr(ab)))
        return ab

    def upload(self, ab):
        ab = _f64(ab).ravel()
        assert ab.size == self.n * (self.bw + 1)
        L.check(L.load().ksh_matrix_upload(self.handle, _dptr(ab)))

    def axpby(self, alpha, beta, other: "BandMatrix"):
        L.check(L.load().ksh_matrix_axpby(self.handle, alpha, beta, other.handle))

    def factor(self):
        L.check(L.load().ksh_factor(self.handle))

    def solve(self, b, max_refine=0, tol=1e-14):
        b = _f64(b)
        x = np.zeros(self.n)
        L.check(L.load().ksh_solve(self.handle, _dptr(b), _dptr(x), max_refine, tol))
        return x

    def matvec(self, x):
        x = _f64(x)
        y = np.zeros(self.n)
        L.check(L.load().ksh_matvec(self.handle, _dptr(x), _dptr(y)))
        return y

    def vector_d(self):
        d = np.zeros(self.n)
        L.check(L.load().ksh_vector_d(self.handle, _dptr(d)))
        return d

    def inertia(self):
        neg = C.c_int64()
        dmin = C.c_double()
        L.check(L.load().ksh_inertia(self.handle, C.byref(neg), C.byref(dmin)))
        return neg.value, dmin.value

    def set_inertia_guard(self, guard):
        L.check(L.load().ksh_set_inertia_guard(self.handle, guard))


class ShellProblem:
    """One single-patch Kirchhoff-Love shell problem resident on the GPU."""

    def __init__(self, patch: GeoPatch):
        lib = L.load()
        self.patch = patch
        self._ctrl = _f64(patch.ctrl).ravel()
        self._fixed = np.ascontiguousarray(patch.fixed, dtype=np.int32)
        self._pxi = _f64(patch.point_xi).ravel()
        self._pF = _f64(patch.point_F).ravel()
        pa = L.Patch(patch.p, patch.nx, patch.ny, _dptr(self._ctrl))
        ma = L.Material(patch.t, patch.E, patch.nu, patch.rho)
        bc = L.BC(self._fixed.ctypes.data_as(C.POINTER(C.c_int)))
        lo = L.Loads(patch.point_xi.shape[0], _dptr(self._pxi), _dptr(self._pF), (C.c_double * 3)(*patch.fsurf))
        self.handle = C.c_void_p()
        L.check(lib.ksh_problem_create(C.byref(pa), C.byref(ma), C.byref(bc), C.byref(lo), C.byref(self.handle)))
        n = C.c_int64()
        bw = C.c_int64()
        L.check(lib.ksh_num_free_dofs(self.handle, C.byref(n)))
        L.check(lib.ksh_bandwidth(self.handle, C.byref(bw)))
        self.nfree, self.bw = n.value, bw.value
        if getattr(patch, "pressure", 0.0) != 0.0:
            self.set_pressure(patch.pressure)

    def set_pressure(self, pressure: float):
        """Follower pressure along a1 x a2 (gsThinShellAssembler::setPressure): F then holds its load
        on the undeformed surface, the residuals and the tangent its follower terms; 0 removes it."""
        L.check(L.load().ksh_problem_set_pressure(self.handle, float(pressure)))
        self.patch.pressure = float(pressure)

    def close(self):
        if self.handle:
            L.load().ksh_problem_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- operator bodies -------------------------------------------------------------------
    def force(self):
        F = np.zeros(self.nfree)
        L.check(L.load().ksh_assemble_force(self.handle, _dptr(F)))
        return F

    def residual(self, u, lam=1.0, conv=L.RES_ALM):
        u = _f64(u)
        R = np.zeros(self.nfree)
        L.check(L.load().ksh_assemble_residual(self.handle, _dptr(u), float(lam), conv, _dptr(R)))
        return R

    def tangent(self, u, precision=L.FP64, out: BandMatrix | None = None) -> BandMatrix:
        K = out if out is not None else BandMatrix(self, precision)
        u = _f64(u)
        L.check(L.load().ksh_assemble_tangent(self.handle, _dptr(u), K.handle))
        return K

    def mass(self, precision=L.FP64) -> BandMatrix:
        M = BandMatrix(self, precision)
        L.check(L.load().ksh_assemble_mass(self.handle, M.handle))
        return M

    # -- drivers ---------------------------------------------------------------------------
    def run_alm(self, nsteps, length, keep_U=False, probe_dof=-1, features=False, **kw):
        o = alm_options(nsteps, length, **kw)
        r, out, U = _alm_result(nsteps, self.nfree, keep_U, probe_dof, features)
        code = L.load().ksh_run_alm(self.handle, C.byref(o), C.byref(r))
        if code not in (L.SUCCESS, L.NOT_CONVERGED):
            L.check(code)
        return _alm_collect(r, out, U, nsteps, self.nfree)

    def _run_eig(self, fn, number, keep_vectors, kw):
        o = eig_options(number, **kw)
        vals = np.zeros(number)
        vecs = np.zeros((number, self.nfree)) if keep_vectors else None
        res = L.EigResult()
        L.check(fn(self.handle, C.byref(o), _dptr(vals), _dptr(vecs) if vecs is not None else None, C.byref(res)))
        return _eig_collect(vals, vecs, res)

    def run_buckling(self, number=10, keep_vectors=False, **kw):
        """solvers/buckling_shell_XML.cpp:212-247 on the device: K_L, F, u = K_L^-1 F, K_NL = K(u),
        gsBucklingSolver(K_L, K_NL).computeSparse(number)."""
        kw.setdefault("mode", L.EIG_BUCKLING)
        return self._run_eig(L.load().ksh_run_buckling, number, keep_vectors, kw)

    def run_modal(self, number=10, keep_vectors=False, **kw):
        """solvers/modal_shell_XML.cpp:233-253 on the device: gsModalSolver(K, M).computeSparse(number)."""
        return self._run_eig(L.load().ksh_run_modal, number, keep_vectors, kw)

    def run_static(self, nsteps, dlambda, precision=L.FP64, max_iter=25, tol=1e-10, max_refine=0,
                   refine_tol=1e-14, verbose=0, keep_U=False, inertia_guard=0.0, refine_accept=0.0):
        o = L.StaticOptions(precision, nsteps, dlambda, max_iter, tol, max_refine, refine_tol, verbose, inertia_guard,
                            refine_accept)
        out = {k: np.zeros(nsteps) for k in ("lambda", "unorm", "residual")}
        out["iterations"] = np.zeros(nsteps, dtype=np.int32)
        out["negatives"] = np.zeros(nsteps, dtype=np.int64)
        U = np.zeros(nsteps * self.nfree) if keep_U else None
        r = L.StaticResult()
        r.lambda_ = _dptr(out["lambda"])
        r.unorm = _dptr(out["unorm"])
        r.residual = _dptr(out["residual"])
        r.iterations = out["iterations"].ctypes.data_as(C.POINTER(C.c_int))
        r.negatives = out["negatives"].ctypes.data_as(C.POINTER(C.c_int64))
        r.U = _dptr(U) if U is not None else None
        code = L.load().ksh_run_static(self.handle, C.byref(o), C.byref(r))
        if code not in (L.SUCCESS, L.NOT_CONVERGED):
            L.check(code)
        k = r.steps_done
        res = {key: val[:k] for key, val in out.items()}
        res.update(steps_done=k, final_status=r.final_status, wall_seconds=r.wall_seconds, jacobians=r.jacobians)
        if U is not None:
            res["U"] = U.reshape(nsteps, self.nfree)[:k]
        return res

    def run_dr(self, keep_U=True, **kw):
        """Dynamic relaxation (solvers/static_shell_XML.cpp:243-275 on the device)."""
        return _dr_run([self], dr_options(**kw), keep_U)[0][0]

    def run_newmark(self, nsteps, U0=None, V0=None, A0=None, keep=True, probe_dof=-1, **kw):
        """Nonlinear Newmark from (U0, V0, A0) (tutorials/nonlinear_shell_dynamic.cpp:155-186 on the device)."""
        o = newmark_options(nsteps, **kw)
        r, out, arrs = _newmark_result(nsteps, self.nfree, keep, probe_dof)
        ics = [None if v is None else _f64(v) for v in (U0, V0, A0)]
        for v in ics:
            if v is not None and v.shape != (self.nfree,):
                raise ValueError("initial states have one entry per free dof")
        code = L.load().ksh_run_newmark(self.handle, C.byref(o), *[(_dptr(v) if v is not None else None) for v in ics],
                                        C.byref(r))
        if code not in (L.SUCCESS, L.NOT_CONVERGED):
            L.check(code)
        return _newmark_collect(r, out, arrs)


# ---- residual-only and mass consumers ---------------------------------------------------------
def dr_options(max_iter=25, tol=1e-6, tol_f=-1.0, tol_u=-1.0, damping=1.0, alpha=2.0, tol_e=1e-6, reset_it=-1,
               verbose=0):
    """ksh_dr_options with gsStaticDR's option names in snake case (gsStaticDR.hpp:20-27,
    gsStaticBase.h:66-72): maxIt, tol, tolF, tolU, damping, alpha, tolE, ResetIt."""
    return L.DrOptions(max_iter, tol, tol_f, tol_u, damping, alpha, tol_e, reset_it, verbose)


def _dr_run(problems, opts, keep_U):
    n = len(problems)
    res = (L.DrResult * n)()
    Us, eks = [], []
    for i, sp in enumerate(problems):
        U = np.zeros(sp.nfree) if keep_U else None
        ek = np.zeros(opts.max_iter)
        res[i].U = _dptr(U) if U is not None else None
        res[i].ek_history = _dptr(ek)
        Us.append(U)
        eks.append(ek)
    handles = (C.c_void_p * n)(*[p.handle for p in problems])
    wall = C.c_double()
    code = L.load().ksh_batch_run_dr(handles, n, C.byref(opts), res, C.byref(wall))
    if code not in (L.SUCCESS, L.NOT_CONVERGED):
        L.check(code)
    out = []
    for i in range(n):
        r = res[i]
        d = dict(iterations=r.iterations, converged=r.final_status == L.SUCCESS, final_status=r.final_status,
                 peaks=r.peaks, residual=r.residual, ek=r.ek, ek0=r.ek0, wall_seconds=r.wall_seconds,
                 residuals=r.residuals, ek_history=eks[i][:r.iterations + 1])
        if Us[i] is not None:
            d["U"] = Us[i]
        out.append(d)
    return out, wall.value


def run_dr_batch(problems, keep_U=False, **kw):
    """Dynamic relaxation of many problems at once (ksh_batch_run_dr); (results, wall seconds)."""
    return _dr_run(problems, dr_options(**kw), keep_U)


def newmark_options(nsteps, dt=1e-2, precision=L.FP64, beta=0.25, gamma=0.5, damping_mass=0.0, max_iter=100,
                    tol_f=1e-3, tol_u=1e-6, max_refine=0, refine_tol=1e-14, verbose=0):
    """ksh_newmark_options (gsDynamicBase.hpp:20-35, gsDynamicNewmark.hpp:20-25: the reference's
    "alpha" and "delta" are beta and gamma here)."""
    return L.NewmarkOptions(precision, nsteps, dt, beta, gamma, damping_mass, max_iter, tol_f, tol_u, max_refine,
                            refine_tol, verbose)


def _newmark_result(nsteps, nfree, keep, probe_dof):
    out = {k: np.zeros(nsteps) for k in ("time", "unorm", "w")}
    out["iterations"] = np.zeros(nsteps, dtype=np.int32)
    arrs = {k: np.zeros((nsteps, nfree)) for k in ("U", "V", "A")} if keep else {}
    r = L.NewmarkResult()
    r.time = _dptr(out["time"])
    r.unorm = _dptr(out["unorm"])
    r.w = _dptr(out["w"])
    r.probe_dof = probe_dof
    r.iterations = out["iterations"].ctypes.data_as(C.POINTER(C.c_int))
    for k, a in arrs.items():
        setattr(r, k, _dptr(a))
    return r, out, arrs


def _newmark_collect(r, out, arrs):
    k = r.steps_done
    res = {key: val[:k] for key, val in out.items()}
    res.update({key: val[:k] for key, val in arrs.items()})
    res.update(steps_done=k, final_status=r.final_status, wall_seconds=r.wall_seconds, jacobians=r.jacobians)
    return res


def run_newmark_batch(problems, nsteps, keep=False, probe_dof=-1, **kw):
    """Newmark time integration of many problems from rest at once (ksh_batch_run_newmark)."""
    o = newmark_options(nsteps, **kw)
    n = len(problems)
    res = (L.NewmarkResult * n)()
    keepers = []
    for i, sp in enumerate(problems):
        r, out, arrs = _newmark_result(nsteps, sp.nfree, keep, probe_dof)
        res[i] = r
        keepers.append((out, arrs))
    handles = (C.c_void_p * n)(*[p.handle for p in problems])
    wall = C.c_double()
    code = L.load().ksh_batch_run_newmark(handles, n, C.byref(o), res, C.byref(wall))
    if code not in (L.SUCCESS, L.NOT_CONVERGED):
        L.check(code)
    return [_newmark_collect(res[i], *keepers[i]) for i in range(n)], wall.value
