"""CPU restatement of the reference's solver loops over the FP64 oracle (TEST INFRASTRUCTURE ONLY).

Used by tests/ as the checker of the GPU drivers and by bench.py's cpu_baseline leg.  Each
method follows the reference file:line it cites; the linear algebra is the oracle's banded
LDL^T (ok_band_ldlt), the assembly is the oracle's element loop.

The ``faithful`` switch reproduces the reference's redundant work: the same tangent is factored
again at gsALMCrisfield.hpp:69/114 and in the determinant stability branch at gsALMBase.hpp:561.
"""
from __future__ import annotations

import time

import numpy as np

from .ok import BandLDLT, Oracle, band_symv


class OracleOps:
    """Jacobian / residual / force of a Patch on the CPU oracle, plus call counters."""

    def __init__(self, patch):
        self.o = Oracle(patch)
        self.n, self.bw = self.o.nfree, self.o.bw
        self.F = self.o.force()
        self.n_jac = self.n_res = self.n_fact = self.n_solve = 0

    def jacobian(self, u):
        self.n_jac += 1
        return self.o.tangent_band(u)

    def residual_alm(self, u, lam):
        self.n_res += 1
        return self.o.residual_alm(u, lam)

    def residual_static(self, u, lam):
        self.n_res += 1
        return self.o.residual_static(u, lam)

    def mass(self):
        return self.o.mass_band()

    def factor(self, ab):
        self.n_fact += 1
        f = BandLDLT(ab, self.n, self.bw)
        if f.info != 0:
            raise _Err(3)
        return f

    def solve(self, fac, b):
        self.n_solve += 1
        return fac.solve(b)


class _Err(Exception):
    def __init__(self, code):
        super().__init__(code)
        self.code = code


# ---------------------------------------------------------------------------------------------
class ALMCrisfield:
    """gsALMBase::_step (gsALMBase.hpp:353-416) with gsALMCrisfield's methods (gsALMCrisfield.hpp:40-472)."""

    def __init__(self, ops: OracleOps, length, max_iter=100, tol_f=1e-3, tol_u=1e-6, scaling=-1.0, angle_method=0,
                 relaxation=1.0, bifurcation_method=0, faithful=False, tol=1e-6):
        self.ops = ops
        self.F = ops.F
        n = ops.n
        self.U = np.zeros(n)
        self.DeltaU = np.zeros(n)
        self.L = self.DeltaL = 0.0
        self.DeltaUold = np.zeros(n)
        self.DeltaLold = 0.0
        self.Uguess = None
        self.arc = length
        self.max_iter, self.tol_f, self.tol_u = max_iter, tol_f, tol_u
        self.phi = scaling
        self.phi_user = scaling != -1
        self.angle = angle_method
        self.relax = relaxation
        self.bif = bifurcation_method
        self.faithful = faithful
        self.indicator, self.negatives = 0.0, 0
        self.stability = self.stability_prev = -1
        self.deltaUt = np.zeros(n)
        self.deltaUbar = np.zeros(n)
        self.deltaU = np.zeros(n)
        self.deltaL = 0.0
        self.eta = 1.0
        self.jac = None
        self.fac = None
        self.num_iter = 0
        self.tol = tol
        self.V = np.zeros(n)

    # gsALMBase::init -> _computeStability(m_U) (gsALMBase.hpp:95-103)
    def initialize(self):
        self._compute_jacobian(self.U)
        self._stability()
        self.stability = -1 if self.indicator >= 0 else 1

    def _compute_jacobian(self, U):
        self.jac = self.ops.jacobian(U)
        self.fac = self.ops.factor(self.jac)

    def _stability(self):
        """Determinant branch of _computeStability (gsALMBase.hpp:557-569,612-614)."""
        if self.bif < 0:
            self.negatives, self.indicator = 0, 0.0
        else:
            if self.faithful:
                self.fac = self.ops.factor(self.jac)
            self.negatives, self.indicator = self.fac.negatives()
        self.stability = 1 if self.indicator < 0 else -1

    def _residual(self):
        self.resVec = self.ops.residual_alm(self.U + self.DeltaU, self.L + self.DeltaL)

    def _norms(self):
        """computeResidualNorms (gsALMBase.hpp:150-183)."""
        if self.num_iter == 0:
            self.basisF = np.linalg.norm((self.L + self.DeltaL) * self.F)
            self.basisU = np.linalg.norm(self.DeltaU)
            self.residueF = np.linalg.norm(self.resVec) / self.basisF
            self.residueU = np.linalg.norm(self.deltaU) / np.linalg.norm(self.DeltaU)
        else:
            self.residueF = np.linalg.norm(self.resVec) / self.basisF
            self.residueU = np.linalg.norm(self.deltaU) / self.basisU

    # gsALMCrisfield::predictor (gsALMCrisfield.hpp:110-164)
    def predictor(self):
        self._compute_jacobian(self.U + self.DeltaU)
        if self.faithful:
            self.fac = self.ops.factor(self.jac)
        self.deltaUt = self.ops.solve(self.fac, self.F)
        FF = self.F @ self.F
        if self.DeltaUold @ self.DeltaUold == 0 and self.DeltaLold * self.DeltaLold == 0:
            self.deltaL = self.arc / np.sqrt(2 * (self.deltaUt @ self.deltaUt))
            self.deltaU = self.deltaUbar + self.deltaL * self.deltaUt
            if not self.phi_user:
                self.phi = np.sqrt((self.deltaUt @ self.deltaUt) / FF)
        else:
            if not self.phi_user:
                self.phi = np.sqrt((self.U @ self.U) / (self.L ** 2 * FF))
            self._lambda_mu()
        self.DeltaU = self.DeltaU + self.deltaU
        self.DeltaL += self.deltaL
        if self.angle in (1, 2):
            self.DeltaUold = self.DeltaU.copy()
            self.DeltaLold = self.DeltaL

    def _lambda_mu(self):
        """computeLambdaMU (gsALMCrisfield.hpp:383-398)."""
        A0 = self.phi ** 2 * (self.F @ self.F)
        dir_ = np.sign(self.DeltaUold @ self.deltaUt + A0 * self.DeltaLold)
        denum = np.sqrt(self.deltaUt @ self.deltaUt + A0)
        mu = self.arc if denum == 0 else self.arc / denum
        self.deltaL = dir_ * mu
        self.deltaU = self.deltaL * self.deltaUt

    def _lambdas(self):
        """computeLambdas + Simple/Modified/Eta/Complex/DOT (gsALMCrisfield.hpp:212-363,400-472)."""
        F = self.F
        A0 = self.phi ** 2 * (F @ F)
        ut, ub, DU = self.deltaUt, self.deltaUbar, self.DeltaU
        a0 = ut @ ut + A0
        b0 = 2 * (ut @ DU + self.DeltaL * A0)
        b1 = 2 * (ub @ ut)
        c0 = DU @ DU + self.DeltaL ** 2 * A0 - self.arc ** 2
        c1 = 2 * (DU @ ub)
        c2 = ub @ ub
        eta = self.eta
        al1, al2, al3 = a0, b0 + eta * b1, c0 + eta * c1 + eta * eta * c2
        disc = al2 ** 2 - 4 * al1 * al3
        if disc >= 0:
            self.eta = 1.0
            lams = [(-al2 + np.sqrt(disc)) / (2 * al1), (-al2 - np.sqrt(disc)) / (2 * al1)]
            self._dot(lams)
            return
        # modified (Lam 1992)
        m1 = b1 * b1 - 4.0 * a0 * c2
        m2 = 2.0 * b0 * b1 - 4.0 * a0 * c1
        m3 = b0 * b0 - 4.0 * a0 * c0
        disc2 = m2 ** 2 - 4.0 * m1 * m3
        if disc2 >= 0:
            e0 = (-m2 + np.sqrt(disc2)) / (2.0 * m1)
            e1 = (-m2 - np.sqrt(disc2)) / (2.0 * m1)
            eta1, eta2 = min(e0, e1), max(e0, e1)
            xi = 0.05 * abs(eta2 - eta1)
            if eta2 < 1.0:
                self.eta = eta2 - xi
            elif eta2 > 1.0 and -m2 / m1 < 1.0:
                self.eta = eta2 + xi
            elif eta1 < 1.0 and -m2 / m1 > 1.0:
                self.eta = eta1 - xi
            elif eta1 > 1.0:
                self.eta = eta1 + xi
        if disc2 >= 0 and self.eta > 0.05:
            al2 = b0 + self.eta * b1
            lam = -al2 / (2 * a0)
            self._dot([lam, lam])
        else:
            self.eta = 1.0
            DeltaUcr = self.DeltaU + self.deltaUbar
            Fint = band_symv(self.jac, self.ops.n, self.ops.bw, self.U + self.DeltaU)
            Lcr = (Fint @ F) / (F @ F)
            DeltaLcr = Lcr - self.L
            arc_cr = np.sqrt(DeltaUcr @ DeltaUcr + A0 * DeltaLcr ** 2)
            mu = self.arc / arc_cr
            self.deltaL = mu * DeltaLcr - self.DeltaL
            self.deltaU = mu * DeltaUcr - self.DeltaU

    def _dot(self, lams):
        base = self.DeltaUold @ self.deltaUt + self.phi ** 2 * self.DeltaLold
        d1, d2 = lams[0] * base, lams[1] * base
        pick = 1 if d1 < d2 else 0
        self.deltaL = lams[pick]
        self.deltaU = self.eta * self.deltaUbar + self.deltaUt * lams[pick]

    # gsALMCrisfield::iteration / quasiNewtonIteration (gsALMCrisfield.hpp:65-100)
    def iteration(self):
        self._compute_jacobian(self.U + self.DeltaU)
        if self.faithful:
            self.fac = self.ops.factor(self.jac)
        self.deltaUt = self.ops.solve(self.fac, self.F)
        self.deltaUbar = self.ops.solve(self.fac, -self.resVec)
        self.eta = 1.0
        lamold = self.deltaL
        self._lambdas()
        if (lamold * self.deltaL < 0 and abs(self.deltaL) <= abs(lamold)) and self.relax != 1.0:
            self.deltaU = self.relax * (self.deltaL * self.deltaUt + self.eta * self.deltaUbar)
            self.deltaL = self.relax * self.deltaL
        self.DeltaU = self.DeltaU + self.deltaU
        self.DeltaL += self.deltaL
        if self.angle == 1:
            self.DeltaUold = self.DeltaU.copy()
            self.DeltaLold = self.DeltaL

    def step(self):
        """gsALMBase::step/_step (gsALMBase.hpp:327-416); returns 0 success, 1 not converged, 3 solver error."""
        try:
            self._step()
            return 0
        except _Err as e:
            return e.code

    def _step(self):
        self.num_iter = 0
        self.DeltaU = np.zeros(self.ops.n)
        self.deltaUbar = np.zeros(self.ops.n)
        self.deltaUt = np.zeros(self.ops.n)
        self.deltaU = np.zeros(self.ops.n)
        self.DeltaL = self.deltaL = 0.0
        self.eta = 1.0
        self.predictor()
        self._residual()
        self._norms()
        self.stability_prev = self.stability
        for it in range(1, self.max_iter):
            self.num_iter = it
            self.iteration()
            self._stability()
            self._residual()
            self._norms()
            if self.residueF < self.tol_f and self.residueU < self.tol_u:
                self.U = self.U + self.DeltaU
                self.L += self.DeltaL
                if self.angle == 0:
                    self.DeltaUold = self.DeltaU.copy()
                    self.DeltaLold = self.DeltaL
                return
            if it == self.max_iter - 1:
                raise _Err(1)

    def set_solution(self, U, L):
        self.U, self.L = U.copy(), L

    # ---- singular points: gsALMBase.hpp:422-446 (_computeSingularPoint), 475-511
    # (_testSingularPoint), 631-735 (extended system), 754-840 (bisection), 843-856 (switchBranch)
    def _power(self, V, its, stop_tol=None):
        """V <- K^-1 V / ||K^-1 V|| with the factor of the current tangent; stop_tol: the
        bisection's early exit ||V - Vold|| < Tol (gsALMBase.hpp:824-837)."""
        Vold = np.zeros_like(V)
        for _ in range(its):
            V = self.ops.solve(self.fac, V)
            V = V / np.linalg.norm(V)
            if stop_tol is not None:
                if np.linalg.norm(V - Vold) < stop_tol:
                    return V, True
                Vold = V
        return V, False

    def test_singular_point(self, test_tol=1e-6, test_it=5):
        """_testSingularPoint(jacobian=false): power iterations from the normalised ones vector
        with the factor of the last counted tangent; a bifurcation point when |V . F| < tol."""
        n = self.ops.n
        self.V, _ = self._power(np.ones(n) / np.sqrt(n), test_it)
        return abs(self.V @ self.F) < test_tol

    def extended_system_solve(self, U, L, tol, sp_fail=1):
        """_extendedSystemSolve / _extendedSystemIteration: Newton on R(U, L) = 0, K(U) W = 0,
        ||W|| = 1 from (U, L, V), the derivative of K along W by a finite difference (eps 1e-8);
        stops when ||K(U + DU) (V + DV)|| < tol.  Returns whether it converged."""
        self.U, self.L = U.copy(), L
        n = self.ops.n
        self._compute_jacobian(self.U)
        DU, DV, DL = np.zeros(n), np.zeros(n), 0.0
        F = self.F
        eps = 1e-8
        for it in range(self.max_iter):
            X = self.U + DU
            R = self.ops.residual_alm(X, self.L + DL)
            self._compute_jacobian(X)
            Ut = self.ops.solve(self.fac, F)
            Ub = self.ops.solve(self.fac, -R)
            W = self.V + DV
            Keps = self.ops.jacobian(X + eps * W)
            h1 = (band_symv(Keps, n, self.ops.bw, Ut) - F) / eps
            h2 = band_symv(self.jac, n, self.ops.bw, W) + (band_symv(Keps, n, self.ops.bw, Ub) + R) / eps
            Vt = self.ops.solve(self.fac, -h1)
            Vb = self.ops.solve(self.fac, -h2)
            nw = np.linalg.norm(W)
            e = W / nw
            dl = -(e @ Vb + nw - 1.0) / (e @ Vt)
            DU = DU + dl * Ut + Ub
            DV = DV + dl * Vt + Vb
            DL += dl
            self._compute_jacobian(self.U + DU)
            self.sp_residual = np.linalg.norm(band_symv(self.jac, n, self.ops.bw, self.V + DV))
            self.sp_iterations = it + 1
            if self.sp_residual < tol:
                self.U = self.U + DU
                self.L += DL
                return True
            if it == self.max_iter - 1:
                if sp_fail == 1:
                    self.U = self.U + DU
                    self.L += DL
                return False
        return False

    def bisection_solve(self, U, L, tol):
        """_bisectionSolve: arc-length steps from the last point whose pivot count did not change,
        halving the arc length when the count changes or the step fails, until
        |min D / min D(start)| < tol; then the eigenvector V by power iterations."""
        self.U, self.L = U.copy(), L
        U_old, L_old = U.copy(), L
        dL = self.arc
        self._compute_jacobian(self.U)
        self._stability()
        ref = self.indicator
        fa = self.negatives
        converged = False
        for _ in range(1, self.max_iter):
            self.U, self.L = U_old.copy(), L_old
            st = self.step()
            self._compute_jacobian(self.U)
            self._stability()
            fb = self.negatives
            if fa == fb and st == 0:
                U_old, L_old = self.U.copy(), self.L
                fa = fb
            else:
                self.arc = self.arc / 2.0
            if abs(self.indicator / ref) < tol:
                converged = True
                break
        self.arc = dL
        n = self.ops.n
        self.V, hit = self._power(np.ones(n) / np.sqrt(n), 5, stop_tol=self.tol)
        return converged or hit

    def switch_branch(self, tau):
        """switchBranch: U += (||V|| / tau) V with V normalised; the previous increment forgotten."""
        self.V = self.V / np.linalg.norm(self.V)
        xi = np.linalg.norm(self.V) / tau
        self.U = self.U + xi * self.V
        self.DeltaLold = 0.0
        self.DeltaUold = np.zeros_like(self.DeltaUold)

    def compute_singular_point(self, U, L, switch=False, test_tol=1e-6, test_it=5, tol_b=0.0, tol_e=1e-10,
                               sp_fail=1, tau=1e3):
        """_computeSingularPoint(U, L, switchBranch, jacobian=false, testPoint=true).  Returns
        0 limit point (the solution is set back to (U, L)), 2 bifurcation point computed,
        3 bifurcation point whose extended iterations did not converge."""
        test = self.test_singular_point(test_tol, test_it)
        self.U, self.L = U.copy(), L
        if not test:
            return 0
        if tol_b != 0:
            self.bisection_solve(self.U, self.L, tol_b)
        conv = self.extended_system_solve(self.U, self.L, tol_e, sp_fail)
        if switch and (conv or sp_fail == 1):
            self.switch_branch(tau)
        self.stability = self.stability_prev
        return 2 if conv else 3

    def set_length(self, length):
        """setLength (gsALMBase.h:118-122): the arc length and its 'previous' value."""
        self.arc = length
        self.arc_prev = length


class ALMRiks(ALMCrisfield):
    """gsALMRiks (gsALMRiks.hpp:22-134,205-212): linearised constraint with phi = 1/n."""

    def __init__(self, ops, length, **kw):
        super().__init__(ops, length, **kw)
        self.Uprev = np.zeros(ops.n)
        self.Lprev = 0.0
        self.arc_prev = length

    def predictor(self):
        self._compute_jacobian(self.U + self.DeltaU)
        if self.faithful:
            self.fac = self.ops.factor(self.jac)
        n = self.ops.n
        self.phi = 0.99999999999 if n == 1 else 1.0 / n
        if np.linalg.norm(self.U - self.Uprev) < 1e-10 and (self.L - self.Lprev) ** 2 < 1e-10:
            DL = 1.0
            self.deltaUt = self.ops.solve(self.fac, self.F)
            den = np.sqrt(self.deltaUt @ self.deltaUt + self.DeltaL * DL)
            self.deltaU = self.deltaUt / den
            self.deltaL = DL / den
        else:
            self.deltaL = (self.L - self.Lprev) / self.arc_prev
            self.deltaU = (self.U - self.Uprev) / self.arc_prev
        self.deltaL *= self.arc
        self.deltaU = self.deltaU * self.arc
        self.DeltaU = self.DeltaU + self.deltaU
        self.DeltaL += self.deltaL

    def iteration(self):
        self._compute_jacobian(self.U + self.DeltaU)
        if self.faithful:
            self.fac = self.ops.factor(self.jac)
        self.deltaUt = self.ops.solve(self.fac, self.F)
        self.deltaUbar = self.ops.solve(self.fac, -self.resVec)
        DU = self.DeltaU
        r = self.phi * (DU @ DU) + (1 - self.phi) * self.DeltaL ** 2 - self.arc ** 2
        self.deltaL = -(r + 2 * self.phi * (DU @ self.deltaUbar)) / (
            2 * (1 - self.phi) * self.DeltaL + 2 * self.phi * (DU @ self.deltaUt))
        self.deltaU = self.deltaUbar + self.deltaL * self.deltaUt
        self.DeltaU = self.DeltaU + self.deltaU
        self.DeltaL += self.deltaL

    def _step(self):
        U0, L0 = self.U.copy(), self.L
        super()._step()
        self.Uprev, self.Lprev = U0, L0
        self.arc_prev = self.arc


def run_crisfield(patch, nsteps, length, max_halvings=10, method=0, ops=None, keep_jac=False, singular_point=0,
                  sp_test_tol=1e-6, sp_test_it=5, sp_tol_b=0.0, sp_tol_e=1e-10, sp_fail=1, perturbation=1e3,
                  length_after_sp=0.0, **kw):
    """Driver loop of benchmarks/benchmark_Roof.cpp:417-512 on the oracle (method 0 Crisfield, 1 Riks).

    ``ops``: any object with the OracleOps interface (default: the shell oracle of ``patch``);
    ``keep_jac``: also return the counted tangent (band) of every converged step;
    ``singular_point``: on a change of stability after a converged step, 1 computeSingularPoint
    from the previous point (solvers/quasistatic_shell_XML.cpp:305-311), 2 the same followed by
    switchBranch (examples/gsThinShell_ArcLength.cpp:1107-1117); the arc length then restarts
    from ``length_after_sp`` (0: the initial length).  ``singular`` records per step 0, or 1 a
    limit point, 2 a computed bifurcation point, 3 one whose extended iterations did not converge."""
    ops = ops if ops is not None else OracleOps(patch)
    alm = (ALMCrisfield if method == 0 else ALMRiks)(ops, length, **kw)
    t0 = time.perf_counter()
    alm.initialize()
    Uold, Lold = np.zeros(ops.n), 0.0
    dL0 = dL = length
    bisected = False
    out = {"lambda": [], "unorm": [], "negatives": [], "indicator": [], "iterations": [], "length": [], "U": [],
           "residue_f": [], "phi_step": [], "singular": []}
    if keep_jac:
        out["jac"] = []
    halvings = 0
    k = 0
    while k < nsteps:
        st = alm.step()
        if st == 1:
            if halvings >= max_halvings:
                break
            halvings += 1
            dL /= 2.0
            alm.set_length(dL)
            alm.set_solution(Uold, Lold)
            bisected = True
            continue
        if st != 0:
            break
        used = dL
        sing = 0
        if singular_point and alm.stability * alm.stability_prev < 0:
            kind = alm.compute_singular_point(Uold, Lold, False, sp_test_tol, sp_test_it, sp_tol_b, sp_tol_e, sp_fail,
                                              perturbation)
            sing = 1 if kind == 0 else kind
            if singular_point == 2:
                alm.switch_branch(perturbation)
            if length_after_sp > 0:
                dL0 = length_after_sp
            dL = dL0
            alm.set_length(dL)
        Uold, Lold = alm.U.copy(), alm.L
        out["lambda"].append(Lold)
        out["unorm"].append(np.linalg.norm(Uold))
        out["negatives"].append(alm.negatives)
        out["indicator"].append(alm.indicator)
        out["iterations"].append(alm.num_iter)
        out["length"].append(used)
        out["U"].append(Uold)
        out["residue_f"].append(alm.residueF)
        out["phi_step"].append(alm.phi)
        out["singular"].append(sing)
        if keep_jac:
            out["jac"].append(np.array(alm.jac, copy=True))
        if not bisected:
            dL = dL0
            alm.set_length(dL)
        bisected = False
        k += 1
    res = {key: np.array(v) for key, v in out.items()}
    res.update(steps_done=len(out["lambda"]), halvings=halvings, wall_seconds=time.perf_counter() - t0,
               phi=alm.phi, jacobians=ops.n_jac, factorizations=ops.n_fact, solves=ops.n_solve,
               residuals=ops.n_res)
    return res


# ---------------------------------------------------------------------------------------------
def run_static_newton(patch, nsteps, dlambda, max_iter=25, tol=1e-10):
    """gsStaticNewton::_solveNonlinear (gsStaticNewton.hpp:141-193,273-316) under
    gsControlDisplacement::step (gsControlDisplacement.h:46-60), static residual lam F - F_int."""
    ops = OracleOps(patch)
    n = ops.n
    t0 = time.perf_counter()
    Uc = np.zeros(n)
    Lc = 0.0
    out = {"lambda": [], "unorm": [], "iterations": [], "negatives": [], "residual": [], "U": []}
    for k in range(nsteps):
        lam = Lc + dlambda
        U = Uc.copy()
        DeltaU = np.zeros(n)
        R = ops.residual_static(U, lam)
        res = np.linalg.norm(R)
        res = 1.0 if res == 0 else res
        res_ini = res
        conv = False
        for it in range(max_iter):
            fac = ops.factor(ops.jacobian(U + DeltaU))
            deltaU = ops.solve(fac, R)
            DeltaU = DeltaU + deltaU
            R = ops.residual_static(U + DeltaU, lam)
            res = np.linalg.norm(R)
            if np.linalg.norm(deltaU) / np.linalg.norm(DeltaU) < tol and res / res_ini < tol:
                conv = True
                break
        if not conv:
            break
        neg, _ = fac.negatives()
        Lc = lam
        Uc = Uc + DeltaU
        out["lambda"].append(Lc)
        out["unorm"].append(np.linalg.norm(Uc))
        out["iterations"].append(it + 1)
        out["negatives"].append(neg)
        out["residual"].append(res)
        out["U"].append(Uc.copy())
    r = {key: np.array(v) for key, v in out.items()}
    r.update(steps_done=len(out["lambda"]), wall_seconds=time.perf_counter() - t0, jacobians=ops.n_jac)
    return r


# ---------------------------------------------------------------------------------------------
def lumped_mass(ops):
    """assembleMass(true): the row sums of the consistent mass matrix (the B-spline basis is a
    partition of unity, so row i sums to the integral of rho t N_i)."""
    n = ops.n
    return band_symv(ops.mass(), n, ops.bw, np.ones(n))


def run_dynamic_relaxation(patch, ops=None, damping=1.0, alpha=2.0, tol=1e-6, tol_f=-1.0, tol_u=-1.0, tol_e=1e-6,
                           max_it=25, reset_it=-1, mass=None):
    """gsStaticDR::_solve (gsStaticDR.hpp:93-137) with _start (:219-266), _iteration (:182-201) and
    _peak (:204-214), dt = 1, the static residual F - F_int(U) at the full load, lumped mass.
    Returns a dict: U, iterations, converged, residual (||R|| / ||R0||), ek history, peaks."""
    ops = ops if ops is not None else OracleOps(patch)
    n = ops.n
    m = lumped_mass(ops) if mass is None else np.asarray(mass, dtype=np.float64)
    minv = 1.0 / m / alpha
    damp = damping * m
    tolF = tol_f if tol_f != -1 else tol
    tolU = tol_u if tol_u != -1 else tol
    dt = 1.0
    F = ops.F

    def res(u):
        return ops.residual_static(u, 1.0)

    U = np.zeros(n)
    # _start (no headstart)
    v = np.zeros(n)
    DU = np.zeros(n)
    R = res(U)
    residual = np.linalg.norm(F)
    if residual == 0:
        residual = 1.0
    residual_ini = residual
    dU = 0.5 * dt * dt * minv * R
    v = dU / dt
    DU = DU + dU
    Ek = v @ (m * v)
    Ek0 = Ek
    eks = [Ek]
    peaks = 0
    converged = False
    it = 1
    reset = 0
    while it != max_it:
        # _iteration
        Ek_prev = Ek
        R = res(U + DU) - damp * v
        residual = np.linalg.norm(R)
        if residual_ini == 0:
            residual_ini = residual
        v = v + dt * minv * R
        dU = dt * v
        DU = DU + dU
        Ek = v @ (m * v)
        if (damping == 0 and Ek_prev > Ek) or reset == reset_it:
            reset = 0
            # _peak
            R = res(U + DU) - damp * v
            dU = -1.5 * dt * v + dt * dt / 2.0 * minv * R
            DU = DU + dU
            v = np.zeros(n)
            Ek = 0.0
            peaks += 1
        eks.append(Ek)
        if residual / residual_ini < tolF and Ek / Ek0 < tol_e and np.linalg.norm(dU) / np.linalg.norm(DU) < tolU:
            U = U + DU
            converged = True
            break
        if it == max_it - 1:
            U = U + DU
            break
        it += 1
        reset += 1
    return {"U": U, "iterations": it, "converged": converged, "residual": residual / residual_ini,
            "ek": np.array(eks), "peaks": peaks, "residuals": ops.n_res}


def run_newmark(patch, nsteps, dt, ops=None, beta=0.25, gamma=0.5, damping_mass=0.0, max_iter=100, tol_f=1e-3,
                tol_u=1e-6, U0=None, V0=None, A0=None):
    """gsDynamicNewmark<T, true>::_step_impl (gsDynamicNewmark.hpp:69-142) under the time loop of
    tutorials/nonlinear_shell_dynamic.cpp:176-186: per step the predictor with A = 0, then Newton
    on M A + C V = R(U) with lhs M + gamma dt C + beta dt^2 K(U); consistent mass, C =
    damping_mass M, R = F - F_int.  ("alpha" and "delta" of the reference are beta and gamma.)"""
    ops = ops if ops is not None else OracleOps(patch)
    n = ops.n
    M = ops.mass()
    bw = ops.bw
    U = np.zeros(n) if U0 is None else np.array(U0, dtype=np.float64)
    V = np.zeros(n) if V0 is None else np.array(V0, dtype=np.float64)
    A = np.zeros(n) if A0 is None else np.array(A0, dtype=np.float64)
    out = {"time": [], "U": [], "V": [], "A": [], "iterations": [], "unorm": []}
    t = 0.0
    status = 0
    for k in range(nsteps):
        Uold, Vold, Aold = U, V, A
        A = np.zeros(n)
        U = Uold + dt * Vold + Aold * (0.5 - beta) * dt * dt + beta * dt * dt * A
        V = Vold + Aold * (1 - gamma) * dt + gamma * dt * A
        R = ops.residual_static(U, 1.0)
        rhs = R - damping_mass * band_symv(M, n, bw, V) - band_symv(M, n, bw, A)
        r0 = np.linalg.norm(rhs)
        r0 = r0 if r0 != 0 else 1.0
        ok = False
        for it in range(max_iter):
            K = ops.jacobian(U)
            lhs = (1.0 + gamma * dt * damping_mass) * M + dt * dt * beta * K
            fac = ops.factor(lhs)
            dA = ops.solve(fac, rhs)
            A = A + dA
            V = V + dA * gamma * dt
            U = U + dA * beta * dt * dt
            an, dan = np.linalg.norm(A), np.linalg.norm(dA)
            upd = dan / an if an != 0 else dan
            R = ops.residual_static(U, 1.0)
            rhs = R - damping_mass * band_symv(M, n, bw, V) - band_symv(M, n, bw, A)
            rn = np.linalg.norm(rhs) / r0
            if upd < tol_u and rn < tol_f:
                ok = True
                break
        if not ok:
            status = 1
            break
        t += dt
        out["time"].append(t)
        out["U"].append(U.copy())
        out["V"].append(V.copy())
        out["A"].append(A.copy())
        out["iterations"].append(it + 1)
        out["unorm"].append(np.linalg.norm(U))
    res = {key: np.array(v) for key, v in out.items()}
    res.update(steps_done=len(out["time"]), status=status, jacobians=ops.n_jac)
    return res
