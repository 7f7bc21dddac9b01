// ksh_solvers.cpp -- C-ABI drivers that run the solver classes on the device operator surface.
//
// ksh_run_alm    : benchmarks/benchmark_Roof.cpp:320-512 (options, initialize, step loop,
//                  halve the arc length and restore (Uold, Lold) on NotConverged/AssemblyError)
// ksh_run_static : examples/example_DisplacementControl.cpp:236-304 (K_L, F, gsStaticNewton with
//                  the load-controlled residual, gsControlDisplacement steps)
// ksh_run_buckling / ksh_run_modal : solvers/buckling_shell_XML.cpp:212-247 and
//                  solvers/modal_shell_XML.cpp:233-253 (matrices, then computeSparse)
#include <chrono>
#include <cstring>
#include <exception>

#include "gs/gsALMCrisfield.h"
#include "gs/gsALMCrisfieldLearned.h"
#include "gs/gsALMRiks.h"
#include "gs/gsEigenSolvers.h"
#include "gs/gsStaticNewton.h"
#include "gs/ksh_ops.h"
#include "ksh_internal.hpp"

using namespace gismo;

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

namespace ksh {

// One arc-length path on the operator surface of `pb`, device calls issued through `exec`
// (direct, or an engine that merges the calls of many problems).
int alm_driver(ksh_problem *pb, gsExec *exec, const ksh_alm_options *o, ksh_alm_result *res)
{
    if (!pb || !o || !res) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    try {
        gsSparseSolver<real_t>::settings().max_refine = o->max_refine;
        gsSparseSolver<real_t>::settings().refine_tol = o->refine_tol;
        if (o->inertia_guard > 0) pb->inertia_guard = o->inertia_guard;
        if (o->refine_accept > 0) pb->refine_accept = o->refine_accept;
        gsShellOps ops(pb, o->precision, exec);
        const index_t n = ops.numDofs();
        gsVector<real_t> Force = ops.force();
        auto Jacobian = ops.jacobian();
        auto ALResidual = ops.alResidual();

        std::unique_ptr<gsALMBase<real_t>> alm;
        gsALMCrisfieldLearned<real_t> *learned = nullptr;
        if (o->method == 0 && o->learned) {
            learned = new gsALMCrisfieldLearned<real_t>(Jacobian, ALResidual, Force, pb, o->learned, exec);
            alm.reset(learned);
        } else if (o->method == 0) alm.reset(new gsALMCrisfield<real_t>(Jacobian, ALResidual, Force));
        else if (o->method == 1) alm.reset(new gsALMRiks<real_t>(Jacobian, ALResidual, Force));
        else return ksh::fail(KSH_OTHER_ERROR, "unknown arc-length method");

        gsOptionList &op = alm->options();
        op.setString("Solver", "SimplicialLDLT");
        op.setInt("BifurcationMethod", o->bifurcation_method);
        op.setReal("Length", o->length);
        if (o->method == 0) {
            op.setInt("AngleMethod", o->angle_method);
            op.setReal("Scaling", o->scaling);
        }
        op.setSwitch("AdaptiveLength", o->adaptive != 0);
        op.setInt("AdaptiveIterations", o->adaptive_iterations);
        op.setReal("Tol", o->tol);
        op.setReal("TolU", o->tol_u);
        op.setReal("TolF", o->tol_f);
        op.setInt("MaxIter", o->max_iter);
        op.setSwitch("Verbose", o->verbose != 0);
        op.setReal("Relaxation", o->relaxation);
        if (o->quasi_iterations > 0) op.setInt("QuasiIterations", o->quasi_iterations);
        op.setSwitch("Quasi", o->quasi != 0 || o->quasi_iterations > 0);
        alm->applyOptions();

        double t0 = now_s();
        alm->initialize();
        real_t Lold = 0;
        gsVector<real_t> Uold = gsVector<real_t>::Zero(n);
        alm->setIndicator(0.0);
        bool bisected = false;
        real_t dL = o->length, dL0 = o->length;
        res->steps_done = 0;
        res->halvings = 0;
        res->final_status = 0;
        res->learned_steps = 0;
        res->fallbacks = 0;
        for (int k = 0; k < o->nsteps; ++k) {
            gsStatus status = alm->step();
            if (learned && learned->usedLearned()) {
                ++res->learned_steps;
                if (status != gsStatus::Success) {
                    // the corrector rejected the learned proposal: redo the step from the tangent predictor
                    ++res->fallbacks;
                    alm->setSolution(Uold, Lold);
                    learned->setLearned(false);
                    status = alm->step();
                    learned->setLearned(true);
                }
            }
            res->final_status = (int)status;
            if (status == gsStatus::NotConverged || status == gsStatus::AssemblyError) {
                if (res->halvings >= o->max_halvings) break;
                ++res->halvings;
                dL = dL / 2.;
                alm->setLength(dL);
                alm->setSolution(Uold, Lold);
                bisected = true;
                k -= 1;
                continue;
            }
            if (status != gsStatus::Success) break;
            Uold = alm->solutionU();
            Lold = alm->solutionL();
            if (learned) learned->record();
            if (res->features && ksh_predictor_features(pb, Uold.data(), Lold, res->features + (size_t)k * 66) != KSH_SUCCESS)
                return ksh::fail(KSH_OTHER_ERROR, ksh_last_error());
            if (res->lambda) res->lambda[k] = Lold;
            if (res->unorm) res->unorm[k] = Uold.norm();
            if (res->indicator) res->indicator[k] = alm->indicator();
            if (res->negatives) res->negatives[k] = alm->negatives();
            if (res->iterations) res->iterations[k] = alm->numIterations();
            if (res->residue_f) res->residue_f[k] = alm->residueF();
            if (res->length) res->length[k] = dL;
            if (res->U) std::memcpy(res->U + (size_t)k * n, Uold.data(), sizeof(double) * n);
            if (res->w && res->probe_dof >= 0 && res->probe_dof < n) res->w[k] = Uold[res->probe_dof];
            res->steps_done = k + 1;
            if (!bisected) {
                dL = dL0;
                alm->setLength(dL);
            }
            bisected = false;
        }
        ksh_synchronize(pb);
        res->wall_seconds = now_s() - t0;
        if (auto *c = dynamic_cast<gsALMCrisfield<real_t> *>(alm.get())) res->phi = c->phi();
        res->jacobians = alm->jacobianCount();
        res->residuals = alm->residualCount();
        res->solves = alm->solveCount();
    } catch (const std::exception &e) {
        return ksh::fail(KSH_OTHER_ERROR, e.what());
    } catch (int code) {
        // the solver classes signal failures outside step()/solve() (e.g. initialize()) with the
        // reference's integer codes: 1 NotConverged, 2 AssemblyError, 3 SolverError
        return ksh::fail(code >= 1 && code <= 3 ? code : KSH_OTHER_ERROR,
                         std::string("solver raised status ") + std::to_string(code) + ": " + ksh_last_error());
    } catch (...) {
        return ksh::fail(KSH_OTHER_ERROR, "unknown exception in the solver driver");
    }
    return res->steps_done == o->nsteps ? KSH_SUCCESS : KSH_NOT_CONVERGED;
}

}  // namespace ksh

extern "C" int ksh_run_alm(ksh_problem *pb, const ksh_alm_options *o, ksh_alm_result *res)
{
    return ksh::alm_driver(pb, gsExec::direct(), o, res);
}

extern "C" int ksh_run_static(ksh_problem *pb, const ksh_static_options *o, ksh_static_result *res)
{
    if (!pb || !o || !res) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    try {
        gsSparseSolver<real_t>::settings().max_refine = o->max_refine;
        gsSparseSolver<real_t>::settings().refine_tol = o->refine_tol;
        if (o->inertia_guard > 0) pb->inertia_guard = o->inertia_guard;
        if (o->refine_accept > 0) pb->refine_accept = o->refine_accept;
        gsShellOps ops(pb, o->precision);
        const index_t n = ops.numDofs();
        double t0 = now_s();
        gsVector<real_t> F = ops.force();
        auto Jacobian = ops.jacobian();
        auto ALResidual = ops.staticALResidual();
        gsSparseMatrix<real_t> K;
        if (!Jacobian(gsVector<real_t>::Zero(n), K)) return ksh::fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
        gsStaticNewton<real_t> NWT(K, F, Jacobian, ALResidual);
        gsOptionList &op = NWT.options();
        op.setInt("maxIt", o->max_iter);
        op.setReal("tol", o->tol);
        op.setInt("verbose", o->verbose);
        op.setString("Solver", "SimplicialLDLT");
        op.setInt("BifurcationMethod", 0);
        NWT.initialize();
        gsControlDisplacement<real_t> control(&NWT);
        res->steps_done = 0;
        res->final_status = 0;
        for (int k = 0; k < o->nsteps; ++k) {
            gsStatus st = control.step(o->dlambda);
            res->final_status = (int)st;
            if (st != gsStatus::Success) break;
            if (res->lambda) res->lambda[k] = control.solutionL();
            if (res->unorm) res->unorm[k] = control.solutionU().norm();
            if (res->iterations) res->iterations[k] = NWT.iterations() + 1;
            if (res->negatives) res->negatives[k] = NWT.negatives();
            if (res->residual) res->residual[k] = NWT.residualNorm();
            if (res->U) std::memcpy(res->U + (size_t)k * n, control.solutionU().data(), sizeof(double) * n);
            res->steps_done = k + 1;
        }
        ksh_synchronize(pb);
        res->wall_seconds = now_s() - t0;
        res->jacobians = NWT.jacobianCount();
    } catch (const std::exception &e) {
        return ksh::fail(KSH_OTHER_ERROR, e.what());
    } catch (int code) {
        // the solver classes signal failures outside step()/solve() (e.g. initialize()) with the
        // reference's integer codes: 1 NotConverged, 2 AssemblyError, 3 SolverError
        return ksh::fail(code >= 1 && code <= 3 ? code : KSH_OTHER_ERROR,
                         std::string("solver raised status ") + std::to_string(code) + ": " + ksh_last_error());
    } catch (...) {
        return ksh::fail(KSH_OTHER_ERROR, "unknown exception in the solver driver");
    }
    return res->steps_done == o->nsteps ? KSH_SUCCESS : KSH_NOT_CONVERGED;
}

namespace {

void eig_options(gsOptionList &op, const ksh_eig_options *o, int solver)
{
    op.setInt("solver", solver);
    op.setInt("ncvFac", o->ncv_fac > 0 ? o->ncv_fac : 3);
    op.setReal("tolerance", o->tol > 0 ? o->tol : 1e-10);
    op.setReal("shift", o->shift);
    op.setInt("Precision", o->precision);
    if (o->max_refine > 0) op.setInt("MaxRefine", o->max_refine);
    if (o->refine_tol > 0) op.setReal("RefineTol", o->refine_tol);
}

int eig_finish(gsEigenProblemBase<real_t> &solver, const ksh_eig_options *o, index_t n, double *values, double *vectors,
               ksh_eig_result *res)
{
    gsStatus st = solver.computeSparse(o->number);
    if (res) *res = solver.result();
    if (st != gsStatus::Success) return ksh::fail(st == gsStatus::NotConverged ? KSH_NOT_CONVERGED : KSH_SOLVER_ERROR, ksh_last_error());
    for (int j = 0; j < o->number; ++j) values[j] = solver.values()[j];
    if (vectors) std::memcpy(vectors, solver.vectors().data(), sizeof(double) * (size_t)n * o->number);
    return KSH_SUCCESS;
}

}  // namespace

extern "C" int ksh_run_buckling(ksh_problem *pb, const ksh_eig_options *o, double *values, double *vectors,
                                ksh_eig_result *res)
{
    if (!pb || !o || !values) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    try {
        gsShellOps ops(pb, KSH_FP64);
        const index_t n = ops.numDofs();
        gsVector<real_t> F = ops.force();
        auto Jacobian = ops.jacobian();
        gsSparseMatrix<real_t> K_L, K_NL;
        if (!Jacobian(gsVector<real_t>::Zero(n), K_L)) return ksh::fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
        // linear solution under F (the reference driver solves it with CGDiagonal; the LDL^T
        // solution is the same vector to that solver's tolerance)
        gsSparseSolver<real_t>::SimplicialLDLT lin;
        lin.compute(K_L);
        if (lin.info() != gsEigen::Success) return ksh::fail(KSH_SOLVER_ERROR, ksh_last_error());
        gsVector<real_t> u = lin.solve(F);
        if (!Jacobian(u, K_NL)) return ksh::fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
        gsBucklingSolver<real_t> solver(K_L, K_NL);
        eig_options(solver.options(), o, 3);
        return eig_finish(solver, o, n, values, vectors, res);
    } catch (const std::exception &e) {
        return ksh::fail(KSH_OTHER_ERROR, e.what());
    } catch (...) {
        return ksh::fail(KSH_OTHER_ERROR, "unknown exception in the buckling driver");
    }
}

extern "C" int ksh_run_modal(ksh_problem *pb, const ksh_eig_options *o, double *values, double *vectors, ksh_eig_result *res)
{
    if (!pb || !o || !values) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    try {
        gsShellOps ops(pb, KSH_FP64);
        const index_t n = ops.numDofs();
        auto Jacobian = ops.jacobian();
        auto Mass = ops.mass();
        gsSparseMatrix<real_t> K, M;
        if (!Jacobian(gsVector<real_t>::Zero(n), K) || !Mass(M)) return ksh::fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
        gsModalSolver<real_t> solver(K, M);
        eig_options(solver.options(), o, 2);
        return eig_finish(solver, o, n, values, vectors, res);
    } catch (const std::exception &e) {
        return ksh::fail(KSH_OTHER_ERROR, e.what());
    } catch (...) {
        return ksh::fail(KSH_OTHER_ERROR, "unknown exception in the modal driver");
    }
}
