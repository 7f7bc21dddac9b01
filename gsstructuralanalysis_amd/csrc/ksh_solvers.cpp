// ksh_solvers.cpp -- C-ABI eigen drivers on the device operator surface.
//
// ksh_run_buckling / ksh_run_modal : solvers/buckling_shell_XML.cpp:212-247 and
//                  solvers/modal_shell_XML.cpp:233-253 (matrices, then computeSparse)
// (the path drivers ksh_run_alm / ksh_batch_run_alm / ksh_run_static are in ksh_path.cpp)
#include <chrono>
#include <cstring>
#include <exception>

#include "gs/gsEigenSolvers.h"
#include "gs/ksh_ops.h"
#include "ksh_internal.hpp"

using namespace gismo;

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
