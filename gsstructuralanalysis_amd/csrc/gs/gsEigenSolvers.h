// gsEigenSolvers.h -- gsEigenProblemBase, gsModalSolver and gsBucklingSolver on the device band.
//
// Restates src/gsEigenSolvers/gsEigenProblemBase.h:40-89 (options), gsEigenProblemBase.hpp:
// 152-216 (computeSparse for the shift modes), gsModalSolver.h:40-57 (A = K, B = M) and
// gsBucklingSolver.h:70-149 (A = K_L, B = K_NL - K_L, or K_NL assembled at u = K_L^-1 F) over
// the compat device matrices.  computeSparse calls ksh_eigen (thick-restart Lanczos on the
// shift-invert operator, solves with the GPU band factor of A - shift B) in place of Spectra.
#pragma once

#include <cstdio>

#include "gsStructuralAnalysisTypes.h"

namespace gismo {

template <class T>
class gsEigenProblemBase {
public:
    gsEigenProblemBase() : m_status(gsStatus::NotStarted) { m_options = defaultOptions(); }
    virtual ~gsEigenProblemBase() = default;

    virtual gsOptionList defaultOptions()
    {
        gsOptionList o;
        o.addSwitch("verbose", "Verbose output", false);
        o.addInt("solver", "0 Cholesky, 1 RegularInverse, 2 ShiftInvert, 3 Buckling, 4 Cayley", 0);
        o.addInt("selectionRule", "Spectra selection rule", 4);
        o.addInt("sortRule", "Spectra sort rule", 4);
        o.addInt("ncvFac", "Lanczos basis = ncvFac * number of eigenvalues", 3);
        o.addReal("tolerance", "Ritz residual tolerance", 1e-10);
        o.addReal("shift", "Shift of the eigenvalue solver", 0.0);
        // device path
        o.addInt("Precision", "factor of A - shift B: 0 FP64, 1 FP32 with FP64 refinement", KSH_FP64);
        o.addInt("MaxRefine", "refinement sweeps per solve (Precision 1)", 8);
        o.addReal("RefineTol", "refinement tolerance (Precision 1)", 1e-13);
        return o;
    }
    gsOptionList &options() { return m_options; }
    virtual void setOptions(gsOptionList &o) { m_options.update(o, gsOptionList::addIfUnknown); }

    /// number lowest-|lambda - shift| eigenpairs of A x = lambda B x
    virtual gsStatus computeSparse(const index_t number = 10)
    {
        if (m_status == gsStatus::AssemblyError) return m_status;
        const int solver = m_options.getInt("solver");
        if (solver == 4) {
            std::fprintf(stderr, "gsEigenProblemBase: the Cayley mode is not available; use 2 or 3\n");
            m_status = gsStatus::NotStarted;
            return m_status;
        }
        ksh_eig_options o{};
        o.mode = solver == 3 ? KSH_EIG_BUCKLING : KSH_EIG_SHIFT_INVERT;
        o.number = (int)number;
        o.ncv_fac = m_options.getInt("ncvFac");
        o.shift = m_options.getReal("shift");
        o.tol = m_options.getReal("tolerance");
        o.precision = m_options.getInt("Precision");
        o.max_refine = o.precision == KSH_FP64 ? 0 : m_options.getInt("MaxRefine");
        o.refine_tol = m_options.getReal("RefineTol");
        const index_t n = m_A.rows();
        m_values = gsVector<T>(number);
        m_vectors.assign((size_t)n * number, T(0));
        const int rc = ksh_eigen(m_A.handle(), m_B.handle(), &o, m_values.data(), m_vectors.data(), &m_result);
        m_status = rc == KSH_SUCCESS ? gsStatus::Success
                                     : (rc == KSH_NOT_CONVERGED ? gsStatus::NotConverged : gsStatus::SolverError);
        if (m_options.getSwitch("verbose"))
            std::fprintf(stderr, "eigen: %d/%d converged, %d restarts, %lld solves\n", m_result.converged, (int)number,
                         m_result.restarts, (long long)m_result.solves);
        return m_status;
    }

    const gsVector<T> &values() const { return m_values; }
    /// eigenvector j (free dofs) = vectors()[j*n .. (j+1)*n)
    const std::vector<T> &vectors() const { return m_vectors; }
    gsStatus status() const { return m_status; }
    const ksh_eig_result &result() const { return m_result; }

protected:
    gsSparseMatrix<T> m_A, m_B;
    gsOptionList m_options;
    gsStatus m_status;
    gsVector<T> m_values;
    std::vector<T> m_vectors;
    ksh_eig_result m_result{};
};

/// free vibration: K x = omega^2 M x (src/gsEigenSolvers/gsModalSolver.h:40-57)
template <class T>
class gsModalSolver : public gsEigenProblemBase<T> {
    using Base = gsEigenProblemBase<T>;

public:
    gsModalSolver(const gsSparseMatrix<T> &stiffness, const gsSparseMatrix<T> &mass)
    {
        m_A = stiffness;
        m_B = mass;
    }

protected:
    using Base::m_A;
    using Base::m_B;
};

/// linear buckling: K_L x = lambda (K_NL - K_L) x (src/gsEigenSolvers/gsBucklingSolver.h:70-149)
template <class T>
class gsBucklingSolver : public gsEigenProblemBase<T> {
    using Base = gsEigenProblemBase<T>;

public:
    /// K_NL assembled by the caller (gsBucklingSolver.h:105-110)
    gsBucklingSolver(const gsSparseMatrix<T> &linear, const gsSparseMatrix<T> &nonlinear)
    {
        m_A = linear;
        m_B = nonlinear - m_A;
        Base::m_options.setInt("solver", 3);
    }

    /// K_NL = dJacobian(K_L^-1 scaling rhs, 0) (gsBucklingSolver.h:70-84,127-149)
    gsBucklingSolver(const gsSparseMatrix<T> &linear, const gsVector<T> &rhs,
                     typename gsStructuralAnalysisOps<T>::Jacobian_t &nonlinear, T scaling = 1.0)
    {
        m_A = linear;
        Base::m_options.setInt("solver", 3);
        typename gsSparseSolver<T>::SimplicialLDLT solver;
        solver.compute(m_A);
        gsVector<T> u = solver.solve(scaling * rhs);
        gsSparseMatrix<T> K_NL;
        try {
            if (!nonlinear(u, K_NL)) throw 2;
        } catch (...) {
            Base::m_status = gsStatus::AssemblyError;
            return;
        }
        m_B = K_NL - m_A;
    }

protected:
    using Base::m_A;
    using Base::m_B;
};

}  // namespace gismo
