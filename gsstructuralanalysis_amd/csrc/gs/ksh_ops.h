// ksh_ops.h -- builds the gsStructuralAnalysisOps callbacks of one shell problem on libksh.so.
//
// These lambdas are what the reference drivers write around gsThinShellAssembler
// (benchmarks/benchmark_Roof.cpp:323-344, tutorials/nonlinear_shell_static.cpp:120-136,
// examples/example_DisplacementControl.cpp:236-254): constructSolution + assemble* + the sign
// convention of the residual.  Here they call the device path through an executor (direct, or a
// batch engine that merges the calls of many problems into one launch); K never leaves the GPU.
#pragma once

#include "gsStructuralAnalysisTypes.h"

namespace gismo {

class gsShellOps {
public:
    gsShellOps(ksh_problem *pb, int precision, gsExec *exec = gsExec::direct())
        : m_pb(pb), m_exec(exec), m_pool(pb, precision, exec)
    {
        int64_t n = 0;
        ksh_num_free_dofs(pb, &n);
        m_n = (index_t)n;
    }

    index_t numDofs() const { return m_n; }
    ksh_problem *problem() const { return m_pb; }
    gsMatrixPool &pool() { return m_pool; }
    gsExec *exec() const { return m_exec; }

    /// F = rhs() of assemble() (load at lambda = 1)
    gsVector<real_t> force() const
    {
        gsVector<real_t> F(m_n);
        if (ksh_assemble_force(m_pb, F.data()) != KSH_SUCCESS) throw std::runtime_error(ksh_last_error());
        return F;
    }

    /// Jacobian_t: K(u)
    gsStructuralAnalysisOps<real_t>::Jacobian_t jacobian()
    {
        return [this](gsVector<real_t> const &u, gsSparseMatrix<real_t> &m) -> bool {
            m = m_pool.acquire<real_t>();
            return m_exec->tangent(m_pb, u.data(), m.handle()) == KSH_SUCCESS;
        };
    }

    /// ALResidual_t of the arc-length solvers: r = F_int(u) - lambda F
    gsStructuralAnalysisOps<real_t>::ALResidual_t alResidual()
    {
        return [this](gsVector<real_t> const &u, const real_t lam, gsVector<real_t> &r) -> bool {
            r.resize(m_n);
            return m_exec->residual(m_pb, u.data(), lam, KSH_RES_ALM, r.data()) == KSH_SUCCESS;
        };
    }

    /// ALResidual_t in the static sign convention used with gsStaticNewton under load
    /// control: R = lambda F - F_int(u)
    gsStructuralAnalysisOps<real_t>::ALResidual_t staticALResidual()
    {
        return [this](gsVector<real_t> const &u, const real_t lam, gsVector<real_t> &r) -> bool {
            r.resize(m_n);
            return m_exec->residual(m_pb, u.data(), lam, KSH_RES_STATIC, r.data()) == KSH_SUCCESS;
        };
    }

    /// Residual_t: R = F - F_int(u)
    gsStructuralAnalysisOps<real_t>::Residual_t residual()
    {
        return [this](gsVector<real_t> const &u, gsVector<real_t> &r) -> bool {
            r.resize(m_n);
            return m_exec->residual(m_pb, u.data(), 1.0, KSH_RES_STATIC, r.data()) == KSH_SUCCESS;
        };
    }

    /// Mass_t: consistent mass matrix
    gsStructuralAnalysisOps<real_t>::Mass_t mass()
    {
        return [this](gsSparseMatrix<real_t> &m) -> bool {
            m = m_pool.acquire<real_t>();
            return ksh_assemble_mass(m_pb, m.handle()) == KSH_SUCCESS;
        };
    }

private:
    ksh_problem *m_pb;
    gsExec *m_exec;
    gsMatrixPool m_pool;
    index_t m_n = 0;
};

}  // namespace gismo
