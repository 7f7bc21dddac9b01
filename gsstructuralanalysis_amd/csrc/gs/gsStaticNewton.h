// gsStaticNewton.h -- static Newton solver, its base class and displacement/load control,
// over the device operator surface.
//
// Same options, constructors, linear / nonlinear solves, residual bookkeeping and status
// mapping as the reference:
//   gsStaticBase           src/gsStaticSolvers/gsStaticBase.h:48-179
//   gsStaticNewton         src/gsStaticSolvers/gsStaticNewton.h:48-143, gsStaticNewton.hpp:19-316
//   gsControlDisplacement  src/gsStaticSolvers/gsControlDisplacement.h:38-86
// gsStaticNewton keeps K_L and F by const reference, as the reference does
// (gsStaticNewton.h:230-231): the caller keeps them alive.
#pragma once

#include "gsStructuralAnalysisTypes.h"

namespace gismo {

template <class T>
class gsStaticBase {
protected:
    typedef typename gsStructuralAnalysisOps<T>::Residual_t Residual_t;
    typedef typename gsStructuralAnalysisOps<T>::ALResidual_t ALResidual_t;
    typedef typename gsStructuralAnalysisOps<T>::Jacobian_t Jacobian_t;
    typedef typename gsStructuralAnalysisOps<T>::dJacobian_t dJacobian_t;

public:
    virtual ~gsStaticBase() {}
    virtual gsStatus solve() = 0;
    virtual void initialize()
    {
        this->reset();
        this->getOptions();
    }
    virtual void initOutput() {}
    virtual void stepOutput(index_t) {}
    virtual void defaultOptions()
    {
        m_options.addReal("tol", "Relative Tolerance", 1e-6);
        m_options.addReal("tolF", "Residual relative tolerance", -1);
        m_options.addReal("tolU", "Solution step relative tolerance", -1);
        m_options.addInt("maxIt", "Maximum number of iterations", 25);
        m_options.addInt("verbose", "Verbose output", 0);
        m_options.addInt("BifurcationMethod", "0: Determinant; 1: Eigenvalue", 1);
        m_options.addString("Solver", "Sparse linear solver", "SimplicialLDLT");
    }
    virtual void getOptions()
    {
        m_tolF = m_options.getReal("tolF") != -1 ? m_options.getReal("tolF") : m_options.getReal("tol");
        m_tolU = m_options.getReal("tolU") != -1 ? m_options.getReal("tolU") : m_options.getReal("tol");
        m_maxIterations = m_options.getInt("maxIt");
        m_verbose = m_options.getInt("verbose");
        m_stabilityMethod = m_options.getInt("BifurcationMethod");
        m_solver = gsSparseSolver<T>::get(m_options.askString("Solver", "SimplicialLDLT"));
    }
    virtual void setOptions(gsOptionList &options) { m_options.update(options, gsOptionList::addIfUnknown); }
    virtual const gsOptionList &options() const { return m_options; }
    virtual gsOptionList &options() { return m_options; }
    virtual gsVector<T> solution() const { return m_U; }
    virtual gsVector<T> update() const { return m_DeltaU; }
    virtual void setDisplacement(const gsVector<T> &displacement)
    {
        m_start = true;
        m_U = displacement;
    }
    virtual void setLoad(const T L) { m_L = L; }
    virtual void setSolution(const gsVector<T> &displacement, const T L)
    {
        this->setDisplacement(displacement);
        this->setLoad(L);
    }
    virtual void setUpdate(const gsVector<T> &update)
    {
        m_headstart = true;
        m_DeltaU = update;
    }
    virtual index_t iterations() const { return m_numIterations; }
    virtual bool converged() const { return m_status == gsStatus::Success; }
    virtual gsStatus status() const { return m_status; }
    virtual index_t numDofs() { return m_dofs; }
    virtual void reset()
    {
        m_U.setZero(m_dofs);
        m_DeltaU.setZero(m_dofs);
        m_deltaU.setZero(m_dofs);
        m_R.setZero(m_dofs);
        m_L = m_DeltaL = m_deltaL = 0.0;
        m_headstart = false;
    }
    virtual T indicator(const gsSparseMatrix<T> &jacMat, T shift = -1e-2)
    {
        _computeStability(jacMat, shift);
        return m_indicator;
    }
    index_t negatives() const { return m_negatives; }
    T residualNorm() const { return m_residual; }

protected:
    // determinant method: factor and take the inertia (gsStaticBase.h:162-179)
    virtual bool _computeStabilityDet(const gsSparseMatrix<T> &jacMat)
    {
        m_solver->compute(jacMat);
        if (m_solver->info() != gsEigen::Success) return false;
        if (auto *s = dynamic_cast<typename gsSparseSolver<T>::SimplicialLDLT *>(m_solver.get())) {
            T dmin = 0;
            m_negatives = (index_t)s->negatives(&dmin);
            m_indicator = dmin;
        }
        return true;
    }
    virtual bool _computeStability(const gsSparseMatrix<T> &jacMat, T)
    {
        return _computeStabilityDet(jacMat);
    }

    mutable gsOptionList m_options;
    T m_L = 0, m_DeltaL = 0, m_deltaL = 0;
    gsVector<T> m_U, m_DeltaU, m_deltaU, m_R;
    T m_residual = 0, m_residualIni = 0, m_residualOld = 0;
    index_t m_numIterations = 0, m_maxIterations = 25;
    bool m_start = false;
    index_t m_headstart = 0;
    index_t m_verbose = 0;
    index_t m_dofs = 0;
    T m_tolF = 1e-6, m_tolU = 1e-6;
    T m_indicator = 0;
    index_t m_negatives = 0;
    mutable typename gsSparseSolver<T>::uPtr m_solver;
    gsStatus m_status = gsStatus::NotStarted;
    index_t m_stabilityMethod = 0;
};

template <class T>
class gsStaticNewton : public gsStaticBase<T> {
    typedef gsStaticBase<T> Base;

public:
    typedef typename Base::Residual_t Residual_t;
    typedef typename Base::ALResidual_t ALResidual_t;
    typedef typename Base::Jacobian_t Jacobian_t;
    typedef typename Base::dJacobian_t dJacobian_t;

    gsStaticNewton(const gsSparseMatrix<T> &linear, const gsVector<T> &force)
        : m_linear(linear), m_force(force), m_nonlinear(nullptr), m_residualFun(nullptr)
    {
        _init();
    }
    gsStaticNewton(const gsSparseMatrix<T> &linear, const gsVector<T> &force, const Jacobian_t &nonlinear,
                   const Residual_t &residual)
        : m_linear(linear), m_force(force), m_nonlinear(nonlinear), m_dnonlinear(nullptr), m_residualFun(residual),
          m_ALresidualFun(nullptr)
    {
        m_dnonlinear = [this](gsVector<T> const &x, gsVector<T> const &, gsSparseMatrix<T> &m) -> bool { return m_nonlinear(x, m); };
        _init();
    }
    gsStaticNewton(const gsSparseMatrix<T> &linear, const gsVector<T> &force, const Jacobian_t &nonlinear,
                   const ALResidual_t &ALResidual)
        : m_linear(linear), m_force(force), m_nonlinear(nonlinear), m_dnonlinear(nullptr), m_residualFun(nullptr),
          m_ALresidualFun(ALResidual)
    {
        this->m_L = 1.0;
        m_residualFun = [this](gsVector<T> const &x, gsVector<T> &result) -> bool { return m_ALresidualFun(x, this->m_L, result); };
        m_dnonlinear = [this](gsVector<T> const &x, gsVector<T> const &, gsSparseMatrix<T> &m) -> bool { return m_nonlinear(x, m); };
        _init();
    }
    gsStaticNewton(const gsSparseMatrix<T> &linear, const gsVector<T> &force, const dJacobian_t &dnonlinear,
                   const Residual_t &residual)
        : m_linear(linear), m_force(force), m_nonlinear(nullptr), m_dnonlinear(dnonlinear), m_residualFun(residual),
          m_ALresidualFun(nullptr)
    {
        _init();
    }

    gsStatus solve() override { return m_NL ? solveNonlinear() : solveLinear(); }
    gsStatus solveLinear()
    {
        return guarded([this] { _solveLinear(); });
    }
    gsStatus solveNonlinear()
    {
        return guarded([this] { _solveNonlinear(); });
    }

    void defaultOptions() override
    {
        Base::defaultOptions();
        this->m_options.setString("Solver", "CGDiagonal");
        this->m_options.addReal("Relaxation", "Relaxation parameter", 1);
    }
    void getOptions() override
    {
        Base::getOptions();
        m_relax = this->m_options.getReal("Relaxation");
    }
    void reset() override
    {
        this->m_dofs = m_force.rows();
        Base::reset();
    }
    void stepOutput(index_t k) override
    {
        gsInfo << "\t" << k << "\t" << this->m_residual << "\t" << this->m_residual / this->m_residualIni << "\t"
               << this->m_DeltaU.norm() << "\t" << m_relax * this->m_deltaU.norm() << "\n";
    }

    index_t jacobianCount() const { return m_nJac; }

protected:
    template <class F> gsStatus guarded(F f)
    {
        try {
            f();
            this->m_status = gsStatus::Success;
        } catch (int errorCode) {
            this->m_status = status_from_code(errorCode);
        } catch (...) {
            this->m_status = gsStatus::OtherError;
        }
        return this->m_status;
    }

    gsVector<T> _solveLinear()
    {
        this->getOptions();
        _factorizeMatrix(m_linear);
        this->m_U = _solveSystem(m_force);
        return this->m_U;
    }

    gsVector<T> _solveNonlinear()
    {
        this->getOptions();
        if (this->m_DeltaU.norm() == 0 && this->m_DeltaU.rows() == 0) {
            this->m_deltaU = this->m_DeltaU = this->_solveLinear();
            this->m_U.setZero();
            this->m_headstart = true;
        }
        _start();
        gsSparseMatrix<T> jacMat;
        for (this->m_numIterations = 0; this->m_numIterations != this->m_maxIterations; ++this->m_numIterations) {
            jacMat = this->_computeJacobian(this->m_U + this->m_DeltaU, this->m_deltaU);
            _factorizeMatrix(jacMat);
            this->m_deltaU = _solveSystem(this->m_R);
            this->m_DeltaU += m_relax * this->m_deltaU;
            this->m_R = this->_computeResidual(this->m_U + this->m_DeltaU);
            this->m_residual = this->m_R.norm();
            if (this->m_verbose > 0) stepOutput(this->m_numIterations);
            this->m_residualOld = this->m_residual;
            if (m_relax * this->m_deltaU.norm() / this->m_DeltaU.norm() < this->m_tolU &&
                this->m_residual / this->m_residualIni < this->m_tolF) {
                this->m_U += this->m_DeltaU;
                // inertia of the last factored tangent (determinant stability indicator)
                if (auto *s = dynamic_cast<typename gsSparseSolver<T>::SimplicialLDLT *>(this->m_solver.get())) {
                    T dmin = 0;
                    this->m_negatives = (index_t)s->negatives(&dmin);
                    this->m_indicator = dmin;
                }
                break;
            } else if (this->m_numIterations + 1 == this->m_maxIterations) {
                this->m_U += this->m_DeltaU;
                throw 1;
            }
        }
        return this->m_U;
    }

    gsVector<T> _computeResidual(const gsVector<T> &U)
    {
        gsVector<T> resVec;
        if (!m_residualFun(U, resVec)) throw 2;
        return resVec;
    }
    gsSparseMatrix<T> _computeJacobian(const gsVector<T> &U, const gsVector<T> &deltaU)
    {
        gsSparseMatrix<T> m;
        ++m_nJac;
        if (!m_dnonlinear(U, deltaU, m)) throw 2;
        return m;
    }
    void _factorizeMatrix(const gsSparseMatrix<T> &jacMat) const
    {
        this->m_solver->compute(jacMat);
        if (this->m_solver->info() != gsEigen::Success) throw 3;
    }
    gsVector<T> _solveSystem(const gsVector<T> &F)
    {
        try {
            return this->m_solver->solve(F);
        } catch (...) {
            throw 3;
        }
    }

    void _init()
    {
        this->reset();
        m_NL = !(m_dnonlinear == nullptr || m_residualFun == nullptr);
        this->m_stabilityMethod = 0;
        this->m_start = false;
        this->m_dofs = m_force.rows();
        this->m_residual = this->m_residualIni = this->m_residualOld = 0;
        defaultOptions();
        this->m_status = gsStatus::NotStarted;
    }

    void _start()
    {
        if (!this->m_headstart) {
            this->m_DeltaU.setZero(this->m_dofs);
            this->m_R = this->_computeResidual(this->m_U);
            this->m_residual = this->m_R.norm();
            if (this->m_residual == 0) this->m_residual = 1;
            this->m_residualIni = this->m_residualOld = this->m_residual;
        } else {
            this->m_R = this->_computeResidual(this->m_U + this->m_DeltaU);
            this->m_residual = this->m_R.norm();
            if (this->m_residual == 0) this->m_residual = 1;
            this->m_residualOld = this->m_residual;
            this->m_residualIni = this->_computeResidual(this->m_U).norm();
            if (this->m_residualIni == 0) this->m_residualIni = 1;
            this->m_headstart = false;
        }
        this->m_deltaU.setZero(this->m_dofs);
    }

    const gsSparseMatrix<T> &m_linear;
    const gsVector<T> &m_force;
    const Jacobian_t m_nonlinear;
    dJacobian_t m_dnonlinear;
    Residual_t m_residualFun;
    const ALResidual_t m_ALresidualFun;
    bool m_NL = false;
    T m_relax = 1;
    index_t m_nJac = 0;
};

template <class T>
class gsControlDisplacement {
public:
    explicit gsControlDisplacement(gsStaticBase<T> *solver) : m_solver(solver), first(true)
    {
        m_U = gsVector<T>::Zero(m_solver->numDofs());
        m_L = 0;
    }
    gsStatus step(T dL)
    {
        m_solver->setLoad(m_L + dL);
        if (!first) m_solver->setDisplacement(m_U);
        else first = false;
        m_status = m_solver->solve();
        if (m_status == gsStatus::Success) {
            m_L += dL;
            m_U += m_solver->update();
        }
        return m_status;
    }
    gsVector<T> &solutionU() { return m_U; }
    T solutionL() { return m_L; }
    void reset() { m_solver->reset(); }
    void setZero()
    {
        m_U.setZero();
        m_L = 0;
        this->reset();
    }

protected:
    gsStaticBase<T> *m_solver;
    bool first;
    gsVector<T> m_U;
    T m_L;
    gsStatus m_status = gsStatus::NotStarted;
};

}  // namespace gismo
