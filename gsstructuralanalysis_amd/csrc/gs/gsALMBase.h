// gsALMBase.h -- arc-length continuation base class over the device operator surface.
//
// Same public interface, options, step logic, convergence norms, error -> gsStatus mapping and
// stability bookkeeping as the reference (src/gsALMSolvers/gsALMBase.h, gsALMBase.hpp):
//   options            gsALMBase.hpp:22-54      getOptions        gsALMBase.hpp:56-93
//   step / _step       gsALMBase.hpp:327-416    residual norms    gsALMBase.hpp:150-183
//   factorize / solve  gsALMBase.hpp:185-211    jacobians         gsALMBase.hpp:213-246
//   stability          gsALMBase.hpp:546-627    singular points   gsALMBase.hpp:421-856
// The determinant stability method takes the negative-pivot count and min(D) of the device
// LDL^T (gsALMHelper.h:30-42 countNegatives semantics).  The eigenvalue method needs Spectra in
// the reference; here it is served by the same inertia (the count of negative eigenvalues equals
// the count of negative pivots) and reported with a warning.
#pragma once

#include <algorithm>
#include <string>

#include "gsStructuralAnalysisTypes.h"

namespace gismo {

template <class T>
inline index_t sign(T val)
{
    return (T(0) < val) - (val < T(0));
}

template <class T>
class gsALMBase {
protected:
    typedef typename gsStructuralAnalysisOps<T>::ALResidual_t ALResidual_t;
    typedef typename gsStructuralAnalysisOps<T>::Jacobian_t Jacobian_t;
    typedef typename gsStructuralAnalysisOps<T>::dJacobian_t dJacobian_t;

public:
    virtual ~gsALMBase() {}

    gsALMBase(const Jacobian_t &Jacobian, const ALResidual_t &ALResidual, const gsVector<T> &Force)
        : m_residualFun(ALResidual), m_forcing(Force)
    {
        m_jacobian = Jacobian;
        m_djacobian = [this](gsVector<T> const &x, gsVector<T> const &, gsSparseMatrix<T> &m) -> bool {
            return m_jacobian(x, m);
        };
        commonInit();
    }

    gsALMBase(const dJacobian_t &dJacobian, const ALResidual_t &ALResidual, const gsVector<T> &Force)
        : m_residualFun(ALResidual), m_forcing(Force)
    {
        m_djacobian = dJacobian;
        commonInit();
    }

    virtual gsStatus status() { return m_status; }
    virtual index_t numDofs() { return m_forcing.size(); }
    virtual T getLength() { return m_arcLength; }
    virtual gsStatus step();

    virtual void initialize(bool stability = true)
    {
        m_initialized = true;
        this->initMethods();
        this->init(stability);
    }

    virtual void setLength(T length)
    {
        m_options.setReal("Length", length);
        m_arcLength = m_arcLength_prev = m_arcLength_ori = m_options.getReal("Length");
    }
    virtual void setLength(T length, bool adaptive)
    {
        this->setLength(length);
        m_adaptiveLength = adaptive;
        m_desiredIterations = 10;
    }
    virtual void setLength(T length, bool adaptive, index_t iterations)
    {
        this->setLength(length);
        m_adaptiveLength = adaptive;
        m_desiredIterations = iterations;
    }

    virtual bool converged() const { return m_converged; }
    virtual index_t numIterations() const { return m_numIterations; }
    virtual T tolerance() const { return m_tolerance; }
    virtual T residue() const { return m_residue; }
    virtual T residueF() const { return m_residueF; }
    virtual T residueU() const { return m_residueU; }
    virtual T indicator() const { return m_indicator; }
    virtual index_t negatives() const { return m_negatives; }
    virtual void setIndicator(T indicator)
    {
        m_indicator = indicator;
        m_stability = this->stability();
    }

    virtual const gsVector<T> &solutionU() const { return m_U; }
    virtual const gsVector<T> &solutionDU() const { return m_DeltaU; }
    virtual T solutionL() const { return m_L; }
    virtual T solutionDL() const { return m_DeltaL; }
    virtual const gsVector<T> &solutionV() const { return m_V; }
    virtual bool isStable() const { return m_stability; }

    virtual void resetStep()
    {
        m_DeltaUold.setZero();
        m_DeltaLold = 0;
    }
    virtual void setInitialGuess(const gsVector<T> &Uguess, const T &Lguess)
    {
        m_Uguess = Uguess;
        m_Lguess = Lguess;
    }
    virtual void setPrevious(const gsVector<T> &Uprev, const T &Lprev)
    {
        m_Uprev = Uprev;
        m_Lprev = Lprev;
        m_DeltaUold = m_U - m_Uprev;
        m_DeltaLold = m_L - m_Lprev;
    }
    virtual void setSolution(const gsVector<T> &U, const T &L)
    {
        m_L = L;
        m_U = U;
    }
    virtual void setSolutionStep(const gsVector<T> &DU, const T &DL)
    {
        m_DeltaUold = DU;
        m_DeltaLold = DL;
    }

    virtual gsOptionList &options() { return m_options; }
    virtual void setOptions(gsOptionList options)
    {
        m_options.update(options, gsOptionList::addIfUnknown);
        this->getOptions();
    }
    virtual void applyOptions() { this->getOptions(); }
    virtual T distance(const gsVector<T> &, const T) const { throw std::logic_error("distance: not implemented"); }

    // singular points
    virtual gsStatus computeSingularPoint(const gsVector<T> &U, const T &L, bool switchBranch = false, bool jacobian = false,
                                          bool testPoint = true);
    virtual bool isBifurcation(bool jacobian = false) { return this->_testSingularPoint(jacobian); }
    virtual bool stabilityChange() const { return m_stability * m_stabilityPrev < 0; }
    virtual gsStatus computeStability(bool jacobian = true, T shift = -1e2);
    virtual index_t stability() const { return (m_indicator < 0) ? 1 : -1; }
    virtual void switchBranch();
    virtual T reduceLength(T fac = 0.5)
    {
        m_arcLength *= fac;
        return m_arcLength;
    }
    virtual T resetLength()
    {
        m_arcLength = m_arcLength_prev = m_arcLength_ori;
        if (m_adaptiveLength) computeLength();
        return m_arcLength;
    }

    // counters for measurement (not part of the reference interface)
    index_t jacobianCount() const { return m_nJac; }
    index_t residualCount() const { return m_nRes; }
    index_t solveCount() const { return m_nSolve; }

    struct bifmethod { enum type { Nothing = -1, Determinant = 0, Eigenvalue = 1 }; };
    struct SPfail { enum type { Without = 0, With = 1 }; };

protected:
    void commonInit()
    {
        m_numIterations = 0;
        this->defaultOptionsBase();
        this->setLength(1e-2);
        m_converged = false;
        m_basisResidualF = 0.0;
        m_basisResidualU = 0.0;
        m_status = gsStatus::NotStarted;
        m_initialized = false;
        m_indicator = 0;
        m_negatives = 0;
        m_stability = m_stabilityPrev = -1;
    }

    virtual void _computeStability(const gsVector<T> &x, bool jacobian = true, T shift = -1e2);
    virtual void _computeSingularPoint(const gsVector<T> &U, const T &L, bool switchBranch, bool jacobian, bool testPoint);
    virtual bool _testSingularPoint(bool jacobian = false);
    virtual void _extendedSystemIteration();
    virtual index_t _bisectionObjectiveFunction(const gsVector<T> &x, bool jacobian = true);
    virtual T _bisectionTerminationFunction(const gsVector<T> &x, bool jacobian = true);
    virtual bool _extendedSystemSolve(const gsVector<T> &U, const T L, const T tol);
    virtual bool _bisectionSolve(const gsVector<T> &U, const T L, const T tol);

    virtual void _step();
    void defaultOptionsBase();
    virtual void getOptions();
    virtual void init(bool stability)
    {
        if (stability) {
            this->_computeStability(m_U);
            m_stability = this->stability();
        }
    }

    virtual void factorizeMatrix(const gsSparseMatrix<T> &M)
    {
        m_solver->compute(M);
        if (m_solver->info() != gsEigen::Success) {
            gsInfo << "Solver error: " << ksh_last_error() << "\n";
            throw 3;
        }
    }
    virtual gsVector<T> solveSystem(const gsVector<T> &F)
    {
        ++m_nSolve;
        try {
            return m_solver->solve(F);
        } catch (...) {
            throw 3;
        }
    }

    virtual gsVector<T> computeResidual(const gsVector<T> &U, const T &L)
    {
        gsVector<T> resVec;
        ++m_nRes;
        if (!m_residualFun(U, L, resVec)) throw 2;
        return resVec;
    }
    virtual void computeResidual() { m_resVec = this->computeResidual(m_U + m_DeltaU, m_L + m_DeltaL); }
    virtual void computeResidualNorms();

    virtual gsSparseMatrix<T> _computeJacobian(const gsVector<T> &U, const gsVector<T> &deltaU)
    {
        gsSparseMatrix<T> m;
        m_note += "J";
        ++m_nJac;
        if (!m_djacobian(U, deltaU, m)) throw 2;
        this->factorizeMatrix(m);
        return m;
    }
    virtual gsSparseMatrix<T> computeJacobian(const gsVector<T> &U, const gsVector<T> &deltaU)
    {
        return this->_computeJacobian(U, deltaU);
    }
    virtual gsSparseMatrix<T> computeJacobian(const gsVector<T> &U)
    {
        gsVector<T> DeltaU = gsVector<T>::Zero(U.rows());
        return this->computeJacobian(U, DeltaU);
    }
    virtual gsSparseMatrix<T> computeJacobian()
    {
        if (m_deltaU.rows() == 0) m_deltaU = gsVector<T>::Zero(m_DeltaU.rows());
        return this->computeJacobian(m_U + m_DeltaU, m_deltaU);
    }

    virtual void computeLength()
    {
        T fac = T(m_desiredIterations) / T(m_numIterations);
        fac = std::min(T(2.0), std::max(T(0.5), fac));
        m_arcLength_prev = m_arcLength;
        m_arcLength = m_arcLength * fac;
    }
    virtual void computeUt() { m_deltaUt = this->solveSystem(m_forcing); }
    virtual void computeUbar() { m_deltaUbar = this->solveSystem(-m_resVec); }

    virtual void initMethods() = 0;
    virtual void initiateStep() = 0;
    virtual void iterationFinish() = 0;
    virtual void quasiNewtonPredictor() = 0;
    virtual void quasiNewtonIteration() = 0;
    virtual void predictor() = 0;
    virtual void predictorGuess() = 0;
    virtual void iteration() = 0;
    virtual void initOutput() = 0;
    virtual void stepOutput() = 0;

protected:
    index_t m_numDof = 0;
    Jacobian_t m_jacobian;
    dJacobian_t m_djacobian;
    const ALResidual_t m_residualFun;
    const gsVector<T> m_forcing;
    mutable typename gsSparseSolver<T>::uPtr m_solver;

    mutable gsOptionList m_options;
    index_t m_numIterations = 0, m_maxIterations = 100, m_desiredIterations = 10;
    T m_arcLength = 0, m_arcLength_prev = 0, m_arcLength_ori = 0;
    bool m_adaptiveLength = false;
    T m_tolerance = 1e-6, m_toleranceF = 1e-3, m_toleranceU = 1e-6;
    bool m_verbose = false, m_initialized = false, m_quasiNewton = false;
    index_t m_quasiNewtonInterval = -1;
    std::string m_note;
    gsStatus m_status;
    bool m_converged = false;
    T m_residue = 0, m_residueF = 0, m_basisResidualF = 0, m_residueU = 0, m_basisResidualU = 0, m_residueL = 0;
    T m_residueKTPhi = 0, m_basisResidualKTPhi = 0;
    T m_indicator = 0;
    index_t m_negatives = 0;
    T m_relax = 1.0;

    gsVector<T> m_DeltaUold;
    T m_DeltaLold = 0;
    gsVector<T> m_U, m_Uprev, m_DeltaU, m_deltaUbar, m_deltaUt, m_deltaU;
    T m_L = 0, m_Lprev = 0, m_DeltaL = 0, m_deltaL = 0;
    gsVector<T> m_deltaLs;
    gsVector<T> m_Uguess;
    T m_Lguess = 0;
    gsSparseMatrix<T> m_jacMat;
    T m_detKT = 0;
    gsVector<T> m_resVec;
    gsVector<T> m_V, m_deltaV, m_deltaVbar, m_deltaVt, m_DeltaV;
    index_t m_stabilityPrev = -1, m_stability = -1;
    index_t m_bifurcationMethod = 0;
    index_t m_SPfail = 1, m_SPTestIt = 5;
    T m_SPTestTol = 1e-6, m_SPCompTolE = 1e-10, m_SPCompTolB = 0, m_tau = 1e3;
    index_t m_nJac = 0, m_nRes = 0, m_nSolve = 0;
};

// ------------------------------------------------------------------------------------------
template <class T>
void gsALMBase<T>::defaultOptionsBase()
{
    m_options.addInt("MaxIter", "Maximum iterations", 100);
    m_options.addReal("Tol", "Tolerance", 1e-6);
    m_options.addReal("TolF", "Tolerance", 1e-3);
    m_options.addReal("TolU", "Tolerance", 1e-6);
    m_options.addReal("Perturbation", "Set Perturbation factor Tau", 1e3);
    m_options.addReal("Length", "Arclength", 1e-2);
    m_options.addSwitch("AdaptiveLength", "Adaptive length", false);
    m_options.addInt("AdaptiveIterations", "Desired iterations for adaptive length", 10);
    m_options.addSwitch("Quasi", "Use Quasi Newton method", false);
    m_options.addInt("QuasiIterations", "Number of iterations for quasi newton method", -1);
    m_options.addInt("BifurcationMethod", "-1: nothing, 0: Determinant; 1: Eigenvalue", bifmethod::Eigenvalue);
    m_options.addInt("SingularPointFailure", "0 = apply solution anyways; 1 = proceed without singular point", SPfail::With);
    m_options.addReal("SingularPointTestTol", "Detection tolerance for singular points", 1e-6);
    m_options.addInt("SingularPointTestIt", "Power iterations of the singular point test", 5);
    m_options.addReal("SingularPointComputeTolE", "Tolerance of the extended iterations", 1e-10);
    m_options.addReal("SingularPointComputeTolB", "Tolerance of the bisection iterations (0: none)", 0);
    m_options.addString("Solver", "Sparse linear solver", "SimplicialLDLT");
    m_options.addSwitch("Verbose", "Verbose output", false);
    m_options.addReal("Relaxation", "Set Relaxation factor alpha", 1.0);
}

template <class T>
void gsALMBase<T>::getOptions()
{
    m_maxIterations = m_options.getInt("MaxIter");
    m_tolerance = m_options.getReal("Tol");
    m_toleranceF = m_options.getReal("TolF");
    m_toleranceU = m_options.getReal("TolU");
    m_tau = m_options.getReal("Perturbation");
    m_adaptiveLength = m_options.getSwitch("AdaptiveLength");
    m_desiredIterations = m_options.getInt("AdaptiveIterations");
    m_quasiNewton = m_options.getSwitch("Quasi");
    m_quasiNewtonInterval = m_options.getInt("QuasiIterations");
    m_bifurcationMethod = m_options.getInt("BifurcationMethod");
    m_solver = gsSparseSolver<T>::get(m_options.getString("Solver"));
    bool isLDLT = dynamic_cast<typename gsSparseSolver<T>::SimplicialLDLT *>(m_solver.get()) != nullptr;
    if (!isLDLT && m_bifurcationMethod == bifmethod::Determinant) {
        gsWarn << "Determinant method cannot be used with solvers other than LDLT. Bifurcation method will be set to 'Eigenvalue'.\n";
        m_bifurcationMethod = bifmethod::Eigenvalue;
    }
    m_verbose = m_options.getSwitch("Verbose");
    m_relax = m_options.getReal("Relaxation");
    m_arcLength = m_arcLength_prev = m_arcLength_ori = m_options.getReal("Length");
    m_SPfail = m_options.getInt("SingularPointFailure");
    m_SPTestTol = m_options.getReal("SingularPointTestTol");
    m_SPTestIt = m_options.getInt("SingularPointTestIt");
    m_SPCompTolE = m_options.getReal("SingularPointComputeTolE");
    m_SPCompTolB = m_options.getReal("SingularPointComputeTolB");
}

template <class T>
void gsALMBase<T>::computeResidualNorms()
{
    if (m_numIterations == 0) {
        m_basisResidualF = ((m_L + m_DeltaL) * m_forcing).norm();
        m_basisResidualU = m_DeltaU.norm();
        m_residueF = m_resVec.norm() / (((m_L + m_DeltaL) * m_forcing).norm());
        m_residueU = m_deltaU.norm() / m_DeltaU.norm();
        m_residueL = m_DeltaL;
    } else {
        m_residueF = m_resVec.norm() / m_basisResidualF;
        m_residueU = m_deltaU.norm() / m_basisResidualU;
        m_residueL = m_deltaL / m_DeltaL;
    }
}

template <class T>
gsStatus gsALMBase<T>::step()
{
    try {
        _step();
        m_status = gsStatus::Success;
    } catch (int errorCode) {
        m_status = status_from_code(errorCode);
    } catch (...) {
        m_status = gsStatus::OtherError;
    }
    return m_status;
}

template <class T>
void gsALMBase<T>::_step()
{
    if (!m_initialized) throw std::logic_error("Arc-Length Method is not initialized! Call initialize()");
    if (m_verbose) initOutput();
    m_converged = false;
    m_numIterations = 0;
    initiateStep();

    bool haveGuess = m_Uguess.rows() != 0 && m_Uguess.cols() != 0 && (m_Uguess - m_U).norm() != 0 && (m_Lguess - m_L) != 0;
    if (haveGuess) predictorGuess();
    else predictor();

    computeResidual();
    computeResidualNorms();
    if (m_verbose) stepOutput();
    if (m_quasiNewton) quasiNewtonPredictor();

    m_stabilityPrev = m_stability;
    for (m_numIterations = 1; m_numIterations < m_maxIterations; ++m_numIterations) {
        if ((!m_quasiNewton) || ((m_quasiNewtonInterval > 0) && (m_numIterations % m_quasiNewtonInterval) < 1e-10))
            quasiNewtonIteration();
        iteration();
        computeStability(false);
        computeResidual();
        computeResidualNorms();
        if (m_verbose) stepOutput();
        if (m_residueF < m_toleranceF && m_residueU < m_toleranceU) {
            iterationFinish();
            if (m_adaptiveLength) computeLength();
            else m_arcLength_prev = m_arcLength;
            break;
        } else if (m_numIterations == m_maxIterations - 1) {
            if (m_verbose) gsInfo << "maximum iterations reached. Solution did not converge\n";
            throw 1;
        }
    }
}

template <class T>
gsStatus gsALMBase<T>::computeStability(bool jacobian, T shift)
{
    try {
        _computeStability(m_U, jacobian, shift);
        m_status = gsStatus::Success;
    } catch (int errorCode) {
        m_status = status_from_code(errorCode);
    } catch (...) {
        m_status = gsStatus::OtherError;
    }
    return m_status;
}

template <class T>
void gsALMBase<T>::_computeStability(const gsVector<T> &x, bool jacobian, T /*shift*/)
{
    if (jacobian) {
        gsVector<T> dx = gsVector<T>::Zero(x.size());
        m_jacMat = this->computeJacobian(x, dx);
        this->factorizeMatrix(m_jacMat);
    }
    if (m_bifurcationMethod == bifmethod::Nothing) {
        m_negatives = 0;
        m_indicator = 0;
    } else {
        auto *s = dynamic_cast<typename gsSparseSolver<T>::SimplicialLDLT *>(m_solver.get());
        if (m_bifurcationMethod == bifmethod::Eigenvalue) {
            static bool warned = false;
            if (!warned) gsWarn << "Eigenvalue stability method: negative count taken from the LDL^T inertia\n";
            warned = true;
            if (!s) throw 3;
        }
        if (!s) {
            gsWarn << "Determinant stability method only works with SimplicialLDLT\n";
            throw 3;
        }
        factorizeMatrix(m_jacMat);
        T dmin = 0;
        m_negatives = (index_t)s->negatives(&dmin);
        m_indicator = dmin;
    }
    m_stability = (m_indicator < 0) ? 1 : -1;
}

// ---- singular points (gsALMBase.hpp:421-856) ------------------------------------------------
template <class T>
void gsALMBase<T>::_computeSingularPoint(const gsVector<T> &U, const T &L, bool switchBranch, bool jacobian, bool testPoint)
{
    bool test = testPoint ? this->_testSingularPoint(jacobian) : true;
    m_U = U;
    m_L = L;
    if (test) {
        if (m_SPCompTolB != 0) this->_bisectionSolve(m_U, m_L, m_SPCompTolB);
        bool converged = this->_extendedSystemSolve(m_U, m_L, m_SPCompTolE);
        if (switchBranch && (converged || m_SPfail == 1)) this->switchBranch();
        m_stability = m_stabilityPrev;
    }
}

template <class T>
gsStatus gsALMBase<T>::computeSingularPoint(const gsVector<T> &U, const T &L, bool switchBranch, bool jacobian, bool testPoint)
{
    try {
        this->_computeSingularPoint(U, L, switchBranch, jacobian, testPoint);
    } catch (int errorCode) {
        m_status = status_from_code(errorCode);
    } catch (...) {
        m_status = gsStatus::OtherError;
    }
    return m_status;
}

template <class T>
bool gsALMBase<T>::_testSingularPoint(bool jacobian)
{
    if (jacobian) {
        m_jacMat = this->computeJacobian(m_U, m_deltaU);
        this->factorizeMatrix(m_jacMat);
    }
    m_V = gsVector<T>::Ones(m_numDof);
    m_V.normalize();
    factorizeMatrix(m_jacMat);
    for (index_t k = 0; k < m_SPTestIt; k++) {
        m_V = this->solveSystem(m_V);
        m_V.normalize();
    }
    T dot = std::fabs(m_V.dot(m_forcing));
    return dot < m_SPTestTol;
}

template <class T>
bool gsALMBase<T>::_extendedSystemSolve(const gsVector<T> &U, const T L, const T tol)
{
    m_U = U;
    m_L = L;
    m_jacMat = this->computeJacobian(m_U, m_deltaU);
    this->factorizeMatrix(m_jacMat);
    m_basisResidualKTPhi = (m_jacMat * m_V).norm();
    m_DeltaV = gsVector<T>::Zero(m_numDof);
    m_DeltaU.setZero();
    m_DeltaL = 0.0;
    m_deltaV = gsVector<T>::Zero(m_numDof);
    m_deltaU.setZero();
    m_deltaL = 0.0;
    for (m_numIterations = 0; m_numIterations < m_maxIterations; ++m_numIterations) {
        _extendedSystemIteration();
        m_DeltaU += m_deltaU;
        m_DeltaL += m_deltaL;
        m_DeltaV += m_deltaV;
        m_jacMat = this->computeJacobian(m_U + m_DeltaU, m_deltaU);
        this->factorizeMatrix(m_jacMat);
        m_residueKTPhi = (m_jacMat * (m_V + m_DeltaV)).norm();
        m_resVec = this->computeResidual(m_U + m_DeltaU, m_L + m_DeltaL);
        computeResidualNorms();
        if (m_residueKTPhi < tol) {
            m_U += m_DeltaU;
            m_L += m_DeltaL;
            break;
        }
        if (m_numIterations == m_maxIterations - 1) {
            if (m_SPfail == 1) {
                m_U += m_DeltaU;
                m_L += m_DeltaL;
            } else {
                m_DeltaV = gsVector<T>::Zero(m_numDof);
                m_DeltaU.setZero();
                m_DeltaL = 0.0;
                m_deltaV = gsVector<T>::Zero(m_numDof);
                m_deltaU.setZero();
                m_deltaL = 0.0;
            }
            return false;
        }
    }
    return true;
}

template <class T>
void gsALMBase<T>::_extendedSystemIteration()
{
    m_resVec = this->computeResidual(m_U + m_DeltaU, m_L + m_DeltaL);
    m_jacMat = this->computeJacobian(m_U + m_DeltaU, m_deltaU);
    this->factorizeMatrix(m_jacMat);
    m_deltaUt = this->solveSystem(m_forcing);
    m_deltaUbar = this->solveSystem(-m_resVec);
    const T eps = 1e-8;
    gsSparseMatrix<T> jacMatEps = this->computeJacobian((m_U + m_DeltaU) + eps * (m_V + m_DeltaV));
    m_note += "J";
    gsVector<T> h1 = (1 / eps) * (jacMatEps * m_deltaUt) - (1 / eps) * m_forcing;
    gsVector<T> h2 = m_jacMat * (m_V + m_DeltaV) + (1 / eps) * (jacMatEps * m_deltaUbar + m_resVec);
    this->factorizeMatrix(m_jacMat);
    m_deltaVt = this->solveSystem(-h1);
    m_deltaVbar = this->solveSystem(-h2);
    gsVector<T> VV = m_V + m_DeltaV;
    T nVV = VV.norm();
    gsVector<T> e = VV / nVV;
    m_deltaL = -(e.dot(m_deltaVbar) + nVV - 1) / e.dot(m_deltaVt);
    m_deltaU = m_deltaL * m_deltaUt + m_deltaUbar;
    m_deltaV = m_deltaL * m_deltaVt + m_deltaVbar;
}

template <class T>
index_t gsALMBase<T>::_bisectionObjectiveFunction(const gsVector<T> &x, bool jacobian)
{
    this->_computeStability(x, jacobian);
    m_note += "(" + std::to_string(m_negatives) + ")";
    return m_negatives;
}

template <class T>
T gsALMBase<T>::_bisectionTerminationFunction(const gsVector<T> &x, bool jacobian)
{
    this->_computeStability(x, jacobian);
    return m_indicator;
}

template <class T>
bool gsALMBase<T>::_bisectionSolve(const gsVector<T> &U, const T L, const T tol)
{
    m_U = U;
    m_L = L;
    gsVector<T> U_old = m_U;
    T L_old = L;
    T dL = m_arcLength;
    bool adaptiveBool = m_adaptiveLength;
    m_adaptiveLength = false;
    bool converged = false;
    T referenceError = _bisectionTerminationFunction(m_U, true);
    index_t fa = _bisectionObjectiveFunction(m_U, false), fb;
    for (index_t k = 1; k < m_maxIterations; ++k) {
        m_U = U_old;
        m_L = L_old;
        gsStatus status = step();
        fb = _bisectionObjectiveFunction(m_U, true);
        if (fa == fb && status == gsStatus::Success) {
            U_old = m_U;
            L_old = m_L;
            fa = fb;
        } else
            m_arcLength = m_arcLength / 2.0;
        T term = _bisectionTerminationFunction(m_U, false);
        if (std::fabs(term / referenceError) < tol) {
            converged = true;
            break;
        }
    }
    m_arcLength = dL;
    m_adaptiveLength = adaptiveBool;
    m_V = gsVector<T>::Ones(m_numDof);
    m_V.normalize();
    gsVector<T> Vold = gsVector<T>::Zero(m_numDof);
    for (index_t k = 0; k < 5; k++) {
        this->factorizeMatrix(m_jacMat);
        m_V = this->solveSystem(m_V);
        m_V.normalize();
        if ((m_V - Vold).norm() < m_tolerance) {
            converged = true;
            break;
        }
        Vold = m_V;
    }
    return converged;
}

template <class T>
void gsALMBase<T>::switchBranch()
{
    m_V.normalize();
    T lenPhi = m_V.norm();
    T xi = lenPhi / m_tau;
    m_DeltaU = xi * m_V;
    m_U = m_U + m_DeltaU;
    m_DeltaLold = 0.0;
    m_DeltaUold.setZero();
}

}  // namespace gismo
