// gsALMRiks.h -- Riks' linearised arc-length method over the device operator surface.
//
// Same steps as the reference (src/gsALMSolvers/gsALMRiks.h, gsALMRiks.hpp):
//   initMethods :22-31, quasi-Newton :37-53, iteration :68-84, initiateStep :86-95,
//   predictor :97-134, predictorGuess :136-170, distance gsALMRiks.h:81-84.
#pragma once

#include "gsALMBase.h"

namespace gismo {

template <class T>
class gsALMRiks : public gsALMBase<T> {
    typedef gsALMBase<T> Base;

public:
    typedef typename Base::ALResidual_t ALResidual_t;
    typedef typename Base::Jacobian_t Jacobian_t;
    typedef typename Base::dJacobian_t dJacobian_t;

    gsALMRiks(const Jacobian_t &Jacobian, const ALResidual_t &ALResidual, const gsVector<T> &Force)
        : Base(Jacobian, ALResidual, Force)
    {
        Base::getOptions();
        initMethods();
    }
    gsALMRiks(const dJacobian_t &dJacobian, const ALResidual_t &ALResidual, const gsVector<T> &Force)
        : Base(dJacobian, ALResidual, Force)
    {
        Base::getOptions();
        initMethods();
    }

    T distance(const gsVector<T> &DeltaU, const T DeltaL) const override
    {
        T nu = DeltaU.norm();
        return std::sqrt(m_phi * nu * nu + (1.0 - m_phi) * DeltaL * DeltaL);
    }

protected:
    void initMethods() override
    {
        this->m_numDof = this->m_forcing.size();
        this->m_DeltaU = this->m_U = gsVector<T>::Zero(this->m_numDof);
        this->m_DeltaL = this->m_L = 0.0;
        this->m_Uprev = gsVector<T>::Zero(this->m_numDof);
        this->m_Lprev = 0.0;
    }
    void quasiNewtonPredictor() override
    {
        this->m_jacMat = this->computeJacobian();
        this->factorizeMatrix(this->m_jacMat);
        this->computeUt();
        this->computeUbar();
    }
    void quasiNewtonIteration() override
    {
        this->m_jacMat = this->computeJacobian();
        this->factorizeMatrix(this->m_jacMat);
        this->computeUt();
    }
    void iteration() override
    {
        this->computeUbar();
        const gsVector<T> &DU = this->m_DeltaU;
        T r = m_phi * DU.dot(DU) + (1 - m_phi) * this->m_DeltaL * this->m_DeltaL - this->m_arcLength * this->m_arcLength;
        this->m_deltaL = -(r + 2 * m_phi * DU.dot(this->m_deltaUbar)) /
                         (2 * (1 - m_phi) * this->m_DeltaL + 2 * m_phi * DU.dot(this->m_deltaUt));
        this->m_deltaU = this->m_deltaUbar + this->m_deltaL * this->m_deltaUt;
        this->m_DeltaU += this->m_deltaU;
        this->m_DeltaL += this->m_deltaL;
    }
    void initiateStep() override
    {
        this->m_DeltaU = this->m_deltaU = gsVector<T>::Zero(this->m_numDof);
        this->m_DeltaL = this->m_deltaL = 0.0;
    }
    void predictor() override
    {
        this->m_jacMat = this->computeJacobian();
        this->factorizeMatrix(this->m_jacMat);
        m_phi = (this->m_numDof == 1) ? 0.99999999999 : 1. / this->m_numDof;
        const T tol = 1e-10;
        if (((this->m_U - this->m_Uprev).norm() < tol) && ((this->m_L - this->m_Lprev) * (this->m_L - this->m_Lprev) < tol)) {
            this->m_note += "predictor\t";
            T DL = 1.;
            this->m_deltaUt = this->solveSystem(this->m_forcing);
            T den = std::sqrt(this->m_deltaUt.dot(this->m_deltaUt) + this->m_DeltaL * DL);
            this->m_deltaU = this->m_deltaUt / den;
            this->m_deltaL = DL / den;
        } else {
            this->m_deltaL = 1. / this->m_arcLength_prev * (this->m_L - this->m_Lprev);
            this->m_deltaU = 1. / this->m_arcLength_prev * (this->m_U - this->m_Uprev);
        }
        this->m_deltaL *= this->m_arcLength;
        this->m_deltaU *= this->m_arcLength;
        this->m_DeltaU += this->m_deltaU;
        this->m_DeltaL += this->m_deltaL;
    }
    void predictorGuess() override
    {
        this->m_jacMat = this->computeJacobian();
        this->factorizeMatrix(this->m_jacMat);
        this->m_deltaUt = this->solveSystem(this->m_forcing);
        const T tol = 1e-10;
        if (((this->m_Uguess - this->m_U).norm() < tol) && ((this->m_Lguess - this->m_L) * (this->m_Lguess - this->m_L) < tol)) {
            T DL = 1.;
            this->m_deltaUt = this->solveSystem(this->m_forcing);
            T den = std::sqrt(this->m_deltaUt.dot(this->m_deltaUt) + this->m_DeltaL * DL);
            this->m_deltaU = this->m_deltaUt / den;
            this->m_deltaL = DL / den;
        } else {
            this->m_deltaL = 1. / this->m_arcLength_prev * (this->m_Lguess - this->m_L);
            this->m_deltaU = 1. / this->m_arcLength_prev * (this->m_Uguess - this->m_U);
        }
        this->m_deltaL *= this->m_arcLength;
        this->m_deltaU *= this->m_arcLength;
        this->m_DeltaU = this->m_deltaU;
        this->m_DeltaL = this->m_deltaL;
        this->m_Uguess.resize(0);
    }
    void iterationFinish() override
    {
        this->m_converged = true;
        this->m_Uprev = this->m_U;
        this->m_Lprev = this->m_L;
        this->m_U += this->m_DeltaU;
        this->m_L += this->m_DeltaL;
    }
    void initOutput() override { gsInfo << "\tIt.\tRes. F\t\t|dU|/|Du|\tL\t(negatives)\n"; }
    void stepOutput() override
    {
        gsInfo << "\t" << this->m_numIterations << "\t" << this->m_residueF << "\t" << this->m_residueU << "\t"
               << (this->m_L + this->m_DeltaL) << "\t(" << this->m_negatives << ")\n";
    }

    T m_phi = 0;
};

}  // namespace gismo
