// gsALMCrisfield.h -- Crisfield's spherical arc-length method over the device operator surface.
//
// Same options, predictor, corrector, root selection and complex-root fallbacks as the
// reference (src/gsALMSolvers/gsALMCrisfield.h, gsALMCrisfield.hpp):
//   options / getOptions     gsALMCrisfield.hpp:22-38     initMethods          :40-49
//   quasi-Newton             gsALMCrisfield.hpp:55-71     iteration            :73-100
//   initiateStep / predictor gsALMCrisfield.hpp:102-164   predictorGuess       :166-191
//   iterationFinish          gsALMCrisfield.hpp:193-206   computeLambdas*      :212-363
//   computeLambdaMU          gsALMCrisfield.hpp:383-398   computeLambdaDOT     :400-472
//   distance                 gsALMCrisfield.h:79-83
// The scalar logic stays FP64 on the host; every K, R, factorization and solve is a device call.
#pragma once

#include "gsALMBase.h"

namespace gismo {

template <class T>
class gsALMCrisfield : public gsALMBase<T> {
    typedef gsALMBase<T> Base;

public:
    typedef typename Base::ALResidual_t ALResidual_t;
    typedef typename Base::Jacobian_t Jacobian_t;
    typedef typename Base::dJacobian_t dJacobian_t;

    gsALMCrisfield(const Jacobian_t &Jacobian, const ALResidual_t &ALResidual, const gsVector<T> &Force)
        : Base(Jacobian, ALResidual, Force)
    {
        defaultOptions();
        getOptions();
        initMethods();
    }
    gsALMCrisfield(const dJacobian_t &dJacobian, const ALResidual_t &ALResidual, const gsVector<T> &Force)
        : Base(dJacobian, ALResidual, Force)
    {
        defaultOptions();
        getOptions();
        initMethods();
    }

    T distance(const gsVector<T> &DeltaU, const T DeltaL) const override
    {
        T A0 = m_phi * m_phi * this->m_forcing.dot(this->m_forcing);
        return std::sqrt(DeltaU.dot(DeltaU) + A0 * DeltaL * DeltaL);
    }
    T phi() const { return m_phi; }

    struct angmethod { enum type { Step = 0, Iteration = 1, Predictor = 2 }; };

protected:
    void defaultOptions()
    {
        this->m_options.addReal("Scaling", "Set Scaling factor Phi", -1);
        this->m_options.addInt("AngleMethod", "0 = Previous step; 1 = Previous iteration", angmethod::Step);
    }
    void getOptions() override
    {
        Base::getOptions();
        m_phi = this->m_options.getReal("Scaling");
        m_phi_user = m_phi != -1;
        m_angleDetermine = this->m_options.getInt("AngleMethod");
    }

    void initMethods() override
    {
        this->m_numDof = this->m_forcing.size();
        this->m_DeltaU = this->m_U = gsVector<T>::Zero(this->m_numDof);
        this->m_DeltaL = this->m_L = 0.0;
        this->m_DeltaUold = gsVector<T>::Zero(this->m_numDof);
        this->m_DeltaLold = 0.0;
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
        m_eta = 1.0;
        T lamold = this->m_deltaL;
        computeLambdas();
        if (((lamold * this->m_deltaL < 0) && (std::fabs(this->m_deltaL) <= std::fabs(lamold))) && this->m_relax != 1.0) {
            this->m_note += "\t relaxated solution!";
            this->m_deltaU = this->m_relax * (this->m_deltaL * this->m_deltaUt + m_eta * this->m_deltaUbar);
            this->m_deltaL = this->m_relax * this->m_deltaL;
        }
        this->m_DeltaU += this->m_deltaU;
        this->m_DeltaL += this->m_deltaL;
        if (m_angleDetermine == angmethod::Iteration) {
            this->m_DeltaUold = this->m_DeltaU;
            this->m_DeltaLold = this->m_DeltaL;
        }
    }

    void initiateStep() override
    {
        this->m_DeltaU = this->m_deltaUbar = this->m_deltaUt = gsVector<T>::Zero(this->m_numDof);
        this->m_DeltaL = this->m_deltaL = 0.0;
        m_eta = 1.0;
    }

    void predictor() override
    {
        this->m_jacMat = this->computeJacobian();
        this->factorizeMatrix(this->m_jacMat);
        this->m_deltaUt = this->solveSystem(this->m_forcing);
        const gsVector<T> &F = this->m_forcing;
        if (this->m_DeltaUold.dot(this->m_DeltaUold) == 0 && this->m_DeltaLold * this->m_DeltaLold == 0) {
            this->m_note += "predictor\t";
            this->m_deltaL = this->m_arcLength / std::sqrt(2 * this->m_deltaUt.dot(this->m_deltaUt));
            this->m_deltaU = this->m_deltaUbar + this->m_deltaL * this->m_deltaUt;
            if (!m_phi_user) m_phi = std::sqrt(this->m_deltaUt.dot(this->m_deltaUt) / F.dot(F));
        } else {
            if (!m_phi_user) m_phi = std::sqrt(this->m_U.dot(this->m_U) / (this->m_L * this->m_L * F.dot(F)));
            computeLambdaMU();
        }
        this->m_DeltaU += this->m_deltaU;
        this->m_DeltaL += this->m_deltaL;
        if (m_angleDetermine == angmethod::Iteration || m_angleDetermine == angmethod::Predictor) {
            this->m_DeltaUold = this->m_DeltaU;
            this->m_DeltaLold = this->m_DeltaL;
        }
    }

    void predictorGuess() override
    {
        this->m_jacMat = this->computeJacobian();
        this->factorizeMatrix(this->m_jacMat);
        this->m_deltaUt = this->solveSystem(this->m_forcing);
        const gsVector<T> &F = this->m_forcing;
        if (!m_phi_user) m_phi = std::sqrt(this->m_deltaUt.dot(this->m_deltaUt) / F.dot(F));
        this->m_DeltaUold = -(this->m_Uguess - this->m_U);
        this->m_DeltaLold = -(this->m_Lguess - this->m_L);
        computeLambdaMU();
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
        if (m_angleDetermine == angmethod::Step) {
            this->m_DeltaUold = this->m_DeltaU;
            this->m_DeltaLold = this->m_DeltaL;
        }
    }

    // ---- load-factor computations -----------------------------------------------------------
    void computeLambdasSimple()
    {
        const gsVector<T> &F = this->m_forcing;
        T A0 = m_phi * m_phi * F.dot(F);
        const gsVector<T> &ut = this->m_deltaUt, &ub = this->m_deltaUbar, &DU = this->m_DeltaU;
        m_a0 = ut.dot(ut) + A0;
        m_b0 = 2 * (ut.dot(DU) + this->m_DeltaL * A0);
        m_b1 = 2 * (ub.dot(ut));
        m_c0 = DU.dot(DU) + this->m_DeltaL * this->m_DeltaL * A0 - this->m_arcLength * this->m_arcLength;
        m_c1 = 2 * (DU.dot(ub));
        m_c2 = ub.dot(ub);
        m_alpha1 = m_a0;
        m_alpha2 = m_b0 + m_eta * m_b1;
        m_alpha3 = m_c0 + m_eta * m_c1 + m_eta * m_eta * m_c2;
        m_discriminant = m_alpha2 * m_alpha2 - 4 * m_alpha1 * m_alpha3;
    }
    void computeLambdasEta()
    {
        m_alpha1 = m_a0;
        m_alpha2 = m_b0 + m_eta * m_b1;
        m_alpha3 = m_c0 + m_eta * m_c1 + m_eta * m_eta * m_c2;
        this->m_deltaLs[0] = (-m_alpha2) / (2 * m_alpha1);
        this->m_deltaLs[1] = (-m_alpha2) / (2 * m_alpha1);
    }
    void computeLambdasModified()
    {
        m_alpha1 = m_b1 * m_b1 - 4.0 * m_a0 * m_c2;
        m_alpha2 = 2.0 * m_b0 * m_b1 - 4.0 * m_a0 * m_c1;
        m_alpha3 = m_b0 * m_b0 - 4.0 * m_a0 * m_c0;
        m_discriminant = m_alpha2 * m_alpha2 - 4.0 * m_alpha1 * m_alpha3;
        if (m_discriminant >= 0) {
            T e0 = (-m_alpha2 + std::sqrt(m_discriminant)) / (2.0 * m_alpha1);
            T e1 = (-m_alpha2 - std::sqrt(m_discriminant)) / (2.0 * m_alpha1);
            T eta1 = std::min(e0, e1), eta2 = std::max(e0, e1);
            T xi = 0.05 * std::fabs(eta2 - eta1);
            // Lam 1992 selection
            if (eta2 < 1.0) m_eta = eta2 - xi;
            else if ((eta2 > 1.0) && (-m_alpha2 / m_alpha1 < 1.0)) m_eta = eta2 + xi;
            else if ((eta1 < 1.0) && (-m_alpha2 / m_alpha1 > 1.0)) m_eta = eta1 - xi;
            else if (eta1 > 1.0) m_eta = eta1 + xi;
        } else if (this->m_verbose) {
            gsInfo << "Discriminant was negative in modified method\n";
        }
    }
    void computeLambdasComplex()
    {
        const gsVector<T> &F = this->m_forcing;
        T A0 = m_phi * m_phi * F.dot(F);
        gsVector<T> DeltaUcr = this->m_DeltaU + this->m_deltaUbar;
        gsVector<T> Fint = this->m_jacMat * (this->m_U + this->m_DeltaU);
        T Lcr = Fint.dot(F) / F.dot(F);
        T DeltaLcr = Lcr - this->m_L;
        T arcLength_cr = std::sqrt(DeltaUcr.dot(DeltaUcr) + A0 * DeltaLcr * DeltaLcr);
        T mu = this->m_arcLength / arcLength_cr;
        this->m_deltaL = mu * DeltaLcr - this->m_DeltaL;
        this->m_deltaU = mu * DeltaUcr - this->m_DeltaU;
    }
    void computeLambdas()
    {
        this->m_deltaLs.setZero(2);
        computeLambdasSimple();
        if (m_discriminant >= 0) {
            m_eta = 1.0;
            this->m_deltaLs[0] = (-m_alpha2 + std::sqrt(m_discriminant)) / (2 * m_alpha1);
            this->m_deltaLs[1] = (-m_alpha2 - std::sqrt(m_discriminant)) / (2 * m_alpha1);
            computeLambdaDOT();
        } else {
            this->m_note += "\tC";
            computeLambdasModified();
            if ((m_discriminant >= 0) && m_eta > 0.05) {
                computeLambdasEta();
                computeLambdaDOT();
            } else {
                m_eta = 1.0;
                computeLambdasComplex();
            }
        }
    }
    void computeLambdaMU()
    {
        const gsVector<T> &F = this->m_forcing;
        T A0 = m_phi * m_phi * F.dot(F);
        index_t dir = sign(this->m_DeltaUold.dot(this->m_deltaUt) + A0 * this->m_DeltaLold);
        T denum = std::sqrt(this->m_deltaUt.dot(this->m_deltaUt) + A0);
        T mu = (denum == 0) ? this->m_arcLength : this->m_arcLength / denum;
        this->m_deltaL = dir * mu;
        this->m_deltaU = this->m_deltaL * this->m_deltaUt;
    }
    void computeLambdaDOT()
    {
        gsVector<T> deltaU1 = m_eta * this->m_deltaUbar + this->m_deltaUt * this->m_deltaLs[0];
        gsVector<T> deltaU2 = m_eta * this->m_deltaUbar + this->m_deltaUt * this->m_deltaLs[1];
        // Ritto-Correa et al. 2008
        T base = this->m_DeltaUold.dot(this->m_deltaUt) + m_phi * m_phi * this->m_DeltaLold;
        T DOT1 = this->m_deltaLs[0] * base, DOT2 = this->m_deltaLs[1] * base;
        if (DOT1 > DOT2) {
            this->m_deltaL = this->m_deltaLs[0];
            this->m_deltaU = deltaU1;
        } else if (DOT1 < DOT2) {
            this->m_deltaL = this->m_deltaLs[1];
            this->m_deltaU = deltaU2;
        } else {
            this->m_deltaL = this->m_deltaLs[0];
            this->m_deltaU = deltaU1;
        }
    }

    void initOutput() override
    {
        gsInfo << "\tIt.\tRes. F\t\t|dU|/|Du|\tdL/DL\t\t|U|\t\tL\t\tDmin\t(negatives)\n";
    }
    void stepOutput() override
    {
        gsInfo << "\t" << this->m_numIterations << "\t" << this->m_residueF << "\t" << this->m_residueU << "\t"
               << this->m_residueL << "\t" << (this->m_U + this->m_DeltaU).norm() << "\t" << (this->m_L + this->m_DeltaL)
               << "\t" << this->m_indicator << "\t(" << this->m_negatives << ")\t" << this->m_note << "\n";
        this->m_note = "";
    }

protected:
    index_t m_angleDetermine = 0;
    T m_phi = -1;
    bool m_phi_user = false;
    T m_eta = 1.0, m_discriminant = 0;
    T m_alpha1 = 0, m_alpha2 = 0, m_alpha3 = 0, m_a0 = 0, m_b0 = 0, m_b1 = 0, m_c0 = 0, m_c1 = 0, m_c2 = 0;
};

}  // namespace gismo
