// gsALMCrisfieldLearned.h -- gsALMCrisfield with a learned predictor.
//
// The reference's hook for a predictor is the protected virtual gsALMCrisfield::predictor()
// (src/gsALMSolvers/gsALMCrisfield.h:85,100; gsALMCrisfield.hpp:111-164); this subclass
// overrides it and leaves everything else -- the corrector iterations, the arc-length roots, the
// residual norms that decide convergence, the stability count -- to gsALMCrisfield unchanged.
//
// Proposal: the next state extrapolated from the last three converged states,
// (U, lambda)_ext = 3 (U, lambda)_t - 3 (U, lambda)_{t-1} + (U, lambda)_{t-2}, corrected by the
// network, which maps the features of the last four converged states (lambda, ||U||, 64
// out-of-plane mode coefficients each; ksh_predictor_features) to the deviation of the next
// state from that extrapolation (64 mode coefficients, 1 load factor).  The increment is scaled
// onto the arc-length sphere ||dU||^2 + phi^2 ||F||^2 dL^2 = ds^2.  Until four states are known,
// after a change of the arc length, and whenever the owner switches it off (the driver does, to
// redo a step that failed from the proposal), the tangent predictor of gsALMCrisfield runs.
#pragma once

#include <deque>

#include "gsALMCrisfield.h"

namespace gismo {

template <class T>
class gsALMCrisfieldLearned : public gsALMCrisfield<T> {
    using Base = gsALMCrisfield<T>;
    static constexpr int kFeat = 66, kCoef = 64;

public:
    gsALMCrisfieldLearned(typename gsStructuralAnalysisOps<T>::Jacobian_t &Jacobian,
                          typename gsStructuralAnalysisOps<T>::ALResidual_t &ALResidual, const gsVector<T> &Force,
                          ksh_problem *pb, ksh_mlp *net, gsExec *exec = gsExec::direct())
        : Base(Jacobian, ALResidual, Force), m_pb(pb), m_net(net), m_exec(exec)
    {
    }

    void setLearned(bool on) { m_on = on; }
    bool usedLearned() const { return m_used; }

    /// record the converged state (the driver calls this after every successful step)
    void record()
    {
        std::vector<double> f(kFeat);
        if (ksh_predictor_features(m_pb, this->m_U.data(), this->m_L, f.data()) != KSH_SUCCESS)
            throw std::runtime_error(ksh_last_error());
        if (!m_len.empty() && this->m_arcLength != m_len.back()) {  // uniform steps only
            m_hist.clear();
            m_Us.clear();
            m_Ls.clear();
            m_len.clear();
        }
        m_hist.push_back(f);
        m_Us.push_back(this->m_U);
        m_Ls.push_back(this->m_L);
        m_len.push_back(this->m_arcLength);
        while (m_hist.size() > 4) m_hist.pop_front();
        while (m_Us.size() > 3) {
            m_Us.pop_front();
            m_Ls.pop_front();
        }
        while (m_len.size() > 4) m_len.pop_front();
    }

protected:
    void predictor() override
    {
        m_used = false;
        if (!m_on || !m_net || m_hist.size() < 4 || m_len.back() != this->m_arcLength) return Base::predictor();
        std::vector<double> x(4 * kFeat), y(kCoef + 1);
        for (int s = 0; s < 4; ++s) std::copy(m_hist[s].begin(), m_hist[s].end(), x.begin() + s * kFeat);
        if (m_exec->predict(m_net, x.data(), y.data()) != KSH_SUCCESS) return Base::predictor();
        // increment to the extrapolated state, corrected in the modal subspace and in the load
        gsVector<T> dU = T(2) * m_Us[2] - T(3) * m_Us[1] + m_Us[0];
        if (ksh_predictor_decode(m_pb, y.data(), 1.0, dU.data()) != KSH_SUCCESS) return Base::predictor();
        T dL = T(2) * m_Ls[2] - T(3) * m_Ls[1] + m_Ls[0] + y[kCoef];
        const gsVector<T> &F = this->m_forcing;
        if (!this->m_phi_user) this->m_phi = std::sqrt(this->m_U.dot(this->m_U) / (this->m_L * this->m_L * F.dot(F)));
        const T phi = this->m_phi > 0 ? this->m_phi : T(0);
        const T nrm = std::sqrt(dU.dot(dU) + phi * phi * F.dot(F) * dL * dL);
        if (!(nrm > 0) || !std::isfinite(nrm)) return Base::predictor();
        const T s = this->m_arcLength / nrm;
        this->m_note += "learned predictor\t";
        this->m_deltaU = s * dU;
        this->m_deltaL = s * dL;
        this->m_DeltaU += this->m_deltaU;
        this->m_DeltaL += this->m_deltaL;
        if (this->m_angleDetermine == Base::angmethod::Iteration || this->m_angleDetermine == Base::angmethod::Predictor) {
            this->m_DeltaUold = this->m_DeltaU;
            this->m_DeltaLold = this->m_DeltaL;
        }
        m_used = true;
    }

private:
    ksh_problem *m_pb;
    ksh_mlp *m_net;
    gsExec *m_exec;
    std::deque<std::vector<double>> m_hist;
    std::deque<gsVector<T>> m_Us;
    std::deque<T> m_Ls, m_len;
    bool m_on = true, m_used = false;
};

}  // namespace gismo
