// ksh_predictor.cpp -- state features of the learned arc-length predictor and the mapping of its
// modal output back to a displacement increment.
//
// A converged state (U, lambda) is described by 66 numbers: lambda, ||U||, and the 64
// least-squares coefficients of the out-of-plane displacement field in the 8 x 8 analytic
// modes sin(a pi xi) sin(b pi eta), a, b = 1..8, sampled at the Greville points of the patch's
// control net.  The predictor's network maps the last four states (264 numbers) to the next
// state's deviation from the quadratic extrapolation of the last three, in the coefficients (64)
// and the load factor (1); the solver class (gs/gsALMCrisfieldLearned.h) adds it to the
// extrapolation and scales the increment onto the arc-length sphere.  No reference counterpart: the reference's
// only predictor hook is gsALMCrisfield::predictor / predictorGuess
// (src/gsALMSolvers/gsALMCrisfield.hpp:111-191).
#include <cmath>
#include <cstring>
#include <vector>

#include "ksh_internal.hpp"

namespace ksh {

constexpr int kModes = 8;
constexpr int kCoef = kModes * kModes;

struct ModalBasis {
    int ncp = 0;
    std::vector<double> Phi;  // [ncp][kCoef]
    std::vector<double> P;    // [kCoef][ncp]: (Phi^T Phi + eps I)^-1 Phi^T
    std::vector<int> zfree;   // [ncp]: free index of the control point's z dof, or -1
};

namespace {

std::vector<double> greville(int p, int n)
{
    // open uniform knots: p+1 zeros, interior j/n, p+1 ones; n + p basis functions
    std::vector<double> t;
    for (int i = 0; i < p; ++i) t.push_back(0.0);
    for (int j = 0; j <= n; ++j) t.push_back((double)j / n);
    for (int i = 0; i < p; ++i) t.push_back(1.0);
    std::vector<double> g(n + p);
    for (int i = 0; i < n + p; ++i) {
        double s = 0.0;
        for (int k = 1; k <= p; ++k) s += t[i + k];
        g[i] = s / p;
    }
    return g;
}

// Cholesky solve of the SPD system G X = B (G m x m, B m x r, in place in B)
void chol_solve(std::vector<double> G, int m, std::vector<double> &B, int r)
{
    for (int j = 0; j < m; ++j) {
        double d = G[(size_t)j * m + j];
        for (int k = 0; k < j; ++k) d -= G[(size_t)j * m + k] * G[(size_t)j * m + k];
        d = std::sqrt(d);
        G[(size_t)j * m + j] = d;
        for (int i = j + 1; i < m; ++i) {
            double s = G[(size_t)i * m + j];
            for (int k = 0; k < j; ++k) s -= G[(size_t)i * m + k] * G[(size_t)j * m + k];
            G[(size_t)i * m + j] = s / d;
        }
    }
    for (int c = 0; c < r; ++c) {
        for (int i = 0; i < m; ++i) {  // forward
            double s = B[(size_t)i * r + c];
            for (int k = 0; k < i; ++k) s -= G[(size_t)i * m + k] * B[(size_t)k * r + c];
            B[(size_t)i * r + c] = s / G[(size_t)i * m + i];
        }
        for (int i = m - 1; i >= 0; --i) {  // backward
            double s = B[(size_t)i * r + c];
            for (int k = i + 1; k < m; ++k) s -= G[(size_t)k * m + i] * B[(size_t)k * r + c];
            B[(size_t)i * r + c] = s / G[(size_t)i * m + i];
        }
    }
}

}  // namespace

const ModalBasis &modal_basis(Problem &pb)
{
    if (pb.modal) return *static_cast<ModalBasis *>(pb.modal.get());
    auto mb = std::make_shared<ModalBasis>();
    const int mx = pb.nx + pb.p, my = pb.ny + pb.p, ncp = mx * my;
    mb->ncp = ncp;
    const std::vector<double> gx = greville(pb.p, pb.nx), gy = greville(pb.p, pb.ny);
    mb->Phi.assign((size_t)ncp * kCoef, 0.0);
    for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i)
            for (int b = 0; b < kModes; ++b)
                for (int a = 0; a < kModes; ++a)
                    mb->Phi[(size_t)(i + mx * j) * kCoef + a + kModes * b] =
                        std::sin((a + 1) * M_PI * gx[i]) * std::sin((b + 1) * M_PI * gy[j]);
    // G = Phi^T Phi + eps I (eps keeps coarse nets with fewer points than modes well posed)
    std::vector<double> G((size_t)kCoef * kCoef, 0.0);
    double tr = 0.0;
    for (int c = 0; c < ncp; ++c)
        for (int r = 0; r < kCoef; ++r)
            for (int s = 0; s < kCoef; ++s) G[(size_t)r * kCoef + s] += mb->Phi[(size_t)c * kCoef + r] * mb->Phi[(size_t)c * kCoef + s];
    for (int r = 0; r < kCoef; ++r) tr += G[(size_t)r * kCoef + r];
    for (int r = 0; r < kCoef; ++r) G[(size_t)r * kCoef + r] += 1e-8 * tr / kCoef;
    mb->P.assign((size_t)kCoef * ncp, 0.0);
    for (int r = 0; r < kCoef; ++r)
        for (int c = 0; c < ncp; ++c) mb->P[(size_t)r * ncp + c] = mb->Phi[(size_t)c * kCoef + r];
    chol_solve(G, kCoef, mb->P, ncp);
    mb->zfree.assign(ncp, -1);
    for (int f = 0; f < pb.nfree; ++f) {
        const int g = pb.h_free2full[f];
        if (g % 3 == 2) mb->zfree[g / 3] = f;
    }
    pb.modal = mb;
    return *mb;
}

void predictor_features(Problem &pb, const double *U, double lambda, double *out)
{
    const ModalBasis &mb = modal_basis(pb);
    std::vector<double> z(mb.ncp, 0.0);
    for (int c = 0; c < mb.ncp; ++c)
        if (mb.zfree[c] >= 0) z[c] = U[mb.zfree[c]];
    double nrm = 0.0;
    for (int f = 0; f < pb.nfree; ++f) nrm += U[f] * U[f];
    out[0] = lambda;
    out[1] = std::sqrt(nrm);
    for (int r = 0; r < kCoef; ++r) {
        double s = 0.0;
        for (int c = 0; c < mb.ncp; ++c) s += mb.P[(size_t)r * mb.ncp + c] * z[c];
        out[2 + r] = s;
    }
}

void modal_to_free(Problem &pb, const double *coef, double scale, double *dU)
{
    const ModalBasis &mb = modal_basis(pb);
    for (int c = 0; c < mb.ncp; ++c) {
        if (mb.zfree[c] < 0) continue;
        double s = 0.0;
        for (int r = 0; r < kCoef; ++r) s += mb.Phi[(size_t)c * kCoef + r] * coef[r];
        dU[mb.zfree[c]] += scale * s;
    }
}

}  // namespace ksh

extern "C" int ksh_predictor_features(ksh_problem *pb, const double *U, double lambda, double *out)
{
    if (!pb || !U || !out) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    ksh::predictor_features(*pb, U, lambda, out);
    return KSH_SUCCESS;
}

extern "C" int ksh_predictor_decode(ksh_problem *pb, const double *coef, double scale, double *dU)
{
    if (!pb || !coef || !dU) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    ksh::modal_to_free(*pb, coef, scale, dU);
    return KSH_SUCCESS;
}

namespace ksh {

// device copies of the modal basis (uploaded on first use; freed with the problem)
int modal_device(Problem &pb)
{
    if (pb.modal_P) return KSH_SUCCESS;
    const ModalBasis &mb = modal_basis(pb);
    const size_t nc = (size_t)mb.ncp * kCoef;
    KSH_HIP(hipMalloc(&pb.modal_P, nc * sizeof(double)));
    KSH_HIP(hipMalloc(&pb.modal_Phi, nc * sizeof(double)));
    KSH_HIP(hipMalloc(&pb.modal_zfree, (size_t)mb.ncp * sizeof(int)));
    KSH_HIP(hipMemcpyAsync(pb.modal_P, mb.P.data(), nc * sizeof(double), hipMemcpyHostToDevice, pb.stream));
    KSH_HIP(hipMemcpyAsync(pb.modal_Phi, mb.Phi.data(), nc * sizeof(double), hipMemcpyHostToDevice, pb.stream));
    KSH_HIP(hipMemcpyAsync(pb.modal_zfree, mb.zfree.data(), (size_t)mb.ncp * sizeof(int), hipMemcpyHostToDevice, pb.stream));
    KSH_HIP(hipStreamSynchronize(pb.stream));
    return KSH_SUCCESS;
}

}  // namespace ksh
