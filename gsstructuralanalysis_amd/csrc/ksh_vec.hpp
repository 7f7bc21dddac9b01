// ksh_vec.hpp -- descriptors of the path engine's batched vector kernels (ksh_vec.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

namespace ksh {

constexpr int kMaxDots = 10;
constexpr int kMaxVecOps = 6;
constexpr int kModalCoef = 64;   // 8 x 8 analytic out-of-plane modes of the learned predictor
constexpr int kFeatures = 2 + kModalCoef;

// growable device buffer (kernel descriptors, staging)
struct DevScratch {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b);
    DevScratch() = default;
    DevScratch(const DevScratch &) = delete;
    DevScratch &operator=(const DevScratch &) = delete;
    ~DevScratch();
};

// out[k] = sum_i a[k][i] b[k][i], k < npairs
struct DotDesc {
    int n = 0, npairs = 0;
    const double *a[kMaxDots] = {};
    const double *b[kMaxDots] = {};
    double *out = nullptr;
};

// dst = a x + b y + c z (a null source contributes 0)
struct VecOp {
    double *dst = nullptr;
    const double *x = nullptr, *y = nullptr, *z = nullptr;
    double a = 0, b = 0, c = 0;
};

struct VecDesc {
    int n = 0, nops = 0;
    VecOp op[kMaxVecOps];
    void add(double *dst, double a, const double *x, double b = 0, const double *y = nullptr, double c = 0,
             const double *z = nullptr)
    {
        VecOp &o = op[nops++];
        o.dst = dst;
        o.x = x;
        o.y = y;
        o.z = z;
        o.a = a;
        o.b = b;
        o.c = c;
    }
};

struct FeatDesc {
    int ncp = 0, nfree = 0;
    const double *P = nullptr;   // [64][ncp] least-squares projector of the modal fit
    const int *zfree = nullptr;  // [ncp] free index of the control point's z dof, or -1
    const double *U = nullptr;   // [nfree]
    double lambda = 0;
    double *out = nullptr;       // [66]
};

struct DecDesc {
    int ncp = 0;
    const double *Phi = nullptr;  // [ncp][64] mode values at the Greville points
    const int *zfree = nullptr;
    const double *coef = nullptr; // [64]
    double scale = 0;
    double *dU = nullptr;         // [nfree]
};

int batch_dots(hipStream_t s, DevScratch &desc_buf, const std::vector<DotDesc> &h);
int batch_vecops(hipStream_t s, DevScratch &desc_buf, const std::vector<VecDesc> &h);
int batch_features(hipStream_t s, DevScratch &desc_buf, const std::vector<FeatDesc> &h);
int batch_decode(hipStream_t s, DevScratch &desc_buf, const std::vector<DecDesc> &h);

// one explicit step of dynamic relaxation on one instance (gsStaticDR.hpp:182-266), dt given:
//   mode 0 (_start)     dU = dt^2 / 2 minv R,              v = dU / dt
//   mode 1 (_iteration) R -= damp v, v += dt minv R,       dU = dt v
//   mode 2 (_peak)      R -= damp v, dU = -1.5 dt v + dt^2 / 2 minv R, v = 0
// then DU += dU, X = U + DU; out = {||R||^2 (after the damping term), v . (m v), ||dU||^2, ||DU||^2}
struct DrDesc {
    int n = 0, mode = 0;
    double dt = 1.0;
    const double *R = nullptr, *m = nullptr, *minv = nullptr, *damp = nullptr, *U = nullptr;
    double *v = nullptr, *dU = nullptr, *DU = nullptr, *X = nullptr;
    double *out = nullptr;
};
int batch_dr(hipStream_t s, DevScratch &desc_buf, const std::vector<DrDesc> &h);

}  // namespace ksh
