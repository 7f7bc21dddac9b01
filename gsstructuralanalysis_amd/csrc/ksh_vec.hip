// ksh_vec.hip -- batched vector kernels of the path engine: dot products, linear combinations,
// and the learned predictor's modal projection / decode.
//
// Every vector of every instance of a batch lives in device memory; the engine's host logic only
// ever sees scalars.  One launch serves all instances of a round:
//   k_dots      one workgroup per instance, up to kMaxDots dot products of length n in one pass;
//               lanes stride over the entries, then a fixed-order reduction (wave shuffles, then
//               the four wave sums in order), so results are bit-reproducible run to run;
//   k_vecops    per instance a short program of 3-term combinations dst = a x + b y + c z,
//               applied in order by each thread to its entry (so an op may read what an earlier op
//               of the same program wrote, and dst may alias x, y or z);
//   k_features  the 66 predictor features of a state: lambda, ||U||, and the 64 least-squares
//               coefficients of the out-of-plane control displacements in the analytic modes
//               (P [64][ncp] times the gathered z values);
//   k_decode    dU[z dof of cp] += scale * sum_r Phi[cp][r] coef[r];
//   k_dr        one explicit dynamic-relaxation update of an instance with its four norms.
// All are HBM / latency bound and tiny next to the assembly and the factor.
#include <hip/hip_runtime.h>

#include "ksh_internal.hpp"
#include "ksh_vec.hpp"

namespace ksh {

namespace {

constexpr int kT = 256;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__global__ __launch_bounds__(kT) void k_dots(const DotDesc *__restrict__ descs)
{
    const DotDesc &D = descs[blockIdx.x];
    const int np = D.npairs;
    double acc[kMaxDots];
#pragma unroll
    for (int k = 0; k < kMaxDots; ++k) acc[k] = 0.0;
    for (int i = threadIdx.x; i < D.n; i += kT) {
#pragma unroll
        for (int k = 0; k < kMaxDots; ++k)
            if (k < np) acc[k] += D.a[k][i] * D.b[k][i];
    }
    __shared__ double part[kMaxDots][kT / 64];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < kMaxDots; ++k) {
        if (k < np) {
            const double s = wave_sum(acc[k]);
            if (lane == 0) part[k][wave] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < np) {
        const int k = threadIdx.x;
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kT / 64; ++w) s += part[k][w];
        D.out[k] = s;
    }
}

__global__ __launch_bounds__(kT) void k_vecops(const VecDesc *__restrict__ descs)
{
    const VecDesc &D = descs[blockIdx.y];
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= D.n) return;
    for (int k = 0; k < D.nops; ++k) {
        const VecOp &o = D.op[k];
        double v = 0.0;
        if (o.x) v = o.a * o.x[i];
        if (o.y) v += o.b * o.y[i];
        if (o.z) v += o.c * o.z[i];
        o.dst[i] = v;
    }
}

constexpr int kChunk = 2048;  // control points staged in LDS per pass

__global__ __launch_bounds__(kT) void k_features(const FeatDesc *__restrict__ descs)
{
    const FeatDesc &D = descs[blockIdx.x];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    __shared__ double z[kChunk];
    __shared__ double red[kT / 64];
    // ||U||
    double acc = 0.0;
    for (int i = threadIdx.x; i < D.nfree; i += kT) acc += D.U[i] * D.U[i];
    acc = wave_sum(acc);
    if (lane == 0) red[wave] = acc;
    // coefficients: wave w owns r = w, w + 4, ... (16 of the 64)
    double coef[kModalCoef / 4];
#pragma unroll
    for (int q = 0; q < kModalCoef / 4; ++q) coef[q] = 0.0;
    for (int c0 = 0; c0 < D.ncp; c0 += kChunk) {
        const int cn = min(kChunk, D.ncp - c0);
        __syncthreads();
        for (int c = threadIdx.x; c < cn; c += kT) {
            const int f = D.zfree[c0 + c];
            z[c] = f >= 0 ? D.U[f] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kModalCoef / 4; ++q) {
            const int r = wave + 4 * q;
            const double *P = D.P + (size_t)r * D.ncp + c0;
            double s = 0.0;
            for (int c = lane; c < cn; c += 64) s += P[c] * z[c];
            coef[q] += s;
        }
    }
#pragma unroll
    for (int q = 0; q < kModalCoef / 4; ++q) {
        const double s = wave_sum(coef[q]);
        if (lane == 0) D.out[2 + wave + 4 * q] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kT / 64; ++w) s += red[w];
        D.out[0] = D.lambda;
        D.out[1] = sqrt(s);
    }
}

__global__ __launch_bounds__(kT) void k_decode(const DecDesc *__restrict__ descs)
{
    const DecDesc &D = descs[blockIdx.y];
    const int c = blockIdx.x * kT + threadIdx.x;
    if (c >= D.ncp) return;
    const int f = D.zfree[c];
    if (f < 0) return;
    const double *phi = D.Phi + (size_t)c * kModalCoef;
    double s = 0.0;
#pragma unroll 8
    for (int r = 0; r < kModalCoef; ++r) s += phi[r] * D.coef[r];
    D.dU[f] += D.scale * s;
}

__global__ __launch_bounds__(kT) void k_dr(const DrDesc *__restrict__ descs)
{
    const DrDesc &D = descs[blockIdx.x];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    __shared__ double red[4][kT / 64];
    const double dt = D.dt;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < D.n; i += kT) {
        double r = D.R[i], v = D.v[i], du;
        if (D.mode == 0) {
            du = 0.5 * dt * dt * D.minv[i] * r;
            v = du / dt;
        } else if (D.mode == 1) {
            r -= D.damp[i] * v;
            v += dt * D.minv[i] * r;
            du = dt * v;
        } else {
            r -= D.damp[i] * v;
            du = -1.5 * dt * v + 0.5 * dt * dt * D.minv[i] * r;
            v = 0.0;
        }
        const double DU = D.DU[i] + du;
        D.v[i] = v;
        D.dU[i] = du;
        D.DU[i] = DU;
        D.X[i] = D.U[i] + DU;
        s[0] += r * r;
        s[1] += v * (D.m[i] * v);
        s[2] += du * du;
        s[3] += DU * DU;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double w = wave_sum(s[q]);
        if (lane == 0) red[q][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < kT / 64; ++w) t += red[threadIdx.x][w];
        D.out[threadIdx.x] = t;
    }
}

template <typename Desc>
int upload_descs(hipStream_t s, DevScratch &buf, const std::vector<Desc> &h)
{
    int rc;
    if ((rc = buf.ensure(h.size() * sizeof(Desc)))) return rc;
    KSH_HIP(hipMemcpyAsync(buf.p, h.data(), h.size() * sizeof(Desc), hipMemcpyHostToDevice, s));
    return KSH_SUCCESS;
}

}  // namespace

int DevScratch::ensure(size_t b)
{
    if (b <= bytes) return KSH_SUCCESS;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    KSH_HIP(hipMalloc(&p, b));
    bytes = b;
    return KSH_SUCCESS;
}

DevScratch::~DevScratch()
{
    if (p) (void)hipFree(p);
}

int batch_dots(hipStream_t s, DevScratch &desc_buf, const std::vector<DotDesc> &h)
{
    if (h.empty()) return KSH_SUCCESS;
    for (const DotDesc &d : h)
        if (d.npairs < 0 || d.npairs > kMaxDots || d.n < 0 || !d.out) return fail(KSH_OTHER_ERROR, "dot descriptor");
    int rc;
    if ((rc = upload_descs(s, desc_buf, h))) return rc;
    hipLaunchKernelGGL(k_dots, dim3((unsigned)h.size()), dim3(kT), 0, s, (const DotDesc *)desc_buf.p);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int batch_vecops(hipStream_t s, DevScratch &desc_buf, const std::vector<VecDesc> &h)
{
    if (h.empty()) return KSH_SUCCESS;
    int nmax = 0;
    for (const VecDesc &d : h) {
        if (d.nops < 0 || d.nops > kMaxVecOps) return fail(KSH_OTHER_ERROR, "vector program too long");
        nmax = std::max(nmax, d.n);
    }
    if (nmax == 0) return KSH_SUCCESS;
    int rc;
    if ((rc = upload_descs(s, desc_buf, h))) return rc;
    hipLaunchKernelGGL(k_vecops, dim3((unsigned)((nmax + kT - 1) / kT), (unsigned)h.size()), dim3(kT), 0, s,
                       (const VecDesc *)desc_buf.p);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int batch_dr(hipStream_t s, DevScratch &desc_buf, const std::vector<DrDesc> &h)
{
    if (h.empty()) return KSH_SUCCESS;
    for (const DrDesc &d : h)
        if (d.n < 0 || d.mode < 0 || d.mode > 2 || !d.out) return fail(KSH_OTHER_ERROR, "dynamic relaxation descriptor");
    int rc;
    if ((rc = upload_descs(s, desc_buf, h))) return rc;
    hipLaunchKernelGGL(k_dr, dim3((unsigned)h.size()), dim3(kT), 0, s, (const DrDesc *)desc_buf.p);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int batch_features(hipStream_t s, DevScratch &desc_buf, const std::vector<FeatDesc> &h)
{
    if (h.empty()) return KSH_SUCCESS;
    int rc;
    if ((rc = upload_descs(s, desc_buf, h))) return rc;
    hipLaunchKernelGGL(k_features, dim3((unsigned)h.size()), dim3(kT), 0, s, (const FeatDesc *)desc_buf.p);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int batch_decode(hipStream_t s, DevScratch &desc_buf, const std::vector<DecDesc> &h)
{
    if (h.empty()) return KSH_SUCCESS;
    int cmax = 0;
    for (const DecDesc &d : h) cmax = std::max(cmax, d.ncp);
    if (cmax == 0) return KSH_SUCCESS;
    int rc;
    if ((rc = upload_descs(s, desc_buf, h))) return rc;
    hipLaunchKernelGGL(k_decode, dim3((unsigned)((cmax + kT - 1) / kT), (unsigned)h.size()), dim3(kT), 0, s,
                       (const DecDesc *)desc_buf.p);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

}  // namespace ksh
