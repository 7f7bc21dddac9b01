// ksh_band.hip -- banded LDL^T factorization, solves, inertia and band SpMV for gfx950.
//
// Replaces the reference's use of gsSparseSolver<T>::SimplicialLDLT (compute / solve /
// vectorD: src/gsALMSolvers/gsALMBase.hpp:186-211,559-562; src/gsStaticSolvers/gsStaticNewton.hpp:215-240;
// src/gsEigenSolvers/gsBucklingSolver.h:131-133).  By Sylvester's law of inertia the number of
// negative pivots is ordering independent, so the band ordering used here reproduces the count
// that gsALMBase::_computeStability takes from Eigen's AMD-ordered factor
// (src/gsALMSolvers/gsALMHelper.h:30-42).
//
// Layout: lower band, column major, ab[j*ld + (i-j)], ld = bw+1.  One workgroup (256 threads)
// factors one matrix; a launch covers a batch of matrices (blockIdx.x = matrix), so the same
// kernel serves one instance or a batch of uneven instances.
//
// Right-looking blocked algorithm, panel width 16:
//   1. panel (16 columns x up to 16+bw rows) is staged in LDS, column major;
//   2. wave 0 factors the 16x16 diagonal block in registers (lane = row, shuffles);
//   3. every thread solves its panel rows against L_dd D (row-wise TRSM) -> L panel;
//   4. trailing symmetric update A -= (L D) L^T of the bw x bw window in 16x16 tiles on
//      v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 (panel operands from LDS, A tiles in HBM).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "ksh_internal.hpp"
#include "ksh_mfma.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {


constexpr int NB = 16;
constexpr int kThreads = 256;

__host__ __device__ inline int panel_rows(int bw) { return ((bw + 15) / 16) * 16 + 32; }

template <typename T>
__device__ void pivot_stats(const T *fac, int n, int bw, double *stats);

template <typename T>
__device__ __forceinline__ void band_ldlt_body(T *fac, int n, int bw, int *info, double *stats, unsigned char *smem)
{
    const int PR = panel_rows(bw);
    T *pan = (T *)smem;                     // [16][PR]
    T *Wdd = pan + (size_t)NB * PR;         // [16][16] d_kk * L_dd[jj][kk]
    T *dv = Wdd + NB * NB;                  // [16]
    T *dinv = dv + NB;                      // [16]
    __shared__ int s_info;
    const size_t ld = (size_t)bw + 1;
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    if (tid == 0) s_info = 0;
    __syncthreads();

    for (int c0 = 0; c0 < n; c0 += NB) {
        const int w = min(NB, n - c0);
        const int rend = min(n - 1, c0 + w - 1 + bw);
        const int nr = rend - c0 + 1;
        // 1. stage the panel (zero outside the band / matrix)
        for (int idx = tid; idx < NB * PR; idx += kThreads) {
            int jj = idx / PR, r = idx % PR;
            T v = 0;
            if (jj < w && r < nr && r >= jj && r - jj <= bw) v = fac[(size_t)(c0 + jj) * ld + (r - jj)];
            if (jj >= w && r == jj) v = 1;  // padding columns of the last panel
            pan[(size_t)jj * PR + r] = v;
        }
        __syncthreads();
        // 2. diagonal block LDL^T in wave 0 (lane r owns row r)
        if (wave == 0) {
            const int r = lane;
            T a[NB];
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) a[kk] = (r < NB && kk <= r) ? pan[(size_t)kk * PR + r] : (T)0;
#pragma unroll
            for (int jj = 0; jj < NB; ++jj) {
                T djj = __shfl(a[jj], jj);
                if (lane == 0 && (djj == (T)0 || !isfinite((double)djj)) && s_info == 0) s_info = c0 + jj + 1;
                T c = a[jj];
                T l = (r > jj) ? c / djj : (T)0;
#pragma unroll
                for (int kk = jj + 1; kk < NB; ++kk) {
                    T ck = __shfl(c, kk);
                    if (r >= kk) a[kk] -= l * ck;
                }
                if (r > jj) a[jj] = l;
            }
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) {
                T dk = __shfl(a[kk], kk);
                if (r < NB) {
                    Wdd[r * NB + kk] = (kk < r) ? dk * a[kk] : (T)0;
                    if (kk <= r) pan[(size_t)kk * PR + r] = a[kk];
                }
            }
            if (r < NB) {
                dv[r] = a[r];
                dinv[r] = (T)1 / a[r];
            }
        }
        __syncthreads();
        // 3. TRSM of the rows below the diagonal block
        for (int r = w + tid; r < nr; r += kThreads) {
            T a[NB];
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) a[kk] = pan[(size_t)kk * PR + r];
#pragma unroll
            for (int jj = 0; jj < NB; ++jj) {
                T x = a[jj];
#pragma unroll
                for (int kk = 0; kk < jj; ++kk) x -= a[kk] * Wdd[jj * NB + kk];
                a[jj] = x * dinv[jj];
            }
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) pan[(size_t)kk * PR + r] = a[kk];
        }
        __syncthreads();
        // write the factored panel back
        for (int idx = tid; idx < NB * PR; idx += kThreads) {
            int jj = idx / PR, r = idx % PR;
            if (jj < w && r < nr && r >= jj && r - jj <= bw) fac[(size_t)(c0 + jj) * ld + (r - jj)] = pan[(size_t)jj * PR + r];
        }
        // 4. trailing update on rows/cols [w, nr) of the panel frame
        const int T_ = (nr - w + 15) / 16;
        const int npair = T_ * (T_ + 1) / 2;
        for (int pidx = wave; pidx < npair; pidx += kThreads / 64) {
            int ti = (int)((sqrtf(8.0f * pidx + 1.0f) - 1.0f) * 0.5f);
            while ((ti + 1) * (ti + 2) / 2 <= pidx) ++ti;
            while (ti * (ti + 1) / 2 > pidx) --ti;
            int tj = pidx - ti * (ti + 1) / 2;
            const int rj0 = w + tj * 16, ri0 = w + ti * 16;
            // accumulator rows = j (A's column), cols = i (A's row): contiguous stores
            typename Mfma16<T>::acc_t acc;
            const int ri = ri0 + lane % 16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int rj = rj0 + Mfma16<T>::row(lane, v);
                T x = 0;
                if (ri < nr && rj < nr && ri >= rj) x = fac[(size_t)(c0 + rj) * ld + (ri - rj)];
                acc[v] = x;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                int kk = 4 * s + lane / 16;
                T wv = -pan[(size_t)kk * PR + rj0 + lane % 16] * dv[kk];
                T lv = pan[(size_t)kk * PR + ri0 + lane % 16];
                acc = Mfma16<T>::run(wv, lv, acc);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int rj = rj0 + Mfma16<T>::row(lane, v);
                if (ri < nr && rj < nr && ri >= rj) fac[(size_t)(c0 + rj) * ld + (ri - rj)] = acc[v];
            }
        }
        __syncthreads();
        if (s_info != 0) break;
    }
    if (tid == 0) *info = s_info;
    __syncthreads();
    if (stats) pivot_stats<T>(fac, n, bw, stats);
}

// negatives, min d, max |d|, min |d| of the factored diagonal (whole workgroup)
template <typename T>
__device__ void pivot_stats(const T *fac, int n, int bw, double *stats)
{
    __shared__ double sneg[kThreads], smin[kThreads], smax[kThreads], sminabs[kThreads];
    const size_t ld = (size_t)bw + 1;
    double c = 0.0, mn = INFINITY, mxa = 0.0, mna = INFINITY;
    for (int i = threadIdx.x; i < n; i += kThreads) {
        double d = (double)fac[(size_t)i * ld];
        if (d < 0.0) c += 1.0;
        mn = fmin(mn, d);
        mxa = fmax(mxa, fabs(d));
        mna = fmin(mna, fabs(d));
    }
    sneg[threadIdx.x] = c;
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mxa;
    sminabs[threadIdx.x] = mna;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sneg[threadIdx.x] += sneg[threadIdx.x + s];
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            sminabs[threadIdx.x] = fmin(sminabs[threadIdx.x], sminabs[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stats[0] = sneg[0];
        stats[1] = smin[0];
        stats[2] = smax[0];
        stats[3] = sminabs[0];
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_band_ldlt(T *fac, int n, int bw, int *info, double *stats)
{
    extern __shared__ __align__(16) unsigned char smem[];
    band_ldlt_body<T>(fac, n, bw, info, stats, smem);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_band_ldlt_batched(const BandDesc *descs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BandDesc &d = descs[blockIdx.x];
    if (!d.active) return;
    band_ldlt_body<T>((T *)d.fac, d.n, d.bw, d.info, d.stats, smem);
}


// negatives / min d / max |d| / min |d| of the factored diagonal (one workgroup)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_band_inertia(const T *fac, int n, int bw, long long *neg, double *stats)
{
    __shared__ long long sneg[kThreads];
    __shared__ double smin[kThreads], smax[kThreads], sminabs[kThreads];
    const size_t ld = (size_t)bw + 1;
    long long c = 0;
    double mn = INFINITY, mxa = 0.0, mna = INFINITY;
    for (int i = threadIdx.x; i < n; i += kThreads) {
        double d = (double)fac[(size_t)i * ld];
        if (d < 0.0) ++c;
        mn = fmin(mn, d);
        mxa = fmax(mxa, fabs(d));
        mna = fmin(mna, fabs(d));
    }
    sneg[threadIdx.x] = c;
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mxa;
    sminabs[threadIdx.x] = mna;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sneg[threadIdx.x] += sneg[threadIdx.x + s];
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            sminabs[threadIdx.x] = fmin(sminabs[threadIdx.x], sminabs[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *neg = sneg[0];
        stats[0] = smin[0];
        stats[1] = smax[0];
        stats[2] = sminabs[0];
    }
}

template <typename TY, typename TX>
__global__ void k_band_axpby(TY *y, const TX *x, size_t count, double alpha, double beta)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) y[i] = (TY)(alpha * (double)y[i] + beta * (double)x[i]);
}

template <typename TD, typename TS>
__global__ void k_convert(TD *dst, const TS *src, size_t count)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = (TD)src[i];
}

__global__ void k_axpy(int n, double a, const double *x, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
__global__ void k_sub(int n, const double *a, const double *b, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}
__global__ __launch_bounds__(kThreads) void k_norm2(int n, const double *x, double *out)
{
    __shared__ double s[kThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kThreads) acc += x[i] * x[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

inline unsigned nblk(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T> size_t factor_smem(int bw)
{
    return sizeof(T) * ((size_t)NB * panel_rows(bw) + NB * NB + 2 * NB);
}

// LDS per CU on gfx950 is 160 KB; the factor kernels also hold static LDS (pivot flag and the
// pivot-statistics reduction, 4 x 256 doubles), which counts against the same budget.
constexpr size_t kStaticFactorLds = 4 * kThreads * sizeof(double) + 64;
constexpr size_t kMaxLds = 160 * 1024 - kStaticFactorLds;

// raise the dynamic-LDS limit of the factor kernels to `need` bytes (once per size increase)
int allow_lds(size_t need)
{
    static size_t granted = 0;
    if (need <= granted) return KSH_SUCCESS;
    const void *fns[] = {(const void *)k_band_ldlt<double>, (const void *)k_band_ldlt<float>,
                         (const void *)k_band_ldlt_batched<double>, (const void *)k_band_ldlt_batched<float>};
    for (const void *f : fns) KSH_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    granted = need;
    return KSH_SUCCESS;
}

static_assert(sizeof(BandDesc) == kBandDescBytes, "BandDesc layout");

}  // namespace

int band_factor(Matrix &K, void *fac, int prec, double *stats_out)
{
    Problem &pb = *K.pb;
    size_t bytes = (size_t)K.n * K.ld * (prec == KSH_FP64 ? sizeof(double) : sizeof(float));
    if (fac != K.ab) {
        if (prec == K.sprec) {
            KSH_HIP(hipMemcpyAsync(fac, K.ab, bytes, hipMemcpyDeviceToDevice, pb.stream));
        } else {
            int rc = band_convert(K.ab, K.sprec, fac, prec, (size_t)K.n * K.ld, pb.stream);
            if (rc) return rc;
        }
    }
    size_t smem = prec == KSH_FP64 ? factor_smem<double>(K.bw) : factor_smem<float>(K.bw);
    if (smem > kMaxLds)
        return fail(KSH_SOLVER_ERROR, "band too wide for the one-workgroup factorization (LDS panel exceeds 160 KB)");
    int rc;
    if ((rc = allow_lds(smem))) return rc;
    double *dstats = K.red + 4;
    if (prec == KSH_FP64) {
        ProfScope ps("band_ldlt", pb.stream);
        hipLaunchKernelGGL(k_band_ldlt<double>, dim3(1), dim3(kThreads), smem, pb.stream, (double *)fac, K.n, K.bw, K.flag,
                           dstats);
    } else {
        ProfScope ps("band_ldlt", pb.stream);
        hipLaunchKernelGGL(k_band_ldlt<float>, dim3(1), dim3(kThreads), smem, pb.stream, (float *)fac, K.n, K.bw, K.flag,
                           dstats);
    }
    KSH_HIP(hipGetLastError());
    int info = 0;
    double st[4];
    KSH_HIP(hipMemcpyAsync(&info, K.flag, sizeof(int), hipMemcpyDeviceToHost, pb.stream));
    KSH_HIP(hipMemcpyAsync(st, dstats, sizeof(st), hipMemcpyDeviceToHost, pb.stream));
    KSH_HIP(hipStreamSynchronize(pb.stream));
    K.info = info;
    if (stats_out)
        for (int i = 0; i < 4; ++i) stats_out[i] = st[i];
    if (info != 0) return fail(KSH_SOLVER_ERROR, "LDL^T: zero or non-finite pivot at column " + std::to_string(info - 1));
    return KSH_SUCCESS;
}

int band_factor_batch(hipStream_t s, void *d_descs, const std::vector<BandDescHost> &h, int prec)
{
    if (h.empty()) return KSH_SUCCESS;
    int maxbw = 0;
    std::vector<BandDesc> dd(h.size());
    for (size_t i = 0; i < h.size(); ++i) {
        dd[i] = BandDesc{h[i].fac, h[i].n, h[i].bw, h[i].info, h[i].stats, h[i].x, h[i].active, 0};
        maxbw = std::max(maxbw, h[i].bw);
    }
    size_t smem = prec == KSH_FP64 ? factor_smem<double>(maxbw) : factor_smem<float>(maxbw);
    if (smem > kMaxLds) return fail(KSH_SOLVER_ERROR, "band too wide for the one-workgroup factorization");
    int rc;
    if ((rc = allow_lds(smem))) return rc;
    KSH_HIP(hipMemcpyAsync(d_descs, dd.data(), sizeof(BandDesc) * dd.size(), hipMemcpyHostToDevice, s));
    ProfScope ps("band_ldlt_batch", s);
    if (prec == KSH_FP64)
        hipLaunchKernelGGL(k_band_ldlt_batched<double>, dim3((unsigned)dd.size()), dim3(kThreads), smem, s,
                           (const BandDesc *)d_descs);
    else
        hipLaunchKernelGGL(k_band_ldlt_batched<float>, dim3((unsigned)dd.size()), dim3(kThreads), smem, s,
                           (const BandDesc *)d_descs);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int band_inertia(Matrix &K, const void *fac, int prec, int64_t *neg, double *dmin, double *dmaxabs, double *dminabs)
{
    Problem &pb = *K.pb;
    long long *dneg = (long long *)K.red;
    double *dstats = K.red + 1;
    if (prec == KSH_FP64)
        hipLaunchKernelGGL(k_band_inertia<double>, dim3(1), dim3(kThreads), 0, pb.stream, (const double *)fac, K.n, K.bw, dneg,
                           dstats);
    else
        hipLaunchKernelGGL(k_band_inertia<float>, dim3(1), dim3(kThreads), 0, pb.stream, (const float *)fac, K.n, K.bw, dneg,
                           dstats);
    KSH_HIP(hipGetLastError());
    double host[4];
    KSH_HIP(hipMemcpyAsync(host, K.red, 4 * sizeof(double), hipMemcpyDeviceToHost, pb.stream));
    KSH_HIP(hipStreamSynchronize(pb.stream));
    long long c;
    memcpy(&c, &host[0], sizeof(c));
    if (neg) *neg = c;
    if (dmin) *dmin = host[1];
    if (dmaxabs) *dmaxabs = host[2];
    if (dminabs) *dminabs = host[3];
    return KSH_SUCCESS;
}

int band_axpby(Matrix &Y, double alpha, double beta, const Matrix &X)
{
    Problem &pb = *Y.pb;
    size_t count = (size_t)Y.n * Y.ld;
    if (Y.sprec == KSH_FP64 && X.sprec == KSH_FP64)
        hipLaunchKernelGGL((k_band_axpby<double, double>), dim3(nblk(count, 256)), dim3(256), 0, pb.stream, (double *)Y.ab,
                           (const double *)X.ab, count, alpha, beta);
    else if (Y.sprec == KSH_FP64)
        hipLaunchKernelGGL((k_band_axpby<double, float>), dim3(nblk(count, 256)), dim3(256), 0, pb.stream, (double *)Y.ab,
                           (const float *)X.ab, count, alpha, beta);
    else if (X.sprec == KSH_FP64)
        hipLaunchKernelGGL((k_band_axpby<float, double>), dim3(nblk(count, 256)), dim3(256), 0, pb.stream, (float *)Y.ab,
                           (const double *)X.ab, count, alpha, beta);
    else
        hipLaunchKernelGGL((k_band_axpby<float, float>), dim3(nblk(count, 256)), dim3(256), 0, pb.stream, (float *)Y.ab,
                           (const float *)X.ab, count, alpha, beta);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int band_convert(const void *src, int src_prec, void *dst, int dst_prec, size_t count, hipStream_t s)
{
    if (src_prec == KSH_FP64 && dst_prec == KSH_FP32)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(count, 256)), dim3(256), 0, s, (float *)dst, (const double *)src,
                           count);
    else if (src_prec == KSH_FP32 && dst_prec == KSH_FP64)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk(count, 256)), dim3(256), 0, s, (double *)dst, (const float *)src,
                           count);
    else
        return fail(KSH_OTHER_ERROR, "band_convert: same precision");
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int vec_axpy(hipStream_t s, int n, double a, const double *x, double *y)
{
    hipLaunchKernelGGL(k_axpy, dim3(nblk(n, 256)), dim3(256), 0, s, n, a, x, y);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int vec_sub(hipStream_t s, int n, const double *a, const double *b, double *out)
{
    hipLaunchKernelGGL(k_sub, dim3(nblk(n, 256)), dim3(256), 0, s, n, a, b, out);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int vec_norm2(hipStream_t s, int n, const double *x, double *scratch, double *host_out)
{
    hipLaunchKernelGGL(k_norm2, dim3(1), dim3(kThreads), 0, s, n, x, scratch);
    KSH_HIP(hipGetLastError());
    double v = 0.0;
    KSH_HIP(hipMemcpyAsync(&v, scratch, sizeof(double), hipMemcpyDeviceToHost, s));
    KSH_HIP(hipStreamSynchronize(s));
    *host_out = sqrt(v);
    return KSH_SUCCESS;
}

}  // namespace ksh
