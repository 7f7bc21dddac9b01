// ksh_band.hip -- inertia of a factored band, band axpby / precision conversion, vector helpers.
//
// The inertia replaces SimplicialLDLT::vectorD() and the negative-pivot count of
// gsALMBase::_computeStability / gsStaticBase (src/gsALMSolvers/gsALMHelper.h:30-42,
// src/gsALMSolvers/gsALMBase.hpp:186-211).  By Sylvester's law of inertia the count does not
// depend on the ordering, so the band ordering reproduces the reference's AMD-ordered count.
// Layout: lower band, column major, ab[j*ld + (i-j)], ld = bw+1.
#include <hip/hip_runtime.h>

#include <cstdint>

#include <cmath>
#include <cstring>

#include "ksh_internal.hpp"
#include "ksh_mfma.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {


constexpr int kThreads = 256;


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

// copy (or convert) of many bands in one launch: a chunk of kCopyBlk entries per block, each
// thread kCopyPer consecutive entries with 16-byte loads and stores (when the chunk is whole and
// the band pointers are 16-byte aligned; elementwise otherwise); each block finds its band by
// binary search over the chunk offsets.  The conversion is the same per entry either way.
constexpr int kCopyPer = 8, kCopyBlk = 256 * kCopyPer;

__global__ __launch_bounds__(256) void k_copy_batched(const CopyDesc *descs, int nband)
{
    const long long blk = blockIdx.x;
    int lo = 0, hi = nband - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (descs[mid].blk0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    const CopyDesc &d = descs[lo];
    const size_t i0 = (size_t)(blk - d.blk0) * kCopyBlk + (size_t)threadIdx.x * kCopyPer;
    if (i0 >= d.count) return;
    const bool f64s = d.src_prec == KSH_FP64, f64d = d.dst_prec == KSH_FP64;
    const bool vec = i0 + kCopyPer <= d.count && ((reinterpret_cast<uintptr_t>(d.src) | reinterpret_cast<uintptr_t>(d.dst)) & 15) == 0;
    double v[kCopyPer];
    if (vec) {
        if (f64s) {
            const double2 *p = reinterpret_cast<const double2 *>((const double *)d.src + i0);
#pragma unroll
            for (int q = 0; q < kCopyPer / 2; ++q) {
                const double2 t = p[q];
                v[2 * q] = t.x;
                v[2 * q + 1] = t.y;
            }
        } else {
            const float4 *p = reinterpret_cast<const float4 *>((const float *)d.src + i0);
#pragma unroll
            for (int q = 0; q < kCopyPer / 4; ++q) {
                const float4 t = p[q];
                v[4 * q] = t.x;
                v[4 * q + 1] = t.y;
                v[4 * q + 2] = t.z;
                v[4 * q + 3] = t.w;
            }
        }
        if (f64d) {
            double2 *p = reinterpret_cast<double2 *>((double *)d.dst + i0);
#pragma unroll
            for (int q = 0; q < kCopyPer / 2; ++q) p[q] = make_double2(v[2 * q], v[2 * q + 1]);
        } else {
            float4 *p = reinterpret_cast<float4 *>((float *)d.dst + i0);
#pragma unroll
            for (int q = 0; q < kCopyPer / 4; ++q)
                p[q] = make_float4((float)v[4 * q], (float)v[4 * q + 1], (float)v[4 * q + 2], (float)v[4 * q + 3]);
        }
        return;
    }
    for (size_t i = i0; i < d.count && i < i0 + kCopyPer; ++i) {
        const double x = f64s ? ((const double *)d.src)[i] : (double)((const float *)d.src)[i];
        if (f64d) ((double *)d.dst)[i] = x;
        else ((float *)d.dst)[i] = (float)x;
    }
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

static_assert(sizeof(BandDesc) == kBandDescBytes, "BandDesc layout");

}  // namespace

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

int band_copy_batch(hipStream_t s, void *d_descs, const std::vector<CopyDesc> &hin)
{
    if (hin.empty()) return KSH_SUCCESS;
    std::vector<CopyDesc> h = hin;
    long long blocks = 0;
    for (auto &d : h) {
        d.blk0 = blocks;
        blocks += (long long)((d.count + kCopyBlk - 1) / kCopyBlk);
    }
    if (blocks >= (1LL << 31)) return fail(KSH_OTHER_ERROR, "band copy of 2^31 blocks or more");
    KSH_HIP(hipMemcpyAsync(d_descs, h.data(), sizeof(CopyDesc) * h.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_copy_batched, dim3((unsigned)blocks), dim3(256), 0, s, (const CopyDesc *)d_descs, (int)h.size());
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
