// ksh_solve_tiled.hip -- L D L^T solve of one large band spread over many workgroups.
//
// For bands too long for the LDS-resident kernel (vector over 160 KB) or wider than its 512
// prefetched columns -- e.g. the 128 x 128 degree-3 plate, n = 51k, bw = 1190, whose band is
// about 0.5 GB in FP64 -- a single workgroup would stream the band at one CU's bandwidth.  Here
// each 64-row step of a sweep is one launch whose workgroups share the rows the step touches:
//   forward, block c0..c0+63 (rows after it already hold b - L y of the earlier blocks):
//     every workgroup solves the unit-lower diagonal block for y (wave 0, lane = row,
//     v_readlane broadcasts; the block is read from L2), workgroup 0 stores z = D^-1 y, and each
//     workgroup takes 64 of the following bw rows: b[r] -= sum_k L[r][c0+k] y_k (lanes = rows,
//     so each load reads 64 consecutive entries of a band column);
//   backward, block c0..c0+63 from the last block down (z of the block already holds
//     D^-1 y - sum over later blocks): every workgroup solves the transposed diagonal block for
//     x, workgroup 0 stores it, and each workgroup takes 64 of the preceding bw rows:
//     z[i] -= sum_k L[c0+k][i] x_k (lanes = k, so each load reads 64 consecutive entries of
//     column i; the 64 x 64 products are summed per row through LDS).
// Rows touched by one launch belong to one workgroup, and the block rows are read but not
// written, so launches on one stream need no other synchronization.  HBM traffic: the band once
// per sweep, 2 n (bw+1) entries per solve.  Replaces SimplicialLDLT::solve
// (src/gsEigenSolvers/gsEigenProblemBase.hpp:152-216 drives it through the shift-invert operator).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ksh_devutil.hpp"
#include "ksh_internal.hpp"

namespace ksh {

namespace {

constexpr int kSB = 64;  // rows per step = threads per workgroup

template <typename T>
__global__ __launch_bounds__(kSB) void k_band_fwd_step(const T *fac, int n, int bw, int c0, double *b, double *z)
{
    __shared__ double ys[kSB];
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    const int lane = threadIdx.x;
    const int w = min(kSB, n - c0), nr = min(n - c0, w + bw);
    // diagonal block: lane ii holds row c0+ii; L[c0+ii][c0+jj] for jj < ii
    T l[kSB];
#pragma unroll
    for (int jj = 0; jj < kSB; ++jj) {
        const bool ok = jj < lane && lane < w && lane - jj <= bw;
        l[jj] = load_if<T>(ok, rs, ((unsigned)(c0 + jj) * ldu + (unsigned)(lane - jj)) * es);
    }
    const T d = load_if<T>(lane < w, rs, (unsigned)(c0 + lane) * ldu * es);
    // this workgroup's rows c0+r, r = w + 64 blockIdx + lane: their band entries do not depend on
    // the diagonal block's solution, so they are loaded together with it
    const int r = w + kSB * (int)blockIdx.x + lane;
    T u[kSB];
#pragma unroll
    for (int kk = 0; kk < kSB; ++kk) {
        const bool ok = kk < w && r < nr && r - kk <= bw;
        u[kk] = load_if<T>(ok, rs, ((unsigned)(c0 + kk) * ldu + (unsigned)(r - kk)) * es);
    }
    const double rb = r < nr ? b[c0 + r] : 0.0;
    double y = lane < w ? b[c0 + lane] : 0.0;
#pragma unroll
    for (int jj = 0; jj < kSB - 1; ++jj) y -= (double)l[jj] * readlane_f64(y, jj);  // l = 0 on and above the diagonal
    ys[lane] = y;
    if (blockIdx.x == 0 && lane < w) z[c0 + lane] = y / (double)d;
    __syncthreads();
    double acc = 0.0;
#pragma unroll
    for (int kk = 0; kk < kSB; ++kk) acc += (double)u[kk] * ys[kk];
    if (r < nr) b[c0 + r] = rb - acc;
}

template <typename T>
__global__ __launch_bounds__(kSB) void k_band_bwd_step(const T *fac, int n, int bw, int c0, double *z, double *x)
{
    __shared__ double prod[kSB][kSB + 1];
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    const int lane = threadIdx.x;
    const int w = min(kSB, n - c0);
    // transposed diagonal block: lane jj holds x_{c0+jj}; L[c0+kk][c0+jj] for kk > jj
    T l[kSB];
#pragma unroll
    for (int kk = 0; kk < kSB; ++kk) {
        const bool ok = kk > lane && kk < w && kk - lane <= bw;
        l[kk] = load_if<T>(ok, rs, ((unsigned)(c0 + lane) * ldu + (unsigned)(kk - lane)) * es);
    }
    // rows i = c0 - 1 - (64 blockIdx + q), q = 0..63, above the block within the band:
    // z[i] -= sum_k L[c0+k][i] x_k; lane = k.  Their entries are loaded with the diagonal block.
    const int lo = max(0, c0 - bw);
    T u[kSB];
#pragma unroll
    for (int q = 0; q < kSB; ++q) {
        const int i = c0 - 1 - (kSB * (int)blockIdx.x + q);
        const bool ok = i >= lo && lane < w && c0 + lane - i <= bw;
        u[q] = load_if<T>(ok, rs, ((unsigned)max(i, 0) * ldu + (unsigned)(c0 + lane - max(i, 0))) * es);
    }
    double xv = lane < w ? z[c0 + lane] : 0.0;
#pragma unroll
    for (int kk = kSB - 1; kk > 0; --kk) xv -= (double)l[kk] * readlane_f64(xv, kk);  // l = 0 on and below the diagonal
    if (blockIdx.x == 0 && lane < w) x[c0 + lane] = xv;
#pragma unroll
    for (int q = 0; q < kSB; ++q) prod[q][lane] = (double)u[q] * xv;  // masked entries loaded as 0
    __syncthreads();
    const int i = c0 - 1 - (kSB * (int)blockIdx.x + lane);
    double s = 0.0;
#pragma unroll 16
    for (int k = 0; k < kSB; ++k) s += prod[lane][k];
    if (i >= lo) z[i] -= s;
}

template <typename T>
int run_solve(hipStream_t s, const T *fac, int n, int bw, double *x, double *z)
{
    for (int c0 = 0; c0 < n; c0 += kSB) {
        const int w = std::min(kSB, n - c0), nr = std::min(n - c0, w + bw);
        const unsigned grid = (unsigned)std::max(1, (nr - w + kSB - 1) / kSB);
        hipLaunchKernelGGL(k_band_fwd_step<T>, dim3(grid), dim3(kSB), 0, s, fac, n, bw, c0, x, z);
    }
    for (int c0 = ((n - 1) / kSB) * kSB; c0 >= 0; c0 -= kSB) {
        const int above = std::min(c0, bw);
        const unsigned grid = (unsigned)std::max(1, (above + kSB - 1) / kSB);
        hipLaunchKernelGGL(k_band_bwd_step<T>, dim3(grid), dim3(kSB), 0, s, fac, n, bw, c0, z, x);
    }
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

}  // namespace

int band_solve_tiled(hipStream_t s, const void *fac, int prec, int n, int bw, double *x, double *z)
{
    const size_t es = prec == KSH_FP64 ? sizeof(double) : sizeof(float);
    if ((size_t)n * (size_t)(bw + 1) * es >= 0x80000000ull)
        return fail(KSH_SOLVER_ERROR, "band of 2 GiB or more: beyond the band kernels' 32-bit offsets");
    return prec == KSH_FP64 ? run_solve<double>(s, (const double *)fac, n, bw, x, z)
                            : run_solve<float>(s, (const float *)fac, n, bw, x, z);
}

}  // namespace ksh
