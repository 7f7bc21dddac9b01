// ksh_spmv.hip -- symmetric band mat-vec y = K x and batched refinement residuals r = b - K x.
//
// K is stored as its lower band, column major: ab[j*ld + (i-j)], ld = bw+1.  A workgroup of 256
// threads owns a block of 64 consecutive rows of one matrix (a launch covers all row blocks of all
// matrices of a batch):
//   lower part  sum_{j<=i} K[i][j] x[j]: lanes = the block's rows, so each band column contributes
//               one contiguous 64-entry run (coalesced); the four waves take every fourth column and
//               issue 8 independent loads per lane before accumulating;
//   upper part  sum_{j>i} K[j][i] x[j] = sum_d ab[i*ld + d] x[i+d]: one wave per 16 rows, lanes along
//               the contiguous column of row i, wave reductions.
// Every band entry is read twice (once for each triangle); the mat-vec is HBM bound at
// 2 n (bw+1) sizeof(T) bytes.  Refinement residual norms are reduced per block, then per matrix in
// block order (deterministic).
#include <hip/hip_runtime.h>

#include "ksh_internal.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {

constexpr int kRows = 64;
constexpr int kT = 256;
constexpr int kU = 8;

// (K x)[i0 + lane] for NR vectors x[0..NR-1] at once: the band entries are loaded once and used
// for every vector.  Returns the products in y (valid in lanes 0..63 of every wave).
template <typename T, int NR>
__device__ __forceinline__ void symv_row_block(const T *__restrict__ ab, int n, int bw, const double *const (&x)[NR], int i0,
                                               double (&y)[NR])
{
    __shared__ double part[NR][4][kRows];
    __shared__ double up[NR][kRows];
    const size_t ld = (size_t)bw + 1;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int i = i0 + lane;
    const bool ok = i < n;
    // lower triangle incl. diagonal
    {
        const int jlo = max(0, i0 - bw), jhi = min(n - 1, i0 + kRows - 1);
        double acc[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = 0.0;
        for (int jb = jlo + wave; jb <= jhi; jb += 4 * kU) {
            double v[kU], xv[NR][kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                int j = jb + 4 * u;
                bool in = ok && j <= jhi && j <= i && i - j <= bw;
                v[u] = in ? (double)ab[(size_t)j * ld + (i - j)] : 0.0;
#pragma unroll
                for (int q = 0; q < NR; ++q) xv[q][u] = (j <= jhi) ? x[q][j] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
#pragma unroll
                for (int q = 0; q < NR; ++q) acc[q] += v[u] * xv[q][u];
        }
#pragma unroll
        for (int q = 0; q < NR; ++q) part[q][wave][lane] = acc[q];
    }
    // strict upper triangle: rows 16*wave .. 16*wave+15 of the block
    for (int rr = 0; rr < 16; ++rr) {
        const int row = i0 + 16 * wave + rr;
        double acc[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = 0.0;
        if (row < n) {
            const int dmax = min(bw, n - 1 - row);
            for (int d0 = 1; d0 <= dmax; d0 += 64 * 4) {
                double v[4], xv[NR][4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int d = d0 + lane + 64 * u;
                    bool in = d <= dmax;
                    v[u] = in ? (double)ab[(size_t)row * ld + d] : 0.0;
#pragma unroll
                    for (int q = 0; q < NR; ++q) xv[q][u] = in ? x[q][row + d] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < NR; ++q) acc[q] += v[u] * xv[q][u];
            }
        }
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            for (int off = 32; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off);
            if (lane == 0) up[q][16 * wave + rr] = acc[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NR; ++q) y[q] = part[q][0][lane] + part[q][1][lane] + part[q][2][lane] + part[q][3][lane] + up[q][lane];
    __syncthreads();  // part/up are reused by the next call
}

template <typename T>
__global__ __launch_bounds__(kT) void k_symv_rows(const T *ab, int n, int bw, const double *x, double *y)
{
    const int i0 = blockIdx.x * kRows;
    const double *xs[1] = {x};
    double v[1];
    symv_row_block<T, 1>(ab, n, bw, xs, i0, v);
    if (threadIdx.x < kRows && i0 + threadIdx.x < n) y[i0 + threadIdx.x] = v[0];
}

__device__ inline int find_block(const RefDesc *d, int n, int blk)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (d[mid].blk0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// r = b - K x on one row block of one matrix (and r2 = b2 - K x2 with the same band loads when the
// descriptor carries a second vector); block partial norms (||r||^2, ||b||^2) -> part (part2)
template <int NR>
__device__ __forceinline__ void refine_rows(const RefDesc &d, int lb)
{
    const int i0 = lb * kRows;
    const double *xs[NR];
    xs[0] = d.x;
    if constexpr (NR > 1) xs[1] = d.x2;
    double kx[NR];
    if (d.sprec == KSH_FP64)
        symv_row_block<double, NR>((const double *)d.ab, d.n, d.bw, xs, i0, kx);
    else
        symv_row_block<float, NR>((const float *)d.ab, d.n, d.bw, xs, i0, kx);
    if (threadIdx.x < 64) {
        const int i = i0 + threadIdx.x;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const double *b = q == 0 ? d.b : d.b2;
            double *r = q == 0 ? d.r : d.r2, *part = q == 0 ? d.part : d.part2;
            double ri = 0.0, bi = 0.0;
            if (i < d.n) {
                bi = b[i];
                ri = bi - kx[q];
                r[i] = ri;
            }
            double rr = ri * ri, bb = bi * bi;
            for (int off = 32; off > 0; off >>= 1) {
                rr += __shfl_xor(rr, off);
                bb += __shfl_xor(bb, off);
            }
            if (threadIdx.x == 0) {
                part[2 * lb] = rr;
                part[2 * lb + 1] = bb;
            }
        }
    }
}

__global__ __launch_bounds__(kT) void k_refine_rows_batched(const RefDesc *descs, int ninst)
{
    const RefDesc &d = descs[find_block(descs, ninst, blockIdx.x)];
    if (!d.active) return;
    const int lb = blockIdx.x - d.blk0;
    if (d.x2)
        refine_rows<2>(d, lb);
    else
        refine_rows<1>(d, lb);
}

// per-matrix norms from the block partials (block order), one workgroup per matrix
__global__ __launch_bounds__(64) void k_refine_norms_batched(const RefDesc *descs)
{
    const RefDesc &d = descs[blockIdx.x];
    if (!d.active) return;
    if (threadIdx.x < 2) {
        const double *part = threadIdx.x == 0 ? d.part : d.part2;
        double *norms = threadIdx.x == 0 ? d.norms : d.norms2;
        if (!part) return;
        double rr = 0.0, bb = 0.0;
        for (int k = 0; k < d.nblk; ++k) {
            rr += part[2 * k];
            bb += part[2 * k + 1];
        }
        norms[0] = rr;
        norms[1] = bb;
    }
}

__global__ __launch_bounds__(kT) void k_refine_update_batched(const RefDesc *descs, int ninst)
{
    const RefDesc &d = descs[find_block(descs, ninst, blockIdx.x)];
    if (!d.active) return;
    const int i = (blockIdx.x - d.blk0) * kRows + threadIdx.x;
    if (threadIdx.x < kRows && i < d.n) d.x[i] += d.r[i];
}

inline unsigned nblocks(int n) { return (unsigned)((n + kRows - 1) / kRows); }

int prefix_blocks(std::vector<RefDesc> &h)
{
    int tot = 0;
    for (auto &d : h) {
        d.blk0 = tot;
        d.nblk = (int)nblocks(d.n);
        tot += d.nblk;
    }
    return tot;
}

}  // namespace

int band_symv(const Matrix &K, const double *x_dev, double *y_dev)
{
    Problem &pb = *K.pb;
    ProfScope ps("band_symv", pb.stream);
    ps.set_work(K.n * (2.0 * K.bw + 1.0) * (K.sprec == KSH_FP64 ? 8.0 : 4.0) + 16.0 * K.n);
    if (K.sprec == KSH_FP64)
        hipLaunchKernelGGL(k_symv_rows<double>, dim3(nblocks(K.n)), dim3(kT), 0, pb.stream, (const double *)K.ab, K.n, K.bw,
                           x_dev, y_dev);
    else
        hipLaunchKernelGGL(k_symv_rows<float>, dim3(nblocks(K.n)), dim3(kT), 0, pb.stream, (const float *)K.ab, K.n, K.bw,
                           x_dev, y_dev);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int refine_residual_batch(hipStream_t s, RefDesc *d_descs, const std::vector<RefDesc> &hin)
{
    if (hin.empty()) return KSH_SUCCESS;
    // two active jobs on the same matrix (the arc-length solves' two right-hand sides) share one
    // descriptor, so that their band is read once for both products
    std::vector<RefDesc> h;
    h.reserve(hin.size());
    std::vector<char> taken(hin.size(), 0);
    for (size_t i = 0; i < hin.size(); ++i) {
        if (taken[i]) continue;
        RefDesc d = hin[i];
        d.b2 = nullptr;
        d.x2 = d.r2 = d.norms2 = d.part2 = nullptr;
        if (d.active)
            for (size_t j = i + 1; j < hin.size(); ++j) {
                const RefDesc &e = hin[j];
                if (taken[j] || !e.active || e.ab != d.ab || e.sprec != d.sprec || e.n != d.n || e.bw != d.bw) continue;
                d.b2 = e.b;
                d.x2 = e.x;
                d.r2 = e.r;
                d.norms2 = e.norms;
                d.part2 = e.part;
                taken[j] = 1;
                break;
            }
        h.push_back(d);
    }
    int tot = prefix_blocks(h);
    KSH_HIP(hipMemcpyAsync(d_descs, h.data(), sizeof(RefDesc) * h.size(), hipMemcpyHostToDevice, s));
    {
        ProfScope ps("refine_residual_batch", s);
        double work = 0.0;  // K read as a full band (2b+1 per row), x and b read, r written (per vector)
        for (const auto &d : h)
            if (d.active) work += d.n * (2.0 * d.bw + 1.0) * (d.sprec == KSH_FP64 ? 8.0 : 4.0) + (d.x2 ? 48.0 : 24.0) * d.n;
        ps.set_work(work);
        hipLaunchKernelGGL(k_refine_rows_batched, dim3((unsigned)tot), dim3(kT), 0, s, (const RefDesc *)d_descs, (int)h.size());
    }
    KSH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_refine_norms_batched, dim3((unsigned)h.size()), dim3(64), 0, s, (const RefDesc *)d_descs);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int refine_update_batch(hipStream_t s, RefDesc *d_descs, const std::vector<RefDesc> &hin)
{
    if (hin.empty()) return KSH_SUCCESS;
    std::vector<RefDesc> h = hin;
    int tot = prefix_blocks(h);
    KSH_HIP(hipMemcpyAsync(d_descs, h.data(), sizeof(RefDesc) * h.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_refine_update_batched, dim3((unsigned)tot), dim3(kT), 0, s, (const RefDesc *)d_descs, (int)h.size());
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

}  // namespace ksh
