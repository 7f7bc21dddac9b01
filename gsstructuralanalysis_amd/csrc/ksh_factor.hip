// ksh_factor.hip -- banded LDL^T factorization for gfx950 (one workgroup of 1024 threads per matrix,
// a launch covers a batch of matrices).
//
// Replaces gsSparseSolver<T>::SimplicialLDLT::compute on the arc-length, Newton and eigen paths
// (src/gsALMSolvers/gsALMBase.hpp:186-211, src/gsStaticSolvers/gsStaticNewton.hpp:215-240,
// src/gsEigenSolvers/gsBucklingSolver.h:131-133).  No pivoting, as there; by Sylvester's law of
// inertia the number of negative pivots does not depend on the ordering, so the band ordering
// reproduces the count gsALMBase::_computeStability takes from the reference's factor
// (src/gsALMSolvers/gsALMHelper.h:30-42).
//
// Layout: lower band, column major, fac[j*ld + (i-j)], ld = bw+1; factored in place into unit L
// (below the diagonal) and D (on it).
//
// Right-looking, panels of NB = 32 columns (16 for bands whose panel does not fit LDS):
//   1. the panel (NB columns x up to NB+bw rows, already updated by earlier panels) is staged in
//      LDS, column major, with buffer loads (masked entries read as 0);
//   2. the panel is factored in sub-panels of 16 columns: wave 0 factors the 16x16 diagonal block
//      in registers (lane = row, v_readlane broadcasts), every thread solves one row below it
//      against L_dd D (row-wise TRSM), and the panel's remaining columns take the sub-panel's
//      rank-16 update;
//   3. the factored panel is written back;
//   4. the trailing bw x bw window takes the symmetric update A -= (L D) L^T in 16x16 tiles on
//      v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64 (panel operands from LDS, A tiles in HBM);
//      each wave walks its tiles with the next tile's A loads issued before the current tile's
//      MFMAs, so the HBM latency overlaps the matrix work.
// Work per matrix: about n bw^2 multiply-adds (tiles), n bw NB (TRSM); HBM traffic about
// 2 n bw^2 / NB entries (the trailing window is read and written once per panel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "ksh_devutil.hpp"
#include "ksh_internal.hpp"
#include "ksh_mfma.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {

constexpr int kFT = 1024;            // threads per workgroup
constexpr int kFW = kFT / 64;        // waves

__host__ __device__ inline int fpanel_rows(int bw, int nb) { return ((bw + nb + 15) / 16) * 16 + 16; }

template <typename T, int NB>
__host__ __device__ inline size_t factor_lds_bytes(int bw)
{
    return sizeof(T) * ((size_t)NB * fpanel_rows(bw, NB) + 16 * 16 + 2 * NB);
}

// negatives, min d, max |d|, min |d| of the factored diagonal (whole workgroup)
template <typename T>
__device__ void pivot_stats_wg(const T *fac, int n, int bw, double *stats)
{
    __shared__ double part[kFW][4];
    const size_t ld = (size_t)bw + 1;
    double c = 0.0, mn = INFINITY, mxa = 0.0, mna = INFINITY;
    for (int i = threadIdx.x; i < n; i += kFT) {
        const double d = (double)fac[(size_t)i * ld];
        c += d < 0.0 ? 1.0 : 0.0;
        mn = fmin(mn, d);
        mxa = fmax(mxa, fabs(d));
        mna = fmin(mna, fabs(d));
    }
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off);
        mn = fmin(mn, __shfl_xor(mn, off));
        mxa = fmax(mxa, __shfl_xor(mxa, off));
        mna = fmin(mna, __shfl_xor(mna, off));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        part[wave][0] = c;
        part[wave][1] = mn;
        part[wave][2] = mxa;
        part[wave][3] = mna;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kFW; ++w) {
            c += part[w][0];
            mn = fmin(mn, part[w][1]);
            mxa = fmax(mxa, part[w][2]);
            mna = fmin(mna, part[w][3]);
        }
        stats[0] = c;
        stats[1] = mn;
        stats[2] = mxa;
        stats[3] = mna;
    }
}

template <typename T, int NB>
__device__ __forceinline__ void band_ldlt_wg(T *fac, int n, int bw, int *info, double *stats, unsigned char *smem)
{
    using M = Mfma16<T>;
    using acc_t = typename M::acc_t;
    const int PR = fpanel_rows(bw, NB);
    T *pan = (T *)smem;                 // [NB][PR]: pan[jj*PR + r] = A / L[c0+r][c0+jj]
    T *Ldd = pan + (size_t)NB * PR;     // [16][16]: Ldd[jj*16 + kk] = L of the current 16x16 diagonal block
    T *dv = Ldd + 16 * 16;              // [NB] pivots of the panel
    T *dinv = dv + NB;                  // [NB] their inverses
    __shared__ int s_info;
    const unsigned ldu = (unsigned)bw + 1;
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * sizeof(T)));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_info = 0;
    __syncthreads();

    for (int c0 = 0; c0 < n; c0 += NB) {
        const int w = min(NB, n - c0);
        const int nr = min(n - c0, w + bw);  // panel rows c0 .. c0+nr-1
        // ---- 1. stage the panel ----
        {
            constexpr int kU = 8;
            const int tot = NB * PR;
            for (int base = 0; base < tot; base += kFT * kU) {
                T v[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int idx = base + tid + kFT * u;
                    const int jj = idx / PR, r = idx - jj * PR;
                    const bool ok = idx < tot && jj < w && r < nr && r >= jj && r - jj <= bw;
                    v[u] = load_if<T>(ok, rs, ((unsigned)(c0 + jj) * ldu + (unsigned)(r - jj)) * (unsigned)sizeof(T));
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int idx = base + tid + kFT * u;
                    const int jj = idx / PR, r = idx - jj * PR;
                    if (idx < tot) pan[idx] = (jj >= w && r == jj) ? T(1) : v[u];  // unit pivots pad the last panel
                }
            }
        }
        lds_barrier();
        // ---- 2.-3. factor the panel in sub-panels of 16 columns ----
        for (int q0 = 0; q0 < NB; q0 += 16) {
            // 2. diagonal 16x16 block (wave 0; lane r < 16 owns row q0+r)
            if (wave == 0) {
                const int r = opaque(lane);
                const int rr = q0 + (r & 15);  // lanes >= 16 shadow a row; their results are discarded
                T a[16];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) a[kk] = pan[(q0 + kk) * PR + rr];
                // Right-looking elimination.  Entries on and above the diagonal of a lane's row
                // take updates too; they are never read, so no per-column lane mask is needed.
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) {
                    const T djj = bcast(a[jj], jj);
                    if (lane == 0) {
                        if (q0 + jj < w && (djj == T(0) || !isfinite((double)djj)) && s_info == 0) s_info = c0 + q0 + jj + 1;
                        dv[q0 + jj] = djj;
                        dinv[q0 + jj] = T(1) / djj;
                    }
                    const T c = a[jj];
                    const T l = c / djj;
#pragma unroll
                    for (int kk = jj + 1; kk < 16; ++kk) a[kk] -= l * bcast(c, kk);
                    a[jj] = l;
                }
                if (r < 16) {
                    const T dr = dv[q0 + r];
#pragma unroll
                    for (int kk = 0; kk < 16; ++kk) {
                        const T lv = kk < r ? a[kk] : T(0);
                        Ldd[r * 16 + kk] = lv;
                        pan[(q0 + kk) * PR + q0 + r] = kk == r ? dr : lv;
                    }
                }
            }
            lds_barrier();
            // 3. TRSM: rows below the block, L[r][q0:q0+16] = A[r][q0:q0+16] (D L_dd^T)^-1
            for (int r = q0 + 16 + tid; r < nr; r += kFT) {
                T a[16];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) a[kk] = pan[(q0 + kk) * PR + r];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) {
                    const T t = a[kk];  // = L[r][q0+kk] d_kk
#pragma unroll
                    for (int jj = kk + 1; jj < 16; ++jj) a[jj] -= t * Ldd[jj * 16 + kk];
                    a[kk] = t * dinv[q0 + kk];
                }
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) pan[(q0 + kk) * PR + r] = a[kk];
            }
            lds_barrier();
            if (q0 + 16 < NB) {
                // update of the panel's remaining columns q0+16 .. NB-1 by this sub-panel:
                // A[r][jj] -= sum_kk L[r][kk] d_kk L[jj][kk], rows r >= jj
                for (int r = q0 + 16 + tid; r < nr; r += kFT) {
                    T l[16];
#pragma unroll
                    for (int kk = 0; kk < 16; ++kk) l[kk] = pan[(q0 + kk) * PR + r] * dv[q0 + kk];
                    for (int jj = q0 + 16; jj < NB; ++jj) {
                        T s = pan[jj * PR + r];
#pragma unroll
                        for (int kk = 0; kk < 16; ++kk) s -= l[kk] * pan[(q0 + kk) * PR + jj];
                        pan[jj * PR + r] = s;  // rows r < jj are overwritten with values never read
                    }
                }
                lds_barrier();
            }
        }
        // ---- 4. write the factored panel back ----
        {
            const int tot = NB * PR;
            for (int idx = tid; idx < tot; idx += kFT) {
                const int jj = idx / PR, r = idx - jj * PR;
                const bool ok = jj < w && r < nr && r >= jj && r - jj <= bw;
                store_if<T>(ok, rs, ((unsigned)(c0 + jj) * ldu + (unsigned)(r - jj)) * (unsigned)sizeof(T), pan[idx]);
            }
        }
        // ---- 5. trailing update of rows/columns [w, nr) of the panel frame ----
        {
            const int nt = (nr - w + 15) / 16;
            const int npair = nt * (nt + 1) / 2;
            const int li = lane & 15, lk = lane >> 4;
            // tile p -> (ti, tj), tj <= ti, row-major over the lower triangle of tiles
            auto decode = [&](int p, int &ri0, int &rj0) {
                int ti = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
                if ((ti + 1) * (ti + 2) / 2 <= p) ++ti;
                if (ti * (ti + 1) / 2 > p) --ti;
                const int tj = p - ti * (ti + 1) / 2;
                ri0 = w + 16 * ti;
                rj0 = w + 16 * tj;
            };
            // accumulator rows = j (A's column), columns = i (A's row): lanes run along a column
            auto tile_io = [&](int ri0, int rj0, int v, bool &ok) -> unsigned {
                const int i = ri0 + li, j = rj0 + M::row(lane, v);
                ok = i < nr && j < nr && i >= j;
                return ((unsigned)(c0 + j) * ldu + (unsigned)(i - j)) * (unsigned)sizeof(T);
            };
            auto tile_load = [&](int p) -> acc_t {
                int ri0, rj0;
                decode(p, ri0, rj0);
                acc_t x;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    bool ok;
                    const unsigned off = tile_io(ri0, rj0, v, ok);
                    x[v] = load_if<T>(ok && p < npair, rs, off);
                }
                return x;
            };
            // a wave keeps the A loads of its next kDepth tiles in flight
            constexpr int kDepth = 4;
            acc_t buf[kDepth];
#pragma unroll
            for (int u = 0; u < kDepth; ++u) buf[u] = tile_load(wave + kFW * u);
            for (int p0 = wave; p0 < npair; p0 += kFW * kDepth) {
#pragma unroll
                for (int u = 0; u < kDepth; ++u) {
                    const int p = p0 + kFW * u;
                    acc_t cur = buf[u];
                    buf[u] = tile_load(p + kFW * kDepth);  // past the last tile every read is masked off
                    int ri0, rj0;
                    decode(p, ri0, rj0);
#pragma unroll
                    for (int s = 0; s < NB / 4; ++s) {
                        const int k = 4 * s + lk;
                        const T wv = -pan[k * PR + rj0 + li] * dv[k];
                        const T lv = pan[k * PR + ri0 + li];
                        cur = M::run(wv, lv, cur);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        bool ok;
                        const unsigned off = tile_io(ri0, rj0, v, ok);
                        store_if<T>(ok && p < npair, rs, off, cur[v]);
                    }
                }
            }
        }
        __syncthreads();  // trailing stores visible to the next panel's staging loads
        if (s_info != 0) break;
    }
    if (tid == 0) *info = s_info;
    __syncthreads();
    if (stats) pivot_stats_wg<T>(fac, n, bw, stats);
}

template <typename T, int NB>
__global__ __launch_bounds__(kFT) void k_band_ldlt_wg(T *fac, int n, int bw, int *info, double *stats)
{
    extern __shared__ __align__(16) unsigned char smem[];
    band_ldlt_wg<T, NB>(fac, n, bw, info, stats, smem);
}

template <typename T, int NB>
__global__ __launch_bounds__(kFT) void k_band_ldlt_wg_batched(const BandDesc *descs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BandDesc &d = descs[blockIdx.x];
    if (!d.active) return;
    band_ldlt_wg<T, NB>((T *)d.fac, d.n, d.bw, d.info, d.stats, smem);
}

// static LDS of the factor kernels (pivot flag, per-wave statistics) counts against the 160 KB
constexpr size_t kFactorStaticLds = kFW * 4 * sizeof(double) + 64;
constexpr size_t kFactorLdsMax = 160 * 1024 - kFactorStaticLds;

template <typename T, int NB>
int allow_factor_lds(size_t need)
{
    static size_t granted = 0;
    if (need <= granted) return KSH_SUCCESS;
    KSH_HIP(hipFuncSetAttribute((const void *)k_band_ldlt_wg<T, NB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    KSH_HIP(hipFuncSetAttribute((const void *)k_band_ldlt_wg_batched<T, NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)need));
    granted = need;
    return KSH_SUCCESS;
}

// panel width for band width bw: 32 when the panel fits LDS, else 16, else 0 (unsupported)
template <typename T>
int panel_width(int bw)
{
    if (factor_lds_bytes<T, 32>(bw) <= kFactorLdsMax) return 32;
    if (factor_lds_bytes<T, 16>(bw) <= kFactorLdsMax) return 16;
    return 0;
}

template <typename T>
int launch_factor(hipStream_t s, int nb, unsigned grid, const BandDesc *d_descs, T *fac, int n, int bw, int *info,
                  double *stats)
{
    int rc;
    if (nb == 32) {
        const size_t smem = factor_lds_bytes<T, 32>(bw);
        if ((rc = allow_factor_lds<T, 32>(smem))) return rc;
        if (d_descs)
            hipLaunchKernelGGL((k_band_ldlt_wg_batched<T, 32>), dim3(grid), dim3(kFT), smem, s, d_descs);
        else
            hipLaunchKernelGGL((k_band_ldlt_wg<T, 32>), dim3(1), dim3(kFT), smem, s, fac, n, bw, info, stats);
    } else {
        const size_t smem = factor_lds_bytes<T, 16>(bw);
        if ((rc = allow_factor_lds<T, 16>(smem))) return rc;
        if (d_descs)
            hipLaunchKernelGGL((k_band_ldlt_wg_batched<T, 16>), dim3(grid), dim3(kFT), smem, s, d_descs);
        else
            hipLaunchKernelGGL((k_band_ldlt_wg<T, 16>), dim3(1), dim3(kFT), smem, s, fac, n, bw, info, stats);
    }
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int check_factor_shape(int n, int bw, int prec, int &nb)
{
    const size_t es = prec == KSH_FP64 ? sizeof(double) : sizeof(float);
    if ((size_t)n * (size_t)(bw + 1) * es >= (size_t)kMaskedOff)
        return fail(KSH_SOLVER_ERROR, "band larger than 2 GiB: not supported by the one-workgroup factorization");
    nb = prec == KSH_FP64 ? panel_width<double>(bw) : panel_width<float>(bw);
    if (nb == 0) return fail(KSH_SOLVER_ERROR, "band too wide for the one-workgroup factorization (LDS panel exceeds 160 KB)");
    return KSH_SUCCESS;
}

}  // namespace

int band_factor(Matrix &K, void *fac, int prec, double *stats_out)
{
    Problem &pb = *K.pb;
    const size_t bytes = (size_t)K.n * K.ld * (prec == KSH_FP64 ? sizeof(double) : sizeof(float));
    int nb, rc;
    if ((rc = check_factor_shape(K.n, K.bw, prec, nb))) return rc;
    if (fac != K.ab) {
        if (prec == K.sprec) {
            KSH_HIP(hipMemcpyAsync(fac, K.ab, bytes, hipMemcpyDeviceToDevice, pb.stream));
        } else if ((rc = band_convert(K.ab, K.sprec, fac, prec, (size_t)K.n * K.ld, pb.stream))) {
            return rc;
        }
    }
    double *dstats = K.red + 4;
    {
        ProfScope ps("band_ldlt", pb.stream);
        rc = prec == KSH_FP64 ? launch_factor<double>(pb.stream, nb, 1, nullptr, (double *)fac, K.n, K.bw, K.flag, dstats)
                              : launch_factor<float>(pb.stream, nb, 1, nullptr, (float *)fac, K.n, K.bw, K.flag, dstats);
    }
    if (rc) return rc;
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
    int maxbw = 0, maxn = 0;
    std::vector<BandDesc> dd(h.size());
    for (size_t i = 0; i < h.size(); ++i) {
        dd[i] = BandDesc{h[i].fac, h[i].n, h[i].bw, h[i].info, h[i].stats, h[i].x, h[i].active, 0};
        if (!h[i].active) continue;
        maxbw = std::max(maxbw, h[i].bw);
        maxn = std::max(maxn, h[i].n);
    }
    int nb, rc;
    for (const auto &x : h)
        if (x.active && (rc = check_factor_shape(x.n, x.bw, prec, nb))) return rc;
    nb = prec == KSH_FP64 ? panel_width<double>(maxbw) : panel_width<float>(maxbw);
    KSH_HIP(hipMemcpyAsync(d_descs, dd.data(), sizeof(BandDesc) * dd.size(), hipMemcpyHostToDevice, s));
    ProfScope ps("band_ldlt_batch", s);
    return prec == KSH_FP64 ? launch_factor<double>(s, nb, (unsigned)dd.size(), (const BandDesc *)d_descs, nullptr, 0, maxbw,
                                                    nullptr, nullptr)
                            : launch_factor<float>(s, nb, (unsigned)dd.size(), (const BandDesc *)d_descs, nullptr, 0, maxbw,
                                                   nullptr, nullptr);
}

}  // namespace ksh
