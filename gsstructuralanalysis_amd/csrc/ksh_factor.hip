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
#include <cstdlib>
#include <cstring>
#include <string>

#include "ksh_devutil.hpp"
#include "ksh_internal.hpp"
#include "ksh_mfma.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {

constexpr int kFT = 1024;            // threads per workgroup
constexpr int kFW = kFT / 64;        // waves

// panel rows in LDS: the frame (nb + bw rows) rounded up, plus the rows a 32-row block past the
// frame edge reads (their products only reach masked-off outputs)
__host__ __device__ inline int fpanel_rows(int bw, int nb) { return ((bw + nb + 15) / 16) * 16 + 32; }

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
    // per-lane part of a trailing tile's band offsets: entry (ri0+li, rj0+row(lane, v)) of the
    // frame at column c0 sits at (c0 ld + rj0 (ld-1) + ri0) + (row (ld-1) + li)
    unsigned lane_off[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) lane_off[v] = ((unsigned)M::row(lane, v) * (ldu - 1) + (unsigned)(lane & 15)) * (unsigned)sizeof(T);
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
        // The lower triangle of the window is cut into 32x32 blocks (bi, bj), bj <= bi, each a 2x2
        // set of 16x16 MFMA tiles sharing operands: A (-d_k L[j][k], rows j of block column bj)
        // and B (L[i][k], rows i of block row bi) come from the LDS panel.  Blocks are dealt to
        // the waves round-robin; a wave loads the A-matrix entries of its next block while it
        // computes the current one.  Tile addresses are a per-lane part fixed for the whole
        // factorization plus a wave-uniform part (SGPR); masked lanes and tiles (edges, the
        // upper tile of a diagonal block, blocks past the wave's last) get kMaskedOff, so the
        // loop is straight-line code.
        {
            const int nbk = (nr - w + 31) / 32;
            const int nblock = nbk * (nbk + 1) / 2;
            const int wv = __builtin_amdgcn_readfirstlane(wave);
            const int li = lane & 15, lk = lane >> 4;
            constexpr unsigned es = sizeof(T);
            auto decode = [&](int p, int &bi, int &bj) {
                int b = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
                if ((b + 1) * (b + 2) / 2 <= p) ++b;
                if (b * (b + 1) / 2 > p) --b;
                bi = b;
                bj = p - b * (b + 1) / 2;
            };
            // accumulator rows = j (A's column), columns = i (A's row): lanes run along a column
            auto tile_off = [&](int p, int bi, int bj, int a, int b, int v, unsigned &soff) -> unsigned {
                const int ri0 = w + 32 * bi + 16 * a, rj0 = w + 32 * bj + 16 * b;
                soff = ((unsigned)c0 * ldu + (unsigned)rj0 * (ldu - 1) + (unsigned)ri0) * es;
                const bool edge = bi == bj || ri0 + 16 > nr;  // wave-uniform
                const int i = ri0 + li, j = rj0 + M::row(lane, v);
                const bool ok = p < nblock && (!edge || (i < nr && i >= j));
                return ok ? lane_off[v] : kMaskedOff;
            };
            auto block_load = [&](int p, acc_t (&x)[2][2]) {
                int bi, bj;
                decode(p, bi, bj);
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            unsigned so;
                            const unsigned vo = tile_off(p, bi, bj, a, b, v, so);
                            x[a][b][v] = load_vs<T>(rs, vo, so);
                        }
            };
            acc_t nxt[2][2];
            block_load(wv, nxt);
            for (int p = wv; p < nblock; p += kFW) {
                acc_t cur[2][2];
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) cur[a][b] = nxt[a][b];
                block_load(p + kFW, nxt);  // past the last block every read is masked off
                int bi, bj;
                decode(p, bi, bj);
                const int ri0 = w + 32 * bi + li, rj0 = w + 32 * bj + li;
                const bool diag = bi == bj;  // its tile (a=0, b=1) lies above the diagonal
#pragma unroll
                for (int s = 0; s < NB / 4; ++s) {
                    const int k = 4 * s + lk;
                    const T nd = -dv[k];
                    const T a0 = nd * pan[k * PR + rj0], a1 = nd * pan[k * PR + rj0 + 16];
                    const T b0 = pan[k * PR + ri0], b1 = pan[k * PR + ri0 + 16];
                    cur[0][0] = M::run(a0, b0, cur[0][0]);
                    cur[1][0] = M::run(a0, b1, cur[1][0]);
                    cur[1][1] = M::run(a1, b1, cur[1][1]);
                    if (!diag) cur[0][1] = M::run(a1, b0, cur[0][1]);
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            unsigned so;
                            const unsigned vo = tile_off(p, bi, bj, a, b, v, so);
                            store_vs<T>(rs, vo, so, cur[a][b][v]);
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
    if ((size_t)n * (size_t)(bw + 1) * es >= ((size_t)1 << 30))
        return fail(KSH_SOLVER_ERROR, "band larger than 1 GiB: not supported by the one-workgroup factorization");
    nb = prec == KSH_FP64 ? panel_width<double>(bw) : panel_width<float>(bw);
    if (nb == 0) return fail(KSH_SOLVER_ERROR, "band too wide for the one-workgroup factorization (LDS panel exceeds 160 KB)");
    return KSH_SUCCESS;
}

}  // namespace

// Default: the multi-workgroup factor (ksh_factor_tiled.hip) for single bands and batches alike.
// KSH_BAND_PATH=wg selects this file's one-workgroup-per-band factor where the bands' panels fit
// LDS (it computes the same factorization with another summation order).
bool band_use_tiled_factor(int n, int bw, int prec)
{
    (void)n;
    const char *e = std::getenv("KSH_BAND_PATH");
    if (!e || std::strcmp(e, "wg") != 0) return true;
    return (prec == KSH_FP64 ? panel_width<double>(bw) : panel_width<float>(bw)) == 0;
}

int band_factor(Matrix &K, void *fac, int prec, double *stats_out)
{
    Problem &pb = *K.pb;
    const size_t bytes = (size_t)K.n * K.ld * (prec == KSH_FP64 ? sizeof(double) : sizeof(float));
    const bool tiled = band_use_tiled_factor(K.n, K.bw, prec);
    int nb = 0, rc;
    if (!tiled && (rc = check_factor_shape(K.n, K.bw, prec, nb))) return rc;
    if (fac != K.ab) {
        if (prec == K.sprec) {
            KSH_HIP(hipMemcpyAsync(fac, K.ab, bytes, hipMemcpyDeviceToDevice, pb.stream));
        } else if ((rc = band_convert(K.ab, K.sprec, fac, prec, (size_t)K.n * K.ld, pb.stream))) {
            return rc;
        }
    }
    double *dstats = K.red + 4;
    if (tiled) {
        ProfScope ps("band_ldlt_tiled", pb.stream);
        ps.set_work((double)K.n * ((double)K.bw * K.bw + 3.0 * K.bw));
        rc = band_factor_tiled_batch(pb.stream, K.desc, {BandDescHost{fac, K.n, K.bw, K.flag, dstats, nullptr, 1}}, prec);
    } else {
        ProfScope ps("band_ldlt", pb.stream);
        ps.set_work((double)K.n * ((double)K.bw * K.bw + 3.0 * K.bw));
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

int band_factor_batch(hipStream_t s, void *d_descs, const std::vector<BandDescHost> &h, int prec, const char *prof_name)
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
    ProfScope ps(prof_name ? prof_name : prec == KSH_FP64 ? "band_ldlt_batch_fp64" : "band_ldlt_batch_fp32", s);
    {
        double work = 0.0;  // n (b^2 + 3b) flop per matrix
        for (const auto &x : h)
            if (x.active) work += (double)x.n * ((double)x.bw * x.bw + 3.0 * x.bw);
        ps.set_work(work);
    }
    if (band_use_tiled_factor(maxn, maxbw, prec)) return band_factor_tiled_batch(s, d_descs, h, prec);
    int nb, rc;
    for (const auto &x : h)
        if (x.active && (rc = check_factor_shape(x.n, x.bw, prec, nb))) return rc;
    nb = prec == KSH_FP64 ? panel_width<double>(maxbw) : panel_width<float>(maxbw);
    KSH_HIP(hipMemcpyAsync(d_descs, dd.data(), sizeof(BandDesc) * dd.size(), hipMemcpyHostToDevice, s));
    return prec == KSH_FP64 ? launch_factor<double>(s, nb, (unsigned)dd.size(), (const BandDesc *)d_descs, nullptr, 0, maxbw,
                                                    nullptr, nullptr)
                            : launch_factor<float>(s, nb, (unsigned)dd.size(), (const BandDesc *)d_descs, nullptr, 0, maxbw,
                                                   nullptr, nullptr);
}

}  // namespace ksh
