// ksh_solve.hip -- forward / diagonal / backward substitution with a factored band L D L^T.
//
// Replaces SimplicialLDLT::solve on the arc-length and Newton paths
// (src/gsALMSolvers/gsALMBase.hpp:559-562, src/gsALMSolvers/gsALMCrisfield.hpp:100-115,
// src/gsStaticSolvers/gsStaticNewton.hpp:215-240).  Layout: lower band, column major,
// fac[j*ld + (i-j)], ld = bw+1, unit L below the diagonal and D on it.
//
// Main kernel (one workgroup per matrix, 1024 threads for an FP32 factor and 512 for FP64; a
// launch covers a batch):
//   * the right-hand side / solution vector stays in LDS for the whole solve, so HBM traffic is
//     the band alone, read once per sweep: 2 n (bw+1) entries per solve;
//   * rows are processed in blocks of 64 and the band entries a block needs are loaded into
//     registers one block ahead (they do not depend on the solution);
//   * a block's product with the solved part splits into "near" entries, which need the block
//     solved just before, and "far" ones, which need only older blocks.  The far sums of the next
//     block are formed while wave 0 runs this block's 64-step recurrence, so the recurrence is off
//     the memory stream; only the near products (64 x 64 entries) sit between two recurrences;
//   * forward block r0..r0+63: lanes = rows; each load instruction reads 64 consecutive entries of
//     one band column; partial sums meet in LDS and wave 0 runs the recurrence with the diagonal
//     block (staged in LDS by rows, read into registers ahead of use), broadcasting through
//     v_readlane;
//   * backward block c0..c0+63: each wave owns 64/W of the block's columns; lanes run along each
//     column's contiguous segment below the block, one wave reduction per block; wave 0 divides by
//     the pivots and runs the transposed recurrence (diagonal block staged by columns).
// Band widths up to 512; wider bands and vectors too long for LDS take the streaming kernel below
// (vector in HBM).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cstdlib>
#include <cstring>

#include "ksh_devutil.hpp"
#include "ksh_internal.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {

constexpr int kRB = 64;          // rows per block
constexpr int kSdLd = kRB + 4;   // diagonal-block row stride: 16-byte aligned rows for vector reads
constexpr int kBandMax = 512;    // widest band of the LDS-resident kernel

// Workgroup shape: 1024 threads (16 waves) for an FP32 factor, 512 for FP64 (its prefetch
// registers fit the 256 VGPRs of two waves per SIMD, not the 128 of four).  Forward: wave w takes
// NK of the 64 near columns (the previous block's) and FK far columns (offsets 65..512 before the
// block).  Backward: wave w owns NK of the block's columns, one near row chunk (the next block's
// rows) and BK far row chunks each.  FC: far registers consumed and refilled per scheduling
// group; RC: diagonal-block entries per recurrence register run.
template <typename T>
struct SolveCfg {
    static constexpr int NT = sizeof(T) == 8 ? 512 : 1024;
    static constexpr int W = NT / 64;
    static constexpr int NK = kRB / W;
    static constexpr int FK = (kBandMax - kRB) / W;
    static constexpr int BK = kBandMax / kRB - 1;
    static constexpr int DQ = kRB * kRB / NT;
    static constexpr int FC = 7;
    static constexpr int RC = sizeof(T) == 8 ? 8 : 16;
};
static_assert(SolveCfg<float>::FK % SolveCfg<float>::FC == 0 && SolveCfg<double>::FK % SolveCfg<double>::FC == 0, "far chunks");
static_assert(kRB + SolveCfg<float>::W * SolveCfg<float>::FK == kBandMax, "far cover");

// Zeros around the vector in LDS, so that the row and column indices of the band products need no
// clamping: forward columns reach back to r0 + 64 - 514 >= -512, backward rows up to
// c0 + 64 (BK + 1) + 63 <= c0 + 575.
constexpr int kXPadF = 512;
constexpr int kXPadB = 640;

template <typename T>
__host__ __device__ inline size_t solve_lds_bytes(int n)
{
    return sizeof(double) * ((size_t)kXPadF + ((n + 1) & ~1) + kXPadB + SolveCfg<T>::W * kRB) + sizeof(T) * kRB * kSdLd;
}

// Sums over the wave's 64 lanes of CW per-lane values (CW a power of two up to 64).  Reduce-
// scatter by halving: at offset 32 the lanes with that bit clear keep the lower half of the
// values and receive their partner's, the others the upper half; then offset 16 on the
// remaining half, and so on down to one value, which a butterfly over the remaining offsets
// completes.  Lane l returns the total of value l / (64 / CW), after CW - 1 + 6 - log2(CW)
// exchanges instead of 6 CW.
template <int CW>
__device__ __forceinline__ double wave_sums(const double (&a)[CW], int lane)
{
    double v[CW];
#pragma unroll
    for (int q = 0; q < CW; ++q) v[q] = a[q];
#pragma unroll
    for (int lvl = 0; (CW >> lvl) > 1; ++lvl) {
        const int off = 32 >> lvl, half = (CW >> lvl) / 2;
        const bool up = (lane & off) != 0;
#pragma unroll
        for (int q = 0; q < half; ++q) {
            const double send = up ? v[q] : v[q + half];
            const double keep = up ? v[q + half] : v[q];
            v[q] = keep + __shfl_xor(send, off);
        }
    }
#pragma unroll
    for (int off = 32 / CW; off > 0; off >>= 1) v[0] += __shfl_xor(v[0], off);
    return v[0];
}

typedef float ksh_f32x4 __attribute__((ext_vector_type(4)));
typedef double ksh_f64x2 __attribute__((ext_vector_type(2)));

// RC consecutive diagonal-block entries from LDS (16-byte aligned) into registers, 16 bytes a read
template <int RC, typename T>
__device__ __forceinline__ void lds_read_run(T (&dst)[RC], const T *src)
{
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int q = 0; q < RC / 2; ++q) {
            const ksh_f64x2 v = reinterpret_cast<const ksh_f64x2 *>(src)[q];
            dst[2 * q] = v.x;
            dst[2 * q + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < RC / 4; ++q) {
            const ksh_f32x4 v = reinterpret_cast<const ksh_f32x4 *>(src)[q];
            dst[4 * q] = v.x;
            dst[4 * q + 1] = v.y;
            dst[4 * q + 2] = v.z;
            dst[4 * q + 3] = v.w;
        }
    }
}

// G band loads of 64 consecutive entries: load u reads elements base + u db + lane, and lanes above
// lim_u = min(e + u de, rl, jl + u djl) read zero (an out-of-range offset).  The scalar state is
// advanced by one add per term and load, so a load costs about six scalar instructions; a load
// whose column lies before the matrix (base < 0) has lim < 0 and byte offset 0.
template <int G, typename T>
__device__ __forceinline__ void band_run(T *dst, __amdgpu_buffer_rsrc_t rs, int lane, unsigned lane_off, int base, int db, int e, int de,
                                         int rl, int jl, int djl)
{
    int sb = base * (int)sizeof(T);
    const int dsb = db * (int)sizeof(T);
#pragma unroll
    for (int u = 0; u < G; ++u) {
        // (two scalar mins: opaque_s keeps the compiler from fusing them into a vector min3)
        const int lim = min(opaque_s(min(e, rl)), jl);
        dst[u] = load_vs<T>(rs, lane <= lim ? lane_off : kMaskedOff, (unsigned)max(sb, 0));
        sb += dsb;
        e += de;
        jl += djl;
    }
}

template <typename T>
__device__ __forceinline__ void band_solve_lds(const T *__restrict__ fac, int n, int bw, double *__restrict__ xg, unsigned char *smem)
{
    using C = SolveCfg<T>;
    constexpr int NT = C::NT, W = C::W, NK = C::NK, FK = C::FK, BK = C::BK, DQ = C::DQ, FC = C::FC, RC = C::RC;
    const unsigned ldu = (unsigned)bw + 1;  // n * ld * sizeof(T) < 2 GiB (checked by the launcher)
    const int ld = bw + 1;
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * sizeof(T)));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const unsigned lane_off = (unsigned)lane * (unsigned)sizeof(T);
    double *xs = (double *)smem + kXPadF;
    double *sp = xs + ((n + 1) & ~1) + kXPadB;  // [W][64] partial sums
    T *sd = (T *)(sp + W * kRB);                // diagonal block, 64 runs of kSdLd
    const int nblk = (n + kRB - 1) / kRB;
    T dg[DQ];

    // diagonal block of rows/columns r0..r0+63 (strict lower part, or with the pivots): thread entry
    // (ii, jj) = L[r0+ii][r0+jj], stored by rows (sd[ii][jj], the forward recurrence's operands of
    // lane ii) or by columns (sd[jj][ii], the backward one's)
    auto diag_load = [&](int r0, bool with_d) {
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int idx = tid + NT * q;
            const int jj = idx >> 6, ii = idx & 63;
            const bool ok = r0 >= 0 && r0 + ii < n && (with_d ? ii >= jj : ii > jj) && ii - jj <= bw;
            dg[q] = load_if<T>(ok, rs, (unsigned)(((r0 + jj) * ldu + (ii - jj)) * sizeof(T)));
        }
    };
    auto diag_store = [&](bool by_rows) {
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int idx = tid + NT * q;
            const int jj = idx >> 6, ii = idx & 63;
            sd[by_rows ? ii * kSdLd + jj : jj * kSdLd + ii] = dg[q];
        }
    };

    for (int i = tid; i < n; i += NT) xs[i] = xg[i];
    for (int i = tid; i < kXPadF; i += NT) xs[i - kXPadF] = 0.0;
    for (int i = n + tid; i < ((n + 1) & ~1) + kXPadB; i += NT) xs[i] = 0.0;

    // ---- forward: L y = b ----
    // Entry (r0 + lane, j), j = r0 - off, is element j (bw + 1) + r0 + lane - j = j bw + r0 + lane;
    // lanes with lane + off > bw (outside the band) or r0 + lane >= n read zero.
    {
        // G loads at offsets off0, off0 + W, ...: column j = r0 - off, lanes up to bw - off (the
        // band) and n - 1 - r0 (the matrix); j < 0 gives jl = 64 j + 63 < 0
        auto fwd_run = [&](T *dst, auto g, int r0, int off0) {
            constexpr int G = decltype(g)::value;
            const int j0 = r0 - off0;
            band_run<G, T>(dst, rs, lane, lane_off, j0 * bw + r0, -W * bw, bw - off0, -W, n - 1 - r0, 64 * j0 + 63, -64 * W);
        };
        // near offsets 1 + w + W k (k < NK) cover 1..64, far offsets 65 + w + W k (k < FK)
        // 65..512.  Blocks 0 and 1 have no entries there (the loads read zeros); they are issued
        // anyway, in the loop's order (near, diagonal block, far), so that the waits the compiler
        // derives for the loop, which also hold on its entry, are the steady-state ones.
        T fn[NK], ff[FK];
        fwd_run(fn, std::integral_constant<int, NK>{}, 0, 1 + wv);
        __builtin_amdgcn_sched_barrier(0);
        diag_load(0, false);
        __builtin_amdgcn_sched_barrier(0);
        fwd_run(ff, std::integral_constant<int, FK>{}, kRB, kRB + 1 + wv);
        double facc = 0.0;  // far sum of the current block
        lds_barrier();
        for (int b = 0; b < nblk; ++b) {
            const int r0 = b * kRB, i = r0 + lane;
            double acc = facc;
#pragma unroll
            for (int k = 0; k < NK; ++k) acc += (double)fn[k] * xs[r0 - 1 - wv - W * k];
            sp[wave * kRB + lane] = acc;
            fwd_run(fn, std::integral_constant<int, NK>{}, r0 + kRB, 1 + wv);
            diag_store(true);
            diag_load(r0 + kRB, false);
            lds_barrier();
            if (wv == 0) {
                double s = 0.0;
#pragma unroll
                for (int w = 0; w < W; ++w) s += sp[w * kRB + lane];
                double y = i < n ? xs[i] - s : 0.0;
                // y_i -= L[i][jj] y_jj for jj = 0..62; L[i][jj] = 0 for jj >= i in the staged block
                const T *row = sd + lane * kSdLd;
                T lr[2][RC];
                lds_read_run<RC>(lr[0], row);
#pragma unroll
                for (int q = 0; q < kRB / RC; ++q) {
                    if (q + 1 < kRB / RC) lds_read_run<RC>(lr[(q + 1) & 1], row + RC * (q + 1));
#pragma unroll
                    for (int t = 0; t < RC; ++t) {
                        const int jj = RC * q + t;
                        if (jj < kRB - 1) y -= (double)lr[q & 1][t] * readlane_f64(y, jj);
                    }
                }
                if (i < n) xs[i] = y;
            }
            // far sum of block b + 1: columns r0 - 1 - w - W k, all solved (wave 0 after its
            // recurrence)
            const double *xw = xs + (r0 - 1 - wv);
            double a = 0.0;
#pragma unroll
            for (int c = 0; c < FK / FC; ++c) {
#pragma unroll
                for (int u = 0; u < FC; ++u) {
                    const int k = FC * c + u;
                    a += (double)ff[k] * xw[-W * k];
                }
                fwd_run(ff + FC * c, std::integral_constant<int, FC>{}, r0 + 2 * kRB, kRB + 1 + wv + W * FC * c);
                __builtin_amdgcn_sched_barrier(0);
            }
            facc = a;
            lds_barrier();
        }
    }

    // ---- diagonal and backward: L^T x = D^-1 y ----
    // Wave w owns columns j = c0 + jj, jj = NK w + c.  Row chunk k of a column holds rows
    // c0 + 64 (k + 1) + lane, element j (bw + 1) + dl + lane with dl = 64 - jj + 64 k; lanes past
    // the band (dl + lane > bw) or the matrix read zero.  Chunk 0 is near (the next block's rows),
    // chunks 1..BK far; the far sums of block b - 1 are formed while wave 0 solves block b.
    {
        // near chunk of the wave's NK columns jj = NK w + u: element (c0 + jj) ld + 64 - jj + lane,
        // lanes up to bw - 64 + jj (the band) and n - 1 - c0 - 64 (the matrix)
        auto bwd_near = [&](T (&dst)[NK], int c0) {
            const int jj0 = NK * wv;
            band_run<NK, T>(dst, rs, lane, lane_off, (c0 + jj0) * ld + kRB - jj0, bw, bw - kRB + jj0, 1, n - 1 - c0 - kRB,
                            c0 >= 0 ? 63 : -1, 0);
        };
        // far chunks 1..BK of column j = c0 + jj: element j ld + dl + lane, dl = 64 (k + 1) - jj, lanes
        // up to min(bw, n - 1 - j) - dl
        auto bwd_far = [&](T (&dst)[BK], int c0, int jj) {
            const int j = c0 + jj, dl = 2 * kRB - jj;
            band_run<BK, T>(dst, rs, lane, lane_off, j * ld + dl, kRB, min(bw, n - 1 - j) - dl, -kRB, 1 << 30, c0 >= 0 ? 63 : -1, 0);
        };
        // the last block has no rows below it, the one before none beyond its near chunk: their
        // loads read zeros, issued in the loop's order as in the forward sweep
        const int cl = (nblk - 1) * kRB;
        T bn[NK], bf[NK][BK];
        double bacc[NK];
        bwd_near(bn, cl);
#pragma unroll
        for (int c = 0; c < NK; ++c) bacc[c] = 0.0;
        __builtin_amdgcn_sched_barrier(0);
        diag_load(cl, true);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < NK; ++c) bwd_far(bf[c], cl - kRB, NK * wv + c);
        for (int b = nblk - 1; b >= 0; --b) {
            const int c0 = b * kRB, i = c0 + lane;
            double a[NK];
#pragma unroll
            for (int c = 0; c < NK; ++c) a[c] = bacc[c] + (double)bn[c] * xs[c0 + kRB + lane];
            const double s = wave_sums<NK>(a, lane);
            if ((lane & (64 / NK - 1)) == 0) sp[NK * wave + lane / (64 / NK)] = s;
            bwd_near(bn, c0 - kRB);
            diag_store(false);
            diag_load(c0 - kRB, true);
            lds_barrier();
            if (wv == 0) {
                T *col = sd + lane * kSdLd;  // col[jj] = L[c0+jj][c0+lane]
                const double dv = (double)col[lane];
                col[lane] = T(0);  // only this lane reads its pivot
                double xv = i < n ? xs[i] / dv - sp[lane] : 0.0;
                // x_i -= L[jj][i] x_jj for jj = 63..1; zeros above the diagonal and the cleared pivot
                T lr[2][RC];
                lds_read_run<RC>(lr[0], col + kRB - RC);
#pragma unroll
                for (int q = 0; q < kRB / RC; ++q) {
                    const int q0 = kRB - RC * (q + 1);  // first entry of this run
                    if (q + 1 < kRB / RC) lds_read_run<RC>(lr[(q + 1) & 1], col + q0 - RC);
#pragma unroll
                    for (int t = RC - 1; t >= 0; --t) {
                        const int jj = q0 + t;
                        if (jj > 0) xv -= (double)lr[q & 1][t] * readlane_f64(xv, jj);
                    }
                }
                if (i < n) xs[i] = xv;
            }
            // far sums of block b - 1: rows c0 + 64 k + lane, k >= 1, all solved
#pragma unroll
            for (int c = 0; c < NK; ++c) {
                double f = 0.0;
#pragma unroll
                for (int k = 0; k < BK; ++k) f += (double)bf[c][k] * xs[c0 + kRB * (k + 1) + lane];
                bacc[c] = f;
                bwd_far(bf[c], c0 - 2 * kRB, NK * wv + c);
                __builtin_amdgcn_sched_barrier(0);
            }
            lds_barrier();
        }
    }
    for (int i = tid; i < n; i += NT) xg[i] = xs[i];
}

template <typename T>
__global__ __launch_bounds__(SolveCfg<T>::NT) void k_band_solve_lds(const T *fac, int n, int bw, double *x)
{
    extern __shared__ __align__(16) unsigned char smem[];
    band_solve_lds<T>(fac, n, bw, x, smem);
}

template <typename T>
__global__ __launch_bounds__(SolveCfg<T>::NT) void k_band_solve_lds_batched(const BandDesc *descs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BandDesc &d = descs[blockIdx.x];
    if (!d.active) return;
    band_solve_lds<T>((const T *)d.fac, d.n, d.bw, d.x, smem);
}

// ---- streaming kernel (vector in HBM; any n) ----
// Same block structure with 256 threads: forward blocks accumulate the left-looking product with
// 16 loads per thread in flight, the diagonal block goes through LDS into registers and wave 0
// runs the recurrence; backward blocks reduce 16 columns per wave.
constexpr int kThreads = 256;
constexpr int SB = 64;
constexpr int SU = 16;  // loads in flight per thread in the forward product


template <typename T>
__device__ __forceinline__ void band_solve_body(const T *__restrict__ fac, int n, int bw, double *__restrict__ x)
{
    __shared__ double sdiag[SB][SB + 1];  // sdiag[jj][ii] = L[r0+ii][r0+jj]
    __shared__ double spart[4][SB];
    const size_t ld = (size_t)bw + 1;
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    // ---- forward: L y = b ----
    for (int r0 = 0; r0 < n; r0 += SB) {
        const int w = min(SB, n - r0);
        for (int idx = tid; idx < SB * SB; idx += kThreads) {
            int jj = idx / SB, ii = idx % SB;
            double v = 0.0;
            if (jj < w && ii < w && ii > jj && ii - jj <= bw) v = (double)fac[(size_t)(r0 + jj) * ld + (ii - jj)];
            sdiag[jj][ii] = v;
        }
        {
            const int i = r0 + lane;
            const int jlo = max(0, i - bw);
            double acc = 0.0;
            const int jbeg = max(0, r0 - bw) + wave;
            for (int base = jbeg; base < r0; base += 4 * SU) {
                double lv[SU], xv[SU];
#pragma unroll
                for (int u = 0; u < SU; ++u) {
                    int j = base + 4 * u;
                    bool ok = lane < w && j < r0 && j >= jlo;
                    lv[u] = ok ? (double)fac[(size_t)j * ld + (i - j)] : 0.0;
                    xv[u] = (j < r0) ? x[j] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < SU; ++u) acc += lv[u] * xv[u];
            }
            spart[wave][lane] = acc;
        }
        __syncthreads();
        if (wave == 0) {
            double row[SB];
#pragma unroll
            for (int jj = 0; jj < SB; ++jj) row[jj] = sdiag[jj][lane];
            double y = 0.0;
            if (lane < w) y = x[r0 + lane] - (spart[0][lane] + spart[1][lane] + spart[2][lane] + spart[3][lane]);
#pragma unroll
            for (int jj = 0; jj < SB; ++jj) {
                double yj = readlane_f64(y, jj);
                if (lane > jj) y -= row[jj] * yj;
            }
            if (lane < w) x[r0 + lane] = y;
        }
        __syncthreads();
    }
    // ---- diagonal ----
    for (int i = tid; i < n; i += kThreads) x[i] /= (double)fac[(size_t)i * ld];
    __syncthreads();
    // ---- backward: L^T x = y ----
    const int nblk = (n + SB - 1) / SB;
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int c0 = bk * SB, w = min(SB, n - c0);
        for (int idx = tid; idx < SB * SB; idx += kThreads) {
            int jj = idx / SB, ii = idx % SB;
            double v = 0.0;
            if (jj < w && ii < w && ii > jj && ii - jj <= bw) v = (double)fac[(size_t)(c0 + jj) * ld + (ii - jj)];
            sdiag[jj][ii] = v;
        }
        double acc[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] = 0.0;
        const int ibase = c0 + w;
        const int iend = min(n, c0 + w + bw);
        for (int i = ibase + lane; i < iend; i += 64) {
            double xi = x[i];
            double lv[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                int jj = 16 * wave + c;
                int j = c0 + jj;
                lv[c] = (jj < w && i - j <= bw) ? (double)fac[(size_t)j * ld + (i - j)] : 0.0;
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) acc[c] += lv[c] * xi;
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            double v = acc[c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) spart[0][16 * wave + c] = v;
        }
        __syncthreads();
        if (wave == 0) {
            double col[SB];  // col[jj] = L[c0+jj][c0+lane]
#pragma unroll
            for (int jj = 0; jj < SB; ++jj) col[jj] = sdiag[lane][jj];
            double xv = (lane < w) ? x[c0 + lane] - spart[0][lane] : 0.0;
#pragma unroll
            for (int jj = SB - 1; jj >= 0; --jj) {
                double xj = readlane_f64(xv, jj);
                if (lane < jj) xv -= col[jj] * xj;
            }
            if (lane < w) x[c0 + lane] = xv;
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_band_solve_stream(const T *fac, int n, int bw, double *x)
{
    band_solve_body<T>(fac, n, bw, x);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_band_solve_stream_batched(const BandDesc *descs)
{
    const BandDesc &d = descs[blockIdx.x];
    if (!d.active) return;
    band_solve_body<T>((const T *)d.fac, d.n, d.bw, d.x);
}

constexpr size_t kSolveLdsMax = 160 * 1024;

int allow_solve_lds(size_t need)
{
    static size_t granted = 0;
    if (need <= granted) return KSH_SUCCESS;
    const void *fns[] = {(const void *)k_band_solve_lds<double>, (const void *)k_band_solve_lds<float>,
                         (const void *)k_band_solve_lds_batched<double>, (const void *)k_band_solve_lds_batched<float>};
    for (const void *f : fns) KSH_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    granted = need;
    return KSH_SUCCESS;
}

// LDS bytes of the LDS-resident kernel for a batch with largest order nmax and band width bwmax,
// or 0 when it does not apply: the vector must fit in LDS, the band width must not exceed the
// columns the block prefetches cover (512 for both precisions), and band byte offsets must stay below
// the 2 GiB masked-read offset.
size_t solve_lds(int nmax, int bwmax, size_t bytes_max, int prec)
{
    const size_t smem = prec == KSH_FP64 ? solve_lds_bytes<double>(nmax) : solve_lds_bytes<float>(nmax);
    if (smem > kSolveLdsMax || bwmax > kBandMax || bytes_max >= 0x7F000000ull) return 0;
    return smem;
}

}  // namespace

static_assert(sizeof(BandDesc) == kBandDescBytes, "BandDesc layout");

int band_solve_batch(hipStream_t s, void *d_descs, const std::vector<BandDescHost> &h, int prec)
{
    if (h.empty()) return KSH_SUCCESS;
    std::vector<BandDesc> dd(h.size());
    int nmax = 0, bwmax = 0;
    size_t bytes_max = 0;
    const size_t es = prec == KSH_FP64 ? sizeof(double) : sizeof(float);
    for (size_t i = 0; i < h.size(); ++i) {
        dd[i] = BandDesc{h[i].fac, h[i].n, h[i].bw, h[i].info, h[i].stats, h[i].x, h[i].active, 0};
        if (!h[i].active) continue;
        nmax = std::max(nmax, h[i].n);
        bwmax = std::max(bwmax, h[i].bw);
        bytes_max = std::max(bytes_max, (size_t)h[i].n * (h[i].bw + 1) * es);
    }
    // Largest bands first: a workgroup's time is about its band's bytes (one CU streams them), and
    // workgroups are dispatched in index order as CUs free up, so with more bands than CUs
    // (two right-hand sides per matrix) the long ones start at once and the short ones fill in
    // behind them (longest-processing-time-first), instead of a second round of long ones.
    std::stable_sort(dd.begin(), dd.end(), [](const BandDesc &a, const BandDesc &b) {
        const size_t ca = a.active ? (size_t)a.n * (size_t)(a.bw + 1) : 0, cb = b.active ? (size_t)b.n * (size_t)(b.bw + 1) : 0;
        return ca > cb;
    });
    KSH_HIP(hipMemcpyAsync(d_descs, dd.data(), sizeof(BandDesc) * dd.size(), hipMemcpyHostToDevice, s));
    const size_t smem = solve_lds(nmax, bwmax, bytes_max, prec);
    ProfScope ps(prec == KSH_FP64 ? "band_solve_batch_fp64" : "band_solve_batch_fp32", s);
    {
        double work = 0.0;  // band read twice, vector in and out
        for (const auto &x : h)
            if (x.active) work += 2.0 * x.n * (x.bw + 1.0) * es + 16.0 * x.n;
        ps.set_work(work);
    }
    if (smem) {
        int rc = allow_solve_lds(smem);
        if (rc) return rc;
        if (prec == KSH_FP64)
            hipLaunchKernelGGL(k_band_solve_lds_batched<double>, dim3((unsigned)dd.size()), dim3(SolveCfg<double>::NT), smem, s,
                               (const BandDesc *)d_descs);
        else
            hipLaunchKernelGGL(k_band_solve_lds_batched<float>, dim3((unsigned)dd.size()), dim3(SolveCfg<float>::NT), smem, s,
                               (const BandDesc *)d_descs);
    } else if (prec == KSH_FP64) {
        hipLaunchKernelGGL(k_band_solve_stream_batched<double>, dim3((unsigned)dd.size()), dim3(kThreads), 0, s,
                           (const BandDesc *)d_descs);
    } else {
        hipLaunchKernelGGL(k_band_solve_stream_batched<float>, dim3((unsigned)dd.size()), dim3(kThreads), 0, s,
                           (const BandDesc *)d_descs);
    }
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

// KSH_BAND_PATH=tiled / wg forces the multi-workgroup / one-workgroup path.  Otherwise bands
// that the LDS-resident kernel takes stay on one workgroup; wider ones go multi-workgroup when
// a step launch (~4 us for 64 rows) costs less than one CU streaming those rows' band entries.
bool band_use_tiled_solve(int n, int bw, int prec)
{
    const char *e = std::getenv("KSH_BAND_PATH");
    if (e && std::strcmp(e, "tiled") == 0) return true;
    if (e && std::strcmp(e, "wg") == 0) return false;
    const size_t es = prec == KSH_FP64 ? sizeof(double) : sizeof(float);
    if (solve_lds(n, bw, (size_t)n * (bw + 1) * es, prec)) return false;
    return (double)(bw + 1) * es >= 3072.0;
}

int band_solve(Matrix &K, const void *fac, int prec, double *x_dev)
{
    Problem &pb = *K.pb;
    if (band_use_tiled_solve(K.n, K.bw, prec)) {
        ProfScope ps("band_solve_tiled", pb.stream);
        ps.set_work(2.0 * K.n * (K.bw + 1.0) * (prec == KSH_FP64 ? 8.0 : 4.0) + 16.0 * K.n);
        return band_solve_tiled(pb.stream, fac, prec, K.n, K.bw, x_dev, K.work + 4 * (size_t)K.n,
                                reinterpret_cast<unsigned *>(K.flag + 1));
    }
    const size_t smem = solve_lds(K.n, K.bw, (size_t)K.n * K.ld * (prec == KSH_FP64 ? sizeof(double) : sizeof(float)), prec);
    ProfScope ps("band_solve", pb.stream);
    ps.set_work(2.0 * K.n * (K.bw + 1.0) * (prec == KSH_FP64 ? 8.0 : 4.0) + 16.0 * K.n);
    if (smem) {
        int rc = allow_solve_lds(smem);
        if (rc) return rc;
        if (prec == KSH_FP64)
            hipLaunchKernelGGL(k_band_solve_lds<double>, dim3(1), dim3(SolveCfg<double>::NT), smem, pb.stream, (const double *)fac, K.n,
                               K.bw, x_dev);
        else
            hipLaunchKernelGGL(k_band_solve_lds<float>, dim3(1), dim3(SolveCfg<float>::NT), smem, pb.stream, (const float *)fac, K.n,
                               K.bw, x_dev);
    } else if (prec == KSH_FP64) {
        hipLaunchKernelGGL(k_band_solve_stream<double>, dim3(1), dim3(kThreads), 0, pb.stream, (const double *)fac, K.n, K.bw,
                           x_dev);
    } else {
        hipLaunchKernelGGL(k_band_solve_stream<float>, dim3(1), dim3(kThreads), 0, pb.stream, (const float *)fac, K.n, K.bw, x_dev);
    }
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

}  // namespace ksh
