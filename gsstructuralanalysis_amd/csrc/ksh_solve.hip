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
//   * rows are processed in blocks of 64, and the band entries a block needs are loaded into
//     registers one block ahead (they do not depend on the solution), so their latency overlaps
//     the reduction and the recurrence of the current block;
//   * forward block r0..r0+63: lanes = rows, wave w of W takes band columns r0-bw+w+Wk; each load
//     instruction reads 64 consecutive entries of one band column; partial products are summed
//     over the waves in LDS, then wave 0 runs the 64-step recurrence with the diagonal block
//     (staged in LDS), broadcasting through v_readlane;
//   * backward block c0..c0+63: each wave owns 64/W consecutive columns; lanes run along each
//     column's contiguous segment below the block, one wave reduction per column; wave 0 divides
//     by the pivots and runs the transposed recurrence.
// Band widths up to 512 (the prefetched columns per block); wider bands and vectors too long for
// LDS take the streaming kernel below (vector in HBM).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "ksh_devutil.hpp"
#include "ksh_internal.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {

constexpr int kRB = 64;       // rows per block
constexpr int kSdLd = kRB + 1;

// Workgroup shape per factor precision: NT threads, PF band entries per thread prefetched for
// the next block (FP32: 1024 x 32, FP64: 512 x 64; both fit 128 / 256 VGPRs without spills and
// prefetch 512 band columns per block).
template <typename T>
struct SolveCfg {
    static constexpr int NT = sizeof(T) == 8 ? 512 : 1024;
    static constexpr int W = NT / 64;          // waves
    static constexpr int PF = 32 * 1024 / NT;  // prefetched entries per thread
    static constexpr int CW = kRB / W;         // backward columns per wave
    static constexpr int PC = PF / CW;         // prefetched backward loads per column
    static constexpr int DQ = kRB * kRB / NT;  // diagonal-block entries per thread
};

// Zeros around the vector in LDS, so that the row and column indices of the band products need no
// clamping: forward columns reach back to r0 - bw >= -512, backward rows up to c0 + 64 + 63 +
// 64 (PC - 1) <= c0 + 575.
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

template <typename T>
__device__ __forceinline__ void band_solve_lds(const T *__restrict__ fac, int n, int bw, double *__restrict__ xg, unsigned char *smem)
{
    const size_t ld = (size_t)bw + 1;
    const unsigned ldu = (unsigned)bw + 1;  // n * ld * sizeof(T) < 4 GiB (checked by the launcher)
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ld * sizeof(T)));
    constexpr int NT = SolveCfg<T>::NT, kWaves = SolveCfg<T>::W, kPF = SolveCfg<T>::PF;
    constexpr int CW = SolveCfg<T>::CW, DQ = SolveCfg<T>::DQ;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    double *xs = (double *)smem + kXPadF;
    double *sp = xs + ((n + 1) & ~1) + kXPadB;  // [kWaves][64] partial sums
    T *sd = (T *)(sp + kWaves * kRB);        // [64][65]: sd[jj*65 + ii] = L[r0+ii][r0+jj]
    const int nblk = (n + kRB - 1) / kRB;
    T pf[kPF];
    T dg[DQ];

    // diagonal block of rows/columns r0..r0+63 (strict lower part, or with the pivots)
    auto diag_load = [&](int r0, bool with_d) {
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int idx = tid + NT * q;
            const int jj = idx >> 6, ii = idx & 63;
            const bool ok = r0 >= 0 && r0 + ii < n && (with_d ? ii >= jj : ii > jj) && ii - jj <= bw;
            dg[q] = load_if<T>(ok, rs, (unsigned)(((r0 + jj) * ldu + (ii - jj)) * sizeof(T)));
        }
    };
    auto diag_store = [&]() {
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int idx = tid + NT * q;
            sd[(idx >> 6) * kSdLd + (idx & 63)] = dg[q];
        }
    };

    for (int i = tid; i < n; i += NT) xs[i] = xg[i];
    for (int i = tid; i < kXPadF; i += NT) xs[i - kXPadF] = 0.0;
    for (int i = n + tid; i < ((n + 1) & ~1) + kXPadB; i += NT) xs[i] = 0.0;

    // ---- forward: L y = b ----
    // Entry (i, j), i = r0 + lane, j = r0 - bw + wv + W k (one band column per wave and k) is at
    // element j (bw + 1) + i - j = j bw + r0 + lane: a per-lane part plus a wave-uniform part,
    // which the scalar unit computes (kMaskedOff for a column outside [0, r0)).  Per-lane masks
    // are needed only where rows leave the band (i - j > bw: the first columns, lane > wv + W k)
    // or the matrix (the last block).
    constexpr int kCh = 8;  // prefetch registers are consumed and refilled in chunks of 8
    auto fwd_load = [&](int r0, int c) {
        const bool row_ok = r0 + lane < n;
#pragma unroll
        for (int u = 0; u < kCh; ++u) {
            const int k = kCh * c + u;
            const int j = r0 - bw + wv + kWaves * k;
            const unsigned so = (j >= 0 && j < r0) ? (unsigned)((j * bw + r0) * (int)sizeof(T)) : kMaskedOff;
            const unsigned off = (unsigned)lane * (unsigned)sizeof(T) + so;
            // lane > wv + W k (i - j > bw) only happens for W k < 64 (a compile-time condition)
            const bool ok = kWaves * k < 64 ? row_ok && lane <= wv + kWaves * k : row_ok;
            pf[k] = load_vs<T>(rs, ok ? off : kMaskedOff, 0);
        }
    };
#pragma unroll
    for (int c = 0; c < kPF / kCh; ++c) fwd_load(0, c);
    diag_load(0, false);
    lds_barrier();
    for (int b = 0; b < nblk; ++b) {
        const int r0 = b * kRB, i = r0 + lane;
        double acc = 0.0;
        // consume a chunk of this block's entries, then refill those registers with the next
        // block's (the scheduling barrier keeps the compiler from hoisting every LDS read and
        // load at once, which would exceed the register budget)
#pragma unroll
        for (int c = 0; c < kPF / kCh; ++c) {
            const double *xw = xs + (r0 - bw + wv);  // column j of k at xw[W k] (zeros before x)
#pragma unroll
            for (int u = 0; u < kCh; ++u) {
                const int k = kCh * c + u;
                acc += (double)pf[k] * xw[kWaves * k];
            }
            fwd_load(r0 + kRB, c);  // past the last block every read is masked off
            __builtin_amdgcn_sched_barrier(0);
        }
        sp[wave * kRB + lane] = acc;
        diag_store();
        diag_load(r0 + kRB, false);
        lds_barrier();
        if (wave == 0) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) s += sp[w * kRB + lane];
            double y = i < n ? xs[i] - s : 0.0;
#pragma unroll 8
            for (int jj = 0; jj < kRB - 1; ++jj) {
                // sd holds zeros on and above the diagonal, so no lane mask is needed
                const double l = (double)sd[jj * kSdLd + lane];
                y -= l * readlane_f64(y, jj);
            }
            if (i < n) xs[i] = y;
        }
        lds_barrier();
    }

    // ---- diagonal and backward: L^T x = D^-1 y ----
    constexpr int kPC = SolveCfg<T>::PC;  // prefetched loads per column
    // column j = c0 + jj, rows j + d, d = (64 - jj) + 64 k + lane: element j (bw + 1) + d, a
    // wave-uniform part plus lane; lanes at or past lim (d > bw or j + d >= n) are masked
    auto bwd_load = [&](int c0, int c) {
        const int jj = CW * wv + c, j = c0 + jj;
        const bool cok = c0 >= 0 && j < n;
#pragma unroll
        for (int k = 0; k < kPC; ++k) {
            const int dl = (kRB - jj) + 64 * k;
            const unsigned so = cok ? (unsigned)((j * (int)ldu + dl) * (int)sizeof(T)) : kMaskedOff;
            unsigned off = (unsigned)lane * (unsigned)sizeof(T) + so;
            const int lim = min(bw - dl + 1, n - j - dl);
            pf[c * kPC + k] = load_vs<T>(rs, lane < lim ? off : kMaskedOff, 0);
        }
    };
#pragma unroll
    for (int c = 0; c < CW; ++c) bwd_load((nblk - 1) * kRB, c);
    diag_load((nblk - 1) * kRB, true);
    for (int b = nblk - 1; b >= 0; --b) {
        const int c0 = b * kRB;
        double a[CW];
        const double *xb = xs + c0 + kRB + lane;  // row j + d of k at xb[64 k] (zeros after x)
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < kPC; ++k) acc += (double)pf[c * kPC + k] * xb[64 * k];
            a[c] = acc;
            bwd_load(c0 - kRB, c);  // before the first block every read is masked off
            __builtin_amdgcn_sched_barrier(0);
        }
        const double s = wave_sums<CW>(a, lane);
        if ((lane & (64 / CW - 1)) == 0) sp[CW * wave + lane / (64 / CW)] = s;
        diag_store();
        diag_load(c0 - kRB, true);
        lds_barrier();
        if (wave == 0) {
            const int i = c0 + lane;
            const double dv = (double)sd[lane * kSdLd + lane];
            sd[lane * kSdLd + lane] = T(0);  // only this lane reads its diagonal entry
            double xv = i < n ? xs[i] / dv - sp[lane] : 0.0;
#pragma unroll 8
            for (int jj = kRB - 1; jj > 0; --jj) {
                // zeros below the diagonal (and the cleared pivot) make the lane mask unnecessary
                const double l = (double)sd[lane * kSdLd + jj];
                xv -= l * readlane_f64(xv, jj);
            }
            if (i < n) xs[i] = xv;
        }
        lds_barrier();
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
// prefetched columns per block (512 for both precisions), and band byte offsets must stay below
// the 2 GiB masked-read offset.
size_t solve_lds(int nmax, int bwmax, size_t bytes_max, int prec)
{
    const int cols = prec == KSH_FP64 ? SolveCfg<double>::W * SolveCfg<double>::PF : SolveCfg<float>::W * SolveCfg<float>::PF;
    const size_t smem = prec == KSH_FP64 ? solve_lds_bytes<double>(nmax) : solve_lds_bytes<float>(nmax);
    if (smem > kSolveLdsMax || bwmax > cols || bwmax > 64 * SolveCfg<float>::PC || bytes_max >= 0x7F000000ull) return 0;
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
