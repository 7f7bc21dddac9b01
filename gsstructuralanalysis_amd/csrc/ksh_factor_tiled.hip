// ksh_factor_tiled.hip -- banded LDL^T factorization of a batch of bands spread over the whole
// GPU: every panel step is a pair of launches covering all bands still being factored, so a
// single large band (the 128 x 128 plate: n 49k, bw 1166) and a batch of uneven ones (the
// imperfection sweep: bw 182..470) both keep all CUs busy, and no panel has to fit LDS.
//
// Replaces SimplicialLDLT::compute on the arc-length, Newton and eigen paths
// (src/gsALMSolvers/gsALMBase.hpp:186-211, src/gsStaticSolvers/gsStaticNewton.hpp:215-240,
// src/gsEigenSolvers/gsBucklingSolver.h:131-133).  No pivoting, as there; by Sylvester's law of
// inertia the number of negative pivots does not depend on the ordering, so the band ordering
// reproduces the count gsALMBase::_computeStability takes (src/gsALMSolvers/gsALMHelper.h:30-42).
//
// Layout: lower band, column major, fac[j*ld + (i-j)], ld = bw+1, factored in place into unit L
// and D.  Right-looking in panels of NB = 64 columns.  Panel step at c0:
//   1. panel (one workgroup of 512 threads per band), two sub-panels of 32 columns: wave 0
//      factors the 32 x 32 diagonal block (lane = row, column jj broadcast with v_readlane),
//      each panel row below it is solved against L_dd D (row-wise TRSM); between the two
//      sub-panels the second one's columns take the first one's rank-32 update (on MFMA in the
//      register-resident form).  (One workgroup per band: a second one would read the diagonal
//      block while the first writes it back.)  The register-resident form (bw <= 480) computes in
//      the storage precision: in FP32 its 101 VGPRs and 80 KB of LDS let a panel workgroup share
//      its CU with a trailing-update workgroup of the same step launch (in FP64, 212 VGPRs and
//      157 KB, one workgroup per CU); the general form (wider bands) computes in FP64.
//   2. trailing update: the lower triangle of the bw x bw trailing window is cut into 32 x 32
//      blocks, one per wave, each a 2 x 2 set of 16 x 16 MFMA tiles (v_mfma_f64_16x16x4f64 /
//      v_mfma_f32_16x16x4f32) over the NB panel columns; panel operands -d_k L[j][k] and
//      L[i][k] come from L2, the block is read, updated and written once.
// Plain schedule: k_band_panel, then k_band_trailing over every block, per step.  Lookahead
// schedule: the blocks of the next panel's columns first (k_band_trailing), then k_band_step,
// whose first workgroups factor the next panel while the others apply the remaining blocks.
// Bands are sorted by order, largest first, so a step's grids cover exactly the bands that are
// still active.  Work per band: about n bw^2 multiply-adds; HBM traffic about 2 n bw^2 / NB
// entries (the trailing window of a step, read and written).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ksh_devutil.hpp"
#include "ksh_internal.hpp"
#include "ksh_mfma.hpp"
#include "ksh_profile.hpp"
#include "ksh_tune.hpp"

namespace ksh {

namespace {

constexpr int kNS = 32;       // sub-panel width (diagonal block factored by one wave)
constexpr int kNB = 2 * kNS;  // panel width (rank of one trailing update)
constexpr int kPT = 512;      // panel kernel threads (TRSM rows per pass)
constexpr int kUW = 4;        // waves per trailing-update workgroup
constexpr int kXr = kNS + 1;  // row stride of the panel update's per-wave LDS region

// Factor of a 32 x 32 diagonal block held in LDS (blk[kk*32 + r], lower triangle, unit pivots
// past w) by one wave: lane r owns row r (lanes >= 32 shadow row r - 32) in registers; at step jj
// column jj (every lane's entry jj) is broadcast with v_readlane, so the wave's FMAs take it as a
// scalar operand and no LDS round trip sits on the chain of dependent steps.  Entries above the
// diagonal take updates too; never read.  On return blk holds L below the diagonal and D on it;
// dval / dinv the pivots and their inverses.  The first bad pivot (zero, not finite, or not
// finite in the storage precision T) is recorded in *info; `clear` resets it first.
template <typename T, typename C>
__device__ __forceinline__ void diag_factor(C *blk, C *dval, C *dinv, int w, int col0, int *info, bool clear)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & (kNS - 1);
    C a[kNS];
#pragma unroll
    for (int kk = 0; kk < kNS; ++kk) a[kk] = blk[kk * kNS + r];
    int bad = 0;
#pragma unroll
    for (int jj = 0; jj < kNS; ++jj) {
        const C cv = a[jj];  // A[r][jj]; lane kk's value is A[kk][jj]
        // the whole column first, into distinct scalar registers: the reads issue back to back and
        // the reciprocal overlaps them, instead of a read / wait / multiply-add triple per entry
        C col[kNS];
#pragma unroll
        for (int kk = jj; kk < kNS; ++kk) col[kk] = bcast(cv, kk);
        const C djj = col[jj];
        if (jj < w && bad == 0 && (djj == C(0) || !isfinite(djj) || !isfinite((double)(T)djj))) bad = col0 + jj + 1;
        const C l = cv * (C(1) / djj);
#pragma unroll
        for (int kk = jj + 1; kk < kNS; ++kk) a[kk] -= l * col[kk];
        if (r > jj) a[jj] = l;
        __builtin_amdgcn_sched_barrier(0);
    }
    if (lane == 0 && info && (clear || (bad != 0 && *info == 0))) *info = bad;
    C ar = C(0);  // a[r] by selects: a lane-dependent index would put a[] in scratch memory
#pragma unroll
    for (int kk = 0; kk < kNS; ++kk) ar = kk == r ? a[kk] : ar;
    dval[r] = ar;
    dinv[r] = C(1) / ar;
#pragma unroll
    for (int kk = 0; kk < kNS; ++kk) blk[kk * kNS + r] = kk <= r ? a[kk] : C(0);
}

// Sub-panel at column cs (width ws = min(32, n - cs)): factor the diagonal block in LDS (wave 0),
// write it back, and solve the rows below it within the band (row-wise TRSM).  Rows cs+32 ..
// cs+63 of the result are also left in `lnext` (for the next sub-panel's update).  The panel's
// arithmetic is FP64 for both storage precisions (an FP32 band is read, and its factor written,
// in FP32; the trailing update, almost all of the work, runs in the storage precision).
template <typename T>
__device__ __forceinline__ void sub_panel(const BandDesc &d, __amdgpu_buffer_rsrc_t rs, int cs, double *blk, double *dinv,
                                          double *dval, double (*lnext)[kNS + 1])
{
    const int n = d.n, bw = d.bw;
    const int w = min(kNS, n - cs), nr = min(n - cs, w + bw);
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const int tid = threadIdx.x, wave = tid >> 6;
    constexpr int kQ = kNS * kNS / kPT;
    // stage the diagonal block (unit pivots pad the last panel)
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int idx = tid + kPT * q, kk = idx / kNS, r = idx % kNS;
        const bool ok = r >= kk && r < w && r - kk <= bw;
        const double v = (double)load_if<T>(ok, rs, ((unsigned)(cs + kk) * ldu + (unsigned)(r - kk)) * es);
        blk[idx] = (r >= w && r == kk) ? 1.0 : v;
    }
    __syncthreads();
    if (wave == 0) diag_factor<T, double>(blk, dval, dinv, w, cs, d.info, cs == 0);
    __syncthreads();
    // write the factored block back (L below the diagonal, D on it)
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int idx = tid + kPT * q, kk = idx / kNS, r = idx % kNS;
        const bool ok = r >= kk && r < w && r - kk <= bw;
        store_if<T>(ok, rs, ((unsigned)(cs + kk) * ldu + (unsigned)(r - kk)) * es, (T)blk[idx]);
    }
    // TRSM: row r of the sub-panel, L[r][0:32] = A[r][0:32] (D L_dd^T)^-1
    for (int r = w + tid; r < nr; r += kPT) {
        double a[kNS];
#pragma unroll
        for (int kk = 0; kk < kNS; ++kk) {
            const bool ok = kk < w && r - kk <= bw;
            a[kk] = (double)load_if<T>(ok, rs, ((unsigned)(cs + kk) * ldu + (unsigned)(r - kk)) * es);
        }
#pragma unroll
        for (int kk = 0; kk < kNS; ++kk) {
            const double t = a[kk];  // = L[r][kk] d_kk
#pragma unroll
            for (int jj = kk + 1; jj < kNS; ++jj) a[jj] -= t * blk[kk * kNS + jj];
            a[kk] = t * dinv[kk];
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int kk = 0; kk < kNS; ++kk) {
            const bool ok = kk < w && r - kk <= bw;
            store_if<T>(ok, rs, ((unsigned)(cs + kk) * ldu + (unsigned)(r - kk)) * es, (T)a[kk]);
        }
        if (r >= kNS && r < 2 * kNS) {
#pragma unroll
            for (int kk = 0; kk < kNS; ++kk) lnext[r - kNS][kk] = a[kk];
        }
    }
}

// Panel of 64 columns at c0 as two sub-panels of 32: the second sub-panel's columns take the
// first one's rank-32 update (A[r][j] -= sum_k L[r][k] d_k L[j][k]) between the two.
template <typename T>
__device__ __forceinline__ void panel_general(const BandDesc &d, int c0)
{
    __shared__ double blk[kNS * kNS];     // blk[kk*32 + r] = A / L[cs+r][cs+kk] for r >= kk, zero above
    __shared__ double dinv[kNS], dval[kNS];
    __shared__ double lnext[kNS][kNS + 1];  // L[c0+32+j][c0+k] of the first sub-panel
    const int n = d.n, bw = d.bw;
    if (c0 >= n) return;
    T *fac = (T *)d.fac;
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    sub_panel<T>(d, rs, c0, blk, dinv, dval, lnext);
    const int c1 = c0 + kNS;
    if (c1 >= n) return;  // uniform: the last panel has one sub-panel
    const int w0 = kNS, w1 = min(kNS, n - c1), nr0 = min(n - c0, w0 + bw);
    __syncthreads();  // lnext complete, TRSM stores visible
    // update of columns c1 .. c1+w1-1, rows r (relative to c0) in [32, nr0)
    for (int r = kNS + (int)threadIdx.x; r < nr0; r += kPT) {
        double l[kNS];
#pragma unroll
        for (int kk = 0; kk < kNS; ++kk) {
            const bool ok = r - kk <= bw;
            l[kk] = (double)load_if<T>(ok, rs, ((unsigned)(c0 + kk) * ldu + (unsigned)(r - kk)) * es) * dval[kk];
        }
        for (int j = 0; j < w1; ++j) {
            const int col = kNS + j;  // relative to c0
            const bool ok = r >= col && r - col <= bw;
            const unsigned off = ((unsigned)(c0 + col) * ldu + (unsigned)(r - col)) * es;
            double acc = (double)load_if<T>(ok, rs, off);
#pragma unroll
            for (int kk = 0; kk < kNS; ++kk) acc -= l[kk] * lnext[j][kk];
            store_if<T>(ok, rs, off, (T)acc);
        }
    }
    __syncthreads();  // updated columns visible to the second sub-panel's loads
    sub_panel<T>(d, rs, c1, blk, dinv, dval, lnext);
}

// Panel of 64 columns at c0 with every panel row below the first diagonal block owned by one
// thread for the whole kernel (bands with bw <= kPT - 32): the thread loads its row's 64 entries
// once (in two halves, each just before use), solves the first sub-panel's part (TRSM), applies the first sub-panel's
// rank-32 update to the second part and solves that (TRSM) in registers; the two diagonal
// blocks go through LDS.  A row's entries are read once and written once.
template <typename T>
__device__ __forceinline__ void panel_rows(const BandDesc &d, int c0)
{
    using C = T;  // arithmetic in the storage precision
    using M = Mfma16<C>;
    __shared__ C blk1[kNS * kNS], blk2[kNS * kNS];
    __shared__ C dval1[kNS], dinv1[kNS], dval2[kNS], dinv2[kNS];
    __shared__ C lnext[kNS][kNS + 1];  // L[c0+32+j][c0+k] of the first sub-panel
    __shared__ C xr[kPT / 64][64 * kXr];  // per wave: its 64 rows, first L d, then the update
    const int n = d.n, bw = d.bw;
    if (c0 >= n) return;
    T *fac = (T *)d.fac;
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    const int tid = threadIdx.x, wave = tid >> 6;
    const int w = min(kNB, n - c0), w1 = min(kNS, w), w2 = w - w1;
    const int nrr = min(n - c0, kNB + bw);  // panel rows (relative to c0)
    const int r = kNS + tid;                // this thread's row
    const bool has_row = r < nrr;
    constexpr int kQ = kNS * kNS / kPT;
    // loads: the first diagonal block, and this thread's row (columns 0..63 of the panel)
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int idx = tid + kPT * q, kk = idx / kNS, rr = idx % kNS;
        const bool ok = rr >= kk && rr < w1 && rr - kk <= bw;
        const C v = load_if<T>(ok, rs, ((unsigned)(c0 + kk) * ldu + (unsigned)(rr - kk)) * es);
        blk1[idx] = (rr >= w1 && rr == kk) ? C(1) : v;
    }
    __syncthreads();
    if (wave == 0) diag_factor<T, C>(blk1, dval1, dinv1, w1, c0, d.info, c0 == 0);
    __syncthreads();
    // the row is loaded in two halves, each just before it is needed, so that no more than 64
    // of its values are in registers at once (and none in the factoring wave during a factor)
    C x[kNB];
#pragma unroll
    for (int k = 0; k < kNS; ++k) {
        const bool ok = has_row && k < w && r - k <= bw;
        x[k] = load_if<T>(ok, rs, ((unsigned)(c0 + k) * ldu + (unsigned)(r - k)) * es);
    }
    // first diagonal block back to the band
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int idx = tid + kPT * q, kk = idx / kNS, rr = idx % kNS;
        const bool ok = rr >= kk && rr < w1 && rr - kk <= bw;
        store_if<T>(ok, rs, ((unsigned)(c0 + kk) * ldu + (unsigned)(rr - kk)) * es, (T)blk1[idx]);
    }
    // TRSM of the first sub-panel: L[r][0:32] = A[r][0:32] (D L_dd^T)^-1
#pragma unroll
    for (int kk = 0; kk < kNS; ++kk) {
        const C tk = x[kk];
#pragma unroll
        for (int jj = kk + 1; jj < kNS; ++jj) x[jj] -= tk * blk1[kk * kNS + jj];
        x[kk] = tk * dinv1[kk];
        __builtin_amdgcn_sched_barrier(0);
    }
    {
        const int ro = opaque(r);  // store masks recomputed, not kept from the loads
#pragma unroll
        for (int k = 0; k < kNS; ++k) {
            const bool ok = has_row && k < w1 && ro - k <= bw;
            store_if<T>(ok, rs, ((unsigned)(c0 + k) * ldu + (unsigned)(ro - k)) * es, (T)x[k]);
        }
    }
    if (w2 <= 0) return;  // uniform: no second sub-panel
#pragma unroll
    for (int k = kNS; k < kNB; ++k) {
        const bool ok = has_row && k < w && r >= k && r - k <= bw;
        x[k] = load_if<T>(ok, rs, ((unsigned)(c0 + k) * ldu + (unsigned)(r - k)) * es);
    }
    if (r < kNB) {
#pragma unroll
        for (int k = 0; k < kNS; ++k) lnext[r - kNS][k] = x[k];
    }
    __syncthreads();
    // rank-32 update of the second part, x[32+j] -= S[r][j], S[r][j] = sum_k L[r][k] d_k L[32+j][k],
    // on MFMA per wave (entries above the diagonal of rows 32..63 are updated too; never
    // read).  The wave's rows go through its LDS region: row i of the wave holds L[r][k] d_k at
    // xr[i*kXr + k], then S[r][j] at xr[i*kXr + j].  S^T tiles (16 j x 16 rows): A = L[32+j][k]
    // from lnext, B = (L d)^T from the region, so that each value is read once per lane (a
    // loop over rows with lnext broadcast to every lane costs 32 x 32 LDS broadcasts per wave).
    {
        C *reg = xr[wave];
        const int lane = tid & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
        for (int k = 0; k < kNS; ++k) reg[lane * kXr + k] = x[k] * dval1[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        C av[2][kNS / 4];
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int s = 0; s < kNS / 4; ++s) av[jt][s] = lnext[16 * jt + li][4 * s + lg];
        const typename M::acc_t zero = {C(0), C(0), C(0), C(0)};
        typename M::acc_t acc[2][4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            acc[0][rt] = zero;
            acc[1][rt] = zero;
#pragma unroll
            for (int s = 0; s < kNS / 4; ++s) {
                const C bv = reg[(16 * rt + li) * kXr + 4 * s + lg];
                acc[0][rt] = M::run(av[0][s], bv, acc[0][rt]);
                acc[1][rt] = M::run(av[1][s], bv, acc[1][rt]);
            }
        }
        // lane holds S[16 rt + li][16 jt + row(lane, v)]: back to the row owners through the region
        // (each row's B reads feed its own tiles, so they precede these writes)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int jt = 0; jt < 2; ++jt)
#pragma unroll
                for (int v = 0; v < 4; ++v) reg[(16 * rt + li) * kXr + 16 * jt + M::row(lane, v)] = acc[jt][rt][v];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
#pragma unroll
        for (int j = 0; j < kNS; ++j) x[kNS + j] -= reg[lane * kXr + j];
    }
    // second diagonal block (rows 32..63) through LDS, unit pivots past w2
    if (r < kNB) {
        const int rr = r - kNS;
#pragma unroll
        for (int kk = 0; kk < kNS; ++kk)
            blk2[kk * kNS + rr] = rr >= w2 ? (rr == kk ? C(1) : C(0)) : (kk <= rr ? x[kNS + kk] : C(0));
    }
    __syncthreads();
    if (wave == 0) diag_factor<T, C>(blk2, dval2, dinv2, w2, c0 + kNS, d.info, false);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int idx = tid + kPT * q, kk = idx / kNS, rr = idx % kNS;
        const bool ok = rr >= kk && rr < w2 && rr - kk <= bw;
        store_if<T>(ok, rs, ((unsigned)(c0 + kNS + kk) * ldu + (unsigned)(rr - kk)) * es, (T)blk2[idx]);
    }
    if (r < kNB) return;  // rows of the second diagonal block are done
    // TRSM of the second sub-panel
#pragma unroll
    for (int kk = 0; kk < kNS; ++kk) {
        const C tk = x[kNS + kk];
#pragma unroll
        for (int jj = kk + 1; jj < kNS; ++jj) x[kNS + jj] -= tk * blk2[kk * kNS + jj];
        x[kNS + kk] = tk * dinv2[kk];
        __builtin_amdgcn_sched_barrier(0);
    }
    const int ro = opaque(r);
#pragma unroll
    for (int j = 0; j < kNS; ++j) {
        const int col = kNS + j;
        const bool ok = has_row && j < w2 && ro - col <= bw;
        store_if<T>(ok, rs, ((unsigned)(c0 + col) * ldu + (unsigned)(ro - col)) * es, (T)x[col]);
    }
}

// lower-triangle block index p -> (bi, bj), bj <= bi
__device__ __forceinline__ void tri_decode(int p, int &bi, int &bj)
{
    int b = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
    if ((b + 1) * (b + 2) / 2 <= p) ++b;
    if (b * (b + 1) / 2 > p) --b;
    bi = b;
    bj = p - b * (b + 1) / 2;
}

// Parts of a step's trailing update (32 x 32 blocks (bi, bj), bj <= bi, of the window):
// kAll every block; kNext block columns 0 and 1 (the next panel's columns); kRest the others.
enum TrailPart { kAll = 0, kNext = 1, kRest = 2 };

// Blocks p[0] < p[1] < ... of the given part (wave-uniform), computed together (independent
// accumulators, the loads of all of them in flight at once): A[i][j] -= sum_k L[i][k] d_k L[j][k]
// over the window rows/columns [w, nr) of the frame at c0.  Indices past the part's block count
// are computed on block p[0]'s operands with their stores masked off.  Each block's MFMA sequence
// (k ascending) is the same for every NBLK, so the result does not depend on it.
template <typename T, int NBLK>
__device__ __forceinline__ void trailing_blocks(const BandDesc &d, int c0, int part, const int (&p)[NBLK])
{
    using M = Mfma16<T>;
    using acc_t = typename M::acc_t;
    const int n = d.n, bw = d.bw;
    if (c0 >= n) return;
    const int w = min(kNB, n - c0), nr = min(n - c0, w + bw);
    const int nbk = (nr - w + 31) / 32;
    const int nblock = part == kAll ? nbk * (nbk + 1) / 2 : part == kNext ? min(nbk, 2) * nbk - (nbk >= 2 ? 1 : 0)
                                                                          : (nbk > 2 ? (nbk - 2) * (nbk - 1) / 2 : 0);
    if (p[0] >= nblock) return;
    T *fac = (T *)d.fac;
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    bool valid[NBLK], diag[NBLK];
    int ri0[NBLK], rj0[NBLK];
    // accumulator rows = j (A's column), columns = i (A's row): lanes run along a column
    acc_t cur[NBLK][2][2];
    unsigned off[NBLK][2][2][4];
#pragma unroll
    for (int s = 0; s < NBLK; ++s) {
        valid[s] = p[s] < nblock;
        const int q = valid[s] ? p[s] : p[0];
        int bi, bj;
        if (part == kNext) {
            bj = q < nbk ? 0 : 1;
            bi = bj == 0 ? q : q - nbk + 1;
        } else {
            tri_decode(q, bi, bj);
            if (part == kRest) {
                bi += 2;
                bj += 2;
            }
        }
        diag[s] = bi == bj;
        ri0[s] = w + 32 * bi + li;
        rj0[s] = w + 32 * bj + li;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = w + 32 * bi + 16 * a + li, j = w + 32 * bj + 16 * b + M::row(lane, v);
                    const bool ok = valid[s] && i < nr && i >= j;
                    off[s][a][b][v] = ok ? ((unsigned)c0 * ldu + (unsigned)j * (ldu - 1) + (unsigned)i) * es : kMaskedOff;
                    cur[s][a][b][v] = load_vs<T>(rs, off[s][a][b][v], 0);
                }
    }
    // panel operands: rows rj (block column bj) and ri (block row bi), k = 4 s + lk, in chunks
    constexpr int kCh = 8 / NBLK;
#pragma unroll
    for (int c = 0; c < kNB / 4 / kCh; ++c) {
        T A0[NBLK][kCh], A1[NBLK][kCh], B0[NBLK][kCh], B1[NBLK][kCh];
#pragma unroll
        for (int u = 0; u < kCh; ++u) {
            const int k = 4 * (kCh * c + u) + lk;
            const unsigned col = ((unsigned)c0 * ldu + (unsigned)k * (ldu - 1)) * es;  // + row * es
            const T nd = -load_if<T>(k < w, rs, (unsigned)(c0 + k) * ldu * es);
#pragma unroll
            for (int s = 0; s < NBLK; ++s) {
                const int rj = rj0[s], ri = ri0[s];
                A0[s][u] = nd * load_if<T>(k < w && rj < nr && rj - k <= bw, rs, col + (unsigned)rj * es);
                A1[s][u] = nd * load_if<T>(k < w && rj + 16 < nr && rj + 16 - k <= bw, rs, col + (unsigned)(rj + 16) * es);
                B0[s][u] = load_if<T>(k < w && ri < nr && ri - k <= bw, rs, col + (unsigned)ri * es);
                B1[s][u] = load_if<T>(k < w && ri + 16 < nr && ri + 16 - k <= bw, rs, col + (unsigned)(ri + 16) * es);
            }
        }
#pragma unroll
        for (int u = 0; u < kCh; ++u)
#pragma unroll
            for (int s = 0; s < NBLK; ++s) {
                cur[s][0][0] = M::run(A0[s][u], B0[s][u], cur[s][0][0]);
                cur[s][1][0] = M::run(A0[s][u], B1[s][u], cur[s][1][0]);
                cur[s][1][1] = M::run(A1[s][u], B1[s][u], cur[s][1][1]);
                if (!diag[s]) cur[s][0][1] = M::run(A1[s][u], B0[s][u], cur[s][0][1]);
            }
    }
#pragma unroll
    for (int s = 0; s < NBLK; ++s)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) store_vs<T>(rs, off[s][a][b][v], 0, cur[s][a][b][v]);
}

// Panel step at c0 of every band in the grid (one workgroup per band).  kRows: the register-resident
// form (band widths up to kPT - 32).
template <typename T, bool kRows>
__global__ __launch_bounds__(kPT) void k_band_panel(const BandDesc *descs, int c0)
{
    if (kRows)
        panel_rows<T>(descs[blockIdx.x], c0);
    else
        panel_general<T>(descs[blockIdx.x], c0);
}

// XCD-aware assignment of a launch's workgroups to (band, index within the band).  Workgroups are
// dispatched to the 8 XCDs round robin (id mod 8), each XCD with its own L2.  Every trailing block
// of a band reads that band's panel, so with 8 or more bands the workgroups of band b take the ids
// h = b (mod 8): all of them run on one XCD and its L2 serves the panel, which otherwise every
// XCD would fetch once per band.  The panel kernels already put band b on workgroup b, the same
// XCD.  With fewer bands than XCDs the plain mapping keeps every XCD busy.  Returns false for
// the padding workgroups of the XCD-aware grid (8 ceil(active / 8) wg of them).
constexpr int kXcds = 8;
__device__ __forceinline__ bool band_of_workgroup(int h, int active, int wg, int &band, int &local)
{
    if (active < kXcds) {
        band = h / wg;
        local = h % wg;
    } else {
        const int s = h / kXcds;
        band = h % kXcds + kXcds * (s / wg);
        local = s % wg;
    }
    return band < active;
}
inline unsigned band_grid(int active, int wg)
{
    return (unsigned)(active < kXcds ? active * wg : ((active + kXcds - 1) / kXcds) * kXcds * wg);
}

// One part of the trailing update at c0, kUW blocks (one per wave) per workgroup
template <typename T>
__global__ __launch_bounds__(64 * kUW) void k_band_trailing(const BandDesc *descs, int c0, int wg_per_band, int part,
                                                            int active)
{
    int band, local;
    if (!band_of_workgroup((int)blockIdx.x, active, wg_per_band, band, local)) return;
    const int p[1] = {__builtin_amdgcn_readfirstlane((int)(local * kUW + (threadIdx.x >> 6)))};
    trailing_blocks<T, 1>(descs[band], c0, part, p);
}

// Lookahead step: workgroups [0, npanel) factor the panel at c0 + NB of their band while the others
// apply the rest of the trailing update at c0 (block columns 2 and up): AW waves of each of these
// workgroups take NBLK blocks each, the other waves leave at once.  The two parts touch disjoint
// columns: the panel reads and writes columns [c0 + NB, c0 + 2 NB), already updated by the kNext
// part; the rest of the update reads the panel at c0 and writes columns from c0 + 2 NB on.
// (The update part starts at workgroup `pad`, npanel rounded up to a multiple of the XCD count,
// so that its XCD-aware mapping lines up with the panels' workgroups.)
template <typename T, bool kRows, int AW, int NBLK>
__global__ __launch_bounds__(kPT) void k_band_step(const BandDesc *descs, int c0, int npanel, int wg_per_band,
                                                    int pad, int active)
{
    const int b = (int)blockIdx.x;
    if (b < npanel) {
        if (kRows)
            panel_rows<T>(descs[b], c0 + kNB);
        else
            panel_general<T>(descs[b], c0 + kNB);
        return;
    }
    if (b < pad) return;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wave >= AW) return;
    int band, local;
    if (!band_of_workgroup(b - pad, active, wg_per_band, band, local)) return;
    int p[NBLK];
#pragma unroll
    for (int s = 0; s < NBLK; ++s) p[s] = (local * NBLK + s) * AW + wave;
    trailing_blocks<T, NBLK>(descs[band], c0, kRest, p);
}

// negatives, min d, max |d|, min |d| of each factored diagonal (one workgroup of 1024 threads per band)
template <typename T>
__global__ __launch_bounds__(1024) void k_pivot_stats(const BandDesc *descs)
{
    __shared__ double part[16][4];
    const BandDesc &dd = descs[blockIdx.x];
    if (!dd.stats) return;
    const T *fac = (const T *)dd.fac;
    const size_t ld = (size_t)dd.bw + 1;
    double c = 0.0, mn = INFINITY, mxa = 0.0, mna = INFINITY;
    for (int i = threadIdx.x; i < dd.n; i += 1024) {
        const double d = (double)fac[(size_t)i * ld];
        c += d < 0.0 ? 1.0 : 0.0;
        mn = fmin(mn, d);
        mxa = fmax(mxa, fabs(d));
        mna = fmin(mna, fabs(d));
    }
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o);
        mn = fmin(mn, __shfl_xor(mn, o));
        mxa = fmax(mxa, __shfl_xor(mxa, o));
        mna = fmin(mna, __shfl_xor(mna, o));
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
        for (int q = 1; q < 16; ++q) {
            c += part[q][0];
            mn = fmin(mn, part[q][1]);
            mxa = fmax(mxa, part[q][2]);
            mna = fmin(mna, part[q][3]);
        }
        dd.stats[0] = c;
        dd.stats[1] = mn;
        dd.stats[2] = mxa;
        dd.stats[3] = mna;
    }
}

// Factor schedules (every one gives the same result bit for bit):
//   0 plain: panel, then the whole trailing update, per step;
//   1, 2 lookahead: per step the next panel's columns first (kNext), then the next panel together
//   with the rest of the update (k_band_step) with AW waves x NBLK blocks per workgroup:
//   1: 8 x 1, 2: 8 x 2 (more blocks in flight per wave).
// The lookahead takes the rest of the update off the panel's critical path but runs it at the
// panel kernel's register budget (2 waves per SIMD); which one is faster depends on the number
// and shape of the bands, so the autotuner picks per shape bucket (KSH_LOOKAHEAD=k forces k).
constexpr int kSchedules = 3;

int schedule_forced()
{
    const char *e = std::getenv("KSH_LOOKAHEAD");
    if (!e || !*e) return -1;
    const int v = std::atoi(e);
    return v >= 0 && v < kSchedules ? v : -1;
}

template <typename T, bool kRows, int AW, int NBLK>
void run_lookahead(hipStream_t s, const BandDesc *d_descs, const std::vector<int> &ns, int maxbw)
{
    const int nbands = (int)ns.size();
    const int nbk = (maxbw + 31) / 32;
    auto active_at = [&](int c0) {
        int a = nbands;
        while (a > 0 && ns[a - 1] <= c0) --a;
        return a;
    };
    const int wg1 = std::max(1, ((nbk >= 2 ? 2 * nbk - 1 : nbk) + kUW - 1) / kUW);
    const int nrest = nbk > 2 ? (nbk - 2) * (nbk - 1) / 2 : 0;
    const int wg2 = (nrest + AW * NBLK - 1) / (AW * NBLK);
    hipLaunchKernelGGL((k_band_panel<T, kRows>), dim3((unsigned)nbands), dim3(kPT), 0, s, d_descs, 0);
    for (int c0 = 0; c0 < ns[0]; c0 += kNB) {
        const int active = active_at(c0), np = active_at(c0 + kNB);
        hipLaunchKernelGGL(k_band_trailing<T>, dim3(band_grid(active, wg1)), dim3(64 * kUW), 0, s, d_descs, c0, wg1,
                           (int)kNext, active);
        const int pad = wg2 > 0 && active >= kXcds ? (np + kXcds - 1) / kXcds * kXcds : np;
        const int grid = pad + (wg2 > 0 ? (int)band_grid(active, wg2) : 0);
        if (grid > 0)
            hipLaunchKernelGGL((k_band_step<T, kRows, AW, NBLK>), dim3((unsigned)grid), dim3(kPT), 0, s, d_descs, c0, np,
                               wg2, pad, active);
    }
}

template <typename T, bool kRows>
void run_tiled_sched(hipStream_t s, const BandDesc *d_descs, const std::vector<int> &ns, int maxbw, int sched)
{
    // ns: orders of the bands, descending (the device array is in the same order)
    const int nbands = (int)ns.size();
    if (sched == 0) {
        const int nbk = (maxbw + 31) / 32;
        const int wg = std::max(1, (nbk * (nbk + 1) / 2 + kUW - 1) / kUW);
        int active = nbands;
        for (int c0 = 0; c0 < ns[0]; c0 += kNB) {
            while (active > 0 && ns[active - 1] <= c0) --active;
            hipLaunchKernelGGL((k_band_panel<T, kRows>), dim3((unsigned)active), dim3(kPT), 0, s, d_descs, c0);
            hipLaunchKernelGGL(k_band_trailing<T>, dim3(band_grid(active, wg)), dim3(64 * kUW), 0, s, d_descs, c0, wg,
                               (int)kAll, active);
        }
    } else if (sched == 1) {
        run_lookahead<T, kRows, 8, 1>(s, d_descs, ns, maxbw);
    } else {
        run_lookahead<T, kRows, 8, 2>(s, d_descs, ns, maxbw);
    }
    hipLaunchKernelGGL(k_pivot_stats<T>, dim3((unsigned)nbands), dim3(1024), 0, s, d_descs);
}

template <typename T>
void run_tiled(hipStream_t s, const BandDesc *d_descs, const std::vector<int> &ns, int maxbw, int sched)
{
    if (maxbw + kNS <= kPT)
        run_tiled_sched<T, true>(s, d_descs, ns, maxbw, sched);
    else
        run_tiled_sched<T, false>(s, d_descs, ns, maxbw, sched);
}

// Captured launch sequences of factor calls (opt-in: KSH_GRAPHS=1).  A call's sequence (about
// 2 n / 64 launches) depends only on the precision, the device descriptor array, the largest band
// width and the orders of the bands; it is captured into a hipGraph the second time that key is
// seen and replayed from then on (the descriptors themselves are uploaded by an ordinary copy
// before each replay).  At most kGraphs sequences are kept; later keys run eagerly.  Off by
// default: replaying changed the batched arc-length throughput by less than 1% (the launches are
// queued asynchronously well ahead of the GPU), and eager launches are what the kernel profiles
// record.
struct GraphKey {
    int prec;
    int sched;
    const void *descs;
    int maxbw;
    std::vector<int> ns;
    bool operator<(const GraphKey &o) const
    {
        if (prec != o.prec) return prec < o.prec;
        if (sched != o.sched) return sched < o.sched;
        if (descs != o.descs) return descs < o.descs;
        if (maxbw != o.maxbw) return maxbw < o.maxbw;
        return ns < o.ns;
    }
};
struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    int seen = 0;
};
constexpr size_t kGraphs = 64;
std::mutex g_graph_mu;
std::map<GraphKey, GraphEntry> g_graphs;
size_t g_captured = 0;

bool graphs_enabled()
{
    const char *e = std::getenv("KSH_GRAPHS");
    return e && std::strcmp(e, "1") == 0;
}

template <typename T>
int run_tiled_graph(hipStream_t s, const BandDesc *d_descs, const std::vector<int> &ns, int maxbw, int prec, int la)
{
    if (!graphs_enabled() || tune_measuring()) {
        run_tiled<T>(s, d_descs, ns, maxbw, la);
        return KSH_SUCCESS;
    }
    std::lock_guard<std::mutex> lk(g_graph_mu);
    GraphKey key{prec, la, d_descs, maxbw, ns};
    auto it = g_graphs.find(key);
    if (it == g_graphs.end()) {
        if (g_graphs.size() < 4 * kGraphs) g_graphs.emplace(key, GraphEntry{nullptr, 1});
        run_tiled<T>(s, d_descs, ns, maxbw, la);
        return KSH_SUCCESS;
    }
    GraphEntry &en = it->second;
    if (!en.exec && ++en.seen >= 2 && g_captured < kGraphs) {
        hipGraph_t g = nullptr;
        KSH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        run_tiled<T>(s, d_descs, ns, maxbw, la);
        KSH_HIP(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&en.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(KSH_OTHER_ERROR, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        ++g_captured;
    }
    if (en.exec) {
        KSH_HIP(hipGraphLaunch(en.exec, s));
    } else {
        run_tiled<T>(s, d_descs, ns, maxbw, la);
    }
    return KSH_SUCCESS;
}

// Schedule: forced by KSH_LOOKAHEAD, else the autotuner's choice for the bucket of (precision,
// largest band width, number of bands, largest order); its timing calls are whole factor calls.
template <typename T>
int run_factor(hipStream_t s, const BandDesc *d_descs, const std::vector<int> &ns, int maxbw, int prec)
{
    const int forced = schedule_forced();
    if (forced >= 0) return run_tiled_graph<T>(s, d_descs, ns, maxbw, prec, forced);
    const long long bucket =
        ((((long long)prec * 64 + tune_bucket(maxbw)) * 64 + tune_bucket((double)ns.size())) * 64) + tune_bucket(ns[0]);
    int rc = KSH_SUCCESS;
    tune_launch("band_factor", bucket, kSchedules, 0, s,
                [&](int v) { rc = run_tiled_graph<T>(s, d_descs, ns, maxbw, prec, v); });
    return rc;
}

}  // namespace

int band_factor_tiled_batch(hipStream_t s, void *d_descs, const std::vector<BandDescHost> &h, int prec)
{
    const size_t es = prec == KSH_FP64 ? sizeof(double) : sizeof(float);
    std::vector<BandDesc> dd;
    for (const auto &x : h) {
        if (!x.active) continue;
        if (x.n < 1) return fail(KSH_SOLVER_ERROR, "empty band");
        if ((size_t)x.n * (size_t)(x.bw + 1) * es >= 0x80000000ull)
            return fail(KSH_SOLVER_ERROR, "band of 2 GiB or more: beyond the band kernels' 32-bit offsets");
        dd.push_back(BandDesc{x.fac, x.n, x.bw, x.info, x.stats, x.x, 1, 0});
    }
    if (dd.empty()) return KSH_SUCCESS;
    std::stable_sort(dd.begin(), dd.end(), [](const BandDesc &a, const BandDesc &b) { return a.n > b.n; });
    std::vector<int> ns(dd.size());
    int maxbw = 0;
    for (size_t i = 0; i < dd.size(); ++i) {
        ns[i] = dd[i].n;
        maxbw = std::max(maxbw, dd[i].bw);
    }
    KSH_HIP(hipMemcpyAsync(d_descs, dd.data(), sizeof(BandDesc) * dd.size(), hipMemcpyHostToDevice, s));
    int rc = prec == KSH_FP64 ? run_factor<double>(s, (const BandDesc *)d_descs, ns, maxbw, prec)
                              : run_factor<float>(s, (const BandDesc *)d_descs, ns, maxbw, prec);
    if (rc) return rc;
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

}  // namespace ksh
