// ksh_assembly.hip -- Gauss-point assembly of the St. Venant-Kirchhoff Kirchhoff-Love shell
// residual R(u), tangent K(u), mass M and load F on tensor B-spline patches (gfx950).
//
// Reference: the hot path is gsKLShell's gsThinShellAssembler (external module, not in the
// reference tree); its call sites are benchmarks/benchmark_Roof.cpp:319-344 (assemble,
// assembleMatrix, assembleVector), solvers/modal_shell_XML.cpp:235-236 (assembleMass) and
// benchmarks/benchmark_Roof.cpp:304-306 (SvK material, analytical thickness integration).
//
// Element tangent = batched B^T D B contraction (see ksh_shell_math.hpp for H_g):
//     K_e^{ik}[A][B] = sum_{g,a} Psi_g[a][A] * ( sum_b H_g[(a,i),(b,k)] Psi_g[b][B] ),
// one 16x16 output tile per component pair (i,k), inner dimension 5*ngp, on
// v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (FP32).  H_g and the internal-force
// vectors f_g come from one thread per Gauss point (state in registers); the element stages run
// one wave per element.  Element matrices are stored dof-major (entry (3A+i, 3B+k) at
// (3B+k)*48 + 3A+i), so the band gather's consecutive threads (consecutive rows of one band
// column) read consecutive addresses; the gather sums in a fixed element order (deterministic,
// no atomics).
//
// Every kernel runs over a batch of instances described by an InstDesc array (a single problem is
// a batch of one): element kernels map blockIdx.x to (instance, element) and entry kernels map a
// thread to (instance, dof / band slot) by binary search over the prefix offsets.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cmath>

#include "ksh_devutil.hpp"
#include "ksh_internal.hpp"
#include "ksh_mfma.hpp"
#include "ksh_profile.hpp"
#include "ksh_shell_math.hpp"
#include "ksh_tune.hpp"

namespace ksh {

namespace {

using shell::ElemArgs;

inline unsigned nblk(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename Get>
__device__ inline int find_inst(int n, long long key, Get get)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (get(mid) <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------------------
__global__ void k_construct_solution(const InstDesc *descs, int ninst, int total)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const InstDesc &d = descs[find_inst(ninst, gid, [&](int m) { return (long long)descs[m].full0; })];
    int g = gid - d.full0;
    int f = d.free_idx[g];
    d.u_full[g] = (f >= 0) ? d.u_free[f] : 0.0;
}

// deformed derivative vectors (a1, a2, a11, a22, a12) at one Gauss point of element (ex, ey)
template <int P>
__device__ __forceinline__ void gp_derivatives(const InstDesc &d, int ex, int ey, const double *tx, const double *ty,
                                               double (&dv)[15])
{
    constexpr int P1 = P + 1, NB = P1 * P1;
#pragma unroll
    for (int r = 0; r < 15; ++r) dv[r] = 0.0;
#pragma unroll
    for (int A = 0; A < NB; ++A) {
        const int ia = A % P1, ib = A / P1, cp = (ex + ia) + d.mx * (ey + ib);
        const double psi[5] = {tx[P1 + ia] * ty[ib], tx[ia] * ty[P1 + ib], tx[2 * P1 + ia] * ty[ib],
                               tx[ia] * ty[2 * P1 + ib], tx[P1 + ia] * ty[P1 + ib]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double x = d.ctrl0[3 * cp + c] + d.u_full[3 * cp + c];
#pragma unroll
            for (int a = 0; a < 5; ++a) dv[3 * a + c] += psi[a] * x;
        }
    }
}

// Residual, stage 1: one thread per Gauss point; its state and the 15 entries of f_g
// (f_int_{A,i} = sum_g phi . f_g), stored entry-major in the H_g buffer (Fg[I * ngp + gp]).
template <int P>
__global__ __launch_bounds__(128) void k_residual_gp(const InstDesc *descs, int ninst, int total)
{
    constexpr int P1 = P + 1, NQ = P1, NG = NQ * NQ;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const InstDesc &d = descs[find_inst(ninst, gid, [&](int m) { return (long long)descs[m].elem0 * NG; })];
    const int gp = gid - d.elem0 * NG;
    const int e = gp / NG, g = gp % NG, ex = e % d.nx, ey = e / d.nx, qx = g % NQ, qy = g / NQ;
    const double *tx = d.tabx + ((size_t)(ex * NQ + qx) * 3) * P1;
    const double *ty = d.taby + ((size_t)(ey * NQ + qy) * 3) * P1;
    double dv[15];
    gp_derivatives<P>(d, ex, ey, tx, ty, dv);
    double st[shell::kState];
    shell::gp_state(dv, d.gpgeo + (size_t)gp * kGeoStride, d.t, st);
    const size_t stride = (size_t)d.nel * NG;
#pragma unroll
    for (int I = 0; I < kFq; ++I) d.Hg[(size_t)I * stride + gp] = shell::f_entry(st, I);
    if (d.pressure != 0.0) {  // -p w (a1 x a2) on the deformed surface: the follower load, as a force on the residual's side
        double n[3];
        shell::cross3(dv, dv + 3, n);
        const double s = -d.pressure * d.gpgeo[(size_t)gp * kGeoStride + 13];
#pragma unroll
        for (int i = 0; i < 3; ++i) d.Hg[(size_t)(kFq + i) * stride + gp] = s * n[i];
    }
}

// Residual, stage 2: one wave per element; lane (A, i) sums Fe[3A+i] = sum_g sum_a
// Psi_g[a][A] f_g[3a+i] (g outer, a inner).  The wave first stages its element's Gauss-point
// resultants (kFq or, with a follower pressure, kFqP rows of NG contiguous values) and its two
// basis tables (NQ x 3 x P1 values each, contiguous per element column / row) in LDS, every load
// in flight at once; the sums then read LDS only.  (Reading them inside the sum cost one memory
// latency per Gauss point.)
template <int P>
__global__ __launch_bounds__(256) void k_residual_tiles(const InstDesc *descs, int ninst, int total_el)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NQ = P1, NG = NQ * NQ, NT = NQ * 3 * P1;
    constexpr int FL = (kFqP * NG + 63) / 64, TL = (2 * NT + 63) / 64;  // loads per lane
    __shared__ double sF[4][kFqP * NG];
    __shared__ double sT[4][2 * NT];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int eg = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + w));
    if (eg >= total_el) return;
    const InstDesc &d = descs[find_inst(ninst, eg, [&](int m) { return (long long)descs[m].elem0; })];
    const int e = eg - d.elem0, ex = e % d.nx, ey = e / d.nx;
    const bool prs = d.pressure != 0.0;
    const int nf = (prs ? kFqP : kFq) * NG;
    const size_t stride = (size_t)d.nel * NG;
    const double *Fg = d.Hg + (size_t)e * NG;
    double vf[FL], vt[TL];
#pragma unroll
    for (int q = 0; q < FL; ++q) {
        const int idx = lane + 64 * q;
        vf[q] = idx < nf ? Fg[(size_t)(idx / NG) * stride + idx % NG] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < TL; ++q) {
        const int idx = lane + 64 * q;
        vt[q] = idx < NT ? d.tabx[(size_t)ex * NT + idx] : idx < 2 * NT ? d.taby[(size_t)ey * NT + idx - NT] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < FL; ++q)
        if (lane + 64 * q < kFqP * NG) sF[w][lane + 64 * q] = vf[q];
#pragma unroll
    for (int q = 0; q < TL; ++q)
        if (lane + 64 * q < 2 * NT) sT[w][lane + 64 * q] = vt[q];
    // the wave reads what its other lanes wrote: LDS accesses of one wave complete in order, and
    // the fence keeps the compiler from moving the reads above the writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    if (lane >= 3 * NB) return;
    const int A = lane / 3, i = lane % 3, ia = A % P1, ib = A / P1;
    const double *F = sF[w];
    // the follower-pressure term is a separate instantiation of the loop, so that the loop without
    // it keeps no branch
    auto sum = [&](auto with_pressure) {
        double acc = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int qx = g % NQ, qy = g / NQ;
            const double *tx = sT[w] + qx * 3 * P1;
            const double *ty = sT[w] + NT + qy * 3 * P1;
            const double psi[5] = {tx[P1 + ia] * ty[ib], tx[ia] * ty[P1 + ib], tx[2 * P1 + ia] * ty[ib],
                                   tx[ia] * ty[2 * P1 + ib], tx[P1 + ia] * ty[P1 + ib]};
#pragma unroll
            for (int a = 0; a < 5; ++a) acc += psi[a] * F[(3 * a + i) * NG + g];
            if constexpr (decltype(with_pressure)::value) acc += tx[ia] * ty[ib] * F[(kFq + i) * NG + g];
        }
        return acc;
    };
    d.Fe[(size_t)e * 3 * NB + lane] = prs ? sum(std::true_type{}) : sum(std::false_type{});
}

// Tangent, stage 1: one thread per Gauss point.  Derivative vectors of the deformed surface,
// the point's state (shell::gp_state) and the 120 entries I <= J of its symmetric 15 x 15 H_g,
// stored entry-major (Hg[q * ngp + gp]) so that a wave's stores are contiguous.  All indices
// into the state are compile-time constants after unrolling, so it lives in registers.
template <int P>
__global__ __launch_bounds__(128) void k_tangent_gp(const InstDesc *descs, int ninst, int total)
{
    constexpr int P1 = P + 1, NQ = P1, NG = NQ * NQ;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const InstDesc &d = descs[find_inst(ninst, gid, [&](int m) { return (long long)descs[m].elem0 * NG; })];
    const int gp = gid - d.elem0 * NG;
    const int e = gp / NG, g = gp % NG, ex = e % d.nx, ey = e / d.nx, qx = g % NQ, qy = g / NQ;
    const double *tx = d.tabx + ((size_t)(ex * NQ + qx) * 3) * P1;
    const double *ty = d.taby + ((size_t)(ey * NQ + qy) * 3) * P1;
    double dv[15];
    gp_derivatives<P>(d, ex, ey, tx, ty, dv);
    double st[shell::kState];
    shell::gp_state(dv, d.gpgeo + (size_t)gp * kGeoStride, d.t, st);
    const size_t stride = (size_t)d.nel * NG;
    double *H = d.Hg + gp;
    int q = 0;
#pragma unroll
    for (int I = 0; I < 15; ++I)
#pragma unroll
        for (int J = I; J < 15; ++J) H[(size_t)(q++) * stride] = shell::h_entry(st, I, J);
}

// index of H_g[I][J] among the stored entries I <= J (row-major upper triangle)
__device__ __forceinline__ int hq_index(int I, int J)
{
    const int r = I < J ? I : J, c = I < J ? J : I;
    return r * 15 - (r * (r - 1)) / 2 + (c - r);
}

// Tangent, stage 2: one wave per element, E elements per workgroup.  The element's H_g (NG x 120 entries) is read once,
// coalesced, into the wave's part of LDS.  K_e^{ik} = Psi^T Z^{ik} on 16 x 16 MFMA tiles
// (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), inner dimension kk = (g, a) over 5 NG; the
// B operand Z^{ik}[kk][B] = sum_b H_g[(a,i),(b,k)] Psi_g[b][B] is formed in the lane that feeds
// it, from H_g (LDS broadcasts) and basis values (tables, L1).  The six tiles i <= k are
// computed, and K_e's lower triangle is stored from them (the tiles i > k as transposes).
template <int P, typename T, int E>
__global__ __launch_bounds__(64 * E) void k_tangent_tiles(const InstDesc *descs, int ninst, int total_el)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NQ = P1, NG = NQ * NQ, NKK = ((5 * NG + 3) / 4) * 4;
    using M = Mfma16<T>;
    const int eg = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * E + (threadIdx.x >> 6)));
    if (eg >= total_el) return;
    const InstDesc &d = descs[find_inst(ninst, eg, [&](int m) { return (long long)descs[m].elem0; })];
    const int e = eg - d.elem0, ex = e % d.nx, ey = e / d.nx;
    const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
    const int ia = li % P1, ib = li / P1;
    const bool inA = li < NB;
    const size_t stride = (size_t)d.nel * NG;
    constexpr int kHs = kHq + 1;  // padded row of the LDS copy
    __shared__ double sH[E][NG][kHs];
    double(*sHw)[kHs] = sH[threadIdx.x >> 6];
    {
        const double *Hb = d.Hg + (size_t)e * NG;
        for (int idx = lane; idx < NG * kHq; idx += 64) {
            const int q = idx / NG, g = idx % NG;  // NG consecutive points of one entry: contiguous
            sHw[g][q] = Hb[(size_t)q * stride + g];
        }
        // the wave reads what its own lanes wrote: the LDS writes must be complete
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
    typename M::acc_t acc[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) acc[t] = typename M::acc_t{0, 0, 0, 0};
    if constexpr (NG % 4 == 0) {
        // inner index kk in steps of one component a (uniform over the wave: every H_g offset
        // below is an immediate) at the point g = 4 gi + kq of the lane's group, whose basis
        // values are formed once per gi instead of once per step
#pragma unroll 1
        for (int gi = 0; gi < NG / 4; ++gi) {
            const int g = 4 * gi + kq, qx = g % NQ, qy = g / NQ;
            const double *tx = d.tabx + ((size_t)(ex * NQ + qx) * 3) * P1;
            const double *ty = d.taby + ((size_t)(ey * NQ + qy) * 3) * P1;
            double psi[5];
            psi[0] = inA ? tx[P1 + ia] * ty[ib] : 0.0;
            psi[1] = inA ? tx[ia] * ty[P1 + ib] : 0.0;
            psi[2] = inA ? tx[2 * P1 + ia] * ty[ib] : 0.0;
            psi[3] = inA ? tx[ia] * ty[2 * P1 + ib] : 0.0;
            psi[4] = inA ? tx[P1 + ia] * ty[P1 + ib] : 0.0;
            const double *Hg = sHw[g];
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                const T aop = (T)psi[a];
                int t = 0;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int k = i; k < 3; ++k, ++t) {
                        double z = 0.0;
#pragma unroll
                        for (int b = 0; b < 5; ++b) z += Hg[hq_index(3 * a + i, 3 * b + k)] * psi[b];
                        acc[t] = M::run(aop, (T)z, acc[t]);
                    }
                __builtin_amdgcn_sched_barrier(0);  // one component's H_g reads in flight at a time
            }
        }
    } else
    for (int s = 0; s < NKK / 4; ++s) {
        const int kk = 4 * s + kq;
        const bool ok = inA && kk < 5 * NG;
        const int g = kk < 5 * NG ? kk / 5 : 0, a = kk % 5, qx = g % NQ, qy = g / NQ;
        const double *tx = d.tabx + ((size_t)(ex * NQ + qx) * 3) * P1;
        const double *ty = d.taby + ((size_t)(ey * NQ + qy) * 3) * P1;
        double psi[5];
        psi[0] = ok ? tx[P1 + ia] * ty[ib] : 0.0;
        psi[1] = ok ? tx[ia] * ty[P1 + ib] : 0.0;
        psi[2] = ok ? tx[2 * P1 + ia] * ty[ib] : 0.0;
        psi[3] = ok ? tx[ia] * ty[2 * P1 + ib] : 0.0;
        psi[4] = ok ? tx[P1 + ia] * ty[P1 + ib] : 0.0;
        const double pa = a == 0 ? psi[0] : a == 1 ? psi[1] : a == 2 ? psi[2] : a == 3 ? psi[3] : psi[4];
        const T aop = (T)pa;
        const double *Hg = sHw[g];
        int t = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = i; k < 3; ++k, ++t) {
                double z = 0.0;
#pragma unroll
                for (int b = 0; b < 5; ++b) z += Hg[hq_index(3 * a + i, 3 * b + k)] * psi[b];
                acc[t] = M::run(aop, ok ? (T)z : T(0), acc[t]);  // padded steps add exact zeros
            }
    }
    // element matrix, local dof 3A + i, column-major with leading dimension kKeLd: only its lower
    // triangle (row >= column), which is all the band gather reads (the local dof order is the
    // global one), so each entry (3A+i, 3B+k) of tile i <= k is stored once, as itself or as its
    // transpose, whichever lies on or below the diagonal
    T *Kout = (T *)d.Ke + (size_t)e * kKeTile;
    int t = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = i; k < 3; ++k, ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = M::row(lane, v);  // A; the lane's column li is B
                const int r1 = 3 * row + i, c1 = 3 * li + k;
                if (r1 >= c1)
                    Kout[c1 * kKeLd + r1] = acc[t][v];
                else if (k != i)
                    Kout[r1 * kKeLd + c1] = acc[t][v];
            }
}

// (e_k x a)_i
__device__ __forceinline__ double ek_cross(int k, const double *a, int i)
{
    if (i == k) return 0.0;
    const int j = 3 - i - k;
    return ((i - k + 3) % 3 == 1) ? -a[j] : a[j];
}

// Follower-pressure part of the element tangent, K_e[r][c] -= dP_r / du_c with
// dP_{A,i} / du_{B,k} = p sum_g w_g N_A (N_B,1 (e_k x a2) + N_B,2 (a1 x e_k))_i (a1, a2 deformed),
// on the lower triangle r >= c that the band gather reads: the follower tangent is not
// symmetric, and its lower triangle is what a lower-triangle LDL^T of the reference's assembled
// matrix reads.  One wave per element, lane r = 3A + i; the element's Gauss-point data (a1, a2,
// p w, N, N,1, N,2) staged in LDS.
template <int P, typename T>
__global__ __launch_bounds__(64) void k_pressure_tiles(const InstDesc *descs, int ninst, int total_el)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NQ = P1, NG = NQ * NQ;
    const int eg = (int)blockIdx.x;
    if (eg >= total_el) return;
    const InstDesc &d = descs[find_inst(ninst, eg, [&](int m) { return (long long)descs[m].elem0; })];
    if (d.pressure == 0.0) return;
    const int e = eg - d.elem0, ex = e % d.nx, ey = e / d.nx;
    const int lane = (int)threadIdx.x;
    __shared__ double sN[NG][3][NB];  // N, N_,1, N_,2 per point
    __shared__ double sA[NG][7];      // a1, a2, p w per point
    for (int idx = lane; idx < NG * NB; idx += 64) {
        const int g = idx / NB, A = idx % NB, qx = g % NQ, qy = g / NQ, ia = A % P1, ib = A / P1;
        const double *tx = d.tabx + ((size_t)(ex * NQ + qx) * 3) * P1;
        const double *ty = d.taby + ((size_t)(ey * NQ + qy) * 3) * P1;
        sN[g][0][A] = tx[ia] * ty[ib];
        sN[g][1][A] = tx[P1 + ia] * ty[ib];
        sN[g][2][A] = tx[ia] * ty[P1 + ib];
    }
    __syncthreads();
    if (lane < NG) {
        const int g = lane;
        double a1[3] = {0.0, 0.0, 0.0}, a2[3] = {0.0, 0.0, 0.0};
        for (int A = 0; A < NB; ++A) {
            const int cp = (ex + A % P1) + d.mx * (ey + A / P1);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double x = d.ctrl0[3 * cp + c] + d.u_full[3 * cp + c];
                a1[c] += sN[g][1][A] * x;
                a2[c] += sN[g][2][A] * x;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            sA[g][c] = a1[c];
            sA[g][3 + c] = a2[c];
        }
        sA[g][6] = d.pressure * d.gpgeo[((size_t)e * NG + g) * kGeoStride + 13];
    }
    __syncthreads();
    if (lane >= 3 * NB) return;
    const int A = lane / 3, i = lane % 3, r = lane;
    T *Kout = (T *)d.Ke + (size_t)e * kKeTile;
    for (int c = 0; c <= r; ++c) {
        const int B = c / 3, k = c % 3;
        double acc = 0.0;
        for (int g = 0; g < NG; ++g)
            acc += sA[g][6] * sN[g][0][A] * (sN[g][1][B] * ek_cross(k, sA[g] + 3, i) - sN[g][2][B] * ek_cross(k, sA[g], i));
        Kout[c * kKeLd + r] = (T)((double)Kout[c * kKeLd + r] - acc);
    }
}

// consistent mass tiles: M_e^{ii}[A][B] = sum_g rho t N_A N_B wdA
template <int P, typename T>
__global__ __launch_bounds__(256) void k_mass_elem(const InstDesc *descs, int ninst)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NQ = P1, NG = NQ * NQ;
    __shared__ double sN[NG][16];
    __shared__ double sW[NG];
    const InstDesc &d = descs[find_inst(ninst, blockIdx.x, [&](int m) { return (long long)descs[m].elem0; })];
    const int e = blockIdx.x - d.elem0, ex = e % d.nx, ey = e / d.nx;
    for (int idx = threadIdx.x; idx < NG * 16; idx += blockDim.x) {
        int g = idx / 16, A = idx % 16;
        double v = 0.0;
        if (A < NB) {
            int qx = g % NQ, qy = g / NQ, ia = A % P1, ib = A / P1;
            v = d.tabx[((size_t)(ex * NQ + qx) * 3) * P1 + ia] * d.taby[((size_t)(ey * NQ + qy) * 3) * P1 + ib];
        }
        sN[g][A] = v;
    }
    for (int g = threadIdx.x; g < NG; g += blockDim.x) sW[g] = d.gpgeo[((size_t)e * NG + g) * kGeoStride];
    __syncthreads();
    T *Kout = (T *)d.Ke + (size_t)e * kKeTile;
    for (int idx = threadIdx.x; idx < 256; idx += blockDim.x) {
        int A = idx / 16, B = idx % 16;
        double acc = 0.0;
        for (int g = 0; g < NG; ++g) acc += d.rho * d.t * sN[g][A] * sN[g][B] * sW[g];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) Kout[(3 * B + k) * kKeLd + 3 * A + i] = (i == k) ? (T)acc : (T)0;
    }
}

// dead surface force per element: Fe[(A,i)] = sum_g N_A f_i wdA
template <int P>
__global__ __launch_bounds__(64) void k_surface_elem(ElemArgs ar, double f0, double f1, double f2, double *Fe)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NQ = P1, NG = NQ * NQ;
    int e = blockIdx.x, ex = e % ar.nx, ey = e / ar.nx;
    double f[3] = {f0, f1, f2};
    for (int idx = threadIdx.x; idx < NB * 3; idx += blockDim.x) {
        int A = idx / 3, i = idx % 3, ia = A % P1, ib = A / P1;
        double acc = 0.0;
        for (int g = 0; g < NG; ++g) {
            int qx = g % NQ, qy = g / NQ;
            double N = ar.tabx[((size_t)(ex * NQ + qx) * 3) * P1 + ia] * ar.taby[((size_t)(ey * NQ + qy) * 3) * P1 + ib];
            acc += N * f[i] * ar.gpgeo[((size_t)e * NG + g) * kGeoStride];
        }
        Fe[(size_t)e * 3 * NB + idx] = acc;
    }
}

// undeformed geometry per Gauss point
template <int P>
__global__ void k_setup_geometry(ElemArgs ar, const double *wq, double E, double nu, int ngp_total, double *gpgeo)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NQ = P1, NG = NQ * NQ;
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= ngp_total) return;
    int e = gid / NG, g = gid % NG, ex = e % ar.nx, ey = e / ar.nx, qx = g % NQ, qy = g / NQ;
    const double *tx = ar.tabx + ((size_t)(ex * NQ + qx) * 3) * P1;
    const double *ty = ar.taby + ((size_t)(ey * NQ + qy) * 3) * P1;
    double dd[5][3] = {};
    for (int A = 0; A < NB; ++A) {
        int ia = A % P1, ib = A / P1;
        int cp = (ex + ia) + ar.mx * (ey + ib);
        double psi[5] = {tx[P1 + ia] * ty[ib], tx[ia] * ty[P1 + ib], tx[2 * P1 + ia] * ty[ib], tx[ia] * ty[2 * P1 + ib],
                         tx[P1 + ia] * ty[P1 + ib]};
        for (int a = 0; a < 5; ++a)
            for (int c = 0; c < 3; ++c) dd[a][c] += psi[a] * ar.ctrl0[3 * cp + c];
    }
    double at[3], a3[3];
    shell::cross3(dd[0], dd[1], at);
    double J0 = sqrt(shell::dot3(at, at));
    for (int i = 0; i < 3; ++i) a3[i] = at[i] / J0;
    double g11 = shell::dot3(dd[0], dd[0]), g22 = shell::dot3(dd[1], dd[1]), g12 = shell::dot3(dd[0], dd[1]);
    double det = g11 * g22 - g12 * g12;
    double ci[2][2] = {{g22 / det, -g12 / det}, {-g12 / det, g11 / det}};
    const int vi[3][2] = {{0, 0}, {1, 1}, {0, 1}};
    double Em = E / (1.0 - nu * nu);
    double C[3][3];
    for (int I = 0; I < 3; ++I)
        for (int J = 0; J < 3; ++J) {
            int al = vi[I][0], be = vi[I][1], ga = vi[J][0], de = vi[J][1];
            C[I][J] = Em * (nu * ci[al][be] * ci[ga][de] + 0.5 * (1.0 - nu) * (ci[al][ga] * ci[be][de] + ci[al][de] * ci[be][ga]));
        }
    double *o = gpgeo + (size_t)gid * kGeoStride;
    o[0] = J0 * wq[g];
    o[1] = C[0][0];
    o[2] = C[1][1];
    o[3] = C[2][2];
    o[4] = C[0][1];
    o[5] = C[0][2];
    o[6] = C[1][2];
    o[7] = g11;
    o[8] = g22;
    o[9] = g12;
    o[10] = shell::dot3(dd[2], a3);
    o[11] = shell::dot3(dd[3], a3);
    o[12] = shell::dot3(dd[4], a3);
    o[13] = wq[g];  // the Gauss weight on the parameter domain (follower pressure)
    o[14] = o[15] = 0.0;
}

// Deterministic gather of element vectors into free dofs (fixed element order ey, ex).
// mode 0: R = sum ; mode 1 (ALM): R = sum - lambda F ; mode 2 (static): R = lambda F - sum ;
// mode 3: R = -sum.  With a follower pressure the element sums hold F_int(u) - P(u), and the ALM
// form adds P(0) (r = F_int - lambda F - (P(u) - P(0)), the reference's Force - lambda Force -
// rhs(u) with Force = F_ext(0)); the static form is its negative.
__global__ void k_vector_gather(const InstDesc *descs, int ninst, int total)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const InstDesc &d = descs[find_inst(ninst, gid, [&](int m) { return (long long)descs[m].dof0; })];
    int I = gid - d.dof0;
    int p = d.p, P1 = p + 1, NB = P1 * P1;
    int g = d.free2full[I], cp = g / 3, c = g % 3, i = cp % d.mx, j = cp / d.mx;
    int exlo = max(i - p, 0), exhi = min(i, d.nx - 1), eylo = max(j - p, 0), eyhi = min(j, d.ny - 1);
    double acc = 0.0;
    for (int ey = eylo; ey <= eyhi; ++ey)
        for (int ex = exlo; ex <= exhi; ++ex) {
            int e = ex + d.nx * ey, A = (i - ex) + P1 * (j - ey);
            acc += d.Fe[(size_t)e * 3 * NB + 3 * A + c];
        }
    const double p0 = d.P0 ? d.P0[I] : 0.0;
    if (d.mode == 1) acc = acc - d.lambda * d.F[I] + p0;
    else if (d.mode == 2) acc = d.lambda * d.F[I] - acc - p0;
    else if (d.mode == 3) acc = -acc;
    d.R[I] = acc;
}

// Deterministic gather of the element matrices into the band, by control-point pairs.  A wave of
// 64 lanes takes 64 consecutive pairs (cpJ, cpI) of one instance: cpI = cpJ + di + mx dj with
// dj in [0, P] and di in [-P, P] (di >= 0 when dj = 0) -- every control point within one
// element's reach of cpJ that comes at or after it in the numbering, so every band slot any
// element reaches belongs to exactly one pair (free dofs are numbered 3 cp + c in increasing order,
// and those pairs are the lower triangle).  The lane sums the pair's 3 x 3 block over the elements
// holding both points, in a fixed element order (rows ey, then columns ex), and stores the slots
// of its free dofs.  Slots no element reaches are never written: the band must hold 0 there (a
// fresh band, or one zeroed by the launcher).  Reads each stored element entry once.
template <int P, typename T>
__global__ __launch_bounds__(256) void k_band_gather(const InstDesc *descs, int ninst)
{
    constexpr int P1 = P + 1, NPAIR = P1 + P * (2 * P + 1);
    const long long chunk = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const InstDesc &d = descs[find_inst(ninst, chunk, [&](int m) { return descs[m].chunk0; })];
    const long long t = (chunk - d.chunk0) * 64 + (threadIdx.x & 63);
    const int my = d.ny + P, ncp = d.mx * my;
    if (t >= (long long)ncp * NPAIR) return;
    const int cpJ = (int)(t / NPAIR), k = (int)(t % NPAIR);
    const int dj = k < P1 ? 0 : 1 + (k - P1) / (2 * P + 1);
    const int di = k < P1 ? k : (k - P1) % (2 * P + 1) - P;
    const int iJ = cpJ % d.mx, jJ = cpJ / d.mx, iI = iJ + di, jI = jJ + dj;
    if (iI < 0 || iI >= d.mx || jI >= my) return;
    const int cpI = iI + d.mx * jI;
    int fI[3], fJ[3];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        fI[c] = d.free_idx[3 * cpI + c];
        fJ[c] = d.free_idx[3 * cpJ + c];
    }
#pragma unroll
    for (int cJ = 0; cJ < 3; ++cJ)
#pragma unroll
        for (int cI = 0; cI < 3; ++cI) any |= fI[cI] >= 0 && fJ[cJ] >= 0 && !(cpI == cpJ && cI < cJ);
    if (!any) return;
    const int ex0 = max(iI, iJ) - P, exhi = min(min(iI, iJ), d.nx - 1);
    const int ey0 = jI - P, eyhi = min(jJ, d.ny - 1);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(d.Ke, (unsigned)((size_t)d.nel * kKeTile * sizeof(T)));
    double acc[3][3];
#pragma unroll
    for (int cJ = 0; cJ < 3; ++cJ)
#pragma unroll
        for (int cI = 0; cI < 3; ++cI) acc[cJ][cI] = 0.0;
#pragma unroll
    for (int ty = 0; ty < P1; ++ty)
#pragma unroll
        for (int tx = 0; tx < P1; ++tx) {
            const int ex = ex0 + tx, ey = ey0 + ty;
            const bool ok = ex >= 0 && ex <= exhi && ey >= 0 && ey <= eyhi;
            const int A = (iI - ex) + P1 * (jI - ey), B = (iJ - ex) + P1 * (jJ - ey);
            const size_t base = (size_t)(ex + d.nx * ey) * kKeTile;
            T v[3][3];
#pragma unroll
            for (int cJ = 0; cJ < 3; ++cJ)
#pragma unroll
                for (int cI = 0; cI < 3; ++cI)
                    v[cJ][cI] = load_if<T>(ok && !(cpI == cpJ && cI < cJ), rs,
                                           (unsigned)((base + (size_t)(3 * B + cJ) * kKeLd + 3 * A + cI) * sizeof(T)));
#pragma unroll
            for (int cJ = 0; cJ < 3; ++cJ)
#pragma unroll
                for (int cI = 0; cI < 3; ++cI) acc[cJ][cI] += (double)v[cJ][cI];
        }
    const long long ld = (long long)d.bw + 1;
#pragma unroll
    for (int cJ = 0; cJ < 3; ++cJ)
#pragma unroll
        for (int cI = 0; cI < 3; ++cI)
            if (fI[cI] >= 0 && fJ[cJ] >= 0 && !(cpI == cpJ && cI < cJ))
                ((T *)d.ab)[(long long)fJ[cJ] * ld + (fI[cI] - fJ[cJ])] = (T)acc[cJ][cI];
}

__global__ void k_point_loads(int npoint, int p, int mx, const int *free_idx, const double *pl_N, const int *pl_span,
                              const double *pl_F, double *F)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int P1 = p + 1;
    for (int l = 0; l < npoint; ++l)
        for (int b = 0; b < P1; ++b)
            for (int a = 0; a < P1; ++a) {
                int cp = (pl_span[2 * l] - p + a) + mx * (pl_span[2 * l + 1] - p + b);
                double N = pl_N[(size_t)l * 2 * P1 + a] * pl_N[(size_t)l * 2 * P1 + P1 + b];
                for (int c = 0; c < 3; ++c) {
                    int f = free_idx[3 * cp + c];
                    if (f >= 0) F[f] += N * pl_F[3 * l + c];
                }
            }
}

ElemArgs host_elem_args(const Problem &pb)
{
    ElemArgs a;
    a.p = pb.p;
    a.nx = pb.nx;
    a.ny = pb.ny;
    a.mx = pb.mx;
    a.my = pb.my;
    a.t = pb.t;
    a.ctrl0 = pb.ctrl0;
    a.u_full = pb.u_full;
    a.tabx = pb.tabx;
    a.taby = pb.taby;
    a.gpgeo = pb.gpgeo;
    return a;
}

struct Totals {
    long long nel = 0, nfree = 0, ndof = 0, slots = 0, chunks = 0;
};

// prefix offsets of a batch
Totals prefix(std::vector<InstDesc> &h)
{
    Totals t;
    for (auto &d : h) {
        d.elem0 = (int)t.nel;
        d.dof0 = (int)t.nfree;
        d.full0 = (int)t.ndof;
        const long long npair = (long long)(d.nx + d.p) * (d.ny + d.p) * ((d.p + 1) + d.p * (2 * d.p + 1));
        d.chunk0 = t.chunks;
        t.chunks += (npair + 63) / 64;
        t.nel += d.nel;
        t.nfree += d.nfree;
        t.ndof += d.ndof;
        t.slots += (long long)d.nfree * (d.bw + 1);
    }
    return t;
}

int upload_descs(hipStream_t s, InstDesc *d_desc, const std::vector<InstDesc> &h)
{
    KSH_HIP(hipMemcpyAsync(d_desc, h.data(), sizeof(InstDesc) * h.size(), hipMemcpyHostToDevice, s));
    return KSH_SUCCESS;
}

int check_batch(const std::vector<InstDesc> &h, bool band)
{
    if (h.empty()) return fail(KSH_OTHER_ERROR, "empty batch");
    for (auto &d : h) {
        if (d.p != h[0].p) return fail(KSH_OTHER_ERROR, "batch mixes spline degrees");
        if (band && d.sprec != h[0].sprec) return fail(KSH_OTHER_ERROR, "batch mixes band precisions");
        if (band && (size_t)d.nel * kKeTile * sizeof(double) >= 0x80000000ull)
            return fail(KSH_OTHER_ERROR, "element matrices of 2 GiB or more: beyond the gather's 32-bit offsets");
    }
    return KSH_SUCCESS;
}

int run_residual(hipStream_t s, InstDesc *d_desc, std::vector<InstDesc> &h, bool construct)
{
    int rc;
    if ((rc = check_batch(h, false))) return rc;
    Totals t = prefix(h);
    const int n = (int)h.size();
    if ((rc = upload_descs(s, d_desc, h))) return rc;
    if (construct)
        hipLaunchKernelGGL(k_construct_solution, dim3(nblk(t.ndof, 256)), dim3(256), 0, s, d_desc, n, (int)t.ndof);
    {
        ProfScope ps("residual_elem", s);
        const long long ng = (h[0].p + 1LL) * (h[0].p + 1LL), ngp = t.nel * ng;
        if (ngp >= (1LL << 31)) return fail(KSH_OTHER_ERROR, "batch has 2^31 or more Gauss points");
        const dim3 gg(nblk(ngp, 128)), ge(nblk(t.nel, 4));
        if (h[0].p == 2) {
            hipLaunchKernelGGL(k_residual_gp<2>, gg, dim3(128), 0, s, d_desc, n, (int)ngp);
            hipLaunchKernelGGL(k_residual_tiles<2>, ge, dim3(256), 0, s, d_desc, n, (int)t.nel);
        } else {
            hipLaunchKernelGGL(k_residual_gp<3>, gg, dim3(128), 0, s, d_desc, n, (int)ngp);
            hipLaunchKernelGGL(k_residual_tiles<3>, ge, dim3(256), 0, s, d_desc, n, (int)t.nel);
        }
    }
    KSH_HIP(hipGetLastError());
    ProfScope ps("vector_gather", s);
    hipLaunchKernelGGL(k_vector_gather, dim3(nblk(t.nfree, 256)), dim3(256), 0, s, d_desc, n, (int)t.nfree);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int run_tangent(hipStream_t s, InstDesc *d_desc, std::vector<InstDesc> &h, bool construct, bool mass)
{
    int rc;
    if ((rc = check_batch(h, true))) return rc;
    Totals t = prefix(h);
    const int n = (int)h.size();
    if ((rc = upload_descs(s, d_desc, h))) return rc;
    if (construct)
        hipLaunchKernelGGL(k_construct_solution, dim3(nblk(t.ndof, 256)), dim3(256), 0, s, d_desc, n, (int)t.ndof);
    const bool f64 = h[0].sprec == KSH_FP64, p2 = h[0].p == 2;
    const dim3 g((unsigned)t.nel), b(256);
    if (mass) {
        if (p2 && f64) hipLaunchKernelGGL((k_mass_elem<2, double>), g, b, 0, s, d_desc, n);
        else if (p2) hipLaunchKernelGGL((k_mass_elem<2, float>), g, b, 0, s, d_desc, n);
        else if (f64) hipLaunchKernelGGL((k_mass_elem<3, double>), g, b, 0, s, d_desc, n);
        else hipLaunchKernelGGL((k_mass_elem<3, float>), g, b, 0, s, d_desc, n);
    } else {
        const long long ng = (h[0].p + 1LL) * (h[0].p + 1LL), ngp = t.nel * ng;
        if (t.nel * ng >= (1LL << 31)) return fail(KSH_OTHER_ERROR, "batch has 2^31 or more Gauss points");
        {
            ProfScope ps("tangent_gp", s);
            ps.set_work((double)ngp * (kHq + kGeoStride) * 8.0);  // H_g written, undeformed geometry read (bytes)
            if (p2) hipLaunchKernelGGL(k_tangent_gp<2>, dim3(nblk(ngp, 128)), dim3(128), 0, s, d_desc, n, (int)ngp);
            else hipLaunchKernelGGL(k_tangent_gp<3>, dim3(nblk(ngp, 128)), dim3(128), 0, s, d_desc, n, (int)ngp);
        }
        ProfScope ps("tangent_elem", s);
        double work = 0.0;  // B^T D B count per element and Gauss point: 72 n_e + 12 n_e^2
        for (const auto &d : h) {
            const double nge = (d.p + 1.0) * (d.p + 1.0), ne = 3.0 * nge;
            work += (double)d.nel * nge * (72.0 * ne + 12.0 * ne * ne);
        }
        ps.set_work(work);
        // variants: 1, 2 or 4 elements (waves) per workgroup; the same arithmetic per element
        const int tn = (int)t.nel;
        auto launch_tiles = [&](int v) {
            const int E = 1 << v;
            const dim3 ge(nblk(t.nel, E)), be(64 * E);
#define KSH_TILES(PP, TT, EE) hipLaunchKernelGGL((k_tangent_tiles<PP, TT, EE>), ge, be, 0, s, d_desc, n, tn)
            if (p2 && f64) { if (E == 1) KSH_TILES(2, double, 1); else if (E == 2) KSH_TILES(2, double, 2); else KSH_TILES(2, double, 4); }
            else if (p2) { if (E == 1) KSH_TILES(2, float, 1); else if (E == 2) KSH_TILES(2, float, 2); else KSH_TILES(2, float, 4); }
            else if (f64) { if (E == 1) KSH_TILES(3, double, 1); else if (E == 2) KSH_TILES(3, double, 2); else KSH_TILES(3, double, 4); }
            else { if (E == 1) KSH_TILES(3, float, 1); else if (E == 2) KSH_TILES(3, float, 2); else KSH_TILES(3, float, 4); }
#undef KSH_TILES
        };
        tune_launch("tangent_tiles", tune_bucket((double)t.nel) * 4 + (p2 ? 0 : 1) * 2 + (f64 ? 1 : 0), 3, 2, s,
                    launch_tiles);
        if (std::any_of(h.begin(), h.end(), [](const InstDesc &d) { return d.pressure != 0.0; })) {
            const dim3 gp((unsigned)t.nel), bp(64);
            if (p2 && f64) hipLaunchKernelGGL((k_pressure_tiles<2, double>), gp, bp, 0, s, d_desc, n, tn);
            else if (p2) hipLaunchKernelGGL((k_pressure_tiles<2, float>), gp, bp, 0, s, d_desc, n, tn);
            else if (f64) hipLaunchKernelGGL((k_pressure_tiles<3, double>), gp, bp, 0, s, d_desc, n, tn);
            else hipLaunchKernelGGL((k_pressure_tiles<3, float>), gp, bp, 0, s, d_desc, n, tn);
        }
    }
    KSH_HIP(hipGetLastError());
    ProfScope ps(mass ? "mass_gather" : "band_gather", s);
    {
        // the lower triangle of each element matrix read (all the gather reads), band slots written
        const double es = f64 ? 8.0 : 4.0;
        double work = 0.0;
        for (const auto &d : h) {
            const double ne = 3.0 * (d.p + 1) * (d.p + 1);
            work += (double)d.nel * ne * (ne + 1.0) / 2.0 * es + (double)d.nfree * (d.bw + 1.0) * es;
        }
        ps.set_work(work);
    }
    // bands that may hold values in slots no element reaches are cleared first
    for (const auto &d : h)
        if (!d.skip_zero) KSH_HIP(hipMemsetAsync(d.ab, 0, (size_t)d.nfree * (d.bw + 1) * (f64 ? 8 : 4), s));
    // variants: 64, 128 or 256 threads per block (one control-point pair per thread either way)
    auto launch_gather = [&](int v) {
        const unsigned wpb = 1u << v, bs = 64u * wpb;
        const dim3 gs((unsigned)((t.chunks + wpb - 1) / wpb)), bb(bs);
        if (p2 && f64) hipLaunchKernelGGL((k_band_gather<2, double>), gs, bb, 0, s, d_desc, n);
        else if (p2) hipLaunchKernelGGL((k_band_gather<2, float>), gs, bb, 0, s, d_desc, n);
        else if (f64) hipLaunchKernelGGL((k_band_gather<3, double>), gs, bb, 0, s, d_desc, n);
        else hipLaunchKernelGGL((k_band_gather<3, float>), gs, bb, 0, s, d_desc, n);
    };
    tune_launch(mass ? "mass_gather" : "band_gather", tune_bucket((double)t.slots) * 4 + (p2 ? 0 : 1) * 2 + (f64 ? 1 : 0), 3, 1,
                s, launch_gather);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// descriptors and batched launchers
// ---------------------------------------------------------------------------------------------
InstDesc make_desc(const Problem &pb)
{
    InstDesc d = {};
    d.p = pb.p;
    d.nx = pb.nx;
    d.ny = pb.ny;
    d.mx = pb.mx;
    d.nel = pb.nel;
    d.nfree = pb.nfree;
    d.bw = pb.bw;
    d.ndof = pb.ndof;
    d.t = pb.t;
    d.rho = pb.rho;
    d.ctrl0 = pb.ctrl0;
    d.tabx = pb.tabx;
    d.taby = pb.taby;
    d.gpgeo = pb.gpgeo;
    d.F = pb.F;
    d.free_idx = pb.free_idx;
    d.free2full = pb.free2full;
    d.u_full = pb.u_full;
    d.Fe = pb.Fe;
    d.Ke = pb.Ke;
    d.Hg = pb.Hg;
    d.pressure = pb.pressure;
    d.P0 = pb.pressure != 0.0 ? pb.P0 : nullptr;
    return d;
}

int batch_construct(hipStream_t s, InstDesc *d_desc, const std::vector<InstDesc> &hin)
{
    std::vector<InstDesc> h = hin;
    Totals t = prefix(h);
    int rc;
    if ((rc = upload_descs(s, d_desc, h))) return rc;
    hipLaunchKernelGGL(k_construct_solution, dim3(nblk(t.ndof, 256)), dim3(256), 0, s, d_desc, (int)h.size(), (int)t.ndof);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

// constructSolution + element residuals + gather, for every instance of the batch
int batch_residual(hipStream_t s, InstDesc *d_desc, const std::vector<InstDesc> &hin)
{
    std::vector<InstDesc> h = hin;
    return run_residual(s, d_desc, h, true);
}

// constructSolution + element tangents + band gather, for every instance of the batch
int batch_tangent(hipStream_t s, InstDesc *d_desc, const std::vector<InstDesc> &hin)
{
    std::vector<InstDesc> h = hin;
    return run_tangent(s, d_desc, h, true, false);
}

// ---------------------------------------------------------------------------------------------
// single-problem launchers (a batch of one)
// ---------------------------------------------------------------------------------------------
int launch_setup_geometry(Problem &pb)
{
    ElemArgs ar = host_elem_args(pb);
    int total = pb.nel * pb.nq * pb.nq;
    if (pb.p == 2)
        hipLaunchKernelGGL(k_setup_geometry<2>, dim3(nblk(total, 128)), dim3(128), 0, pb.stream, ar, pb.wq, pb.E, pb.nu,
                           total, pb.gpgeo);
    else
        hipLaunchKernelGGL(k_setup_geometry<3>, dim3(nblk(total, 128)), dim3(128), 0, pb.stream, ar, pb.wq, pb.E, pb.nu,
                           total, pb.gpgeo);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int launch_construct_solution(Problem &pb, const double *u_free_dev)
{
    InstDesc d = make_desc(pb);
    d.u_free = u_free_dev;
    return batch_construct(pb.stream, pb.d_desc, {d});
}

int launch_residual(Problem &pb, double lambda, int conv, double *R_dev)
{
    InstDesc d = make_desc(pb);
    d.lambda = lambda;
    d.mode = (conv == KSH_RES_ALM) ? 1 : 2;
    d.R = R_dev;
    std::vector<InstDesc> h = {d};
    return run_residual(pb.stream, pb.d_desc, h, false);
}

// residual with caller-chosen partial-result storage (u_full: ndof, Hg: kFq ngp, or kFqP ngp with a follower
// pressure, Fe: 3 nb nel doubles)
int run_residual_into(Problem &pb, const double *u_free_dev, double lambda, int conv, double *R_dev, double *u_full,
                      double *Hg, double *Fe)
{
    InstDesc d = make_desc(pb);
    d.u_free = u_free_dev;
    d.u_full = u_full;
    d.Hg = Hg;
    d.Fe = Fe;
    d.lambda = lambda;
    d.mode = (conv == KSH_RES_ALM) ? 1 : 2;
    d.R = R_dev;
    std::vector<InstDesc> h = {d};
    return run_residual(pb.stream, pb.d_desc, h, true);
}

// y += x
__global__ void k_add(int n, const double *x, double *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// P(0), the follower pressure's load on the undeformed surface, into pb.P0 (the residual kernels at
// u = 0, where F_int = 0, gathered as -(F_int - P)), then F += P0: the reference's Force =
// assemble() rhs at zero displacement holds it (examples/gsThinShell_ArcLength.cpp:963-965)
int launch_pressure_rest(Problem &pb)
{
    KSH_HIP(hipMemsetAsync(pb.u_free, 0, sizeof(double) * pb.nfree, pb.stream));
    InstDesc d = make_desc(pb);
    d.u_free = pb.u_free;
    d.mode = 3;
    d.R = pb.P0;
    d.P0 = nullptr;
    std::vector<InstDesc> h = {d};
    int rc;
    if ((rc = run_residual(pb.stream, pb.d_desc, h, true))) return rc;
    hipLaunchKernelGGL(k_add, dim3(nblk(pb.nfree, 256)), dim3(256), 0, pb.stream, pb.nfree, (const double *)pb.P0, pb.F);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int launch_external_force(Problem &pb)
{
    ElemArgs ar = host_elem_args(pb);
    KSH_HIP(hipMemsetAsync(pb.F, 0, sizeof(double) * pb.nfree, pb.stream));
    if (pb.fsurf[0] != 0.0 || pb.fsurf[1] != 0.0 || pb.fsurf[2] != 0.0) {
        if (pb.p == 2)
            hipLaunchKernelGGL(k_surface_elem<2>, dim3(pb.nel), dim3(64), 0, pb.stream, ar, pb.fsurf[0], pb.fsurf[1],
                               pb.fsurf[2], pb.Fe);
        else
            hipLaunchKernelGGL(k_surface_elem<3>, dim3(pb.nel), dim3(64), 0, pb.stream, ar, pb.fsurf[0], pb.fsurf[1],
                               pb.fsurf[2], pb.Fe);
        KSH_HIP(hipGetLastError());
        InstDesc d = make_desc(pb);
        d.mode = 0;
        d.R = pb.F;
        std::vector<InstDesc> h = {d};
        prefix(h);
        int rc;
        if ((rc = upload_descs(pb.stream, pb.d_desc, h))) return rc;
        hipLaunchKernelGGL(k_vector_gather, dim3(nblk(pb.nfree, 256)), dim3(256), 0, pb.stream, pb.d_desc, 1, pb.nfree);
        KSH_HIP(hipGetLastError());
    }
    if (pb.npoint > 0) {
        hipLaunchKernelGGL(k_point_loads, dim3(1), dim3(1), 0, pb.stream, pb.npoint, pb.p, pb.mx, pb.free_idx, pb.pl_N,
                           pb.pl_span, pb.pl_F, pb.F);
        KSH_HIP(hipGetLastError());
    }
    return KSH_SUCCESS;
}

int launch_tangent(Problem &pb, Matrix &K)
{
    InstDesc d = make_desc(pb);
    d.ab = K.ab;
    d.sprec = K.sprec;
    d.skip_zero = K.ab_clean ? 1 : 0;
    std::vector<InstDesc> h = {d};
    const int rc = run_tangent(pb.stream, pb.d_desc, h, false, false);
    K.ab_clean = rc == KSH_SUCCESS;
    return rc;
}

int launch_mass(Problem &pb, Matrix &M)
{
    InstDesc d = make_desc(pb);
    d.ab = M.ab;
    d.sprec = M.sprec;
    d.skip_zero = M.ab_clean ? 1 : 0;
    std::vector<InstDesc> h = {d};
    const int rc = run_tangent(pb.stream, pb.d_desc, h, false, true);
    M.ab_clean = rc == KSH_SUCCESS;
    return rc;
}

}  // namespace ksh
