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
// written, so launches on one stream need no other synchronization.  An opt-in form runs the
// same steps in one cooperative launch with a grid barrier between them (see persist_requested).  HBM traffic: the band once
// per sweep, 2 n (bw+1) entries per solve.  Replaces SimplicialLDLT::solve
// (src/gsEigenSolvers/gsEigenProblemBase.hpp:152-216 drives it through the shift-invert operator).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "ksh_devutil.hpp"
#include "ksh_internal.hpp"

namespace ksh {

namespace {

constexpr int kSB = 64;  // rows per step = threads per workgroup

// Band entries of one step: they do not depend on the solution, so a step's loads can be issued
// before the data it combines them with is final (before the previous step's barrier, in the
// one-launch form).
template <typename T>
struct StepRegs {
    T l[kSB];  // diagonal block (forward: row lane; backward: column lane)
    T u[kSB];  // this workgroup's off-diagonal rows / columns
    T d;       // pivot (forward)
};

// forward, block c0: lane ii holds row c0+ii of the diagonal block, L[c0+ii][c0+jj] for jj < ii;
// the workgroup's rows c0+r, r = w + 64 wg + lane, their entries L[c0+r][c0+kk]
template <typename T>
__device__ __forceinline__ void fwd_load_diag(const T *fac, int n, int bw, int c0, StepRegs<T> &q)
{
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    const int lane = threadIdx.x;
    const int w = min(kSB, n - c0);
#pragma unroll
    for (int jj = 0; jj < kSB; ++jj) {
        const bool ok = jj < lane && lane < w && lane - jj <= bw;
        q.l[jj] = load_if<T>(ok, rs, ((unsigned)(c0 + jj) * ldu + (unsigned)(lane - jj)) * es);
    }
    q.d = load_if<T>(lane < w, rs, (unsigned)(c0 + lane) * ldu * es);
}

template <typename T>
__device__ __forceinline__ void fwd_load_rows(const T *fac, int n, int bw, int c0, int wg, StepRegs<T> &q)
{
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    const int lane = threadIdx.x;
    const int w = min(kSB, n - c0), nr = min(n - c0, w + bw);
    const int r = w + kSB * wg + lane;
#pragma unroll
    for (int kk = 0; kk < kSB; ++kk) {
        const bool ok = kk < w && r < nr && r - kk <= bw;
        q.u[kk] = load_if<T>(ok, rs, ((unsigned)(c0 + kk) * ldu + (unsigned)(r - kk)) * es);
    }
}

template <typename T>
__device__ __forceinline__ void fwd_load(const T *fac, int n, int bw, int c0, int wg, StepRegs<T> &q)
{
    fwd_load_diag<T>(fac, n, bw, c0, q);
    fwd_load_rows<T>(fac, n, bw, c0, wg, q);
}

template <typename T>
__device__ __forceinline__ void fwd_compute(int n, int bw, int c0, double *b, double *z, int wg, const StepRegs<T> &q)
{
    __shared__ double ys[kSB];
    const int lane = threadIdx.x;
    const int w = min(kSB, n - c0), nr = min(n - c0, w + bw);
    const int r = w + kSB * wg + lane;
    const double rb = r < nr ? b[c0 + r] : 0.0;
    double y = lane < w ? b[c0 + lane] : 0.0;
#pragma unroll
    for (int jj = 0; jj < kSB - 1; ++jj) y -= (double)q.l[jj] * readlane_f64(y, jj);  // l = 0 on and above the diagonal
    ys[lane] = y;
    if (wg == 0 && lane < w) z[c0 + lane] = y / (double)q.d;
    __syncthreads();
    double acc = 0.0;
#pragma unroll
    for (int kk = 0; kk < kSB; ++kk) acc += (double)q.u[kk] * ys[kk];
    if (r < nr) b[c0 + r] = rb - acc;
}

template <typename T>
__global__ __launch_bounds__(kSB) void k_band_fwd_step(const T *fac, int n, int bw, int c0, double *b, double *z)
{
    StepRegs<T> q;
    fwd_load<T>(fac, n, bw, c0, (int)blockIdx.x, q);
    fwd_compute<T>(n, bw, c0, b, z, (int)blockIdx.x, q);
}

// backward, block c0: lane jj holds column c0+jj of the transposed diagonal block,
// L[c0+kk][c0+jj] for kk > jj; rows i = c0 - 1 - (64 wg + q) above the block, L[c0+k][i] (lane = k)
template <typename T>
__device__ __forceinline__ void bwd_load_diag(const T *fac, int n, int bw, int c0, StepRegs<T> &q)
{
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    const int lane = threadIdx.x;
    const int w = min(kSB, n - c0);
#pragma unroll
    for (int kk = 0; kk < kSB; ++kk) {
        const bool ok = kk > lane && kk < w && kk - lane <= bw;
        q.l[kk] = load_if<T>(ok, rs, ((unsigned)(c0 + lane) * ldu + (unsigned)(kk - lane)) * es);
    }
}

template <typename T>
__device__ __forceinline__ void bwd_load_rows(const T *fac, int n, int bw, int c0, int wg, StepRegs<T> &q)
{
    const unsigned ldu = (unsigned)bw + 1;
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = band_rsrc(fac, (unsigned)((size_t)n * ldu * es));
    const int lane = threadIdx.x;
    const int w = min(kSB, n - c0);
    const int lo = max(0, c0 - bw);
#pragma unroll
    for (int qq = 0; qq < kSB; ++qq) {
        const int i = c0 - 1 - (kSB * wg + qq);
        const bool ok = i >= lo && lane < w && c0 + lane - i <= bw;
        q.u[qq] = load_if<T>(ok, rs, ((unsigned)max(i, 0) * ldu + (unsigned)(c0 + lane - max(i, 0))) * es);
    }
}

template <typename T>
__device__ __forceinline__ void bwd_load(const T *fac, int n, int bw, int c0, int wg, StepRegs<T> &q)
{
    bwd_load_diag<T>(fac, n, bw, c0, q);
    bwd_load_rows<T>(fac, n, bw, c0, wg, q);
}

template <typename T>
__device__ __forceinline__ void bwd_compute(int n, int bw, int c0, double *z, double *x, int wg, const StepRegs<T> &q)
{
    __shared__ double prod[kSB][kSB + 1];
    const int lane = threadIdx.x;
    const int w = min(kSB, n - c0);
    const int lo = max(0, c0 - bw);
    double xv = lane < w ? z[c0 + lane] : 0.0;
#pragma unroll
    for (int kk = kSB - 1; kk > 0; --kk) xv -= (double)q.l[kk] * readlane_f64(xv, kk);  // l = 0 on and below the diagonal
    if (wg == 0 && lane < w) x[c0 + lane] = xv;
#pragma unroll
    for (int qq = 0; qq < kSB; ++qq) prod[qq][lane] = (double)q.u[qq] * xv;  // masked entries loaded as 0
    __syncthreads();
    const int i = c0 - 1 - (kSB * wg + lane);
    double s = 0.0;
#pragma unroll 16
    for (int k = 0; k < kSB; ++k) s += prod[lane][k];
    if (i >= lo) z[i] -= s;
}

template <typename T>
__global__ __launch_bounds__(kSB) void k_band_bwd_step(const T *fac, int n, int bw, int c0, double *z, double *x)
{
    StepRegs<T> q;
    bwd_load<T>(fac, n, bw, c0, (int)blockIdx.x, q);
    bwd_compute<T>(n, bw, c0, z, x, (int)blockIdx.x, q);
}

// Barrier of every workgroup of a cooperative launch (all resident at once): bar[0] counts
// arrivals, bar[1] is the generation, bar[2] an error flag (all three zero before the launch).  A
// wait that exceeds kSpinLimit polls (about 0.1 s; a step takes microseconds) sets the flag, and
// every workgroup then leaves at its next barrier: the launcher reports the failure instead of the
// GPU waiting forever.  Returns false on that error.
constexpr unsigned kSpinLimit = 1u << 22;
__device__ __forceinline__ bool grid_barrier(unsigned *bar, unsigned nwg, unsigned &gen)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned target = gen + 1;
        __threadfence();  // this workgroup's stores visible at device scope before it arrives
        if (atomicAdd(&bar[0], 1u) == nwg - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned it = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++it > kSpinLimit || __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (!ok) __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();  // the other workgroups' stores visible here (this CU's L1 invalidated)
        s_ok = ok;
    }
    __syncthreads();
    gen += 1;
    return s_ok != 0;
}

// The whole solve in one cooperative launch: the same steps as the stepped launches, workgroup g
// taking the part blockIdx g of each step (idle where a step has fewer parts), a grid barrier
// between dependent steps instead of a launch boundary.  A step's off-diagonal band entries are
// loaded before the barrier that precedes it, so their latency overlaps the wait.
template <typename T>
__global__ __launch_bounds__(kSB) void k_band_solve_persist(const T *fac, int n, int bw, double *x, double *z,
                                                            unsigned *bar)
{
    const unsigned nwg = gridDim.x;
    const int g = (int)blockIdx.x;
    const int last = ((n - 1) / kSB) * kSB;
    auto fwd_parts = [&](int c0) {
        const int w = min(kSB, n - c0), nr = min(n - c0, w + bw);
        return max(1, (nr - w + kSB - 1) / kSB);
    };
    auto bwd_parts = [&](int c0) { return max(1, (min(c0, bw) + kSB - 1) / kSB); };
    unsigned gen = 0;
    StepRegs<T> q;
    // the rows' entries of the next step are loaded before the barrier; the diagonal block after it
    // (both held across the barrier would not fit the registers)
    if (g < fwd_parts(0)) fwd_load_rows<T>(fac, n, bw, 0, g, q);
    for (int c0 = 0; c0 < n; c0 += kSB) {
        if (g < fwd_parts(c0)) {
            fwd_load_diag<T>(fac, n, bw, c0, q);
            fwd_compute<T>(n, bw, c0, x, z, g, q);
        }
        if (c0 + kSB < n) {
            if (g < fwd_parts(c0 + kSB)) fwd_load_rows<T>(fac, n, bw, c0 + kSB, g, q);
        } else if (g < bwd_parts(last)) {
            bwd_load_rows<T>(fac, n, bw, last, g, q);
        }
        if (!grid_barrier(bar, nwg, gen)) return;
    }
    for (int c0 = last; c0 >= 0; c0 -= kSB) {
        if (g < bwd_parts(c0)) {
            bwd_load_diag<T>(fac, n, bw, c0, q);
            bwd_compute<T>(n, bw, c0, z, x, g, q);
        }
        if (c0 > 0) {
            if (g < bwd_parts(c0 - kSB)) bwd_load_rows<T>(fac, n, bw, c0 - kSB, g, q);
            if (!grid_barrier(bar, nwg, gen)) return;
        }
    }
}

// The one-launch form is opt-in (KSH_SOLVE_PERSIST=1): on the C4 plate it took 1.07 s to 10 modal
// pairs against 0.83 s for the stepped launches.  Its barrier's device-scope release and acquire
// (the eight XCDs' L2s are not coherent with each other) cost more per step than the launch
// boundary they replace.
bool persist_requested()
{
    const char *e = std::getenv("KSH_SOLVE_PERSIST");
    return e && std::strcmp(e, "1") == 0;
}

// One cooperative launch (grid = the most parts any step has); false if the runtime refuses the
// cooperative launch (the caller then uses the stepped launches)
template <typename T>
int run_solve_persist(hipStream_t s, const T *fac, int n, int bw, double *x, double *z, unsigned *bar, bool &done)
{
    done = false;
    const int parts = std::max(1, (std::min(bw, std::max(n - kSB, 0)) + kSB - 1) / kSB);
    KSH_HIP(hipMemsetAsync(bar, 0, 3 * sizeof(unsigned), s));
    void *args[] = {(void *)&fac, (void *)&n, (void *)&bw, (void *)&x, (void *)&z, (void *)&bar};
    const hipError_t e = hipLaunchCooperativeKernel((const void *)k_band_solve_persist<T>, dim3((unsigned)parts), dim3(kSB),
                                                    args, 0, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return KSH_SUCCESS;
    }
    unsigned err = 0;
    KSH_HIP(hipMemcpyAsync(&err, bar + 2, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    KSH_HIP(hipStreamSynchronize(s));
    if (err) return fail(KSH_SOLVER_ERROR, "band solve: a grid barrier of the cooperative launch timed out");
    done = true;
    return KSH_SUCCESS;
}

template <typename T>
int run_solve(hipStream_t s, const T *fac, int n, int bw, double *x, double *z, unsigned *bar)
{
    if (bar && persist_requested()) {
        bool done = false;
        int rc = run_solve_persist<T>(s, fac, n, bw, x, z, bar, done);
        if (rc || done) return rc;
    }
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

int band_solve_tiled(hipStream_t s, const void *fac, int prec, int n, int bw, double *x, double *z, unsigned *bar)
{
    const size_t es = prec == KSH_FP64 ? sizeof(double) : sizeof(float);
    if ((size_t)n * (size_t)(bw + 1) * es >= 0x80000000ull)
        return fail(KSH_SOLVER_ERROR, "band of 2 GiB or more: beyond the band kernels' 32-bit offsets");
    return prec == KSH_FP64 ? run_solve<double>(s, (const double *)fac, n, bw, x, z, bar)
                            : run_solve<float>(s, (const float *)fac, n, bw, x, z, bar);
}

}  // namespace ksh
