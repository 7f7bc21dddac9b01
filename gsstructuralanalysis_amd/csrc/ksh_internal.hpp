// ksh_internal.hpp -- device-side data layout of one shell problem and one band matrix.
//
// HBM layout (one problem):
//   ctrl0     [ncp][3]            undeformed control net, FP64
//   free_idx  [ndof]              full dof -> free index or -1
//   free2full [nfree]             free index -> full dof
//   tabx/taby [n][nq][3][p+1]     1D basis values / 1st / 2nd parametric derivatives at the
//                                 Gauss points of each element column / row
//   wq        [nq*nq]             tensor Gauss weight of a point inside an element (incl. 1/(nx ny))
//   gpgeo     [nel][nq*nq][16]    undeformed geometry per Gauss point: wdA, SvK matrix C (6),
//                                 metric A_ab (3), curvature B_ab (3)
//   u_full    [ndof]              scratch: constructSolution of the current free vector
//   Hg        [120][nel*ngp]     per-Gauss-point H_g of the tangent (upper triangle, FP64)
//   Ke        [nel][48][48]       element matrices, local dof 3A+i, column major (FP64 or FP32)
//   Fe        [nel][3*nb]         element internal-force vectors
//   F         [nfree]             external load (lambda = 1)
// Band matrix: lower band, column major, ab[J*(bw+1) + (I-J)], FP64 or FP32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ksh.h"

namespace ksh {

constexpr int kGeoStride = 16;   // doubles per Gauss point in gpgeo
constexpr int kTile = 16;        // element tile edge (basis functions per element padded to 16)
constexpr int kKeTile = 9 * kTile * kTile;  // element matrix: 48 x 48 (3 components x 16 basis functions)
constexpr int kKeLd = 3 * kTile;             // its leading dimension: entry (3A+i, 3B+k) at (3B+k) kKeLd + 3A+i
constexpr int kHq = 120;         // stored entries of a Gauss point's symmetric 15x15 H_g

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define KSH_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::ksh::fail(KSH_OTHER_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// One instance of a batched launch (device-resident array; a single-instance call is a batch of
// one).  Prefix offsets let every block of a launch over several instances find its instance.
struct InstDesc {
    int p, nx, ny, mx, nel, nfree, bw, ndof;
    int elem0;              // first element of this instance in the launch
    int dof0;               // first free dof
    int full0;              // first full dof
    int skip_zero;          // 1: band slots no element reaches already hold 0 in ab and are not stored
    long long chunk0;       // band gather: first 64-thread chunk of this instance
    double t, rho, lambda;
    int mode, sprec;        // residual gather mode (0 sum, 1 ALM, 2 static); band storage precision
    const double *ctrl0, *tabx, *taby, *gpgeo, *F, *u_free;
    const int *free_idx, *free2full;
    double *u_full, *Fe, *R;
    void *Ke, *ab;
    double *Hg;             // per-Gauss-point H_g (upper triangle, kHq entries), entry-major
    double pressure;        // follower pressure along a1 x a2 (0: none)
    const double *P0;       // its load on the undeformed surface (free dofs; with a pressure)
};
// entries of a Gauss point's residual record: f_g (15), then -p w (a1 x a2) with a pressure (3)
constexpr int kFq = 15, kFqP = 18;

struct Problem;
struct Matrix;

// batched launchers over (problem, matrix / vectors) lists, all on stream s
int batch_construct(hipStream_t s, InstDesc *d_desc, const std::vector<InstDesc> &h);
int batch_residual(hipStream_t s, InstDesc *d_desc, const std::vector<InstDesc> &h);
int batch_tangent(hipStream_t s, InstDesc *d_desc, const std::vector<InstDesc> &h);
InstDesc make_desc(const Problem &pb);

struct Problem {
    int p = 0, nx = 0, ny = 0, mx = 0, my = 0, nq = 0, nb = 0, ne = 0;
    int ncp = 0, ndof = 0, nfree = 0, bw = 0, nel = 0, ngp = 0;
    double t = 0, E = 0, nu = 0, rho = 0;
    int npoint = 0;
    double fsurf[3] = {0, 0, 0};
    double pressure = 0.0;         // follower pressure along a1 x a2 (ksh_problem_set_pressure)
    double *P0 = nullptr;          // its load on the undeformed surface, free dofs (device)
    int device = 0;
    hipStream_t stream = nullptr;
    double inertia_guard = 0.5;    // pivot-count certificate threshold given to new matrices
    double refine_accept = 1e-6;   // FP32-factor solves above this relative residual go FP64
    std::vector<int> h_free2full;  // host copy of free2full
    std::shared_ptr<void> modal;   // learned-predictor modal basis (built on first use)
    double *modal_P = nullptr;     // device copies of the modal basis: [64][ncp] projector,
    double *modal_Phi = nullptr;   // [ncp][64] mode values, [ncp] z-dof map
    int *modal_zfree = nullptr;

    double *ctrl0 = nullptr;
    int *free_idx = nullptr;
    int *free2full = nullptr;
    double *tabx = nullptr, *taby = nullptr;
    double *wq = nullptr;
    double *gpgeo = nullptr;
    double *u_full = nullptr;
    double *u_free = nullptr;      // device staging of host free vectors
    double *r_free = nullptr;      // device staging of host results
    void *Ke = nullptr;            // FP64 or FP32 element tiles (sized for FP64)
    double *Hg = nullptr;          // [kHq][ngp] per-Gauss-point H_g of the last tangent (FP64)
    double *Fe = nullptr;
    double *F = nullptr;
    double *pl_N = nullptr;        // [npoint][2][p+1] basis values at the load points
    int *pl_span = nullptr;        // [npoint][2] knot spans
    double *pl_F = nullptr;        // [npoint][3] load vectors
    InstDesc *d_desc = nullptr;    // descriptor of this problem for single-instance launches
};

// Pivot-count certificate of FP32 factors: power steps z <- z - M^-1 K z = M^-1 E z (M = K + E the
// matrix whose exact LDL^T the FP32 factor is) from a fixed pseudo-random start estimate
// rho(M^-1 E).  A start with a component along every eigenvector lets the steps see a mode that
// the solves' right-hand sides barely touch (the near-null mode at a singular point).
constexpr int kPowerSteps = 3;
inline void power_start_vector(double *h, size_t n)
{
    uint64_t st = 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i < n; ++i) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        h[i] = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
}

struct Matrix {
    Problem *pb = nullptr;
    int n = 0, bw = 0, prec = KSH_FP64;
    int sprec = KSH_FP64;          // storage precision of the assembled values (KSH_FP64 / KSH_FP32)
    int fprec = KSH_FP64;          // precision of the LDL^T factor
    size_t ld = 0;                 // bw + 1
    void *ab = nullptr;            // assembled values (sprec)
    void *fac = nullptr;           // L and D after factorization (fprec)
    double *fac64 = nullptr;       // FP64 refactorization (near-singular FP32 pivots)
    bool use64 = false;            // solves of the current factor use fac64
    double *u_at = nullptr;        // free vector the matrix was assembled at (FP32 storage)
    bool has_u = false;
    bool factored = false;
    bool ab_clean = false;         // every band slot that no element reaches holds 0 in ab
    bool stalled = false;          // a solve of the current FP32 factor fell back to FP64 (batch engine)
    hipEvent_t early_done = nullptr;  // recorded after the last early FP64 factor into fac64 (batch engine)
    double *rc64 = nullptr;        // FP64 copy factored for a deferred pivot count (batch engine)
    hipEvent_t rc_done = nullptr;  // recorded after the last deferred count's factor of rc64
    uint64_t version = 0;          // bumped on every change of ab
    uint64_t factored_version = ~0ull;
    int info = 0;                  // 0 or 1-based index of a zero / non-finite pivot
    double guard = 0.5;            // FP32 pivot count trusted while rho <= guard (guard <= 0: always)
    double accept = 1e-6;          // refined FP32 solves above accept ||b|| are redone in FP64
    double rho = 0.0;              // largest residual contraction seen in refinement sweeps of this
    int rho_obs = 0;               // FP32 factor (estimate of rho(M^-1 E), M = K + E) and their count
    double rho_pw = -1.0;          // power-step estimate of rho(M^-1 E) for the current FP32 factor
                                   // (-1: not taken); the count certificate uses max(rho, rho_pw)
    double fstats[4] = {0, 0, 0, 0};  // pivot statistics of the current factor (negatives, min d, max|d|, min|d|)
    bool has_stats = false;
    // device workspace for solves and inertia
    double *work = nullptr;        // 5*n doubles (the last n: tiled-solve workspace)
    int *flag = nullptr;           // small device int buffer
    void *desc = nullptr;          // device descriptor of this band for batched kernels (one BandDesc)
    double *red = nullptr;         // reduction scratch
};

// kernels / launchers (ksh_assembly.hip, ksh_band.hip)
int launch_setup_geometry(Problem &pb);
int launch_external_force(Problem &pb);
int launch_pressure_rest(Problem &pb);  // pb.P0 = P(0) of the follower pressure, F += P0
int launch_construct_solution(Problem &pb, const double *u_free_dev);
int launch_residual(Problem &pb, double lambda, int conv, double *R_dev);
int launch_tangent(Problem &pb, Matrix &K);
int launch_mass(Problem &pb, Matrix &M);

// stats_out (optional, 4 doubles): negatives, min d, max |d|, min |d| of the new factor
int band_factor(Matrix &K, void *fac, int prec, double *stats_out = nullptr);

// one matrix of a batched factor / solve launch; `active` = 0 skips it
struct BandDescHost {
    void *fac;
    int n, bw;
    int *info;
    double *stats;
    double *x;
    int active;
};
// device-side form of BandDescHost (blockIdx.x of a batched factor / solve launch = matrix)
struct BandDesc {
    void *fac;       // factor storage (in place)
    int n, bw;
    int *info;       // 0 or 1-based zero-pivot column
    double *stats;   // [4]: negatives, min d, max |d|, min |d|
    double *x;       // solve: right-hand side in, solution out
    int active;      // 0: skip this matrix
    int pad_;
};
// prof_name: name of the launch in the HIP-event profile (default band_ldlt_batch_fp64 / _fp32)
int band_factor_batch(hipStream_t s, void *d_descs, const std::vector<BandDescHost> &h, int prec,
                      const char *prof_name = nullptr);
// factor of a batch of bands over the whole GPU (ksh_factor_tiled.hip): panel steps of two
// launches covering every band still active; d_descs holds h.size() device BandDescs
int band_factor_tiled_batch(hipStream_t s, void *d_descs, const std::vector<BandDescHost> &h, int prec);
// batched refinement: r = b - K x with norms (||r||^2, ||b||^2), then x += r on active matrices
struct RefDesc {
    const void *ab;
    int sprec, n, bw, active;
    const double *b;
    double *x, *r, *norms;
    double *part;          // 2 doubles per 64-row block (partial norms)
    int blk0, nblk;        // set by the launcher: first row block in the launch, row blocks
    // a second vector on the same matrix (set by the launcher when two jobs share ab; else null)
    const double *b2;
    double *x2, *r2, *norms2, *part2;
};
int refine_residual_batch(hipStream_t s, RefDesc *d_descs, const std::vector<RefDesc> &h);
int refine_update_batch(hipStream_t s, RefDesc *d_descs, const std::vector<RefDesc> &h);
// FP64 refactorization of K's assembled values into fac64 (K.use64 = true afterwards)
int refactor_fp64(Matrix &K);
// allocates K.fac64 if it is not there yet
int ensure_fac64(Matrix &K);
// allocates K.rc64 (the deferred pivot count's FP64 copy) if it is not there yet
int ensure_rc64(Matrix &K);
int band_solve_batch(hipStream_t s, void *d_descs, const std::vector<BandDescHost> &h, int prec);
constexpr size_t kBandDescBytes = 48;
int band_solve(Matrix &K, const void *fac, int prec, double *x_dev);
// solve of one large band over many workgroups (ksh_solve_tiled.hip); z: n doubles of workspace;
// bar: 3 device unsigneds for the one-launch form's grid barrier (nullptr: stepped launches)
int band_solve_tiled(hipStream_t s, const void *fac, int prec, int n, int bw, double *x, double *z, unsigned *bar);
// path of a single-band factor / solve: 1 = multi-workgroup (tiled), 0 = one workgroup
bool band_use_tiled_factor(int n, int bw, int prec);
bool band_use_tiled_solve(int n, int bw, int prec);
int band_symv(const Matrix &K, const double *x_dev, double *y_dev);
int band_inertia(Matrix &K, const void *fac, int prec, int64_t *neg, double *dmin, double *dmaxabs,
                 double *dminabs);
int band_axpby(Matrix &Y, double alpha, double beta, const Matrix &X);
int band_convert(const void *src, int src_prec, void *dst, int dst_prec, size_t count, hipStream_t s);
// copies (with conversion between FP64 and FP32 where the precisions differ) of many bands in
// one launch; d_descs: device room for h.size() CopyDescs
struct CopyDesc {
    const void *src;
    void *dst;
    size_t count;      // entries
    int src_prec, dst_prec;
    long long blk0;    // set by the launcher
};
int band_copy_batch(hipStream_t s, void *d_descs, const std::vector<CopyDesc> &h);
int vec_axpy(hipStream_t s, int n, double a, const double *x, double *y);
int vec_sub(hipStream_t s, int n, const double *a, const double *b, double *out);
int vec_norm2(hipStream_t s, int n, const double *x, double *scratch, double *host_out);

}  // namespace ksh

struct ksh_problem : ksh::Problem {};
struct ksh_matrix : ksh::Matrix {};
