// ksh_api.cpp -- C-ABI of libksh.so: problem set-up, matrices, LDL^T factor / solve / inertia.
//
// Host work here is limited to set-up tables (1D B-spline values at the Gauss points, Gauss
// rule, dof maps, load-point basis values) and orchestration; every per-Gauss-point,
// per-element and per-band-entry computation runs in the HIP kernels of ksh_assembly.hip and
// ksh_band.hip.  There is no CPU execution path for assembly or factorization.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ksh_internal.hpp"

namespace ksh {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

namespace {

// Gauss-Legendre rule on [-1,1] (Newton iteration on the Legendre recurrence).
void gauss_rule(int nq, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(nq, 0.0);
    w.assign(nq, 0.0);
    for (int i = 0; i < (nq + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (nq + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= nq; ++k) {
                double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            if (nq == 1) { p0 = 1.0; p1 = z; }
            dp = nq * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= nq; ++k) {
            double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = pk;
        }
        dp = nq * (z * p1 - p0) / (z * z - 1.0);
        x[i] = -z;
        x[nq - 1 - i] = z;
        w[i] = w[nq - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// Values, first and second derivatives of the p+1 basis functions non-zero on knot span
// `span` at x: out[d*(p+1)+a].  Cox-de Boor triangle for degrees 0..p, then the derivative
// rule N'_{i,q} = q (N_{i,q-1}/(u_{i+q}-u_i) - N_{i+1,q-1}/(u_{i+q+1}-u_{i+1})) applied twice.
void basis_ders(int p, const std::vector<double> &U, int span, double x, double *out)
{
    // tri[q][k] = N_{span-q+k, q}(x), k = 0..q
    double tri[8][8] = {};
    tri[0][0] = 1.0;
    for (int q = 1; q <= p; ++q)
        for (int k = 0; k <= q; ++k) {
            int i = span - q + k;
            double v = 0.0;
            if (k >= 1) {  // N_{i,q-1} = tri[q-1][k-1]
                double den = U[i + q] - U[i];
                if (den > 0) v += (x - U[i]) / den * tri[q - 1][k - 1];
            }
            if (k <= q - 1) {  // N_{i+1,q-1} = tri[q-1][k]
                double den = U[i + q + 1] - U[i + 1];
                if (den > 0) v += (U[i + q + 1] - x) / den * tri[q - 1][k];
            }
            tri[q][k] = v;
        }
    auto deriv = [&](int q, const double *lower, int nl, double *res) {
        // lower: degree q-1 functions N_{span-q+1+k, q-1}, k = 0..q-1 ; res: degree q derivative
        for (int k = 0; k <= q; ++k) {
            int i = span - q + k;
            double v = 0.0;
            if (k >= 1) {
                double den = U[i + q] - U[i];
                if (den > 0) v += q * lower[k - 1] / den;
            }
            if (k <= nl - 1) {
                double den = U[i + q + 1] - U[i + 1];
                if (den > 0) v -= q * lower[k] / den;
            }
            res[k] = v;
        }
    };
    const int p1 = p + 1;
    for (int a = 0; a <= p; ++a) out[a] = tri[p][a];
    double d1p[8] = {}, d1pm1[8] = {}, d2p[8] = {};
    deriv(p, tri[p - 1], p, d1p);
    if (p >= 2) {
        deriv(p - 1, tri[p - 2], p - 1, d1pm1);  // first derivatives of degree p-1 functions
        deriv(p, d1pm1, p, d2p);
    }
    for (int a = 0; a <= p; ++a) {
        out[p1 + a] = d1p[a];
        out[2 * p1 + a] = d2p[a];
    }
}

std::vector<double> uniform_knots(int p, int n)
{
    std::vector<double> U;
    for (int i = 0; i <= p; ++i) U.push_back(0.0);
    for (int i = 1; i < n; ++i) U.push_back((double)i / n);
    for (int i = 0; i <= p; ++i) U.push_back(1.0);
    return U;
}

template <typename T> T *dalloc(size_t count, int &rc)
{
    void *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
        rc = fail(KSH_OTHER_ERROR, std::string("hipMalloc: ") + hipGetErrorString(e));
        return nullptr;
    }
    return (T *)p;
}

size_t elem_bytes(int prec) { return prec == KSH_FP64 ? sizeof(double) : sizeof(float); }
int storage_prec(int mode) { return mode == KSH_FP32 ? KSH_FP32 : KSH_FP64; }
int factor_prec(int mode) { return mode == KSH_FP64 ? KSH_FP64 : KSH_FP32; }

int upload_free(Problem &pb, const double *host, double *dev)
{
    KSH_HIP(hipMemcpyAsync(dev, host, sizeof(double) * pb.nfree, hipMemcpyHostToDevice, pb.stream));
    return KSH_SUCCESS;
}

void destroy_problem(Problem *pb)
{
    if (!pb) return;
    void *ptrs[] = {pb->ctrl0, pb->free_idx, pb->free2full, pb->tabx, pb->taby, pb->wq, pb->gpgeo,
                    pb->u_full, pb->u_free, pb->r_free, pb->Ke, pb->Hg, pb->Fe, pb->F, pb->P0, pb->pl_N, pb->pl_span, pb->pl_F,
                    pb->d_desc, pb->modal_P, pb->modal_Phi, pb->modal_zfree};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (pb->stream) (void)hipStreamDestroy(pb->stream);
}

void destroy_matrix(Matrix *K)
{
    if (!K) return;
    if (K->early_done) {  // an early FP64 factor of a batch engine may still be writing fac64
        (void)hipEventSynchronize(K->early_done);
        (void)hipEventDestroy(K->early_done);
        K->early_done = nullptr;
    }
    if (K->rc_done) {  // a deferred pivot count of a batch engine may still be factoring rc64
        (void)hipEventSynchronize(K->rc_done);
        (void)hipEventDestroy(K->rc_done);
        K->rc_done = nullptr;
    }
    void *ptrs[] = {K->ab, K->fac, K->fac64, K->rc64, K->u_at, K->work, K->flag, K->red, K->desc};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

// FP64 refactorization at the assembled point, used when an FP32 pivot is too small to trust
int inertia_fp64(Matrix &K, int64_t *neg, double *dmin)
{
    Problem &pb = *K.pb;
    size_t count = (size_t)K.n * K.ld;
    int rc = 0;
    if (!K.fac64) {
        K.fac64 = dalloc<double>(count, rc);
        if (!K.fac64) return rc;
    }
    if (K.early_done) KSH_HIP(hipStreamWaitEvent(pb.stream, K.early_done, 0));
    Matrix tmp = K;
    tmp.prec = KSH_FP64;
    tmp.sprec = KSH_FP64;
    tmp.fprec = KSH_FP64;
    tmp.ab = K.fac64;
    tmp.ab_clean = false;  // fac64 holds the last factor: every slot is written
    if (K.sprec == KSH_FP64) {
        KSH_HIP(hipMemcpyAsync(K.fac64, K.ab, count * sizeof(double), hipMemcpyDeviceToDevice, pb.stream));
    } else if (K.has_u) {
        if ((rc = launch_construct_solution(pb, K.u_at))) return rc;
        if ((rc = launch_tangent(pb, tmp))) return rc;
    } else {
        if ((rc = band_convert(K.ab, K.sprec, K.fac64, KSH_FP64, count, pb.stream))) return rc;
    }
    double st[4];
    if ((rc = band_factor(tmp, K.fac64, KSH_FP64, st))) return rc;
    K.use64 = true;
    for (int i = 0; i < 4; ++i) K.fstats[i] = st[i];
    K.has_stats = true;
    if (neg) *neg = (int64_t)st[0];
    if (dmin) *dmin = st[1];
    return KSH_SUCCESS;
}

}  // namespace

int refactor_fp64(Matrix &K) { return inertia_fp64(K, nullptr, nullptr); }

// kPowerSteps power steps z <- z - M^-1 K z = M^-1 E z with the FP32 factor M = K + E of K, from the
// fixed pseudo-random start; the contraction of the last step estimates rho(M^-1 E), the quantity
// the pivot-count certificate compares with the guard.
int power_contraction(Matrix &K)
{
    hipStream_t s = K.pb->stream;
    const int n = K.n;
    double *z = K.work + 2 * (size_t)n, *y = K.work + 3 * (size_t)n;
    std::vector<double> h(n);
    power_start_vector(h.data(), (size_t)n);
    KSH_HIP(hipMemcpyAsync(z, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    int rc;
    double prev = 0.0;
    if ((rc = vec_norm2(s, n, z, K.red, &prev))) return rc;
    K.rho_pw = 0.0;
    for (int k = 0; k < kPowerSteps; ++k) {
        if ((rc = band_symv(K, z, y)) || (rc = band_solve(K, K.fac, KSH_FP32, y)) || (rc = vec_sub(s, n, z, y, z))) return rc;
        double nz = 0.0;
        if ((rc = vec_norm2(s, n, z, K.red, &nz))) return rc;
        const double ratio = prev > 0 ? nz / prev : 0.0;
        if (!std::isfinite(ratio)) {
            K.rho_pw = INFINITY;
            break;
        }
        K.rho_pw = ratio;  // the last step's contraction
        prev = nz;
        if (nz == 0) break;
    }
    return KSH_SUCCESS;
}

int ensure_fac64(Matrix &K)
{
    int rc = 0;
    if (!K.fac64 && !(K.fac64 = dalloc<double>((size_t)K.n * K.ld, rc))) return rc;
    return KSH_SUCCESS;
}

int ensure_rc64(Matrix &K)
{
    int rc = 0;
    if (!K.rc64 && !(K.rc64 = dalloc<double>((size_t)K.n * K.ld, rc))) return rc;
    return KSH_SUCCESS;
}

}  // namespace ksh

using namespace ksh;

extern "C" {

const char *ksh_last_error(void) { return g_err.c_str(); }

int ksh_device_count(int *count)
{
    int c = 0;
    KSH_HIP(hipGetDeviceCount(&c));
    *count = c;
    return KSH_SUCCESS;
}

int ksh_problem_create(const ksh_patch *patch, const ksh_material *mat, const ksh_bc *bc, const ksh_loads *loads,
                       ksh_problem **out)
{
    if (!patch || !mat || !bc || !out || !patch->ctrl || !bc->fixed) return fail(KSH_OTHER_ERROR, "null argument");
    if (patch->p != 2 && patch->p != 3) return fail(KSH_OTHER_ERROR, "degree must be 2 or 3");
    if (patch->nx < 1 || patch->ny < 1) return fail(KSH_OTHER_ERROR, "need at least one element per direction");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(KSH_OTHER_ERROR, "no HIP device: libksh runs on the GPU only");
    ksh_problem *pb = new ksh_problem();
    Problem &P = *pb;
    P.p = patch->p;
    P.nx = patch->nx;
    P.ny = patch->ny;
    P.mx = P.nx + P.p;
    P.my = P.ny + P.p;
    P.nq = P.p + 1;
    P.nb = P.nq * P.nq;
    P.ne = 3 * P.nb;
    P.ncp = P.mx * P.my;
    P.ndof = 3 * P.ncp;
    P.nel = P.nx * P.ny;
    P.ngp = P.nel * P.nq * P.nq;
    P.t = mat->t;
    P.E = mat->E;
    P.nu = mat->nu;
    P.rho = mat->rho;
    (void)hipGetDevice(&P.device);
    int rc = 0;
    auto bail = [&](int code) {
        destroy_problem(pb);
        delete pb;
        return code;
    };
    if (hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(KSH_OTHER_ERROR, "hipStreamCreate failed"));

    // dof maps and bandwidth
    std::vector<int> free_idx(P.ndof), free2full;
    for (int g = 0; g < P.ndof; ++g) {
        if (bc->fixed[g]) free_idx[g] = -1;
        else {
            free_idx[g] = (int)free2full.size();
            free2full.push_back(g);
        }
    }
    P.nfree = (int)free2full.size();
    if (P.nfree == 0) return bail(fail(KSH_OTHER_ERROR, "no free dofs"));
    P.h_free2full = free2full;
    int bw = 0;
    for (int ey = 0; ey < P.ny; ++ey)
        for (int ex = 0; ex < P.nx; ++ex) {
            int lo = 1 << 30, hi = -1;
            for (int b = 0; b < P.nq; ++b)
                for (int a = 0; a < P.nq; ++a)
                    for (int c = 0; c < 3; ++c) {
                        int f = free_idx[3 * ((ex + a) + P.mx * (ey + b)) + c];
                        if (f < 0) continue;
                        lo = std::min(lo, f);
                        hi = std::max(hi, f);
                    }
            if (hi >= 0) bw = std::max(bw, hi - lo);
        }
    P.bw = bw;

    // 1D tables at the Gauss points
    std::vector<double> gx, gw;
    gauss_rule(P.nq, gx, gw);
    auto table = [&](int n) {
        std::vector<double> U = uniform_knots(P.p, n), tab((size_t)n * P.nq * 3 * P.nq);
        for (int e = 0; e < n; ++e)
            for (int q = 0; q < P.nq; ++q) {
                double x = (e + 0.5 * (1.0 + gx[q])) / n;
                basis_ders(P.p, U, e + P.p, x, &tab[((size_t)e * P.nq + q) * 3 * P.nq]);
            }
        return tab;
    };
    std::vector<double> tabx = table(P.nx), taby = table(P.ny);
    std::vector<double> wq(P.nq * P.nq);
    for (int qy = 0; qy < P.nq; ++qy)
        for (int qx = 0; qx < P.nq; ++qx) wq[qx + P.nq * qy] = gw[qx] * gw[qy] * 0.25 / ((double)P.nx * (double)P.ny);

    // loads
    P.npoint = loads ? loads->npoint : 0;
    if (loads)
        for (int i = 0; i < 3; ++i) P.fsurf[i] = loads->fsurf[i];
    std::vector<double> plN((size_t)std::max(P.npoint, 1) * 2 * P.nq);
    std::vector<int> plspan((size_t)std::max(P.npoint, 1) * 2);
    std::vector<double> plF((size_t)std::max(P.npoint, 1) * 3);
    {
        std::vector<double> Ux = uniform_knots(P.p, P.nx), Uy = uniform_knots(P.p, P.ny);
        for (int l = 0; l < P.npoint; ++l) {
            double xi = loads->point_xi[2 * l], eta = loads->point_xi[2 * l + 1];
            if (!(xi >= 0.0 && xi <= 1.0 && eta >= 0.0 && eta <= 1.0))
                return bail(fail(KSH_OTHER_ERROR, "point load outside the parameter domain"));
            int ex = std::min((int)std::floor(xi * P.nx), P.nx - 1), ey = std::min((int)std::floor(eta * P.ny), P.ny - 1);
            double bx[3 * 8], by[3 * 8];
            basis_ders(P.p, Ux, ex + P.p, xi, bx);
            basis_ders(P.p, Uy, ey + P.p, eta, by);
            for (int a = 0; a < P.nq; ++a) {
                plN[(size_t)l * 2 * P.nq + a] = bx[a];
                plN[(size_t)l * 2 * P.nq + P.nq + a] = by[a];
            }
            plspan[2 * l] = ex + P.p;
            plspan[2 * l + 1] = ey + P.p;
            for (int c = 0; c < 3; ++c) plF[3 * l + c] = loads->point_F[3 * l + c];
        }
    }

    // device buffers
    if (!(P.ctrl0 = dalloc<double>((size_t)P.ncp * 3, rc))) return bail(rc);
    if (!(P.free_idx = dalloc<int>(P.ndof, rc))) return bail(rc);
    if (!(P.free2full = dalloc<int>(P.nfree, rc))) return bail(rc);
    if (!(P.tabx = dalloc<double>(tabx.size(), rc))) return bail(rc);
    if (!(P.taby = dalloc<double>(taby.size(), rc))) return bail(rc);
    if (!(P.wq = dalloc<double>(wq.size(), rc))) return bail(rc);
    if (!(P.gpgeo = dalloc<double>((size_t)P.ngp * kGeoStride, rc))) return bail(rc);
    if (!(P.u_full = dalloc<double>(P.ndof, rc))) return bail(rc);
    if (!(P.u_free = dalloc<double>(P.nfree, rc))) return bail(rc);
    if (!(P.r_free = dalloc<double>(P.nfree, rc))) return bail(rc);
    if (!(P.Ke = dalloc<double>((size_t)P.nel * kKeTile, rc))) return bail(rc);
    if (!(P.Hg = dalloc<double>((size_t)P.ngp * kHq, rc))) return bail(rc);
    if (!(P.Fe = dalloc<double>((size_t)P.nel * P.ne, rc))) return bail(rc);
    if (!(P.F = dalloc<double>(P.nfree, rc))) return bail(rc);
    if (!(P.pl_N = dalloc<double>(plN.size(), rc))) return bail(rc);
    if (!(P.pl_span = dalloc<int>(plspan.size(), rc))) return bail(rc);
    if (!(P.pl_F = dalloc<double>(plF.size(), rc))) return bail(rc);
    if (!(P.d_desc = dalloc<InstDesc>(1, rc))) return bail(rc);
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, P.stream) == hipSuccess;
    };
    bool ok = up(P.ctrl0, patch->ctrl, sizeof(double) * P.ncp * 3) && up(P.free_idx, free_idx.data(), sizeof(int) * P.ndof) &&
              up(P.free2full, free2full.data(), sizeof(int) * P.nfree) && up(P.tabx, tabx.data(), sizeof(double) * tabx.size()) &&
              up(P.taby, taby.data(), sizeof(double) * taby.size()) && up(P.wq, wq.data(), sizeof(double) * wq.size()) &&
              up(P.pl_N, plN.data(), sizeof(double) * plN.size()) && up(P.pl_span, plspan.data(), sizeof(int) * plspan.size()) &&
              up(P.pl_F, plF.data(), sizeof(double) * plF.size());
    if (!ok) return bail(fail(KSH_OTHER_ERROR, "upload failed"));
    if (hipMemsetAsync(P.u_full, 0, sizeof(double) * P.ndof, P.stream) != hipSuccess) return bail(fail(KSH_OTHER_ERROR, "memset"));
    if ((rc = launch_setup_geometry(P))) return bail(rc);
    if ((rc = launch_external_force(P))) return bail(rc);
    if (hipStreamSynchronize(P.stream) != hipSuccess) return bail(fail(KSH_OTHER_ERROR, "set-up kernels failed"));
    *out = pb;
    return KSH_SUCCESS;
}

int ksh_problem_set_pressure(ksh_problem *pb, double pressure)
{
    if (!pb) return fail(KSH_OTHER_ERROR, "null argument");
    if (!std::isfinite(pressure)) return fail(KSH_OTHER_ERROR, "pressure must be finite");
    int rc = 0;
    if (pressure != 0.0 && !pb->P0 && !(pb->P0 = dalloc<double>((size_t)pb->nfree, rc))) return rc;
    pb->pressure = pressure;
    if ((rc = launch_external_force(*pb))) return rc;                       // the dead loads
    if (pressure != 0.0 && (rc = launch_pressure_rest(*pb))) return rc;   // P0, and F += P0
    KSH_HIP(hipStreamSynchronize(pb->stream));
    return KSH_SUCCESS;
}

int ksh_problem_destroy(ksh_problem *pb)
{
    if (!pb) return KSH_SUCCESS;
    (void)hipStreamSynchronize(pb->stream);
    destroy_problem(pb);
    delete pb;
    return KSH_SUCCESS;
}

int ksh_num_free_dofs(const ksh_problem *pb, int64_t *n)
{
    if (!pb || !n) return fail(KSH_OTHER_ERROR, "null argument");
    *n = pb->nfree;
    return KSH_SUCCESS;
}

int ksh_bandwidth(const ksh_problem *pb, int64_t *bw)
{
    if (!pb || !bw) return fail(KSH_OTHER_ERROR, "null argument");
    *bw = pb->bw;
    return KSH_SUCCESS;
}

int ksh_synchronize(ksh_problem *pb)
{
    KSH_HIP(hipStreamSynchronize(pb->stream));
    return KSH_SUCCESS;
}

int ksh_assemble_force(ksh_problem *pb, double *F_out)
{
    if (!pb || !F_out) return fail(KSH_OTHER_ERROR, "null argument");
    KSH_HIP(hipMemcpyAsync(F_out, pb->F, sizeof(double) * pb->nfree, hipMemcpyDeviceToHost, pb->stream));
    KSH_HIP(hipStreamSynchronize(pb->stream));
    return KSH_SUCCESS;
}

int ksh_assemble_residual_dev(ksh_problem *pb, const double *u_free_dev, double lambda, int conv, double *R_out_dev)
{
    if (!pb || !u_free_dev || !R_out_dev) return fail(KSH_OTHER_ERROR, "null argument");
    if (conv != KSH_RES_STATIC && conv != KSH_RES_ALM) return fail(KSH_OTHER_ERROR, "unknown residual convention");
    int rc;
    if ((rc = launch_construct_solution(*pb, u_free_dev))) return fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
    if ((rc = launch_residual(*pb, lambda, conv, R_out_dev))) return fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
    return KSH_SUCCESS;
}

int ksh_assemble_residual(ksh_problem *pb, const double *u_free, double lambda, int conv, double *R_out)
{
    if (!pb || !u_free || !R_out) return fail(KSH_OTHER_ERROR, "null argument");
    int rc;
    if ((rc = upload_free(*pb, u_free, pb->u_free))) return rc;
    if ((rc = ksh_assemble_residual_dev(pb, pb->u_free, lambda, conv, pb->r_free))) return rc;
    KSH_HIP(hipMemcpyAsync(R_out, pb->r_free, sizeof(double) * pb->nfree, hipMemcpyDeviceToHost, pb->stream));
    KSH_HIP(hipStreamSynchronize(pb->stream));
    for (int i = 0; i < pb->nfree; ++i)
        if (!std::isfinite(R_out[i])) return fail(KSH_ASSEMBLY_ERROR, "non-finite residual");
    return KSH_SUCCESS;
}

int ksh_matrix_create(ksh_problem *pb, int prec, ksh_matrix **out)
{
    if (!pb || !out) return fail(KSH_OTHER_ERROR, "null argument");
    if (prec != KSH_FP64 && prec != KSH_FP32_IR && prec != KSH_FP32) return fail(KSH_OTHER_ERROR, "unknown precision");
    ksh_matrix *K = new ksh_matrix();
    K->pb = pb;
    K->n = pb->nfree;
    K->bw = pb->bw;
    K->ld = (size_t)pb->bw + 1;
    K->prec = prec;
    K->sprec = storage_prec(prec);
    K->fprec = factor_prec(prec);
    K->guard = pb->inertia_guard;
    K->accept = pb->refine_accept;
    int rc = 0;
    size_t count = (size_t)K->n * K->ld, eb = elem_bytes(K->sprec), fb = elem_bytes(K->fprec);
    auto bail = [&](int code) {
        destroy_matrix(K);
        delete K;
        return code;
    };
    if (!(K->ab = dalloc<unsigned char>(count * eb, rc))) return bail(rc);
    if (!(K->fac = dalloc<unsigned char>(count * fb, rc))) return bail(rc);
    if (!(K->work = dalloc<double>((size_t)5 * K->n, rc))) return bail(rc);
    if (!(K->flag = dalloc<int>(4, rc))) return bail(rc);
    if (!(K->red = dalloc<double>(8, rc))) return bail(rc);
    if (!(K->desc = dalloc<unsigned char>(kBandDescBytes, rc))) return bail(rc);
    if (K->sprec == KSH_FP32 && !(K->u_at = dalloc<double>(K->n, rc))) return bail(rc);
    // zeroed before return: the matrix may next be written on another stream (a batch engine's),
    // which must not race with this clear
    if (hipMemsetAsync(K->ab, 0, count * eb, pb->stream) != hipSuccess || hipStreamSynchronize(pb->stream) != hipSuccess)
        return bail(fail(KSH_OTHER_ERROR, "memset"));
    K->ab_clean = true;
    *out = K;
    return KSH_SUCCESS;
}

int ksh_matrix_destroy(ksh_matrix *K)
{
    if (!K) return KSH_SUCCESS;
    (void)hipStreamSynchronize(K->pb->stream);
    destroy_matrix(K);
    delete K;
    return KSH_SUCCESS;
}

int ksh_matrix_dims(const ksh_matrix *K, int64_t *n, int64_t *bw, int *prec)
{
    if (!K) return fail(KSH_OTHER_ERROR, "null argument");
    if (n) *n = K->n;
    if (bw) *bw = K->bw;
    if (prec) *prec = K->prec;
    return KSH_SUCCESS;
}

int ksh_assemble_tangent_dev(ksh_problem *pb, const double *u_free_dev, ksh_matrix *K)
{
    if (!pb || !u_free_dev || !K) return fail(KSH_OTHER_ERROR, "null argument");
    if (K->pb != pb) return fail(KSH_OTHER_ERROR, "matrix belongs to another problem");
    int rc;
    if ((rc = launch_construct_solution(*pb, u_free_dev))) return fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
    if ((rc = launch_tangent(*pb, *K))) return fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
    if (K->sprec == KSH_FP32) {
        KSH_HIP(hipMemcpyAsync(K->u_at, u_free_dev, sizeof(double) * K->n, hipMemcpyDeviceToDevice, pb->stream));
        K->has_u = true;
    }
    ++K->version;
    return KSH_SUCCESS;
}

int ksh_assemble_tangent(ksh_problem *pb, const double *u_free, ksh_matrix *K)
{
    if (!pb || !u_free || !K) return fail(KSH_OTHER_ERROR, "null argument");
    int rc;
    if ((rc = upload_free(*pb, u_free, pb->u_free))) return rc;
    return ksh_assemble_tangent_dev(pb, pb->u_free, K);
}

int ksh_assemble_mass(ksh_problem *pb, ksh_matrix *M)
{
    if (!pb || !M) return fail(KSH_OTHER_ERROR, "null argument");
    int rc;
    if ((rc = launch_mass(*pb, *M))) return fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
    M->has_u = false;
    ++M->version;
    return KSH_SUCCESS;
}

int ksh_matrix_download(const ksh_matrix *K, double *ab_out)
{
    if (!K || !ab_out) return fail(KSH_OTHER_ERROR, "null argument");
    size_t count = (size_t)K->n * K->ld;
    hipStream_t s = K->pb->stream;
    if (K->sprec == KSH_FP64) {
        KSH_HIP(hipMemcpyAsync(ab_out, K->ab, count * sizeof(double), hipMemcpyDeviceToHost, s));
        KSH_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<float> tmp(count);
        KSH_HIP(hipMemcpyAsync(tmp.data(), K->ab, count * sizeof(float), hipMemcpyDeviceToHost, s));
        KSH_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < count; ++i) ab_out[i] = tmp[i];
    }
    return KSH_SUCCESS;
}

int ksh_matrix_upload(ksh_matrix *K, const double *ab_in)
{
    if (!K || !ab_in) return fail(KSH_OTHER_ERROR, "null argument");
    size_t count = (size_t)K->n * K->ld;
    hipStream_t s = K->pb->stream;
    if (K->sprec == KSH_FP64) {
        KSH_HIP(hipMemcpyAsync(K->ab, ab_in, count * sizeof(double), hipMemcpyHostToDevice, s));
        KSH_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<float> tmp(ab_in, ab_in + count);
        KSH_HIP(hipMemcpyAsync(K->ab, tmp.data(), count * sizeof(float), hipMemcpyHostToDevice, s));
        KSH_HIP(hipStreamSynchronize(s));
    }
    K->has_u = false;
    K->ab_clean = false;
    ++K->version;
    return KSH_SUCCESS;
}

int ksh_matrix_axpby(ksh_matrix *Y, double alpha, double beta, const ksh_matrix *X)
{
    if (!Y || !X) return fail(KSH_OTHER_ERROR, "null argument");
    if (Y->n != X->n || Y->bw != X->bw) return fail(KSH_OTHER_ERROR, "matrix shapes differ");
    int rc;
    Y->ab_clean = false;
    if ((rc = band_axpby(*Y, alpha, beta, *X))) return rc;
    Y->has_u = false;
    ++Y->version;
    return KSH_SUCCESS;
}

int ksh_matrix_copy(ksh_matrix *dst, const ksh_matrix *src)
{
    if (!dst || !src) return fail(KSH_OTHER_ERROR, "null argument");
    if (dst->n != src->n || dst->bw != src->bw) return fail(KSH_OTHER_ERROR, "matrix shapes differ");
    size_t count = (size_t)dst->n * dst->ld;
    hipStream_t s = dst->pb->stream;
    dst->ab_clean = false;  // src may belong to another problem, whose unreached slots differ
    if (dst->sprec == src->sprec) {
        KSH_HIP(hipMemcpyAsync(dst->ab, src->ab, count * elem_bytes(dst->sprec), hipMemcpyDeviceToDevice, s));
    } else {
        int rc;
        if ((rc = band_convert(src->ab, src->sprec, dst->ab, dst->sprec, count, s))) return rc;
    }
    dst->has_u = src->has_u && dst->sprec == KSH_FP32 && src->sprec == KSH_FP32;
    if (dst->has_u) KSH_HIP(hipMemcpyAsync(dst->u_at, src->u_at, sizeof(double) * dst->n, hipMemcpyDeviceToDevice, s));
    ++dst->version;
    return KSH_SUCCESS;
}

int ksh_matvec_dev(const ksh_matrix *K, const double *x_dev, double *y_dev)
{
    if (!K || !x_dev || !y_dev) return fail(KSH_OTHER_ERROR, "null argument");
    return band_symv(*K, x_dev, y_dev);
}

int ksh_matvec(const ksh_matrix *K, const double *x, double *y)
{
    if (!K || !x || !y) return fail(KSH_OTHER_ERROR, "null argument");
    hipStream_t s = K->pb->stream;
    double *xd = K->work, *yd = K->work + K->n;
    KSH_HIP(hipMemcpyAsync(xd, x, sizeof(double) * K->n, hipMemcpyHostToDevice, s));
    int rc;
    if ((rc = band_symv(*K, xd, yd))) return rc;
    KSH_HIP(hipMemcpyAsync(y, yd, sizeof(double) * K->n, hipMemcpyDeviceToHost, s));
    KSH_HIP(hipStreamSynchronize(s));
    return KSH_SUCCESS;
}

int ksh_factor(ksh_matrix *K)
{
    if (!K) return fail(KSH_OTHER_ERROR, "null argument");
    if (K->factored && K->factored_version == K->version) return K->info ? KSH_SOLVER_ERROR : KSH_SUCCESS;
    K->use64 = false;
    K->rho = 0.0;
    K->rho_obs = 0;
    K->rho_pw = -1.0;
    double st[4] = {0, 0, 0, 0};
    int rc = band_factor(*K, K->fac, K->fprec, st);
    K->factored = true;
    K->factored_version = K->version;
    for (int i = 0; i < 4; ++i) K->fstats[i] = st[i];
    K->has_stats = rc == KSH_SUCCESS;
    if (K->fprec == KSH_FP32 && rc != KSH_SUCCESS) {
        // a zero or non-finite FP32 pivot: an FP64 factor of the same matrix serves this version
        rc = inertia_fp64(*K, nullptr, nullptr);
        K->info = (rc == KSH_SUCCESS) ? 0 : 1;
    }
    return rc;
}

int ksh_solve_dev(ksh_matrix *K, const double *b_dev, double *x_dev, int max_refine, double tol)
{
    if (!K || !b_dev || !x_dev) return fail(KSH_OTHER_ERROR, "null argument");
    if (!K->factored || K->factored_version != K->version) return fail(KSH_SOLVER_ERROR, "matrix is not factored");
    if (K->info) return fail(KSH_SOLVER_ERROR, "factorization failed");
    hipStream_t s = K->pb->stream;
    int n = K->n, rc;
    const void *F = K->use64 ? (const void *)K->fac64 : K->fac;
    const int fp = K->use64 ? KSH_FP64 : K->fprec;
    if (x_dev != b_dev) KSH_HIP(hipMemcpyAsync(x_dev, b_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if ((rc = band_solve(*K, F, fp, x_dev))) return rc;
    if (max_refine > 0) {
        double *bcopy = K->work + 2 * n, *r = K->work + 3 * n;
        if (x_dev == b_dev) return fail(KSH_OTHER_ERROR, "refinement needs distinct b and x");
        double bnorm = 0.0, rprev = 0.0;
        if ((rc = vec_norm2(s, n, b_dev, K->red, &bnorm))) return rc;
        KSH_HIP(hipMemcpyAsync(bcopy, b_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        for (int it = 0; it < max_refine; ++it) {
            // r = b - K x  (FP64 arithmetic on the stored matrix values)
            if ((rc = band_symv(*K, x_dev, r))) return rc;
            if ((rc = vec_sub(s, n, bcopy, r, r))) return rc;
            double rn = 0.0;
            if ((rc = vec_norm2(s, n, r, K->red, &rn))) return rc;
            if (!(rn <= tol * bnorm) && !std::isfinite(rn)) return fail(KSH_SOLVER_ERROR, "refinement diverged");
            if (fp == KSH_FP32 && bnorm > 0 && std::isfinite(rn)) {
                // residual contraction of this sweep: an estimate of rho(M^-1 E) for the count
                K->rho = std::max(K->rho, it == 0 ? rn / bnorm : (rprev > 0 ? rn / rprev : 0.0));
                ++K->rho_obs;
            }
            rprev = rn;
            if (rn <= tol * bnorm) break;
            if (it == max_refine - 1 && fp == KSH_FP32 && rn > K->accept * bnorm) {
                // refinement stalled on the FP32 factor: solve with an FP64 factor instead
                if ((rc = inertia_fp64(*K, nullptr, nullptr))) return rc;
                KSH_HIP(hipMemcpyAsync(x_dev, bcopy, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                return band_solve(*K, K->fac64, KSH_FP64, x_dev);
            }
            if ((rc = band_solve(*K, F, fp, r))) return rc;
            if ((rc = vec_axpy(s, n, 1.0, r, x_dev))) return rc;
        }
    }
    return KSH_SUCCESS;
}

int ksh_solve(ksh_matrix *K, const double *b, double *x, int max_refine, double tol)
{
    if (!K || !b || !x) return fail(KSH_OTHER_ERROR, "null argument");
    hipStream_t s = K->pb->stream;
    double *bd = K->work, *xd = K->work + K->n;
    KSH_HIP(hipMemcpyAsync(bd, b, sizeof(double) * K->n, hipMemcpyHostToDevice, s));
    int rc;
    if ((rc = ksh_solve_dev(K, bd, xd, max_refine, tol))) return rc;
    KSH_HIP(hipMemcpyAsync(x, xd, sizeof(double) * K->n, hipMemcpyDeviceToHost, s));
    KSH_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < K->n; ++i)
        if (!std::isfinite(x[i])) return fail(KSH_SOLVER_ERROR, "non-finite solution");
    return KSH_SUCCESS;
}

int ksh_set_inertia_guard(ksh_matrix *K, double guard)
{
    if (!K) return fail(KSH_OTHER_ERROR, "null argument");
    K->guard = guard;
    return KSH_SUCCESS;
}

int ksh_inertia(ksh_matrix *K, int64_t *n_negative, double *d_min)
{
    if (!K) return fail(KSH_OTHER_ERROR, "null argument");
    int rc;
    if ((rc = ksh_factor(K))) return rc;
    if (K->fprec == KSH_FP32 && !K->use64 && K->guard > 0) {
        if (K->rho_pw < 0 && (rc = power_contraction(*K))) return rc;
        if (std::max(K->rho, K->rho_pw) > K->guard && (rc = inertia_fp64(*K, nullptr, nullptr))) return rc;
    }
    int64_t neg = 0;
    double dmin = 0, dmaxabs = 0, dminabs = 0;
    if (K->has_stats) {  // statistics reduced by the factorization kernel itself
        neg = (int64_t)K->fstats[0];
        dmin = K->fstats[1];
    } else if (K->use64) {
        Matrix tmp = *K;
        tmp.sprec = tmp.fprec = KSH_FP64;
        if ((rc = band_inertia(tmp, K->fac64, KSH_FP64, &neg, &dmin, &dmaxabs, &dminabs))) return rc;
    } else {
        if ((rc = band_inertia(*K, K->fac, K->fprec, &neg, &dmin, &dmaxabs, &dminabs))) return rc;
    }
    if (n_negative) *n_negative = neg;
    if (d_min) *d_min = dmin;
    return KSH_SUCCESS;
}

int ksh_vector_d(ksh_matrix *K, double *d_out)
{
    if (!K || !d_out) return fail(KSH_OTHER_ERROR, "null argument");
    int rc;
    if ((rc = ksh_factor(K))) return rc;
    hipStream_t s = K->pb->stream;
    const int fp = K->use64 ? KSH_FP64 : K->fprec;
    const void *F = K->use64 ? (const void *)K->fac64 : K->fac;
    size_t eb = elem_bytes(fp);
    std::vector<unsigned char> tmp((size_t)K->n * eb);
    KSH_HIP(hipMemcpy2DAsync(tmp.data(), eb, F, K->ld * eb, eb, K->n, hipMemcpyDeviceToHost, s));
    KSH_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < K->n; ++i)
        d_out[i] = fp == KSH_FP64 ? ((double *)tmp.data())[i] : (double)((float *)tmp.data())[i];
    return KSH_SUCCESS;
}

}  // extern "C"
