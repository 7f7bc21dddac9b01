// ksh_eigen.cpp -- generalized symmetric eigenpairs (A x = lambda B x) nearest a shift, by
// thick-restart Lanczos on a shift-invert operator built from the GPU band factor.
//
// Replaces gsEigenProblemBase::computeSparse_impl (src/gsEigenSolvers/gsEigenProblemBase.hpp:
// 152-216, Spectra's SymGEigsShiftSolver) for gsModalSolver (A = K, B = M,
// src/gsEigenSolvers/gsModalSolver.h:46-51) and gsBucklingSolver (A = K_L, B = K_NL - K_L,
// src/gsEigenSolvers/gsBucklingSolver.h:100-105,127-149).
//
// Operator S and inner product W (S is self-adjoint in <x,y>_W = x^T W y):
//   shift-invert (B SPD):        S = (A - s B)^-1 B,  W = B,  lambda = s + 1/theta
//   buckling, s = 0 (A SPD):     S = A^-1 B,          W = A,  lambda = 1/theta
//   buckling, s != 0 (A SPD):    S = (A - s B)^-1 A,  W = A,  lambda = s theta / (theta - 1)
// The eigenvalues nearest s are the Ritz values theta of largest magnitude.  Each Lanczos step
// costs one product W v (GPU band mat-vec) and one solve with the factor of A - s B (GPU);
// S v = solve(W v) reuses the W v that the W-inner products need, and every vector carries
// its W-image through the orthogonalizations, so no further products are needed.
//
// Thick restart (Wu & Simon, SIAM J. Matrix Anal. Appl. 22 (2000) 602-616): after m steps the
// projected matrix T (arrowhead + tridiagonal) is diagonalized, the p most wanted Ritz pairs
// are kept as the new basis with their coupling to the last Lanczos vector, and the expansion
// resumes.  Full reorthogonalization (two passes, in the W inner product) throughout.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "ksh_internal.hpp"

namespace ksh {

namespace {

// eigen-decomposition of a small dense symmetric matrix (cyclic Jacobi); a is m x m row major,
// overwritten; vectors v (m x m, column j = eigenvector j)
void jacobi_eig(std::vector<double> &a, int m, std::vector<double> &w, std::vector<double> &v)
{
    v.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) v[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const double x = a[(size_t)i * m + j] * a[(size_t)i * m + j];
                tot += x;
                if (i != j) off += x;
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = a[(size_t)p * m + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = a[(size_t)p * m + p], aqq = a[(size_t)q * m + q];
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < m; ++k) {  // rotate columns p, q
                    const double akp = a[(size_t)k * m + p], akq = a[(size_t)k * m + q];
                    a[(size_t)k * m + p] = c * akp - s * akq;
                    a[(size_t)k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {  // rotate rows p, q
                    const double apk = a[(size_t)p * m + k], aqk = a[(size_t)q * m + k];
                    a[(size_t)p * m + k] = c * apk - s * aqk;
                    a[(size_t)q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double vkp = v[(size_t)k * m + p], vkq = v[(size_t)k * m + q];
                    v[(size_t)k * m + p] = c * vkp - s * vkq;
                    v[(size_t)k * m + q] = s * vkp + c * vkq;
                }
            }
    }
    w.resize(m);
    for (int i = 0; i < m; ++i) w[i] = a[(size_t)i * m + i];
}

double dot(const double *x, const double *y, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}

struct Op {
    ksh_matrix *C;  // factored A - s B (or A)
    ksh_matrix *W;  // inner-product matrix
    int max_refine;
    double refine_tol;
    int64_t solves = 0, products = 0;
    int apply_W(const double *x, double *y)
    {
        ++products;
        return ksh_matvec(W, x, y);
    }
    int solve(const double *b, double *x)
    {
        ++solves;
        return ksh_solve(C, b, x, max_refine, refine_tol);
    }
};

}  // namespace

}  // namespace ksh

using namespace ksh;

extern "C" int ksh_eigen(ksh_matrix *A, ksh_matrix *B, const ksh_eig_options *o, double *values, double *vectors,
                         ksh_eig_result *res)
{
    if (!A || !B || !o || !values) return fail(KSH_OTHER_ERROR, "null argument");
    if (A->n != B->n || A->pb != B->pb) return fail(KSH_OTHER_ERROR, "A and B of different problems");
    const int n = A->n, k = o->number;
    const int m = std::min(n, std::max(o->ncv_fac * k, 2 * k + 1));
    if (k <= 0 || k >= n) return fail(KSH_OTHER_ERROR, "number of eigenpairs out of range");
    const auto t0 = std::chrono::steady_clock::now();
    const double s = o->shift;
    const bool buckling = o->mode == KSH_EIG_BUCKLING;
    int rc;

    // C = A - s B in the requested factor precision; W = B (shift-invert) or A (buckling)
    ksh_matrix *C = nullptr;
    if ((rc = ksh_matrix_create(static_cast<ksh_problem *>(A->pb), o->precision, &C))) return rc;
    struct Guard {
        ksh_matrix *C;
        ~Guard() { ksh_matrix_destroy(C); }
    } guard{C};
    if ((rc = ksh_matrix_copy(C, A))) return rc;
    if (s != 0.0 && (rc = ksh_matrix_axpby(C, 1.0, -s, B))) return rc;
    if ((rc = ksh_factor(C))) return rc;
    Op op{C, buckling ? A : B, o->max_refine, o->refine_tol};
    // S = C^-1 M with M = B (shift-invert; buckling at s = 0) or A (buckling, s != 0).  Except
    // for buckling at s = 0, M = W and S v = C^-1 (W v) reuses the stored W-image.
    const bool reuse = !(buckling && s == 0.0);
    ksh_matrix *M = B;

    // basis V (m+1 vectors) and W-images WV
    std::vector<double> V((size_t)(m + 1) * n), WV((size_t)(m + 1) * n), w(n), Ww(n), Mv(n);
    auto vec = [&](std::vector<double> &X, int j) { return X.data() + (size_t)j * n; };
    std::mt19937_64 rng(o->seed ? o->seed : 20261202ull);
    std::normal_distribution<double> nd;
    for (int i = 0; i < n; ++i) w[i] = nd(rng);
    // start vector in the range of S (removes components of a singular W's null space)
    if (reuse) {
        if ((rc = op.apply_W(w.data(), Ww.data())) || (rc = op.solve(Ww.data(), w.data()))) return rc;
    } else {
        if ((rc = ksh_matvec(M, w.data(), Mv.data())) || (rc = op.solve(Mv.data(), w.data()))) return rc;
    }
    if ((rc = op.apply_W(w.data(), Ww.data()))) return rc;
    {
        const double nrm = std::sqrt(std::max(dot(w.data(), Ww.data(), n), 0.0));
        if (!(nrm > 0.0)) return fail(KSH_SOLVER_ERROR, "start vector has zero W-norm");
        for (int i = 0; i < n; ++i) {
            vec(V, 0)[i] = w[i] / nrm;
            vec(WV, 0)[i] = Ww[i] / nrm;
        }
    }

    // projected matrix T (m x m), thick-restart state: p kept Ritz values with couplings
    std::vector<double> T((size_t)m * m, 0.0);
    int p = 0;  // kept vectors at the start of the expansion (0 on the first cycle)
    int restarts = 0, converged = 0;
    std::vector<double> theta, Y;
    std::vector<int> order;
    double beta_m = 0.0;
    const int max_restarts = o->max_restarts > 0 ? o->max_restarts : 1000;
    for (;;) {
        // expand from column p to m-1
        for (int j = p; j < m; ++j) {
            // w = S v_j
            if (reuse) {
                if ((rc = op.solve(vec(WV, j), w.data()))) return rc;
            } else {
                if ((rc = ksh_matvec(M, vec(V, j), Mv.data())) || (rc = op.solve(Mv.data(), w.data()))) return rc;
            }
            if ((rc = op.apply_W(w.data(), Ww.data()))) return rc;
            // full reorthogonalization against v_0..v_j (two passes), recording T's column j
            for (int pass = 0; pass < 2; ++pass)
                for (int i = 0; i <= j; ++i) {
                    const double c = dot(w.data(), vec(WV, i), n);
                    if (pass == 0) {
                        T[(size_t)i * m + j] = c;
                        T[(size_t)j * m + i] = c;
                    } else {
                        T[(size_t)i * m + j] += c;
                        T[(size_t)j * m + i] = T[(size_t)i * m + j];
                    }
                    const double *vi = vec(V, i), *wvi = vec(WV, i);
                    for (int q = 0; q < n; ++q) {
                        w[q] -= c * vi[q];
                        Ww[q] -= c * wvi[q];
                    }
                }
            const double beta = std::sqrt(std::max(dot(w.data(), Ww.data(), n), 0.0));
            if (!(beta > 0.0)) return fail(KSH_SOLVER_ERROR, "Lanczos breakdown (invariant subspace)");
            for (int q = 0; q < n; ++q) {
                vec(V, j + 1)[q] = w[q] / beta;
                vec(WV, j + 1)[q] = Ww[q] / beta;
            }
            if (j + 1 < m) {
                T[(size_t)(j + 1) * m + j] = beta;
                T[(size_t)j * m + (j + 1)] = beta;
            } else {
                beta_m = beta;
            }
        }
        // Rayleigh-Ritz on T
        std::vector<double> Tc = T;
        jacobi_eig(Tc, m, theta, Y);
        order.resize(m);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int a, int b) { return std::fabs(theta[a]) > std::fabs(theta[b]); });
        converged = 0;
        for (int r = 0; r < k; ++r) {
            const int c = order[r];
            const double resid = std::fabs(beta_m * Y[(size_t)(m - 1) * m + c]);
            if (resid <= o->tol * std::max(std::fabs(theta[c]), 1e-300)) ++converged;
        }
        if (converged == k || restarts >= max_restarts) break;
        // thick restart: keep the p most wanted Ritz vectors
        p = std::min(m - 1, k + (m - k) / 2);
        std::vector<double> U((size_t)p * n, 0.0), WU((size_t)p * n, 0.0);
        for (int r = 0; r < p; ++r) {
            const int c = order[r];
            double *u = U.data() + (size_t)r * n, *wu = WU.data() + (size_t)r * n;
            for (int i = 0; i < m; ++i) {
                const double y = Y[(size_t)i * m + c];
                const double *vi = vec(V, i), *wvi = vec(WV, i);
                for (int q = 0; q < n; ++q) {
                    u[q] += y * vi[q];
                    wu[q] += y * wvi[q];
                }
            }
        }
        std::memcpy(vec(V, p), vec(V, m), sizeof(double) * n);  // the residual direction v_m
        std::memcpy(vec(WV, p), vec(WV, m), sizeof(double) * n);
        for (int r = 0; r < p; ++r) {
            std::memcpy(vec(V, r), U.data() + (size_t)r * n, sizeof(double) * n);
            std::memcpy(vec(WV, r), WU.data() + (size_t)r * n, sizeof(double) * n);
        }
        std::fill(T.begin(), T.end(), 0.0);
        for (int r = 0; r < p; ++r) {
            const int c = order[r];
            T[(size_t)r * m + r] = theta[c];
            const double sc = beta_m * Y[(size_t)(m - 1) * m + c];
            T[(size_t)r * m + p] = sc;
            T[(size_t)p * m + r] = sc;
        }
        ++restarts;
    }

    // eigenvalues of the pencil from the k wanted Ritz values, ascending
    std::vector<std::pair<double, int>> lam(k);
    for (int r = 0; r < k; ++r) {
        const double th = theta[order[r]];
        double l;
        if (!buckling) l = s + 1.0 / th;
        else if (s == 0.0) l = 1.0 / th;
        else l = s * th / (th - 1.0);
        lam[r] = {l, order[r]};
    }
    std::sort(lam.begin(), lam.end());
    for (int r = 0; r < k; ++r) values[r] = lam[r].first;
    if (vectors) {
        for (int r = 0; r < k; ++r) {
            double *x = vectors + (size_t)r * n;
            std::fill(x, x + n, 0.0);
            const int c = lam[r].second;
            for (int i = 0; i < m; ++i) {
                const double y = Y[(size_t)i * m + c];
                const double *vi = vec(V, i);
                for (int q = 0; q < n; ++q) x[q] += y * vi[q];
            }
        }
    }
    if (res) {
        res->converged = converged;
        res->restarts = restarts;
        res->solves = op.solves;
        res->products = op.products;
        res->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return converged == k ? KSH_SUCCESS : fail(KSH_NOT_CONVERGED, "eigen-solver: not all pairs converged");
}
