// ksh_assembly.hip -- Gauss-point assembly of the St. Venant-Kirchhoff Kirchhoff-Love shell
// residual R(u), tangent K(u), mass M and load F on one tensor B-spline patch (gfx950).
//
// Reference: the hot path is gsKLShell's gsThinShellAssembler (external module, not in the
// reference tree); its call sites are benchmarks/benchmark_Roof.cpp:319-344 (assemble,
// assembleMatrix, assembleVector), solvers/modal_shell_XML.cpp:235-236 (assembleMass) and
// benchmarks/benchmark_Roof.cpp:304-306 (SvK material, analytical thickness integration).
//
// Formulation (Kiendl et al. 2009): eps = (a_ab - A_ab)/2, kappa = b_ab - B_ab,
// n = t C:eps, m = t^3/12 C:kappa.  Every first variation of eps and kappa with respect to a
// control-point displacement u_{A,i} is linear in the 15-vector
//     phi_{A,i} = (N_A,1 e_i, N_A,2 e_i, N_A,11 e_i, N_A,22 e_i, N_A,12 e_i),
// and every second variation is a bilinear form in two such vectors.  Hence per Gauss point
//     f_int_{A,i} = phi_{A,i} . f_g             (f_g  in R^15)
//     K_{(A,i),(B,k)} = phi_{A,i}^T H_g phi_{B,k}  (H_g in R^{15x15}, symmetric)
// with H_g = B^T diag(tC, t^3/12 C) B + G_n + G_m (B: 6x15 strain-variation operator,
// G: the second variations of the strains weighted by n and m).  The element tangent is then
// the batched B^T D B contraction
//     K_e^{ik}[A][B] = sum_{g,a} Psi_g[a][A] * ( sum_b H_g[(a,i),(b,k)] Psi_g[b][B] ),
// a 16x16 output tile per component pair (i,k) with inner dimension 5*ngp, run on
// v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (FP32).  The element tiles are then
// gathered into the band in a fixed element order (deterministic, no atomics).
#include <hip/hip_runtime.h>

#include <cmath>

#include "ksh_internal.hpp"
#include "ksh_mfma.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {

constexpr int kState = 88;

__device__ inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// state layout (doubles)
enum {
    S_A1 = 0, S_A2 = 3, S_A3 = 6, S_J = 9, S_NV = 10, S_MV = 13, S_DM = 16, S_DB = 22, S_MU = 28,
    S_Q = 31, S_Q1 = 34, S_Q2 = 43, S_C1 = 52, S_C2 = 61, S_TM = 70, S_TA = 76, S_MA = 82, S_G = 83
};

// symmetric 3x3 from 6 stored entries (00,11,22,01,02,12)
__device__ inline double sym3(const double *c, int r, int s)
{
    if (r == s) return c[r];
    int lo = r < s ? r : s, hi = r < s ? s : r;
    return lo == 0 ? (hi == 1 ? c[3] : c[4]) : c[5];
}

// Per-Gauss-point state from the deformed derivative vectors dv = (a1,a2,a11,a22,a12) and the
// undeformed geometry geo (wdA, C[6], A_ab[3], B_ab[3]).  All stress-like quantities carry wdA.
__device__ void gp_state(const double *dv, const double *geo, double t, double *s)
{
    const double *a1 = dv, *a2 = dv + 3, *a11 = dv + 6, *a22 = dv + 9, *a12 = dv + 12;
    double at[3], a3[3];
    cross3(a1, a2, at);
    double j = sqrt(dot3(at, at));
    for (int i = 0; i < 3; ++i) a3[i] = at[i] / j;
    double wdA = geo[0];
    const double *C = geo + 1;
    double ev[3] = {0.5 * (dot3(a1, a1) - geo[7]), 0.5 * (dot3(a2, a2) - geo[8]), dot3(a1, a2) - geo[9]};
    double kv[3] = {dot3(a11, a3) - geo[10], dot3(a22, a3) - geo[11], 2.0 * (dot3(a12, a3) - geo[12])};
    double fm = t * wdA, fb = t * t * t / 12.0 * wdA;
    for (int q = 0; q < 6; ++q) {
        s[S_DM + q] = fm * C[q];
        s[S_DB + q] = fb * C[q];
    }
    double nv[3], mv[3];
    for (int r = 0; r < 3; ++r) {
        nv[r] = fm * (sym3(C, r, 0) * ev[0] + sym3(C, r, 1) * ev[1] + sym3(C, r, 2) * ev[2]);
        mv[r] = fb * (sym3(C, r, 0) * kv[0] + sym3(C, r, 1) * kv[1] + sym3(C, r, 2) * kv[2]);
    }
    double mu[3] = {mv[0], mv[1], 2.0 * mv[2]};
    double mh[3];
    for (int i = 0; i < 3; ++i) mh[i] = mu[0] * a11[i] + mu[1] * a22[i] + mu[2] * a12[i];
    // P = I - a3 a3^T ; Q1 = -P S(a2)/j ; Q2 = P S(a1)/j
    double S1[3][3] = {{0, -a1[2], a1[1]}, {a1[2], 0, -a1[0]}, {-a1[1], a1[0], 0}};
    double S2[3][3] = {{0, -a2[2], a2[1]}, {a2[2], 0, -a2[0]}, {-a2[1], a2[0], 0}};
    double Q1[3][3], Q2[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double p1 = 0.0, p2 = 0.0;
            for (int l = 0; l < 3; ++l) {
                double Prl = (r == l ? 1.0 : 0.0) - a3[r] * a3[l];
                p1 += Prl * S2[l][c];
                p2 += Prl * S1[l][c];
            }
            Q1[r][c] = -p1 / j;
            Q2[r][c] = p2 / j;
        }
    double ma = dot3(mh, a3);
    for (int i = 0; i < 3; ++i) {
        s[S_A1 + i] = a1[i];
        s[S_A2 + i] = a2[i];
        s[S_A3 + i] = a3[i];
        s[S_NV + i] = nv[i];
        s[S_MV + i] = mv[i];
        s[S_MU + i] = mu[i];
        s[S_Q + i] = (mh[i] - ma * a3[i]) / j;
    }
    s[S_J] = j;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            s[S_Q1 + 3 * r + c] = Q1[r][c];
            s[S_Q2 + 3 * r + c] = Q2[r][c];
        }
    const double *aab[3] = {a11, a22, a12};
    for (int r = 0; r < 3; ++r)
        for (int i = 0; i < 3; ++i) {
            double c1 = 0.0, c2 = 0.0;
            for (int l = 0; l < 3; ++l) {
                c1 += aab[r][l] * Q1[l][i];
                c2 += aab[r][l] * Q2[l][i];
            }
            s[S_C1 + 3 * r + i] = c1;
            s[S_C2 + 3 * r + i] = c2;
        }
    double t1[3], t2[3];
    cross3(a2, mh, t1);
    cross3(mh, a1, t2);
    for (int i = 0; i < 3; ++i) {
        s[S_TM + i] = t1[i];
        s[S_TM + 3 + i] = t2[i];
    }
    cross3(a2, a3, t1);
    cross3(a3, a1, t2);
    for (int i = 0; i < 3; ++i) {
        s[S_TA + i] = t1[i];
        s[S_TA + 3 + i] = t2[i];
    }
    s[S_MA] = ma;
    s[S_G + 0] = dot3(a1, a1);
    s[S_G + 1] = dot3(a2, a2);
    s[S_G + 2] = dot3(a1, a2);
}

// column I = 3a+i of the 6x15 strain-variation operator B (rows: eps11, eps22, 2eps12,
// kap11, kap22, 2kap12)
__device__ inline void bcol(const double *s, int I, double *m, double *b)
{
    int a = I / 3, i = I % 3;
    m[0] = (a == 0) ? s[S_A1 + i] : 0.0;
    m[1] = (a == 1) ? s[S_A2 + i] : 0.0;
    m[2] = (a == 0) ? s[S_A2 + i] : ((a == 1) ? s[S_A1 + i] : 0.0);
    for (int r = 0; r < 3; ++r) {
        double f = (r == 2) ? 2.0 : 1.0;
        double v = 0.0;
        if (a == 0) v = s[S_C1 + 3 * r + i];
        else if (a == 1) v = s[S_C2 + 3 * r + i];
        else if (a == 2 + r) v = s[S_A3 + i];
        b[r] = f * v;
    }
}

// H_g[I][J]
__device__ double h_entry(const double *s, int I, int J)
{
    double mI[3], bI[3], mJ[3], bJ[3];
    bcol(s, I, mI, bI);
    bcol(s, J, mJ, bJ);
    double h = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) h += mI[r] * sym3(s + S_DM, r, c) * mJ[c] + bI[r] * sym3(s + S_DB, r, c) * bJ[c];
    int a = I / 3, i = I % 3, b = J / 3, k = J % 3;
    if (a < 2 && b < 2) {
        double j2 = s[S_J] * s[S_J];
        // second variation of eps weighted by n
        if (i == k) h += (a == 0 && b == 0) ? s[S_NV] : ((a == 1 && b == 1) ? s[S_NV + 1] : s[S_NV + 2]);
        // m^ . d2 a3 : cross-product part
        const double *q = s + S_Q;
        double Sq = 0.0;  // S(q)[i][k]
        if (i != k) {
            int l = 3 - i - k;
            double sgn = ((i == 0 && k == 1) || (i == 1 && k == 2) || (i == 2 && k == 0)) ? -1.0 : 1.0;
            Sq = sgn * q[l];
        }
        if (a == 0 && b == 1) h -= Sq;
        if (a == 1 && b == 0) h += Sq;
        int Ip = 3 * a + i, Jp = 3 * b + k;
        const double *tm = s + S_TM, *ta = s + S_TA;
        double ma = s[S_MA];
        h -= (tm[Ip] * ta[Jp] + ta[Ip] * tm[Jp]) / j2;
        double TT;
        const double *a1 = s + S_A1, *a2 = s + S_A2;
        double dik = (i == k) ? 1.0 : 0.0;
        if (a == 0 && b == 0) TT = s[S_G + 1] * dik - a2[i] * a2[k];
        else if (a == 1 && b == 1) TT = s[S_G + 0] * dik - a1[i] * a1[k];
        else if (a == 0) TT = -s[S_G + 2] * dik + a1[i] * a2[k];
        else TT = -s[S_G + 2] * dik + a2[i] * a1[k];
        h -= ma * TT / j2;
        h += 3.0 * ma * ta[Ip] * ta[Jp] / j2;
    } else if (a >= 2 && b < 2) {
        const double *Q = (b == 0) ? s + S_Q1 : s + S_Q2;
        h += s[S_MU + a - 2] * Q[3 * i + k];
    } else if (a < 2 && b >= 2) {
        const double *Q = (a == 0) ? s + S_Q1 : s + S_Q2;
        h += s[S_MU + b - 2] * Q[3 * k + i];
    }
    return h;
}

__device__ inline double f_entry(const double *s, int I)
{
    double m[3], b[3];
    bcol(s, I, m, b);
    return m[0] * s[S_NV] + m[1] * s[S_NV + 1] + m[2] * s[S_NV + 2] + b[0] * s[S_MV] + b[1] * s[S_MV + 1] +
           b[2] * s[S_MV + 2];
}

struct ElemArgs {
    int p, nx, ny, mx, my;
    double t;
    const double *ctrl0, *u_full, *tabx, *taby, *gpgeo;
};

// Psi[(g,a)][A] for a in {,1 ,2 ,11 ,22 ,12}; zero padded to [NKK][16]
template <int P>
__device__ void load_psi(const ElemArgs &ar, int ex, int ey, double (*sPsi)[16], int nkk)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NQ = P1, NG = NQ * NQ;
    for (int idx = threadIdx.x; idx < nkk * 16; idx += blockDim.x) {
        int kk = idx / 16, A = idx % 16;
        double v = 0.0;
        if (kk < 5 * NG && A < NB) {
            int g = kk / 5, a = kk % 5, qx = g % NQ, qy = g / NQ, ia = A % P1, ib = A / P1;
            const double *tx = ar.tabx + ((size_t)(ex * NQ + qx) * 3) * P1;
            const double *ty = ar.taby + ((size_t)(ey * NQ + qy) * 3) * P1;
            switch (a) {
            case 0: v = tx[P1 + ia] * ty[ib]; break;
            case 1: v = tx[ia] * ty[P1 + ib]; break;
            case 2: v = tx[2 * P1 + ia] * ty[ib]; break;
            case 3: v = tx[ia] * ty[2 * P1 + ib]; break;
            default: v = tx[P1 + ia] * ty[P1 + ib]; break;
            }
        }
        sPsi[kk][A] = v;
    }
}

// deformed control points + Psi + derivative vectors + per-GP state, shared by R and K kernels
template <int P>
__device__ void element_states(const ElemArgs &ar, int ex, int ey, double (*sX)[3], double (*sPsi)[16],
                               double (*sDv)[15], double (*sSt)[kState], int nkk)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NG = P1 * P1;
    for (int idx = threadIdx.x; idx < NB * 3; idx += blockDim.x) {
        int A = idx / 3, c = idx % 3, a = A % P1, b = A / P1;
        int cp = (ex + a) + ar.mx * (ey + b);
        sX[A][c] = ar.ctrl0[3 * cp + c] + ar.u_full[3 * cp + c];
    }
    load_psi<P>(ar, ex, ey, sPsi, nkk);
    __syncthreads();
    for (int idx = threadIdx.x; idx < NG * 15; idx += blockDim.x) {
        int g = idx / 15, r = idx % 15, a = r / 3, c = r % 3;
        double acc = 0.0;
        for (int A = 0; A < NB; ++A) acc += sPsi[5 * g + a][A] * sX[A][c];
        sDv[g][r] = acc;
    }
    __syncthreads();
    int e = ex + ar.nx * ey;
    for (int g = threadIdx.x; g < NG; g += blockDim.x)
        gp_state(sDv[g], ar.gpgeo + ((size_t)e * NG + g) * kGeoStride, ar.t, sSt[g]);
    __syncthreads();
}

// ---------------------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void k_residual_elem(ElemArgs ar, double *Fe)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NG = P1 * P1, NKK = ((5 * NG + 3) / 4) * 4;
    __shared__ double sX[NB][3];
    __shared__ double sPsi[NKK][16];
    __shared__ double sDv[NG][15];
    __shared__ double sSt[NG][kState];
    __shared__ double sF[NG][15];
    int e = blockIdx.x, ex = e % ar.nx, ey = e / ar.nx;
    element_states<P>(ar, ex, ey, sX, sPsi, sDv, sSt, NKK);
    for (int idx = threadIdx.x; idx < NG * 15; idx += blockDim.x) {
        int g = idx / 15, I = idx % 15;
        sF[g][I] = f_entry(sSt[g], I);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < NB * 3; idx += blockDim.x) {
        int A = idx / 3, i = idx % 3;
        double acc = 0.0;
        for (int g = 0; g < NG; ++g)
            for (int a = 0; a < 5; ++a) acc += sPsi[5 * g + a][A] * sF[g][3 * a + i];
        Fe[(size_t)e * 3 * NB + idx] = acc;
    }
}


// Element tangent: per component row i, H rows of component i -> Z^i = H Psi -> 3 MFMA tiles.
template <int P, typename T>
__global__ __launch_bounds__(256) void k_tangent_elem(ElemArgs ar, T *Ke)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NG = P1 * P1, NKK = ((5 * NG + 3) / 4) * 4;
    __shared__ double sX[NB][3];
    __shared__ double sPsi[NKK][16];
    __shared__ double sDv[NG][15];
    __shared__ double sSt[NG][kState];
    __shared__ double sH[NG][5][15];
    __shared__ T sZ[NKK][3][16];
    __shared__ T sPsiT[NKK][16];
    int e = blockIdx.x, ex = e % ar.nx, ey = e / ar.nx;
    element_states<P>(ar, ex, ey, sX, sPsi, sDv, sSt, NKK);
    for (int idx = threadIdx.x; idx < NKK * 16; idx += blockDim.x) sPsiT[idx / 16][idx % 16] = (T)sPsi[idx / 16][idx % 16];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    T *Kout = Ke + (size_t)e * kKeTile;
    for (int i = 0; i < 3; ++i) {
        for (int idx = threadIdx.x; idx < NG * 5 * 15; idx += blockDim.x) {
            int g = idx / 75, r = idx % 75, a = r / 15, J = r % 15;
            sH[g][a][J] = h_entry(sSt[g], 3 * a + i, J);
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < NKK * 48; idx += blockDim.x) {
            int kk = idx / 48, col = idx % 48, k = col / 16, B = col % 16;
            double z = 0.0;
            if (kk < 5 * NG && B < NB) {
                int g = kk / 5, a = kk % 5;
                for (int b = 0; b < 5; ++b) z += sH[g][a][3 * b + k] * sPsi[5 * g + b][B];
            }
            sZ[kk][k][B] = (T)z;
        }
        __syncthreads();
        if (wave < 3) {
            const int k = wave;
            typename Mfma16<T>::acc_t acc = {0, 0, 0, 0};
            for (int st = 0; st < NKK / 4; ++st) {
                int kk = 4 * st + lane / 16;
                acc = Mfma16<T>::run(sPsiT[kk][lane % 16], sZ[kk][k][lane % 16], acc);
            }
            T *tile = Kout + (i * 3 + k) * 256;
            for (int v = 0; v < 4; ++v) tile[Mfma16<T>::row(lane, v) * 16 + lane % 16] = acc[v];
        }
        __syncthreads();
    }
}

// consistent mass tiles: M_e^{ii}[A][B] = sum_g rho t N_A N_B wdA
template <int P, typename T>
__global__ __launch_bounds__(256) void k_mass_elem(ElemArgs ar, double rho, T *Ke)
{
    constexpr int P1 = P + 1, NB = P1 * P1, NQ = P1, NG = NQ * NQ;
    __shared__ double sN[NG][16];
    __shared__ double sW[NG];
    int e = blockIdx.x, ex = e % ar.nx, ey = e / ar.nx;
    for (int idx = threadIdx.x; idx < NG * 16; idx += blockDim.x) {
        int g = idx / 16, A = idx % 16;
        double v = 0.0;
        if (A < NB) {
            int qx = g % NQ, qy = g / NQ, ia = A % P1, ib = A / P1;
            v = ar.tabx[((size_t)(ex * NQ + qx) * 3) * P1 + ia] * ar.taby[((size_t)(ey * NQ + qy) * 3) * P1 + ib];
        }
        sN[g][A] = v;
    }
    for (int g = threadIdx.x; g < NG; g += blockDim.x) sW[g] = ar.gpgeo[((size_t)e * NG + g) * kGeoStride];
    __syncthreads();
    T *Kout = Ke + (size_t)e * kKeTile;
    for (int idx = threadIdx.x; idx < 256; idx += blockDim.x) {
        int A = idx / 16, B = idx % 16;
        double acc = 0.0;
        for (int g = 0; g < NG; ++g) acc += rho * ar.t * sN[g][A] * sN[g][B] * sW[g];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) Kout[(i * 3 + k) * 256 + idx] = (i == k) ? (T)acc : (T)0;
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
    double d[5][3] = {};
    for (int A = 0; A < NB; ++A) {
        int ia = A % P1, ib = A / P1;
        int cp = (ex + ia) + ar.mx * (ey + ib);
        double psi[5] = {tx[P1 + ia] * ty[ib], tx[ia] * ty[P1 + ib], tx[2 * P1 + ia] * ty[ib], tx[ia] * ty[2 * P1 + ib],
                         tx[P1 + ia] * ty[P1 + ib]};
        for (int a = 0; a < 5; ++a)
            for (int c = 0; c < 3; ++c) d[a][c] += psi[a] * ar.ctrl0[3 * cp + c];
    }
    double at[3], a3[3];
    cross3(d[0], d[1], at);
    double J0 = sqrt(dot3(at, at));
    for (int i = 0; i < 3; ++i) a3[i] = at[i] / J0;
    double g11 = dot3(d[0], d[0]), g22 = dot3(d[1], d[1]), g12 = dot3(d[0], d[1]);
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
    o[10] = dot3(d[2], a3);
    o[11] = dot3(d[3], a3);
    o[12] = dot3(d[4], a3);
    o[13] = o[14] = o[15] = 0.0;
}

__global__ void k_construct_solution(int ndof, const int *free_idx, const double *u_free, double *u_full)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ndof) return;
    int f = free_idx[g];
    u_full[g] = (f >= 0) ? u_free[f] : 0.0;
}

// Deterministic gather of element vectors into free dofs (fixed element order ey, ex).
// mode 0: out = sum ; mode 1 (ALM): out = sum - lambda F ; mode 2 (static): out = lambda F - sum
__global__ void k_vector_gather(int nfree, int p, int nx, int ny, int mx, const int *free2full, const double *Fe,
                                const double *F, double lambda, int mode, double *out)
{
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nfree) return;
    int P1 = p + 1, NB = P1 * P1;
    int g = free2full[I], cp = g / 3, c = g % 3, i = cp % mx, j = cp / mx;
    int exlo = max(i - p, 0), exhi = min(i, nx - 1), eylo = max(j - p, 0), eyhi = min(j, ny - 1);
    double acc = 0.0;
    for (int ey = eylo; ey <= eyhi; ++ey)
        for (int ex = exlo; ex <= exhi; ++ex) {
            int e = ex + nx * ey, A = (i - ex) + P1 * (j - ey);
            acc += Fe[(size_t)e * 3 * NB + 3 * A + c];
        }
    if (mode == 1) acc = acc - lambda * F[I];
    else if (mode == 2) acc = lambda * F[I] - acc;
    out[I] = acc;
}

// Deterministic gather of element tiles into the lower band: one thread per band slot.
template <typename T>
__global__ void k_band_gather(int n, int bw, int p, int nx, int ny, int mx, const int *free2full, const T *Ke, T *ab)
{
    size_t slot = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t ld = (size_t)bw + 1;
    if (slot >= (size_t)n * ld) return;
    int J = (int)(slot / ld), d = (int)(slot % ld), I = J + d;
    double acc = 0.0;
    if (I < n) {
        int P1 = p + 1;
        int gI = free2full[I], gJ = free2full[J];
        int cpI = gI / 3, cI = gI % 3, cpJ = gJ / 3, cJ = gJ % 3;
        int iI = cpI % mx, jI = cpI / mx, iJ = cpJ % mx, jJ = cpJ / mx;
        int exlo = max(max(iI, iJ) - p, 0), exhi = min(min(iI, iJ), nx - 1);
        int eylo = max(max(jI, jJ) - p, 0), eyhi = min(min(jI, jJ), ny - 1);
        for (int ey = eylo; ey <= eyhi; ++ey)
            for (int ex = exlo; ex <= exhi; ++ex) {
                int e = ex + nx * ey;
                int A = (iI - ex) + P1 * (jI - ey), B = (iJ - ex) + P1 * (jJ - ey);
                acc += (double)Ke[(size_t)e * kKeTile + (cI * 3 + cJ) * 256 + A * 16 + B];
            }
    }
    ab[slot] = (T)acc;
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

ElemArgs elem_args(const Problem &pb)
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

inline unsigned nblk(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
int launch_setup_geometry(Problem &pb)
{
    ElemArgs ar = elem_args(pb);
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
    hipLaunchKernelGGL(k_construct_solution, dim3(nblk(pb.ndof, 256)), dim3(256), 0, pb.stream, pb.ndof, pb.free_idx,
                       u_free_dev, pb.u_full);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int launch_residual(Problem &pb, double lambda, int conv, double *R_dev)
{
    ElemArgs ar = elem_args(pb);
    {
        ProfScope ps("residual_elem", pb.stream);
        if (pb.p == 2)
            hipLaunchKernelGGL(k_residual_elem<2>, dim3(pb.nel), dim3(256), 0, pb.stream, ar, pb.Fe);
        else
            hipLaunchKernelGGL(k_residual_elem<3>, dim3(pb.nel), dim3(256), 0, pb.stream, ar, pb.Fe);
    }
    KSH_HIP(hipGetLastError());
    int mode = (conv == KSH_RES_ALM) ? 1 : 2;
    ProfScope ps("vector_gather", pb.stream);
    hipLaunchKernelGGL(k_vector_gather, dim3(nblk(pb.nfree, 256)), dim3(256), 0, pb.stream, pb.nfree, pb.p, pb.nx, pb.ny,
                       pb.mx, pb.free2full, pb.Fe, pb.F, lambda, mode, R_dev);
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int launch_external_force(Problem &pb)
{
    ElemArgs ar = elem_args(pb);
    KSH_HIP(hipMemsetAsync(pb.F, 0, sizeof(double) * pb.nfree, pb.stream));
    if (pb.fsurf[0] != 0.0 || pb.fsurf[1] != 0.0 || pb.fsurf[2] != 0.0) {
        if (pb.p == 2)
            hipLaunchKernelGGL(k_surface_elem<2>, dim3(pb.nel), dim3(64), 0, pb.stream, ar, pb.fsurf[0], pb.fsurf[1],
                               pb.fsurf[2], pb.Fe);
        else
            hipLaunchKernelGGL(k_surface_elem<3>, dim3(pb.nel), dim3(64), 0, pb.stream, ar, pb.fsurf[0], pb.fsurf[1],
                               pb.fsurf[2], pb.Fe);
        KSH_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_vector_gather, dim3(nblk(pb.nfree, 256)), dim3(256), 0, pb.stream, pb.nfree, pb.p, pb.nx,
                           pb.ny, pb.mx, pb.free2full, pb.Fe, pb.F, 0.0, 0, pb.F);
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
    ElemArgs ar = elem_args(pb);
    size_t slots = (size_t)K.n * K.ld;
    if (K.sprec == KSH_FP64) {
        double *Ke = (double *)pb.Ke;
        {
            ProfScope ps("tangent_elem", pb.stream);
            if (pb.p == 2)
                hipLaunchKernelGGL((k_tangent_elem<2, double>), dim3(pb.nel), dim3(256), 0, pb.stream, ar, Ke);
            else
                hipLaunchKernelGGL((k_tangent_elem<3, double>), dim3(pb.nel), dim3(256), 0, pb.stream, ar, Ke);
        }
        KSH_HIP(hipGetLastError());
        ProfScope ps("band_gather", pb.stream);
        hipLaunchKernelGGL(k_band_gather<double>, dim3(nblk(slots, 256)), dim3(256), 0, pb.stream, K.n, K.bw, pb.p, pb.nx,
                           pb.ny, pb.mx, pb.free2full, Ke, (double *)K.ab);
    } else {
        float *Ke = (float *)pb.Ke;
        {
            ProfScope ps("tangent_elem", pb.stream);
            if (pb.p == 2)
                hipLaunchKernelGGL((k_tangent_elem<2, float>), dim3(pb.nel), dim3(256), 0, pb.stream, ar, Ke);
            else
                hipLaunchKernelGGL((k_tangent_elem<3, float>), dim3(pb.nel), dim3(256), 0, pb.stream, ar, Ke);
        }
        KSH_HIP(hipGetLastError());
        ProfScope ps("band_gather", pb.stream);
        hipLaunchKernelGGL(k_band_gather<float>, dim3(nblk(slots, 256)), dim3(256), 0, pb.stream, K.n, K.bw, pb.p, pb.nx,
                           pb.ny, pb.mx, pb.free2full, Ke, (float *)K.ab);
    }
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

int launch_mass(Problem &pb, Matrix &M)
{
    ElemArgs ar = elem_args(pb);
    size_t slots = (size_t)M.n * M.ld;
    if (M.sprec == KSH_FP64) {
        double *Ke = (double *)pb.Ke;
        if (pb.p == 2)
            hipLaunchKernelGGL((k_mass_elem<2, double>), dim3(pb.nel), dim3(256), 0, pb.stream, ar, pb.rho, Ke);
        else
            hipLaunchKernelGGL((k_mass_elem<3, double>), dim3(pb.nel), dim3(256), 0, pb.stream, ar, pb.rho, Ke);
        KSH_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_band_gather<double>, dim3(nblk(slots, 256)), dim3(256), 0, pb.stream, M.n, M.bw, pb.p, pb.nx,
                           pb.ny, pb.mx, pb.free2full, Ke, (double *)M.ab);
    } else {
        float *Ke = (float *)pb.Ke;
        if (pb.p == 2)
            hipLaunchKernelGGL((k_mass_elem<2, float>), dim3(pb.nel), dim3(256), 0, pb.stream, ar, pb.rho, Ke);
        else
            hipLaunchKernelGGL((k_mass_elem<3, float>), dim3(pb.nel), dim3(256), 0, pb.stream, ar, pb.rho, Ke);
        KSH_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_band_gather<float>, dim3(nblk(slots, 256)), dim3(256), 0, pb.stream, M.n, M.bw, pb.p, pb.nx,
                           pb.ny, pb.mx, pb.free2full, Ke, (float *)M.ab);
    }
    KSH_HIP(hipGetLastError());
    return KSH_SUCCESS;
}

}  // namespace ksh
