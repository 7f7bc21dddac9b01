// ksh_shell_math.hpp -- per-Gauss-point Kirchhoff-Love / St. Venant-Kirchhoff kernel algebra
// (device functions shared by the residual and tangent kernels of ksh_assembly.hip).
//
// Every first variation of the membrane strain eps = (a_ab - A_ab)/2 and of the bending strain
// kappa = b_ab - B_ab with respect to a control-point displacement u_{A,i} is linear in the
// 15-vector phi_{A,i} = (N_A,1 e_i, N_A,2 e_i, N_A,11 e_i, N_A,22 e_i, N_A,12 e_i); every second
// variation is a bilinear form in two of them.  gp_state() gathers the per-point quantities,
// h_entry() gives H_g[I][J] (K_{(A,i),(B,k)} = phi^T H_g phi) and f_entry() gives f_g[I]
// (f_int_{A,i} = phi . f_g).  Formulation: Kiendl et al. 2009, CMAME 198:3902.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "ksh_internal.hpp"

namespace ksh {
namespace shell {

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

}  // namespace shell
}  // namespace ksh
