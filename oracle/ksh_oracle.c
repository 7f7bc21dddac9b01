/*
 * ksh_oracle.c -- FP64 CPU oracle (test infrastructure only; see ksh_oracle.h).
 *
 * Written in the plainest form that states the mathematics: per Gauss point the
 * first variations of the membrane and bending strains are formed per element
 * dof, the second variations per dof pair, and everything is summed in FP64
 * with -ffp-contract=off.  Nothing here is shared with the HIP path, which
 * uses a different (factorised) formulation of the same tangent.
 */
#include "ksh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OK_MAXP 4
#define OK_MAXP1 (OK_MAXP + 1)
#define OK_MAXNB (OK_MAXP1 * OK_MAXP1)
#define OK_MAXNE (3 * OK_MAXNB)

/* ------------------------------------------------------------------------- */
/* small vector helpers                                                       */
/* ------------------------------------------------------------------------- */
static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* ------------------------------------------------------------------------- */
/* splines and quadrature                                                     */
/* ------------------------------------------------------------------------- */
void ok_knots_uniform(int p, int n, double *U)
{
    int k = 0;
    for (int i = 0; i <= p; ++i) U[k++] = 0.0;
    for (int i = 1; i < n; ++i) U[k++] = (double)i / (double)n;
    for (int i = 0; i <= p; ++i) U[k++] = 1.0;
}

int ok_find_span(int p, int n, double x)
{
    int e = (int)floor(x * n);
    if (e < 0) e = 0;
    if (e > n - 1) e = n - 1;
    return e + p;
}

/*
 * Non-zero basis functions N_{span-p..span} and their derivatives up to order nd
 * at x (The NURBS Book, Piegl & Tiller, algorithm A2.3).  ders[k*(p+1)+j].
 */
void ok_ders_basis(int p, const double *U, int span, double x, int nd, double *ders)
{
    double ndu[OK_MAXP1][OK_MAXP1], left[OK_MAXP1], right[OK_MAXP1], a[2][OK_MAXP1];
    ndu[0][0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = x - U[span + 1 - j];
        right[j] = U[span + j] - x;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            ndu[j][r] = right[r + 1] + left[j - r];
            double tmp = ndu[r][j - 1] / ndu[j][r];
            ndu[r][j] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        ndu[j][j] = saved;
    }
    for (int j = 0; j <= p; ++j) ders[j] = ndu[j][p];
    for (int r = 0; r <= p; ++r) {
        int s1 = 0, s2 = 1;
        a[0][0] = 1.0;
        for (int k = 1; k <= nd; ++k) {
            double d = 0.0;
            int rk = r - k, pk = p - k;
            if (r >= k) {
                a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
                d = a[s2][0] * ndu[rk][pk];
            }
            int j1 = (rk >= -1) ? 1 : -rk;
            int j2 = (r - 1 <= pk) ? k - 1 : p - r;
            for (int j = j1; j <= j2; ++j) {
                a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
                d += a[s2][j] * ndu[rk + j][pk];
            }
            if (r <= pk) {
                a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
                d += a[s2][k] * ndu[r][pk];
            }
            ders[k * (p + 1) + r] = d;
            int t = s1; s1 = s2; s2 = t;
        }
    }
    double fac = p;
    for (int k = 1; k <= nd; ++k) {
        for (int j = 0; j <= p; ++j) ders[k * (p + 1) + j] *= fac;
        fac *= (double)(p - k);
    }
}

/* Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_nq. */
void ok_gauss_legendre(int nq, double *x, double *w)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (nq + 1) / 2; ++i) {
        double z = cos(pi * (i + 0.75) / (nq + 0.5)), z1, pp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= nq; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = nq * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-15) break;
        }
        {   /* final derivative at the converged node */
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= nq; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = nq * (z * p1 - p2) / (z * z - 1.0);
        }
        x[i] = -z;
        x[nq - 1 - i] = z;
        w[i] = w[nq - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

/* ------------------------------------------------------------------------- */
/* dof maps                                                                   */
/* ------------------------------------------------------------------------- */
int ok_free_map(const ok_problem *pb, int *free_idx)
{
    int ndof = 3 * pb->mx * pb->my, nf = 0;
    for (int g = 0; g < ndof; ++g) free_idx[g] = pb->fixed[g] ? -1 : nf++;
    return nf;
}

int ok_bandwidth(const ok_problem *pb, const int *free_idx)
{
    int p1 = pb->p + 1, bw = 0;
    for (int ey = 0; ey < pb->ny; ++ey)
        for (int ex = 0; ex < pb->nx; ++ex) {
            int lo = 1 << 30, hi = -1;
            for (int b = 0; b < p1; ++b)
                for (int a = 0; a < p1; ++a)
                    for (int c = 0; c < 3; ++c) {
                        int g = 3 * ((ex + a) + pb->mx * (ey + b)) + c;
                        int f = free_idx[g];
                        if (f < 0) continue;
                        if (f < lo) lo = f;
                        if (f > hi) hi = f;
                    }
            if (hi >= 0 && hi - lo > bw) bw = hi - lo;
        }
    return bw;
}

/* ------------------------------------------------------------------------- */
/* element kernel                                                             */
/* ------------------------------------------------------------------------- */
enum { OK_W = 1, OK_F = 2, OK_K = 4, OK_FEXT = 8, OK_M = 16 };

typedef struct {
    int p, p1, nb, ne, nq;
    double gx[OK_MAXP1], gw[OK_MAXP1];
} ok_ctx;

static void ctx_init(const ok_problem *pb, ok_ctx *c, double **Ux, double **Uy)
{
    c->p = pb->p;
    c->p1 = pb->p + 1;
    c->nb = c->p1 * c->p1;
    c->ne = 3 * c->nb;
    c->nq = pb->p + 1;
    *Ux = (double *)malloc(sizeof(double) * (pb->nx + 2 * pb->p + 1));
    *Uy = (double *)malloc(sizeof(double) * (pb->ny + 2 * pb->p + 1));
    ok_knots_uniform(pb->p, pb->nx, *Ux);
    ok_knots_uniform(pb->p, pb->ny, *Uy);
    ok_gauss_legendre(c->nq, c->gx, c->gw);
}

/*
 * One element.  Local dof r = 3*A + i with A = a + (p+1)*b the local basis
 * function (a along x, b along y) and i the displacement component.
 *   We : strain energy, fe : internal force, Ke : tangent (ne x ne, row major),
 *   fx : external (dead surface) force, Me : consistent mass.
 */
static void ok_element(const ok_problem *pb, const ok_ctx *c, const double *Ux, const double *Uy,
                       const double *u_full, int ex, int ey, int want,
                       double *We, double *fe, double *Ke, double *fx, double *Me)
{
    const int p = c->p, p1 = c->p1, nb = c->nb, ne = c->ne, nq = c->nq;
    double P[OK_MAXNB][3], X[OK_MAXNB][3];
    for (int b = 0; b < p1; ++b)
        for (int a = 0; a < p1; ++a) {
            int A = a + p1 * b, cp = (ex + a) + pb->mx * (ey + b);
            for (int i = 0; i < 3; ++i) {
                P[A][i] = pb->ctrl[3 * cp + i];
                X[A][i] = P[A][i] + (u_full ? u_full[3 * cp + i] : 0.0);
            }
        }
    if (want & OK_W) *We = 0.0;
    if (want & OK_F) memset(fe, 0, sizeof(double) * ne);
    if (want & OK_K) memset(Ke, 0, sizeof(double) * ne * ne);
    if (want & OK_FEXT) memset(fx, 0, sizeof(double) * ne);
    if (want & OK_M) memset(Me, 0, sizeof(double) * ne * ne);

    const double t = pb->t;
    const double Em = pb->E / (1.0 - pb->nu * pb->nu), nu = pb->nu;

    for (int qy = 0; qy < nq; ++qy)
        for (int qx = 0; qx < nq; ++qx) {
            double xi = (ex + 0.5 * (1.0 + c->gx[qx])) / pb->nx;
            double eta = (ey + 0.5 * (1.0 + c->gx[qy])) / pb->ny;
            double wq = c->gw[qx] * c->gw[qy] * 0.25 / ((double)pb->nx * (double)pb->ny);
            double dx[3 * OK_MAXP1], dy[3 * OK_MAXP1];
            ok_ders_basis(p, Ux, ex + p, xi, 2, dx);
            ok_ders_basis(p, Uy, ey + p, eta, 2, dy);
            double N[OK_MAXNB], N1[OK_MAXNB], N2[OK_MAXNB], N11[OK_MAXNB], N22[OK_MAXNB], N12[OK_MAXNB];
            for (int b = 0; b < p1; ++b)
                for (int a = 0; a < p1; ++a) {
                    int A = a + p1 * b;
                    N[A] = dx[a] * dy[b];
                    N1[A] = dx[p1 + a] * dy[b];
                    N2[A] = dx[a] * dy[p1 + b];
                    N11[A] = dx[2 * p1 + a] * dy[b];
                    N22[A] = dx[a] * dy[2 * p1 + b];
                    N12[A] = dx[p1 + a] * dy[p1 + b];
                }
            /* undeformed and deformed covariant bases and their derivatives */
            double G1[3] = {0}, G2[3] = {0}, G11[3] = {0}, G22[3] = {0}, G12[3] = {0};
            double a1[3] = {0}, a2[3] = {0}, a11[3] = {0}, a22[3] = {0}, a12[3] = {0};
            for (int A = 0; A < nb; ++A)
                for (int i = 0; i < 3; ++i) {
                    G1[i] += N1[A] * P[A][i];
                    G2[i] += N2[A] * P[A][i];
                    G11[i] += N11[A] * P[A][i];
                    G22[i] += N22[A] * P[A][i];
                    G12[i] += N12[A] * P[A][i];
                    a1[i] += N1[A] * X[A][i];
                    a2[i] += N2[A] * X[A][i];
                    a11[i] += N11[A] * X[A][i];
                    a22[i] += N22[A] * X[A][i];
                    a12[i] += N12[A] * X[A][i];
                }
            double Gt[3], G3[3];
            cross3(G1, G2, Gt);
            double J0 = sqrt(dot3(Gt, Gt));
            for (int i = 0; i < 3; ++i) G3[i] = Gt[i] / J0;
            double dA = J0 * wq;

            if (want & (OK_FEXT | OK_M)) {
                for (int A = 0; A < nb; ++A)
                    for (int i = 0; i < 3; ++i) {
                        if (want & OK_FEXT) fx[3 * A + i] += N[A] * pb->fsurf[i] * dA;
                        if (want & OK_M)
                            for (int B = 0; B < nb; ++B)
                                Me[(3 * A + i) * ne + 3 * B + i] += pb->rho * t * N[A] * N[B] * dA;
                    }
            }
            if (!(want & (OK_W | OK_F | OK_K))) continue;

            /* undeformed metric, its inverse and curvature */
            double g11 = dot3(G1, G1), g22 = dot3(G2, G2), g12 = dot3(G1, G2);
            double det = g11 * g22 - g12 * g12;
            double c11 = g22 / det, c22 = g11 / det, c12 = -g12 / det;
            double B11 = dot3(G11, G3), B22 = dot3(G22, G3), B12 = dot3(G12, G3);
            /* SvK constitutive matrix in Voigt form [11, 22, 12] (engineering shear) */
            double D[3][3];
            {
                double ci[2][2] = {{c11, c12}, {c12, c22}};
                int vi[3][2] = {{0, 0}, {1, 1}, {0, 1}};
                for (int I = 0; I < 3; ++I)
                    for (int J = 0; J < 3; ++J) {
                        int al = vi[I][0], be = vi[I][1], ga = vi[J][0], de = vi[J][1];
                        D[I][J] = Em * (nu * ci[al][be] * ci[ga][de] +
                                        0.5 * (1.0 - nu) * (ci[al][ga] * ci[be][de] + ci[al][de] * ci[be][ga]));
                    }
            }
            /* deformed state */
            double at[3], a3[3];
            cross3(a1, a2, at);
            double j = sqrt(dot3(at, at));
            for (int i = 0; i < 3; ++i) a3[i] = at[i] / j;
            double ev[3] = {0.5 * (dot3(a1, a1) - g11), 0.5 * (dot3(a2, a2) - g22), dot3(a1, a2) - g12};
            double kv[3] = {dot3(a11, a3) - B11, dot3(a22, a3) - B22, 2.0 * (dot3(a12, a3) - B12)};
            double Dm = t, Db = t * t * t / 12.0;
            double nv[3], mv[3];
            for (int I = 0; I < 3; ++I) {
                nv[I] = Dm * (D[I][0] * ev[0] + D[I][1] * ev[1] + D[I][2] * ev[2]);
                mv[I] = Db * (D[I][0] * kv[0] + D[I][1] * kv[1] + D[I][2] * kv[2]);
            }
            if (want & OK_W)
                *We += 0.5 * (nv[0] * ev[0] + nv[1] * ev[1] + nv[2] * ev[2] +
                              mv[0] * kv[0] + mv[1] * kv[1] + mv[2] * kv[2]) * dA;
            if (!(want & (OK_F | OK_K))) continue;

            /* first variations per local dof */
            double de[OK_MAXNE][3], dk[OK_MAXNE][3], dat[OK_MAXNE][3], da3[OK_MAXNE][3], dj[OK_MAXNE];
            for (int A = 0; A < nb; ++A)
                for (int i = 0; i < 3; ++i) {
                    int r = 3 * A + i;
                    double e[3] = {0, 0, 0};
                    e[i] = 1.0;
                    double v1[3] = {N1[A] * e[0], N1[A] * e[1], N1[A] * e[2]};
                    double v2[3] = {N2[A] * e[0], N2[A] * e[1], N2[A] * e[2]};
                    de[r][0] = N1[A] * a1[i];
                    de[r][1] = N2[A] * a2[i];
                    de[r][2] = N1[A] * a2[i] + N2[A] * a1[i];
                    double t1[3], t2[3];
                    cross3(v1, a2, t1);
                    cross3(a1, v2, t2);
                    for (int k = 0; k < 3; ++k) dat[r][k] = t1[k] + t2[k];
                    dj[r] = dot3(a3, dat[r]);
                    for (int k = 0; k < 3; ++k) da3[r][k] = (dat[r][k] - a3[k] * dj[r]) / j;
                    dk[r][0] = N11[A] * a3[i] + dot3(a11, da3[r]);
                    dk[r][1] = N22[A] * a3[i] + dot3(a22, da3[r]);
                    dk[r][2] = 2.0 * (N12[A] * a3[i] + dot3(a12, da3[r]));
                }
            if (want & OK_F)
                for (int r = 0; r < ne; ++r)
                    fe[r] += (nv[0] * de[r][0] + nv[1] * de[r][1] + nv[2] * de[r][2] +
                              mv[0] * dk[r][0] + mv[1] * dk[r][1] + mv[2] * dk[r][2]) * dA;
            if (!(want & OK_K)) continue;

            /* tangent: material part + second variations */
            double Dde[OK_MAXNE][3], Ddk[OK_MAXNE][3];
            for (int s = 0; s < ne; ++s)
                for (int I = 0; I < 3; ++I) {
                    Dde[s][I] = Dm * (D[I][0] * de[s][0] + D[I][1] * de[s][1] + D[I][2] * de[s][2]);
                    Ddk[s][I] = Db * (D[I][0] * dk[s][0] + D[I][1] * dk[s][1] + D[I][2] * dk[s][2]);
                }
            for (int A = 0; A < nb; ++A)
                for (int i = 0; i < 3; ++i) {
                    int r = 3 * A + i;
                    for (int B = 0; B < nb; ++B)
                        for (int k = 0; k < 3; ++k) {
                            int s = 3 * B + k;
                            double val = 0.0;
                            val += de[r][0] * Dde[s][0] + de[r][1] * Dde[s][1] + de[r][2] * Dde[s][2];
                            val += dk[r][0] * Ddk[s][0] + dk[r][1] * Ddk[s][1] + dk[r][2] * Ddk[s][2];
                            /* second variation of the membrane strain */
                            if (i == k)
                                val += nv[0] * N1[A] * N1[B] + nv[1] * N2[A] * N2[B] +
                                       nv[2] * (N1[A] * N2[B] + N2[A] * N1[B]);
                            /* second variation of the normal: d2a3 */
                            double ei[3] = {0, 0, 0}, ek[3] = {0, 0, 0}, eik[3];
                            ei[i] = 1.0;
                            ek[k] = 1.0;
                            cross3(ei, ek, eik);
                            double cAB = N1[A] * N2[B] - N2[A] * N1[B];
                            double d2at[3] = {cAB * eik[0], cAB * eik[1], cAB * eik[2]};
                            double a3d2at = dot3(a3, d2at), datdat = dot3(dat[r], dat[s]);
                            double d2a3[3];
                            for (int l = 0; l < 3; ++l)
                                d2a3[l] = d2at[l] / j - (dat[r][l] * dj[s] + dat[s][l] * dj[r]) / (j * j) -
                                          a3[l] * (datdat / (j * j) + a3d2at / j) +
                                          3.0 * a3[l] * dj[r] * dj[s] / (j * j);
                            double d2b11 = N11[A] * da3[s][i] + N11[B] * da3[r][k] + dot3(a11, d2a3);
                            double d2b22 = N22[A] * da3[s][i] + N22[B] * da3[r][k] + dot3(a22, d2a3);
                            double d2b12 = N12[A] * da3[s][i] + N12[B] * da3[r][k] + dot3(a12, d2a3);
                            val += mv[0] * d2b11 + mv[1] * d2b22 + mv[2] * 2.0 * d2b12;
                            Ke[r * ne + s] += val * dA;
                        }
                }
        }
}

/* ------------------------------------------------------------------------- */
/* global assembly                                                            */
/* ------------------------------------------------------------------------- */
static int local_to_global(const ok_problem *pb, int p1, int ex, int ey, int r)
{
    int A = r / 3, i = r % 3, a = A % p1, b = A / p1;
    return 3 * ((ex + a) + pb->mx * (ey + b)) + i;
}

double ok_energy(const ok_problem *pb, const double *u_full)
{
    ok_ctx c;
    double *Ux, *Uy, W = 0.0;
    ctx_init(pb, &c, &Ux, &Uy);
    int nel = pb->nx * pb->ny;
    double *We = (double *)malloc(sizeof(double) * nel);
#pragma omp parallel for schedule(dynamic, 4)
    for (int e = 0; e < nel; ++e)
        ok_element(pb, &c, Ux, Uy, u_full, e % pb->nx, e / pb->nx, OK_W, &We[e], 0, 0, 0, 0);
    for (int e = 0; e < nel; ++e) W += We[e];
    free(We);
    free(Ux);
    free(Uy);
    return W;
}

void ok_internal_force(const ok_problem *pb, const double *u_full, double *fint)
{
    ok_ctx c;
    double *Ux, *Uy;
    ctx_init(pb, &c, &Ux, &Uy);
    int nel = pb->nx * pb->ny, ne = c.ne;
    double *fe = (double *)malloc(sizeof(double) * nel * ne);
#pragma omp parallel for schedule(dynamic, 4)
    for (int e = 0; e < nel; ++e)
        ok_element(pb, &c, Ux, Uy, u_full, e % pb->nx, e / pb->nx, OK_F, 0, fe + (size_t)e * ne, 0, 0, 0);
    memset(fint, 0, sizeof(double) * 3 * pb->mx * pb->my);
    for (int e = 0; e < nel; ++e)
        for (int r = 0; r < ne; ++r) fint[local_to_global(pb, c.p1, e % pb->nx, e / pb->nx, r)] += fe[(size_t)e * ne + r];
    free(fe);
    free(Ux);
    free(Uy);
}

void ok_external_force(const ok_problem *pb, double *fext)
{
    ok_ctx c;
    double *Ux, *Uy;
    ctx_init(pb, &c, &Ux, &Uy);
    int nel = pb->nx * pb->ny, ne = c.ne;
    memset(fext, 0, sizeof(double) * 3 * pb->mx * pb->my);
    double fx[OK_MAXNE];
    if (pb->fsurf[0] != 0.0 || pb->fsurf[1] != 0.0 || pb->fsurf[2] != 0.0)
        for (int e = 0; e < nel; ++e) {
            ok_element(pb, &c, Ux, Uy, 0, e % pb->nx, e / pb->nx, OK_FEXT, 0, 0, 0, fx, 0);
            for (int r = 0; r < ne; ++r) fext[local_to_global(pb, c.p1, e % pb->nx, e / pb->nx, r)] += fx[r];
        }
    for (int l = 0; l < pb->npl; ++l) {
        double xi = pb->pl_xi[2 * l], eta = pb->pl_xi[2 * l + 1];
        int sx = ok_find_span(pb->p, pb->nx, xi), sy = ok_find_span(pb->p, pb->ny, eta);
        double dx[3 * OK_MAXP1], dy[3 * OK_MAXP1];
        ok_ders_basis(pb->p, Ux, sx, xi, 0, dx);
        ok_ders_basis(pb->p, Uy, sy, eta, 0, dy);
        for (int b = 0; b <= pb->p; ++b)
            for (int a = 0; a <= pb->p; ++a) {
                int cp = (sx - pb->p + a) + pb->mx * (sy - pb->p + b);
                for (int i = 0; i < 3; ++i) fext[3 * cp + i] += dx[a] * dy[b] * pb->pl_F[3 * l + i];
            }
    }
    free(Ux);
    free(Uy);
}

/*
 * Follower pressure of one element: pe[3A+i] = p sum_g w_g N_A (a1 x a2)_i and
 * Kpe[(3A+i)*ne + 3B+k] = d pe[3A+i] / d u_{B,k}
 *                       = p sum_g w_g N_A (N_B,1 (e_k x a2) + N_B,2 (a1 x e_k))_i,
 * with a1, a2 the deformed covariant base vectors and w_g the Gauss weight on the
 * parameter domain (p (a1 x a2) dxi deta = p n da on the deformed surface).
 */
static void ok_pressure_element(const ok_problem *pb, const ok_ctx *c, const double *Ux, const double *Uy,
                                const double *u_full, int ex, int ey, double *pe, double *Kpe)
{
    const int p = c->p, p1 = c->p1, nb = c->nb, ne = c->ne, nq = c->nq;
    double X[OK_MAXNB][3];
    for (int b = 0; b < p1; ++b)
        for (int a = 0; a < p1; ++a) {
            int A = a + p1 * b, cp = (ex + a) + pb->mx * (ey + b);
            for (int i = 0; i < 3; ++i) X[A][i] = pb->ctrl[3 * cp + i] + (u_full ? u_full[3 * cp + i] : 0.0);
        }
    if (pe) memset(pe, 0, sizeof(double) * ne);
    if (Kpe) memset(Kpe, 0, sizeof(double) * ne * ne);
    for (int qy = 0; qy < nq; ++qy)
        for (int qx = 0; qx < nq; ++qx) {
            double xi = (ex + 0.5 * (1.0 + c->gx[qx])) / pb->nx;
            double eta = (ey + 0.5 * (1.0 + c->gx[qy])) / pb->ny;
            double wq = c->gw[qx] * c->gw[qy] * 0.25 / ((double)pb->nx * (double)pb->ny);
            double dx[3 * OK_MAXP1], dy[3 * OK_MAXP1];
            ok_ders_basis(p, Ux, ex + p, xi, 1, dx);
            ok_ders_basis(p, Uy, ey + p, eta, 1, dy);
            double N[OK_MAXNB], N1[OK_MAXNB], N2[OK_MAXNB];
            for (int b = 0; b < p1; ++b)
                for (int a = 0; a < p1; ++a) {
                    int A = a + p1 * b;
                    N[A] = dx[a] * dy[b];
                    N1[A] = dx[p1 + a] * dy[b];
                    N2[A] = dx[a] * dy[p1 + b];
                }
            double a1[3] = {0}, a2[3] = {0}, n[3];
            for (int A = 0; A < nb; ++A)
                for (int i = 0; i < 3; ++i) {
                    a1[i] += N1[A] * X[A][i];
                    a2[i] += N2[A] * X[A][i];
                }
            cross3(a1, a2, n);
            const double s = pb->pressure * wq;
            if (pe)
                for (int A = 0; A < nb; ++A)
                    for (int i = 0; i < 3; ++i) pe[3 * A + i] += s * N[A] * n[i];
            if (Kpe)
                for (int k = 0; k < 3; ++k) {
                    double ek[3] = {0.0, 0.0, 0.0}, c2[3], c1[3];
                    ek[k] = 1.0;
                    cross3(ek, a2, c2); /* d(a1 x a2) / d a1_k */
                    cross3(a1, ek, c1); /* d(a1 x a2) / d a2_k */
                    for (int A = 0; A < nb; ++A)
                        for (int i = 0; i < 3; ++i)
                            for (int B = 0; B < nb; ++B)
                                Kpe[(3 * A + i) * ne + 3 * B + k] += s * N[A] * (N1[B] * c2[i] + N2[B] * c1[i]);
                }
        }
}

void ok_pressure_force(const ok_problem *pb, const double *u_full, double *fp)
{
    ok_ctx c;
    double *Ux, *Uy;
    ctx_init(pb, &c, &Ux, &Uy);
    int nel = pb->nx * pb->ny, ne = c.ne;
    memset(fp, 0, sizeof(double) * 3 * pb->mx * pb->my);
    double pe[OK_MAXNE];
    for (int e = 0; e < nel; ++e) {
        ok_pressure_element(pb, &c, Ux, Uy, u_full, e % pb->nx, e / pb->nx, pe, 0);
        for (int r = 0; r < ne; ++r) fp[local_to_global(pb, c.p1, e % pb->nx, e / pb->nx, r)] += pe[r];
    }
    free(Ux);
    free(Uy);
}

void ok_pressure_tangent_dense(const ok_problem *pb, const double *u_full, double *Kp)
{
    ok_ctx c;
    double *Ux, *Uy;
    ctx_init(pb, &c, &Ux, &Uy);
    int nel = pb->nx * pb->ny, ne = c.ne, ndof = 3 * pb->mx * pb->my;
    memset(Kp, 0, sizeof(double) * (size_t)ndof * ndof);
    double *Kpe = (double *)malloc(sizeof(double) * ne * ne);
    for (int e = 0; e < nel; ++e) {
        int ex = e % pb->nx, ey = e / pb->nx;
        ok_pressure_element(pb, &c, Ux, Uy, u_full, ex, ey, 0, Kpe);
        for (int r = 0; r < ne; ++r)
            for (int q = 0; q < ne; ++q)
                Kp[(size_t)local_to_global(pb, c.p1, ex, ey, r) * ndof + local_to_global(pb, c.p1, ex, ey, q)] +=
                    Kpe[r * ne + q];
    }
    free(Kpe);
    free(Ux);
    free(Uy);
}

static void assemble_matrix(const ok_problem *pb, const double *u_full, int which, const int *free_idx,
                            int nfree, int bw, double *out, int band)
{
    ok_ctx c;
    double *Ux, *Uy;
    ctx_init(pb, &c, &Ux, &Uy);
    int nel = pb->nx * pb->ny, ne = c.ne, ndof = 3 * pb->mx * pb->my;
    double *Ke = (double *)malloc(sizeof(double) * (size_t)nel * ne * ne);
#pragma omp parallel for schedule(dynamic, 2)
    for (int e = 0; e < nel; ++e) {
        double *k = Ke + (size_t)e * ne * ne;
        if (which == OK_K)
            ok_element(pb, &c, Ux, Uy, u_full, e % pb->nx, e / pb->nx, OK_K, 0, 0, k, 0, 0);
        else
            ok_element(pb, &c, Ux, Uy, 0, e % pb->nx, e / pb->nx, OK_M, 0, 0, 0, 0, k);
    }
    if (band)
        memset(out, 0, sizeof(double) * (size_t)nfree * (bw + 1));
    else
        memset(out, 0, sizeof(double) * (size_t)ndof * ndof);
    for (int e = 0; e < nel; ++e) {
        const double *k = Ke + (size_t)e * ne * ne;
        int gl[OK_MAXNE];
        for (int r = 0; r < ne; ++r) gl[r] = local_to_global(pb, c.p1, e % pb->nx, e / pb->nx, r);
        for (int r = 0; r < ne; ++r)
            for (int s = 0; s < ne; ++s) {
                if (!band) {
                    out[(size_t)gl[r] * ndof + gl[s]] += k[r * ne + s];
                } else {
                    int I = free_idx[gl[r]], J = free_idx[gl[s]];
                    if (I < 0 || J < 0 || I < J) continue;
                    out[(size_t)J * (bw + 1) + (I - J)] += k[r * ne + s];
                }
            }
    }
    free(Ke);
    free(Ux);
    free(Uy);
}

void ok_tangent_dense(const ok_problem *pb, const double *u_full, double *K)
{
    assemble_matrix(pb, u_full, OK_K, 0, 0, 0, K, 0);
}
void ok_mass_dense(const ok_problem *pb, double *M) { assemble_matrix(pb, 0, OK_M, 0, 0, 0, M, 0); }
void ok_tangent_band(const ok_problem *pb, const int *free_idx, int nfree, int bw, const double *u_full,
                     double *ab)
{
    assemble_matrix(pb, u_full, OK_K, free_idx, nfree, bw, ab, 1);
}
void ok_mass_band(const ok_problem *pb, const int *free_idx, int nfree, int bw, double *ab)
{
    assemble_matrix(pb, 0, OK_M, free_idx, nfree, bw, ab, 1);
}

/* ------------------------------------------------------------------------- */
/* banded LDL^T                                                               */
/* ------------------------------------------------------------------------- */
#define AB(i, j) ab[(size_t)(j) * (bw + 1) + ((i) - (j))]

/* Right-looking, column by column; returns 0, or j+1 for a zero/non-finite pivot. */
int ok_band_ldlt(int n, int bw, double *ab)
{
    double *col = (double *)malloc(sizeof(double) * (bw + 1));
    for (int j = 0; j < n; ++j) {
        double d = AB(j, j);
        if (d == 0.0 || !isfinite(d)) {
            free(col);
            return j + 1;
        }
        int iend = j + bw < n - 1 ? j + bw : n - 1;
        for (int i = j + 1; i <= iend; ++i) {
            col[i - j] = AB(i, j);
            AB(i, j) = col[i - j] / d;
        }
        for (int k = j + 1; k <= iend; ++k) {
            double lk = AB(k, j);
            if (lk == 0.0) continue;
            for (int i = k; i <= iend; ++i) AB(i, k) -= AB(i, j) * col[k - j];
        }
    }
    free(col);
    return 0;
}

void ok_band_solve(int n, int bw, const double *ab, double *x)
{
    for (int j = 0; j < n; ++j) {
        int iend = j + bw < n - 1 ? j + bw : n - 1;
        double xj = x[j];
        for (int i = j + 1; i <= iend; ++i) x[i] -= AB(i, j) * xj;
    }
    for (int j = 0; j < n; ++j) x[j] /= AB(j, j);
    for (int j = n - 1; j >= 0; --j) {
        int iend = j + bw < n - 1 ? j + bw : n - 1;
        double s = x[j];
        for (int i = j + 1; i <= iend; ++i) s -= AB(i, j) * x[i];
        x[j] = s;
    }
}

int ok_band_negatives(int n, int bw, const double *ab, double *dmin)
{
    int neg = 0;
    double m = INFINITY;
    for (int j = 0; j < n; ++j) {
        double d = AB(j, j);
        if (d < 0.0) ++neg;
        if (d < m) m = d;
    }
    if (dmin) *dmin = m;
    return neg;
}

void ok_band_symv(int n, int bw, const double *ab, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int j = 0; j < n; ++j) {
        int iend = j + bw < n - 1 ? j + bw : n - 1;
        y[j] += AB(j, j) * x[j];
        for (int i = j + 1; i <= iend; ++i) {
            y[i] += AB(i, j) * x[j];
            y[j] += AB(i, j) * x[i];
        }
    }
}
#undef AB
