/*
 * ksh_oracle.h -- FP64 CPU oracle for the Kirchhoff-Love shell equilibrium path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libksh.so, the package's
 * Python modules) links, loads or calls this code.  It is used by tests/ as the
 * checker, by __graft_entry__.smoke() as the checker, and by bench.py's
 * cpu_baseline leg as the timed CPU port.
 *
 * What it restates (the reference's hot path lives in the external gsKLShell
 * module, which is not in the reference tree; call sites are):
 *   - K(u)  = assembleMatrix(mp_def) / matrix()      benchmarks/benchmark_Roof.cpp:323-332
 *   - R(u)  = assembleVector(mp_def) / rhs()         benchmarks/benchmark_Roof.cpp:334-344
 *   - F, K_L = assemble()                            benchmarks/benchmark_Roof.cpp:319-321
 *   - M      = assembleMass()                        solvers/modal_shell_XML.cpp:235-236
 *   - SvK material, analytical thickness integration benchmarks/benchmark_Roof.cpp:304-306
 *   - p+1 Gauss-Legendre points per direction        filedata/options/solver_options.xml:11-14
 *   - LDL^T with the D vector and its negative count src/gsALMSolvers/gsALMBase.hpp:559-562,612
 *                                                    src/gsALMSolvers/gsALMHelper.h:30-42
 * The shell formulation is the published Kirchhoff-Love / St. Venant-Kirchhoff
 * isogeometric shell (Kiendl et al. 2009, CMAME 198:3902): membrane strain
 * eps = (a_ab - A_ab)/2, bending strain kappa = b_ab - B_ab, n = t C:eps,
 * m = t^3/12 C:kappa with C on the undeformed contravariant metric.
 *
 * Parity pinning (all in tests/): energy -> internal force and internal force
 * -> tangent by central finite differences, tangent symmetry, rigid-body
 * invariance of the residual, six zero-energy modes of the free K_L, plate
 * closed forms (clamped / simply supported, Timoshenko), scipy BSpline and
 * numpy leggauss for basis and quadrature, scipy for LDL^T solves and inertia.
 */
#ifndef KSH_ORACLE_H
#define KSH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int p;               /* spline degree in both directions              */
    int nx, ny;          /* elements per direction                        */
    int mx, my;          /* control points per direction (n + p)          */
    const double *ctrl;  /* (mx*my) x 3 undeformed control net, cp = i+mx*j */
    double t, E, nu, rho;
    const int *fixed;    /* 3*mx*my flags, 1 = homogeneous Dirichlet dof  */
    int npl;             /* point loads                                   */
    const double *pl_xi; /* 2*npl parametric coordinates                  */
    const double *pl_F;  /* 3*npl load vectors                            */
    double fsurf[3];     /* dead surface force per unit undeformed area   */
    double pressure;     /* follower pressure along a1 x a2 (0: none)     */
} ok_problem;

/* spline / quadrature primitives */
void ok_knots_uniform(int p, int n, double *U);
int  ok_find_span(int p, int n, double x);
void ok_ders_basis(int p, const double *U, int span, double x, int nd, double *ders);
void ok_gauss_legendre(int nq, double *x, double *w);

/* free-dof map: returns nfree, fills free_idx[ndof] (-1 for fixed) */
int  ok_free_map(const ok_problem *pb, int *free_idx);
int  ok_bandwidth(const ok_problem *pb, const int *free_idx);

/* assembly on full dofs (u_full has 3*mx*my entries) */
double ok_energy(const ok_problem *pb, const double *u_full);
void   ok_internal_force(const ok_problem *pb, const double *u_full, double *fint);
void   ok_external_force(const ok_problem *pb, double *fext);
void   ok_tangent_dense(const ok_problem *pb, const double *u_full, double *K);
void   ok_mass_dense(const ok_problem *pb, double *M);
/* follower pressure (gsThinShellAssembler::setPressure, examples/gsThinShell_ArcLength.cpp:956-957,
 * solvers/static_shell_XML.cpp:236-237): P(u)_{A,i} = p sum_g w_g N_A (a1 x a2)_i on the deformed
 * surface, and its derivative dP/du (full dofs, row-major ndof x ndof, not symmetric) */
void   ok_pressure_force(const ok_problem *pb, const double *u_full, double *fp);
void   ok_pressure_tangent_dense(const ok_problem *pb, const double *u_full, double *Kp);
/* lower band of the free-dof tangent / mass: ab[J*(bw+1) + (I-J)] */
void   ok_tangent_band(const ok_problem *pb, const int *free_idx, int nfree, int bw,
                       const double *u_full, double *ab);
void   ok_mass_band(const ok_problem *pb, const int *free_idx, int nfree, int bw, double *ab);

/* banded LDL^T (no pivoting), solves, inertia, symmetric band mat-vec */
int    ok_band_ldlt(int n, int bw, double *ab);
void   ok_band_solve(int n, int bw, const double *ab, double *x);
int    ok_band_negatives(int n, int bw, const double *ab, double *dmin);
void   ok_band_symv(int n, int bw, const double *ab, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
