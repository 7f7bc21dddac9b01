/*
 * ksh.h -- C-ABI of the MI355X Kirchhoff-Love shell equilibrium path (libksh.so).
 *
 * Plain C: pointers and sizes only.  Every call returns an int status whose
 * values are the ordinals of the reference's gsStatus
 * (src/gsStructuralAnalysisTools/gsStructuralAnalysisTypes.h:22-30):
 *   0 Success, 1 NotConverged, 2 AssemblyError, 3 SolverError, 4 NotStarted,
 *   5 OtherError.  ksh_last_error() returns a message for the calling thread.
 *
 * Vectors passed as `const double *` / `double *` are HOST arrays unless the
 * name ends in _dev (device pointers on the problem's device).  All work of one
 * problem is issued on that problem's HIP stream; handles are not thread-safe.
 *
 * Dof numbering: control point cp = i + mx*j (i along xi, j along eta),
 * full dof g = 3*cp + c; free dofs are the non-fixed full dofs in increasing g.
 * Matrices are the symmetric free-dof matrices stored as a lower band with
 * half-bandwidth bw: ab[J*(bw+1) + (I-J)], 0 <= I-J <= bw.
 */
#ifndef KSH_H
#define KSH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ksh_status {
    KSH_SUCCESS = 0,
    KSH_NOT_CONVERGED = 1,
    KSH_ASSEMBLY_ERROR = 2,
    KSH_SOLVER_ERROR = 3,
    KSH_NOT_STARTED = 4,
    KSH_OTHER_ERROR = 5
};

/* residual sign conventions of the two operator typedefs */
enum ksh_residual_conv {
    KSH_RES_STATIC = 0, /* R = lambda F - F_int(u)   (Residual_t, gsStaticNewton)  */
    KSH_RES_ALM = 1     /* r = F_int(u) - lambda F   (ALResidual_t, gsALMBase)     */
};

/* matrix / factor precision */
enum ksh_precision {
    KSH_FP64 = 0,     /* FP64 assembly, FP64 band LDL^T                                   */
    KSH_FP32_IR = 1,  /* FP64 assembly, FP32 band LDL^T on MFMA, refinement residuals in
                         FP64 against the FP64 K; the FP32 pivot count is used only where
                         the refinement certifies it (see ksh_inertia), else an FP64 factor
                         of the same matrix counts                                        */
    KSH_FP32 = 2      /* FP32 element GEMMs and FP32 band (K itself rounded to FP32),
                         FP32 LDL^T, refinement residuals in FP64 arithmetic on K32       */
};

typedef struct ksh_problem ksh_problem;
typedef struct ksh_matrix ksh_matrix;
typedef struct ksh_mlp ksh_mlp;

typedef struct {
    int p;              /* degree in both directions (2..4)                    */
    int nx, ny;         /* elements per direction, open uniform knots on [0,1] */
    const double *ctrl; /* (nx+p)*(ny+p) x 3 undeformed control points         */
} ksh_patch;

typedef struct {
    double t, E, nu, rho; /* thickness, Young's modulus, Poisson, density      */
} ksh_material;           /* St. Venant-Kirchhoff, analytical thickness integr. */

typedef struct {
    const int *fixed;   /* 3*(nx+p)*(ny+p) flags, 1 = homogeneous Dirichlet dof */
} ksh_bc;

typedef struct {
    int npoint;
    const double *point_xi; /* 2*npoint parametric points                      */
    const double *point_F;  /* 3*npoint load vectors                           */
    double fsurf[3];        /* dead surface force per unit undeformed area      */
} ksh_loads;

const char *ksh_last_error(void);
int ksh_device_count(int *count);

/* ---- problem ------------------------------------------------------------- */
/* Replaces the gsThinShellAssembler<3,real_t,true> construction of the drivers
 * (benchmarks/benchmark_Roof.cpp:309-314): patch + material + BCs + loads. */
int ksh_problem_create(const ksh_patch *patch, const ksh_material *mat, const ksh_bc *bc,
                       const ksh_loads *loads, ksh_problem **out);
int ksh_problem_destroy(ksh_problem *pb);
/* Follower pressure p along a1 x a2 of the deformed surface (p (a1 x a2) dxi deta = p n da):
 * gsThinShellAssembler::setPressure with a constant function
 * (examples/gsThinShell_ArcLength.cpp:849,956-957; solvers/static_shell_XML.cpp:145-149,236-237).
 * F becomes the dead loads plus the pressure's load on the undeformed surface P(0) (the drivers'
 * Force = assemble() rhs, examples/gsThinShell_ArcLength.cpp:963-965); the residuals become
 * r = F_int(u) - lambda F - (P(u) - P(0)) (ALM, the driver's Force - lambda Force - rhs(u),
 * :997-1005) and its negative (static); the tangent K(u) - dP/du, of which the band keeps the
 * lower triangle (what a lower-triangle LDL^T of the assembled matrix reads).  p = 0 removes it.
 * Every engine and entry point that assembles K or R of the problem uses it. */
int ksh_problem_set_pressure(ksh_problem *pb, double pressure);
int ksh_num_free_dofs(const ksh_problem *pb, int64_t *n);
int ksh_bandwidth(const ksh_problem *pb, int64_t *bw);
int ksh_synchronize(ksh_problem *pb);

/* F = rhs() of assemble() at u = 0: external load on the free dofs
 * (benchmarks/benchmark_Roof.cpp:319-321). */
int ksh_assemble_force(ksh_problem *pb, double *F_out);

/* Residual_t / ALResidual_t body: constructSolution + assembleVector + the
 * driver's sign convention (benchmarks/benchmark_Roof.cpp:335-344,
 * tutorials/nonlinear_shell_static.cpp:129-136,
 * examples/example_DisplacementControl.cpp:245-254). */
int ksh_assemble_residual(ksh_problem *pb, const double *u_free, double lambda, int conv,
                          double *R_out);
int ksh_assemble_residual_dev(ksh_problem *pb, const double *u_free_dev, double lambda, int conv,
                              double *R_out_dev);
/* No counterpart in the reference: ksh_assemble_residual_dev for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops).  Same arguments and result, bit for bit;
 * its partial results go to one fixed 48 MiB device array of this library per GPU instead of the
 * problem's workspace, and nothing is allocated or freed.  Problems whose partial results exceed
 * the array (about 20k degree-3 elements) are rejected with KSH_OTHER_ERROR.  Calls that overlap
 * on one GPU race on the array. */
int ksh_assemble_residual_dev_noalloc(ksh_problem *pb, const double *u_free_dev, double lambda, int conv,
                                      double *R_out_dev);

/* ---- matrices -------------------------------------------------------------- */
int ksh_matrix_create(ksh_problem *pb, int prec, ksh_matrix **out);
int ksh_matrix_destroy(ksh_matrix *K);
int ksh_matrix_dims(const ksh_matrix *K, int64_t *n, int64_t *bw, int *prec);
/* Jacobian_t body: constructSolution + assembleMatrix + matrix()
 * (benchmarks/benchmark_Roof.cpp:323-332); K stays device resident. */
int ksh_assemble_tangent(ksh_problem *pb, const double *u_free, ksh_matrix *K);
int ksh_assemble_tangent_dev(ksh_problem *pb, const double *u_free_dev, ksh_matrix *K);
/* assembleMass() + matrix() (solvers/modal_shell_XML.cpp:235-236). */
int ksh_assemble_mass(ksh_problem *pb, ksh_matrix *M);
/* band <-> host (tests, Eigen-compat downloads); values in FP64 */
int ksh_matrix_download(const ksh_matrix *K, double *ab_out);
int ksh_matrix_upload(ksh_matrix *K, const double *ab_in);
/* Y = alpha*Y + beta*X (A - sigma B: src/gsEigenSolvers/gsEigenProblemBase.hpp:32,105;
 * B -= A: src/gsEigenSolvers/gsBucklingSolver.h:104,144) */
int ksh_matrix_axpby(ksh_matrix *Y, double alpha, double beta, const ksh_matrix *X);
int ksh_matrix_copy(ksh_matrix *dst, const ksh_matrix *src);
/* y = K x (SpMV used at src/gsALMSolvers/gsALMBase.hpp:639,663,720-721 and
 * src/gsALMSolvers/gsALMCrisfield.hpp:316) */
int ksh_matvec(const ksh_matrix *K, const double *x, double *y);
int ksh_matvec_dev(const ksh_matrix *K, const double *x_dev, double *y_dev);

/* ---- LDL^T: replaces gsSparseSolver<T>::SimplicialLDLT ---------------------
 * compute()  src/gsALMSolvers/gsALMBase.hpp:186-198
 * solve()    src/gsALMSolvers/gsALMBase.hpp:200-211
 * vectorD()  src/gsALMSolvers/gsALMBase.hpp:559-562, src/gsStaticSolvers/gsStaticBase.h:176-177
 * The factor lives beside the matrix; precision is the matrix's precision. */
int ksh_factor(ksh_matrix *K);
int ksh_solve(ksh_matrix *K, const double *b, double *x, int max_refine, double tol);
int ksh_solve_dev(ksh_matrix *K, const double *b_dev, double *x_dev, int max_refine, double tol);
/* countNegatives(vectorD()) and min(D) (src/gsALMSolvers/gsALMHelper.h:30-42,
 * src/gsALMSolvers/gsALMBase.hpp:612-614).  FP32 factors M = K + E: the count of M equals
 * the count of K whenever rho(M^-1 E) < 1 (K(t) = M (I - t M^-1 E) stays nonsingular on
 * t in [0, 1]).  rho is estimated from the residual contractions of refinement sweeps with
 * M (the solves of this factor, else three power steps here); when the estimate exceeds the
 * matrix's guard (default 0.5) the count comes from an FP64 factor of K. */
int ksh_inertia(ksh_matrix *K, int64_t *n_negative, double *d_min);
int ksh_vector_d(ksh_matrix *K, double *d_out);
/* guard > 0: certificate threshold on the rho estimate; guard <= 0: FP32 counts used as they are */
int ksh_set_inertia_guard(ksh_matrix *K, double guard);

/* ---- measurement --------------------------------------------------------------
 * No counterpart in the reference (it only accumulates wall time inside its lambdas,
 * benchmarks/benchmark_Roof.cpp:316-343).  When enabled, every launch of the named
 * kernels ("tangent_elem", "band_gather", "residual_elem", "vector_gather", "band_ldlt",
 * "band_ldlt_batch", "band_solve", "band_solve_batch", "band_symv", "refine_residual_batch")
 * is bracketed by HIP events on its stream. */
int ksh_profile_enable(int on);
int ksh_profile_reset(void);
int ksh_profile_query(const char *kernel, int64_t *launches, double *total_ms);
/* Algorithmic work summed over the profiled launches of `kernel`: flops for the MFMA-bound
 * kernels (band_ldlt*: n (b^2 + 3b) per matrix; tangent_elem: n_el n_g (72 n_e + 12 n_e^2)),
 * bytes for the HBM-bound ones (band_solve*: 2 n (b+1) sizeof(factor) + 16 n per matrix;
 * refine_residual_batch / band_symv: n (2b+1) sizeof(K) + 24 n; band_gather: element tiles
 * read + band written). */
int ksh_profile_query_work(const char *kernel, double *work);
/* Launch-shape autotuner: the variant (0-based) the last launch of `kernel` used and the number of
 * variants ("tangent_tiles": 1, 2, 4 elements per workgroup; "band_gather": 128, 256, 512
 * threads per block); -1 / 0 if it has not run.  No counterpart in the reference. */
int ksh_tune_choice(const char *kernel, int64_t *variant, int64_t *variants);

/* ---- solver drivers ---------------------------------------------------------
 * The path engine: the algorithms of gsALMCrisfield / gsALMRiks (arc length) and of
 * gsStaticNewton under gsControlDisplacement (load control) over this library's device
 * operators, inside the driver loop of the reference's programs (halve the arc length and
 * retry from the last converged point on NotConverged / AssemblyError).  Every vector stays
 * in device memory; one host thread drives any number of problems in batched launches. */
typedef struct {
    int method;              /* 0 Crisfield (gsALMCrisfield), 1 Riks (gsALMRiks)         */
    int precision;           /* KSH_FP64 | KSH_FP32_IR                                   */
    int nsteps;              /* converged arc-length steps to take                        */
    double length;           /* arc length ("Length")                                    */
    int max_iter;            /* "MaxIter"                                                */
    double tol, tol_u, tol_f;/* "Tol", "TolU", "TolF"                                    */
    double scaling;          /* Crisfield "Scaling" (phi), -1 = automatic                */
    int angle_method;        /* Crisfield "AngleMethod"                                  */
    int bifurcation_method;  /* "BifurcationMethod": -1 nothing, 0 determinant            */
    double relaxation;       /* "Relaxation"                                             */
    int quasi, quasi_iterations;
    int adaptive, adaptive_iterations;
    int max_refine;          /* refinement sweeps per solve (FP32_IR: FP64 residuals)     */
    double refine_tol;       /* stop refinement at ||b - K x|| <= refine_tol ||b||         */
    int max_halvings;        /* driver: halve the arc length on failure, retry            */
    int verbose;
    double inertia_guard;    /* FP32 factors: pivot-count certificate threshold on the    */
                             /* estimated rho(M^-1 E) (ksh_inertia); 0: the library       */
                             /* default 0.5; < 0: FP32 counts used as they are            */
    struct ksh_mlp *learned; /* Crisfield only: learned predictor network (NULL: tangent   */
                             /* predictor); a step that fails from the learned proposal   */
                             /* is redone from the tangent predictor                      */
    double refine_accept;    /* FP32 factors: a solve whose residual is still above        */
                             /* refine_accept ||b|| after max_refine sweeps is redone      */
                             /* with an FP64 factor (0: the library default, 1e-6)         */
    int learned_mode;        /* with `learned`: 0 network-corrected extrapolation, 1 the    */
                             /* quadratic extrapolation alone (network output taken as 0)  */
    const double *lengths;   /* ksh_batch_run_alm: arc length of each problem (NULL: all    */
                             /* take `length`)                                              */
    /* singular points: on a change of stability after a converged step the driver calls
     * computeSingularPoint from the previous converged point (gsALMBase.hpp:422-446): the
     * singular-point test, optionally a bisection on the pivot count, the extended system for
     * the critical point; then, with singular_point = 2, switchBranch (gsALMBase.hpp:843-856).
     * The arc length restarts from length_after_sp (the drivers' "Length2" / dL).            */
    int singular_point;      /* 0 off; 1 solvers/quasistatic_shell_XML.cpp:305-311; 2 with  */
                             /* the branch switch, examples/gsThinShell_ArcLength.cpp:1107-1117 */
    double sp_test_tol;      /* "SingularPointTestTol" (0: 1e-6)                          */
    int sp_test_it;          /* "SingularPointTestIt" (0: 5)                              */
    double sp_tol_e;         /* "SingularPointComputeTolE" (0: 1e-10)                     */
    double sp_tol_b;         /* "SingularPointComputeTolB" (0: no bisection)              */
    double perturbation;     /* "Perturbation" tau of switchBranch (0: 1e3)               */
    int sp_fail;             /* "SingularPointFailure": 1 (the reference's default) keeps  */
                             /* the last extended iterate when they do not converge, 0     */
                             /* returns to the start point                                 */
    double length_after_sp;  /* arc length after a singular point (0: `length`)           */
} ksh_alm_options;

typedef struct {
    int steps_done;
    int halvings;
    int final_status;        /* gsStatus ordinal of the last step                        */
    double *lambda;          /* [nsteps] load factor after each converged step (optional) */
    double *unorm;           /* [nsteps] ||U||                                            */
    double *indicator;       /* [nsteps] min D of the counted tangent                     */
    int64_t *negatives;      /* [nsteps] negative pivots of the counted tangent           */
    int *iterations;         /* [nsteps] corrector iterations                             */
    double *residue_f;       /* [nsteps] ||r|| / ||lambda F|| at convergence              */
    double *length;          /* [nsteps] arc length of the step                           */
    double *U;               /* [nsteps*n] converged displacements (optional)             */
    double *w;               /* [nsteps] value of U at the dof `probe_dof` (optional)     */
    int probe_dof;
    double wall_seconds;
    double phi;              /* Crisfield scaling in use at the end                       */
    int64_t jacobians, residuals, solves;
    double *features;        /* [nsteps*66] predictor features of each converged state (opt.) */
    int64_t learned_steps;   /* steps that started from the learned proposal              */
    int64_t fallbacks;       /* of those, steps redone from the tangent predictor         */
    int *singular;           /* [nsteps] 0; singular point handled after the step: 1 limit */
                             /* point (solution set back), 2 bifurcation point computed,   */
                             /* 3 bifurcation, extended iterations not converged (opt.)    */
} ksh_alm_result;

/* benchmarks/benchmark_Roof.cpp:320-512: options, initialize, step loop with halving. */
int ksh_run_alm(ksh_problem *pb, const ksh_alm_options *opt, ksh_alm_result *res);

/* Many independent problems (uneven sizes, same degree) each run ksh_run_alm's loop; every
 * round advances each live path by one predictor or corrector iteration in batched launches
 * (assembly over the union of all elements, one LDL^T / solve workgroup per band).  Each
 * path's numerics are those of its single-problem run.  res[count]; wall_seconds is the whole
 * batch's wall time.  No counterpart in the reference, which runs one problem per
 * process (the closest is APALM's path-interval parallelism, src/gsALMSolvers/gsAPALM.hpp:251-430). */
int ksh_batch_run_alm(ksh_problem *const *pbs, int count, const ksh_alm_options *opt, ksh_alm_result *res,
                      double *wall_seconds);

typedef struct {
    int precision;
    int nsteps;              /* load steps of dlambda (gsControlDisplacement::step)       */
    double dlambda;
    int max_iter;            /* "maxIt"                                                  */
    double tol;              /* "tol"                                                    */
    int max_refine;
    double refine_tol;
    int verbose;
    double inertia_guard;    /* as in ksh_alm_options                                      */
    double refine_accept;    /* as in ksh_alm_options                                      */
} ksh_static_options;

typedef struct {
    int steps_done;
    int final_status;
    double *lambda;          /* [nsteps]                                                 */
    double *unorm;           /* [nsteps]                                                 */
    int *iterations;         /* [nsteps]                                                 */
    int64_t *negatives;      /* [nsteps]                                                 */
    double *residual;        /* [nsteps] final ||R||                                      */
    double *U;               /* [nsteps*n] (optional)                                     */
    double wall_seconds;
    int64_t jacobians;
} ksh_static_result;

/* examples/example_DisplacementControl.cpp:236-304 with a load-controlled residual: Newton
 * (src/gsStaticSolvers/gsStaticNewton.hpp:142-193) for each load step of
 * gsControlDisplacement::step (src/gsStaticSolvers/gsControlDisplacement.h:46-60). */
int ksh_run_static(ksh_problem *pb, const ksh_static_options *opt, ksh_static_result *res);

/* ---- residual-only and mass consumers -----------------------------------------
 * Dynamic relaxation: gsStaticDR (src/gsStaticSolvers/gsStaticDR.hpp:93-266) over the static
 * residual F - F_int(U) at the full load, with the lumped mass (assembleMass(true), the row
 * sums of the consistent mass) as solvers/static_shell_XML.cpp:243-275 builds it; explicit
 * steps dt = 1 from U = 0, viscous damping c m or kinetic damping (c = 0: velocities reset at
 * each kinetic-energy peak).  Option names are gsStaticDR's and gsStaticBase's. */
typedef struct {
    int max_iter;            /* "maxIt"                                                   */
    double tol;              /* "tol"                                                     */
    double tol_f, tol_u;     /* "tolF", "tolU" (-1: tol)                                  */
    double damping;          /* "damping" c                                               */
    double alpha;            /* "alpha": fictitious mass alpha m                          */
    double tol_e;            /* "tolE": kinetic energy over the first one                  */
    int reset_it;            /* "ResetIt" (-1: never)                                     */
    int verbose;
} ksh_dr_options;

typedef struct {
    int iterations;
    int final_status;        /* KSH_SUCCESS | KSH_NOT_CONVERGED (maximum iterations)       */
    int peaks;               /* kinetic-damping resets                                    */
    double residual;         /* ||R|| / ||R_0|| at the end                                 */
    double ek, ek0;          /* kinetic energy at the end and after the first step         */
    double *U;               /* [n] solution (optional)                                   */
    double *ek_history;      /* [max_iter] kinetic energy per iteration (optional)        */
    double wall_seconds;
    int64_t residuals;       /* residual evaluations                                      */
} ksh_dr_result;

int ksh_run_dr(ksh_problem *pb, const ksh_dr_options *opt, ksh_dr_result *res);
/* many problems at once, one residual launch per round for all; res[count] */
int ksh_batch_run_dr(ksh_problem *const *pbs, int count, const ksh_dr_options *opt, ksh_dr_result *res,
                     double *wall_seconds);

/* Nonlinear Newmark: gsDynamicNewmark<T, true> (src/gsDynamicSolvers/gsDynamicNewmark.hpp:69-142)
 * under the time loop of tutorials/nonlinear_shell_dynamic.cpp:176-186: per step the predictor
 * with A = 0, then Newton on M A + C V = F - F_int(U) with the tangent M + gamma dt C + beta dt^2
 * K(U) factored on the GPU; consistent mass, C = damping_mass M. */
typedef struct {
    int precision;           /* KSH_FP64 | KSH_FP32_IR (factor of the iteration matrix)    */
    int nsteps;
    double dt;               /* "DT"                                                      */
    double beta, gamma;      /* the reference's "alpha" (0.25) and "delta" (0.5)           */
    double damping_mass;     /* C = damping_mass M                                        */
    int max_iter;            /* "MaxIter"                                                 */
    double tol_f, tol_u;     /* "TolF", "TolU"                                            */
    int max_refine;
    double refine_tol;
    int verbose;
} ksh_newmark_options;

typedef struct {
    int steps_done;
    int final_status;        /* KSH_SUCCESS | KSH_NOT_CONVERGED (a step's Newton loop)     */
    double *time;            /* [nsteps] (optional)                                       */
    double *unorm;           /* [nsteps] ||U||                                            */
    int *iterations;         /* [nsteps] Newton iterations                                */
    double *U, *V, *A;       /* [nsteps*n] states after each step (optional)              */
    double *w;               /* [nsteps] U at probe_dof (optional)                        */
    int probe_dof;
    double wall_seconds;
    int64_t jacobians;
} ksh_newmark_result;

/* U0, V0, A0: initial state [n] (NULL: zero) */
int ksh_run_newmark(ksh_problem *pb, const ksh_newmark_options *opt, const double *U0, const double *V0,
                    const double *A0, ksh_newmark_result *res);
/* many problems from rest (U = V = A = 0) at once; res[count] */
int ksh_batch_run_newmark(ksh_problem *const *pbs, int count, const ksh_newmark_options *opt, ksh_newmark_result *res,
                          double *wall_seconds);

/* ---- eigen-solvers --------------------------------------------------------------
 * Generalized symmetric eigenpairs A x = lambda B x nearest `shift`, by thick-restart Lanczos
 * on a shift-invert operator whose solves use the GPU band factor of A - shift B.  Replaces
 * gsEigenProblemBase::computeSparse (src/gsEigenSolvers/gsEigenProblemBase.hpp:152-216,
 * Spectra SymGEigsShiftSolver) under gsModalSolver (A = K, B = M; mode SHIFT_INVERT, B SPD)
 * and gsBucklingSolver (A = K_L, B = K_NL - K_L; mode BUCKLING, A SPD). */
enum ksh_eig_mode {
    KSH_EIG_SHIFT_INVERT = 0, /* S = (A - s B)^-1 B, B-inner product ("solver" 0, 1, 2)   */
    KSH_EIG_BUCKLING = 1      /* S = A^-1 B (s = 0) or (A - s B)^-1 A, A-inner product (3) */
};

typedef struct {
    int mode;                /* ksh_eig_mode                                              */
    int number;              /* eigenpairs wanted ("number" of computeSparse)             */
    int ncv_fac;             /* Lanczos basis size = ncv_fac * number ("ncvFac")         */
    double shift;            /* "shift"                                                   */
    double tol;              /* "tolerance": Ritz residual <= tol |theta|                 */
    int max_restarts;        /* 0: 1000 (Spectra's maxit of computeSparse_impl)           */
    int precision;           /* factor of A - shift B: KSH_FP64 | KSH_FP32_IR            */
    int max_refine;          /* refinement sweeps per solve (FP32_IR)                     */
    double refine_tol;
    uint64_t seed;           /* start vector (0: fixed default)                          */
} ksh_eig_options;

typedef struct {
    int converged;           /* pairs meeting tol                                         */
    int restarts;
    int64_t solves, products;/* operator solves and W mat-vecs                            */
    double wall_seconds;
} ksh_eig_result;

/* values[number] ascending; vectors[number*n] (optional, W-normalized, free dofs) */
int ksh_eigen(ksh_matrix *A, ksh_matrix *B, const ksh_eig_options *opt, double *values, double *vectors,
              ksh_eig_result *res);

/* solvers/buckling_shell_XML.cpp:212-247: K_L, F = assemble(); u = K_L^-1 F; K_NL = K(u);
 * gsBucklingSolver(K_L, K_NL).computeSparse(number) (opt->mode is taken as BUCKLING). */
int ksh_run_buckling(ksh_problem *pb, const ksh_eig_options *opt, double *values, double *vectors,
                     ksh_eig_result *res);

/* solvers/modal_shell_XML.cpp:233-253: K = assemble().matrix(), M = assembleMass();
 * gsModalSolver(K, M).computeSparse(number) (opt->mode is taken as SHIFT_INVERT). */
int ksh_run_modal(ksh_problem *pb, const ksh_eig_options *opt, double *values, double *vectors,
                  ksh_eig_result *res);

/* ---- learned predictor network ------------------------------------------------------
 * The MLP of the learned arc-length predictor (no reference counterpart; it enters the solver
 * through the protected predictor() of gsALMCrisfield, src/gsALMSolvers/gsALMCrisfield.h:85,100).
 * widths[4] = {264, 256, 256, 65}: GELU hidden layers, linear output; BF16 products with FP32
 * accumulation on MFMA, FP32 master weights, Adam.  Parameters are packed as W0, W1, W2
 * (row-major [in][out]) then b0, b1, b2.  Host arrays, row-major samples. */
int ksh_mlp_create(const int *widths, uint64_t seed, ksh_mlp **out);
int ksh_mlp_destroy(ksh_mlp *m);
int ksh_mlp_num_params(const ksh_mlp *m, int64_t *n);
int ksh_mlp_get_params(const ksh_mlp *m, float *out);
int ksh_mlp_set_params(ksh_mlp *m, const float *in);
int ksh_mlp_forward(ksh_mlp *m, const float *x, int batch, float *y);
/* gradient of the mean squared error (over batch x outputs) with respect to the parameters */
int ksh_mlp_gradients(ksh_mlp *m, const float *x, const float *target, int batch, float *grads, double *loss);
/* minibatch Adam over the samples (full batches per epoch, shuffled with `seed`) */
int ksh_mlp_train(ksh_mlp *m, const float *x, const float *target, int nsamples, int epochs, int batch, double lr,
                  double weight_decay, uint64_t seed, double *loss_per_epoch);
/* input / output normalization used by ksh_mlp_predict: x_n = (x - in_mean) / in_std,
 * y = y_n out_std + out_mean */
int ksh_mlp_set_normalization(ksh_mlp *m, const double *in_mean, const double *in_std, const double *out_mean,
                              const double *out_std);
int ksh_mlp_predict(ksh_mlp *m, const double *x, int batch, double *y);

/* The predictor's 66 features of a converged state (U on the free dofs, lambda): lambda, ||U||,
 * and the 64 least-squares coefficients of the out-of-plane displacement in the modes
 * sin(a pi xi) sin(b pi eta), a, b = 1..8, at the Greville points of the control net. */
int ksh_predictor_features(ksh_problem *pb, const double *U, double lambda, double *out66);
/* dU[free] += scale * (out-of-plane field of the 64 mode coefficients coef64) */
int ksh_predictor_decode(ksh_problem *pb, const double *coef64, double scale, double *dU);

#ifdef __cplusplus
}
#endif
#endif
