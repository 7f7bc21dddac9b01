// ksh_path.cpp -- the path engine: arc-length (Crisfield, Riks) and load-controlled Newton paths of
// many shell problems at once, driven from one host thread with every vector in device memory.
//
// Design.  Each instance is a small state machine whose per-step scalars (load factor, arc length,
// Crisfield root choice, convergence norms, pivot counts) live in host arrays; its vectors (U, the
// step increment, the tangent and residual solutions, ...) are slices of one device allocation.
// A round advances every live instance by one predictor or one corrector iteration, with the same
// sequence of batched launches for all of them:
//   A  learned proposals (instances starting a step with the learned predictor): MLP inference,
//      the extrapolated increment plus its modal correction, scaled onto the arc-length sphere,
//      and its residual -- the instance then joins this round as a corrector;
//   B  tangent K(u) + LDL^T of every instance that needs one;
//   C  solves: K^-1 F (every instance), K^-1 (-r) (correctors);
//   D  dot products -> host scalar logic (roots, root choice) -> vector updates;
//   E  residual at the updated point -> norms -> convergence, step completion or failure.
// The scalar logic follows the reference's algorithms (cited at each function): the Crisfield
// quadratic constraint with its root selection and complex-root fallbacks, the Riks linearised
// constraint, gsStaticNewton's convergence test, and benchmark_Roof's driver loop (halve the arc
// length and retry from the last converged point when a step does not converge).
// Singular points (gsALMBase.hpp:422-856): an instance whose stability changes in a step leaves
// the step loop for a small state machine -- the singular-point test (power iterations), the
// optional bisection on the pivot count (ordinary steps of the same engine), the extended system
// for the critical point and the branch switch -- advanced one stage per round, batched with the
// other instances at the same stage, while the rest of the batch goes on stepping.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <vector>

#include "ksh_batch.hpp"
#include "ksh_internal.hpp"
#include "ksh_profile.hpp"

namespace ksh {

int modal_device(Problem &pb);
int mlp_io_widths(const ksh_mlp *m, int *in, int *out);

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int sgn(double v) { return (0.0 < v) - (v < 0.0); }  // gsALMHelper.h:21-25

// device vectors of one path instance (slices of one allocation)
enum VecId {
    V_U, V_DU, V_DUOLD, V_DU_STEP, V_UT, V_UB, V_R, V_NR, V_X, V_UH1, V_UH2, V_UPREV, V_T1, V_T2, V_BASE_COUNT,
    // singular points (allocated when they are handled): the critical mode V and its increment,
    // W = V + DV, the driver's previous point, the bisection's last point, the extended system's
    // mode solves, two products and the perturbed point
    V_SPV = V_BASE_COUNT, V_SPDV, V_SPW, V_SPUOLD, V_SPBU, V_SPVT, V_SPVB, V_SPP, V_SPQ, V_SPY, V_COUNT
};

struct Slab {
    double *base = nullptr;
    ~Slab()
    {
        if (base) (void)hipFree(base);
    }
};

struct MatrixOwner {
    std::vector<ksh_matrix *> ms;
    ~MatrixOwner()
    {
        for (ksh_matrix *K : ms)
            if (K) ksh_matrix_destroy(K);
    }
};

enum class Phase { Start, Correct, Singular, Done };

// stages of computeSingularPoint (+ switchBranch) for one instance
enum class Sp { None, Test, BisectStart, BisectStep, BisectEval, BisectVector, ExtStart, ExtIter, Finish };

constexpr int kHist = 4;  // converged states the learned predictor reads

struct Inst {
    Problem *pb = nullptr;
    Matrix *K = nullptr;
    int n = 0;
    ksh_alm_result *res = nullptr;
    double *v[V_COUNT] = {};
    double FF = 0;                     // F . F
    // reference state (gsALMBase / gsALMCrisfield / gsALMRiks members)
    double L = 0, DL = 0, DLold = 0, dl = 0, eta = 1, phi = 0;
    double arc = 0, arc0 = 0, arc_prev = 0, Lprev = 0;
    double prev_dnorm = 0;             // Riks: ||U - Uprev|| of the current point
    bool phi_user = false, prev_zero = true;
    double basisF = 0, basisU = 0, residueF = 0, residueU = 0;
    int it = 0;
    Phase ph = Phase::Start;
    bool learned_step = false, learned_off = false, bisected = false;
    int k = 0, halvings = 0, status = KSH_SUCCESS;
    int64_t negatives = 0;
    double indicator = 0;
    int stability = -1, stability_prev = -1;
    // learned history: features of the last converged states, load factors, arc lengths
    std::deque<std::array<double, kFeatures>> hist;
    double Lh1 = 0, Lh2 = 0;
    std::deque<double> lens;
    double dl_raw = 0;
    // per-round scratch
    bool riks_first = false, complex_root = false;
    double unorm_new = 0, DUDU = 0, last_arc = 0;
    int last_fail = KSH_SUCCESS;
    int64_t jac = 0, resid = 0, solves = 0;
    // singular point: stage, K(U + eps W) of the extended system, the driver's previous load
    // factor, the bisection's reference pivot / count / last load factor / saved arc length, the
    // extended system's load increment, iteration counters, outcomes
    Sp sp = Sp::None;
    Matrix *Keps = nullptr;
    double sp_Lold = 0, sp_Lbu = 0, sp_ref = 0, sp_arc = 0, sp_DL = 0;
    int64_t sp_fa = 0;
    int sp_k = 0, sp_kind = 0;
    bool sp_test = false, sp_conv = false, sp_step_ok = false;
    // what the step's record shows (a singular point is recorded after its procedure)
    int rec_it = 0;
    double rec_resF = 0, rec_arc = 0;
};

class AlmEngine {
public:
    AlmEngine(const ksh_alm_options &o) : m_o(o) {}

    int run(ksh_problem *const *pbs, int count, ksh_alm_result *res, double *wall)
    {
        int rc;
        if ((rc = validate())) return rc;
        if ((rc = m_dev.init())) return rc;
        const double t0 = now_s();
        if ((rc = setup(pbs, count, res))) return rc;
        if ((rc = initialize())) return rc;
        for (;;) {
            if ((rc = take_counts(false))) return rc;
            std::vector<Inst *> live;
            for (Inst &I : m_in)
                if (I.ph != Phase::Done) live.push_back(&I);
            if (live.empty()) break;
            const double r0 = now_s();
            if ((rc = round(live))) return rc;
            prof_count("engine_rounds", 1);
            prof_count("engine_round_us", (long long)((now_s() - r0) * 1e6));
        }
        if ((rc = take_counts(true))) return rc;
        if ((rc = m_dev.synchronize())) return rc;
        const double t1 = now_s();
        if (wall) *wall = t1 - t0;
        int worst = KSH_SUCCESS;
        for (Inst &I : m_in) {
            ksh_alm_result &r = *I.res;
            r.steps_done = I.k;
            r.halvings = I.halvings;
            r.final_status = I.status == KSH_NOT_CONVERGED && I.last_fail ? I.last_fail : I.status;
            r.wall_seconds = t1 - t0;
            r.phi = I.phi;
            r.jacobians = I.jac;
            r.residuals = I.resid;
            r.solves = I.solves;
            if (I.status != KSH_SUCCESS && I.status != KSH_NOT_CONVERGED && worst == KSH_SUCCESS) worst = I.status;
        }
        if (worst != KSH_SUCCESS) return worst;
        for (Inst &I : m_in)
            if (I.k != m_o.nsteps) return KSH_NOT_CONVERGED;
        return KSH_SUCCESS;
    }

private:
    int validate()
    {
        const ksh_alm_options &o = m_o;
        if (o.method != 0 && o.method != 1) return fail(KSH_OTHER_ERROR, "unknown arc-length method");
        if (o.precision != KSH_FP64 && o.precision != KSH_FP32_IR && o.precision != KSH_FP32)
            return fail(KSH_OTHER_ERROR, "unknown precision");
        if (o.nsteps < 0 || o.max_iter < 2 || (!o.lengths && !(o.length > 0))) return fail(KSH_OTHER_ERROR, "bad step options");
        if (o.quasi || o.quasi_iterations > 0)
            return fail(KSH_OTHER_ERROR, "quasi-Newton iterations are not supported by the path engine");
        if (o.bifurcation_method != -1 && o.bifurcation_method != 0)
            return fail(KSH_OTHER_ERROR, "BifurcationMethod: only -1 (nothing) and 0 (determinant) are supported");
        if (o.singular_point < 0 || o.singular_point > 2) return fail(KSH_OTHER_ERROR, "singular_point: 0, 1 or 2");
        if (o.singular_point) {
            if (o.sp_test_tol < 0 || o.sp_test_it < 0 || o.sp_tol_e < 0 || o.sp_tol_b < 0 || o.perturbation < 0 ||
                o.length_after_sp < 0 || (o.sp_fail != 0 && o.sp_fail != 1))
                return fail(KSH_OTHER_ERROR, "bad singular-point options");
            m_sp_test_tol = o.sp_test_tol > 0 ? o.sp_test_tol : 1e-6;
            m_sp_test_it = o.sp_test_it > 0 ? o.sp_test_it : 5;
            m_sp_tol_e = o.sp_tol_e > 0 ? o.sp_tol_e : 1e-10;
            m_sp_tau = o.perturbation > 0 ? o.perturbation : 1e3;
        }
        if (o.learned) {
            if (o.method != 0) return fail(KSH_OTHER_ERROR, "the learned predictor serves the Crisfield method only");
            int in = 0, out = 0, rc;
            if ((rc = mlp_io_widths(o.learned, &in, &out))) return rc;
            if (in != kHist * kFeatures || out != kModalCoef + 1)
                return fail(KSH_OTHER_ERROR, "learned predictor network must map " + std::to_string(kHist * kFeatures) +
                                                 " inputs to " + std::to_string(kModalCoef + 1) + " outputs");
        }
        return KSH_SUCCESS;
    }

    int setup(ksh_problem *const *pbs, int count, ksh_alm_result *res)
    {
        m_in.resize(count);
        size_t tot = 0, nmax = 0;
        for (int i = 0; i < count; ++i) {
            tot += (size_t)pbs[i]->nfree;
            nmax = std::max(nmax, (size_t)pbs[i]->nfree);
        }
        const int nvec = m_o.singular_point ? V_COUNT : V_BASE_COUNT;
        const size_t nones = m_o.singular_point ? nmax : 0;
        const size_t bytes = (tot * nvec + (size_t)count * kModalCoef + nones) * sizeof(double);
        KSH_HIP(hipMalloc(&m_slab.base, bytes));
        KSH_HIP(hipMemsetAsync(m_slab.base, 0, bytes, m_dev.stream()));
        double *p = m_slab.base;
        m_coef = p + tot * nvec;
        if (nones) {  // the power iterations' start vector (gsALMBase.hpp:485, 822)
            m_ones = m_coef + (size_t)count * kModalCoef;
            std::vector<double> one(nones, 1.0);
            KSH_HIP(hipMemcpyAsync(m_ones, one.data(), nones * sizeof(double), hipMemcpyHostToDevice, m_dev.stream()));
            KSH_HIP(hipStreamSynchronize(m_dev.stream()));
        }
        for (int i = 0; i < count; ++i) {
            Inst &I = m_in[i];
            I.pb = pbs[i];
            I.n = pbs[i]->nfree;
            I.res = &res[i];
            for (int q = 0; q < nvec; ++q) {
                I.v[q] = p;
                p += I.n;
            }
            if (m_o.inertia_guard != 0) I.pb->inertia_guard = m_o.inertia_guard > 0 ? m_o.inertia_guard : 0.0;
            if (m_o.refine_accept > 0) I.pb->refine_accept = m_o.refine_accept;
            ksh_matrix *K = nullptr;
            int rc;
            if ((rc = ksh_matrix_create(static_cast<ksh_problem *>(I.pb), m_o.precision, &K))) return rc;
            m_mats.ms.push_back(K);
            I.K = K;
            I.arc = I.arc0 = I.arc_prev = m_o.lengths ? m_o.lengths[i] : m_o.length;
            if (!(I.arc > 0)) return fail(KSH_OTHER_ERROR, "arc lengths must be positive");
            I.phi = m_o.scaling;
            I.phi_user = m_o.scaling != -1;
            if (m_o.learned && (rc = modal_device(*I.pb))) return rc;
            if (res[i].features && (rc = modal_device(*I.pb))) return rc;
        }
        std::vector<DotDesc> dd;
        for (Inst &I : m_in) dd.push_back(dot1(I, I.pb->F, I.pb->F));
        std::vector<double> out;
        int rc;
        if ((rc = m_dev.dots(dd, out))) return rc;
        for (size_t i = 0; i < m_in.size(); ++i) m_in[i].FF = out[i];
        return KSH_SUCCESS;
    }

    static DotDesc dot1(const Inst &I, const double *a, const double *b)
    {
        DotDesc d;
        d.n = I.n;
        d.npairs = 1;
        d.a[0] = a;
        d.b[0] = b;
        return d;
    }
    static void add_pair(DotDesc &d, const double *a, const double *b)
    {
        d.a[d.npairs] = a;
        d.b[d.npairs] = b;
        ++d.npairs;
    }

    // gsALMBase::initialize -> init(stability) -> _computeStability(U = 0) (gsALMBase.hpp:95-103):
    // the tangent at the start point, its factor and its pivot count
    int initialize()
    {
        std::vector<TangentJob> tj;
        for (Inst &I : m_in) {
            tj.push_back(TangentJob{I.pb, I.K, I.v[V_U]});
            ++I.jac;
        }
        int rc;
        if ((rc = m_dev.tangent(tj))) return rc;
        for (Inst &I : m_in) {
            if (I.K->info) {
                I.status = fail(KSH_SOLVER_ERROR, "LDL^T of the initial tangent: zero or non-finite pivot");
                I.ph = Phase::Done;
                continue;
            }
            count_pivots(I);
            // the drivers reset the indicator after initialize() (setIndicator(0)): stability -1
            I.indicator = 0.0;
            I.stability = -1;
            if (m_o.nsteps == 0) I.ph = Phase::Done;
        }
        return KSH_SUCCESS;
    }

    // negatives and min D of the matrix's current factor (gsALMBase.hpp:557-569,612-614)
    void count_pivots(Inst &I)
    {
        if (m_o.bifurcation_method < 0) {
            I.negatives = 0;
            I.indicator = 0.0;
            return;
        }
        I.negatives = (int64_t)I.K->fstats[0];
        I.indicator = I.K->fstats[1];
    }

    bool learned_usable(const Inst &I) const
    {
        return m_o.learned && !I.learned_off && I.sp == Sp::None && (int)I.hist.size() >= kHist && !I.lens.empty() &&
               I.lens.back() == I.arc;
    }

    void begin_step(Inst &I)
    {
        I.it = 0;
        I.DL = I.dl = 0.0;
        I.eta = 1.0;
        I.complex_root = false;
        I.learned_step = false;
    }

    int round(std::vector<Inst *> &all)
    {
        int rc;
        // instances at a singular point advance one stage of its procedure
        std::vector<Inst *> live, sing;
        // host wall time of each phase (profiling counters; GPU waits included)
        double tq = now_s();
        auto mark = [&](const char *name) {
            const double t = now_s();
            prof_count(name, (long long)((t - tq) * 1e6));
            tq = t;
        };
        for (Inst *I : all) (I->ph == Phase::Singular ? sing : live).push_back(I);
        if (!sing.empty() && (rc = singular_round(sing))) return rc;
        // ---- A: learned proposals ------------------------------------------------------------
        std::vector<Inst *> lp;
        for (Inst *I : live)
            if (I->ph == Phase::Start) {
                begin_step(*I);
                if (learned_usable(*I)) lp.push_back(I);
            }
        if (!lp.empty() && (rc = learned_predictor(lp))) return rc;
        mark("round_predictor_us");
        // ---- B: tangents and their factors ---------------------------------------------------
        std::vector<TangentJob> tj;
        std::vector<Inst *> act;
        for (Inst *I : live) {
            if (I->ph == Phase::Done) continue;
            tj.push_back(TangentJob{I->pb, I->K, I->ph == Phase::Start ? I->v[V_U] : I->v[V_X]});
            ++I->jac;
            act.push_back(I);
        }
        if (act.empty()) return KSH_SUCCESS;
        if ((rc = m_dev.tangent(tj))) return rc;
        mark("round_tangent_us");
        {
            std::vector<Inst *> ok;
            for (Inst *I : act) {
                if (I->K->info) {  // gsALMBase::factorizeMatrix throws 3: SolverError ends the path
                    I->status = fail(KSH_SOLVER_ERROR, "LDL^T: zero or non-finite pivot");
                    I->ph = Phase::Done;
                } else {
                    ok.push_back(I);
                }
            }
            act.swap(ok);
        }
        // ---- C: solves -------------------------------------------------------------------------
        std::vector<VecDesc> neg;
        std::vector<SolveJob> sj;
        for (Inst *I : act) {
            const bool start = I->ph == Phase::Start;
            I->riks_first = false;
            if (start && m_o.method == 1) {
                // gsALMRiks::predictor (gsALMRiks.hpp:98-134): the tangent solve only at the first
                // point of a path (U == Uprev, L == Lprev)
                I->riks_first = I->prev_dnorm < 1e-10 && (I->L - I->Lprev) * (I->L - I->Lprev) < 1e-10;
                if (!I->riks_first) continue;
            }
            sj.push_back(SolveJob{I->K, I->pb->F, I->v[V_UT], m_o.max_refine, m_o.refine_tol, KSH_SUCCESS, I});
            if (!start) {
                VecDesc d;
                d.n = I->n;
                d.add(I->v[V_NR], -1.0, I->v[V_R]);
                neg.push_back(d);
                sj.push_back(SolveJob{I->K, I->v[V_NR], I->v[V_UB], m_o.max_refine, m_o.refine_tol, KSH_SUCCESS, I});
            }
        }
        if ((rc = m_dev.vecops(neg))) return rc;
        if ((rc = m_dev.solve(sj))) return rc;
        mark("round_solve_us");
        for (const SolveJob &q : sj) {
            Inst *I = static_cast<Inst *>(q.tag);
            ++I->solves;
            if (q.rc != KSH_SUCCESS && I->ph != Phase::Done) {
                I->status = q.rc;
                I->ph = Phase::Done;
            }
        }
        {
            std::vector<Inst *> ok;
            for (Inst *I : act)
                if (I->ph != Phase::Done) ok.push_back(I);
            act.swap(ok);
        }
        // ---- D: scalar logic and vector updates -------------------------------------------------
        if ((rc = update(act))) return rc;
        mark("round_update_us");
        // ---- E: residual, norms, convergence ----------------------------------------------------
        rc = residual_and_check(act, false);
        mark("round_residual_us");
        return rc;
    }

    // ---- learned predictor (A) -------------------------------------------------------------------
    // The proposal is the quadratic extrapolation of the last three converged states,
    // 3 x_t - 3 x_t-1 + x_t-2, corrected by the network in the 64 out-of-plane mode coefficients
    // and the load factor, and scaled onto ||dU||^2 + phi^2 ||F||^2 dL^2 = ds^2.  It replaces the
    // tangent predictor of gsALMCrisfield (gsALMCrisfield.hpp:111-164); the corrector is unchanged.
    int learned_predictor(std::vector<Inst *> &lp)
    {
        int rc;
        const int cnt = (int)lp.size();
        const int in = kHist * kFeatures, out = kModalCoef + 1;
        std::vector<double> X((size_t)cnt * in), Y((size_t)cnt * out);
        for (int i = 0; i < cnt; ++i)
            for (int s = 0; s < kHist; ++s)
                std::copy(lp[i]->hist[s].begin(), lp[i]->hist[s].end(), X.begin() + (size_t)i * in + s * kFeatures);
        if (m_o.learned_mode == 1) {
            std::fill(Y.begin(), Y.end(), 0.0);  // extrapolation only: the network's correction left out
        } else if ((rc = ksh_mlp_predict(m_o.learned, X.data(), cnt, Y.data()))) {
            return rc;
        }
        std::vector<double> coef((size_t)cnt * kModalCoef);
        for (int i = 0; i < cnt; ++i)
            std::copy(Y.begin() + (size_t)i * out, Y.begin() + (size_t)i * out + kModalCoef, coef.begin() + (size_t)i * kModalCoef);
        KSH_HIP(hipMemcpyAsync(m_coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, m_dev.stream()));
        std::vector<VecDesc> vd;
        std::vector<DecDesc> dec;
        for (int i = 0; i < cnt; ++i) {
            Inst &I = *lp[i];
            VecDesc d;
            d.n = I.n;
            d.add(I.v[V_DU_STEP], 2.0, I.v[V_U], -3.0, I.v[V_UH1], 1.0, I.v[V_UH2]);
            vd.push_back(d);
            DecDesc e;
            e.ncp = I.pb->ncp;
            e.Phi = I.pb->modal_Phi;
            e.zfree = I.pb->modal_zfree;
            e.coef = m_coef + (size_t)i * kModalCoef;
            e.scale = 1.0;
            e.dU = I.v[V_DU_STEP];
            dec.push_back(e);
            I.dl_raw = 2.0 * I.L - 3.0 * I.Lh1 + I.Lh2 + Y[(size_t)i * out + kModalCoef];
        }
        if ((rc = m_dev.vecops(vd)) || (rc = batch_decode(m_dev.stream(), m_decdesc, dec))) return rc;
        std::vector<DotDesc> dd;
        for (Inst *I : lp) {
            DotDesc d = dot1(*I, I->v[V_DU_STEP], I->v[V_DU_STEP]);
            add_pair(d, I->v[V_U], I->v[V_U]);
            dd.push_back(d);
        }
        std::vector<double> dv;
        if ((rc = m_dev.dots(dd, dv))) return rc;
        std::vector<Inst *> used;
        vd.clear();
        for (int i = 0; i < cnt; ++i) {
            Inst &I = *lp[i];
            const double rr = dv[2 * i], UU = dv[2 * i + 1];
            if (!I.phi_user) I.phi = std::sqrt(UU / (I.L * I.L * I.FF));
            const double ph = I.phi > 0 ? I.phi : 0.0;
            const double nrm = std::sqrt(rr + ph * ph * I.FF * I.dl_raw * I.dl_raw);
            if (!(nrm > 0) || !std::isfinite(nrm)) continue;  // the tangent predictor takes this step
            const double s = I.arc / nrm;
            I.dl = s * I.dl_raw;
            I.DL = I.dl;
            VecDesc d;
            d.n = I.n;
            d.add(I.v[V_DU_STEP], s, I.v[V_DU_STEP]);
            d.add(I.v[V_DU], 1.0, I.v[V_DU_STEP]);
            d.add(I.v[V_X], 1.0, I.v[V_U], 1.0, I.v[V_DU]);
            if (m_o.angle_method == 1 || m_o.angle_method == 2) {
                d.add(I.v[V_DUOLD], 1.0, I.v[V_DU]);
                I.DLold = I.DL;
            }
            vd.push_back(d);
            I.learned_step = true;
            ++I.res->learned_steps;
            used.push_back(&I);
        }
        if ((rc = m_dev.vecops(vd))) return rc;
        if (used.empty()) return KSH_SUCCESS;
        return residual_and_check(used, true);
    }

    // ---- D ------------------------------------------------------------------------------------
    int update(std::vector<Inst *> &act)
    {
        int rc;
        std::vector<DotDesc> dd;
        for (Inst *I : act) {
            double *U = I->v[V_U], *DU = I->v[V_DU], *Ut = I->v[V_UT], *Ub = I->v[V_UB], *O = I->v[V_DUOLD];
            DotDesc d;
            d.n = I->n;
            if (I->ph == Phase::Start) {
                if (m_o.method == 0) {
                    add_pair(d, Ut, Ut);
                    add_pair(d, O, Ut);
                    add_pair(d, U, U);
                } else if (I->riks_first) {
                    add_pair(d, Ut, Ut);
                }
            } else if (m_o.method == 0) {
                add_pair(d, Ut, Ut);
                add_pair(d, Ut, DU);
                add_pair(d, Ub, Ut);
                add_pair(d, DU, DU);
                add_pair(d, DU, Ub);
                add_pair(d, Ub, Ub);
                add_pair(d, O, Ut);
            } else {
                add_pair(d, DU, DU);
                add_pair(d, DU, Ub);
                add_pair(d, DU, Ut);
            }
            dd.push_back(d);
        }
        std::vector<double> dv;
        if ((rc = m_dev.dots(dd, dv))) return rc;
        std::vector<VecDesc> vd;
        std::vector<Inst *> cx;
        size_t o = 0;
        for (size_t i = 0; i < act.size(); ++i) {
            Inst &I = *act[i];
            const double *g = dv.data() + o;
            o += (size_t)dd[i].npairs;
            VecDesc d;
            d.n = I.n;
            if (I.ph == Phase::Start) {
                if (m_o.method == 0) crisfield_predictor(I, g, d);
                else riks_predictor(I, g, d);
            } else {
                count_pivots(I);  // the counted tangent is the one of this iteration (gsALMBase.hpp:386-390)
                I.stability = I.indicator < 0 ? 1 : -1;
                if (m_o.method == 0) {
                    if (!crisfield_corrector(I, g, d)) {
                        cx.push_back(&I);
                        continue;
                    }
                } else {
                    riks_corrector(I, g, d);
                }
            }
            vd.push_back(d);
        }
        if (!cx.empty()) {
            if ((rc = complex_roots(cx, vd))) return rc;
        }
        return m_dev.vecops(vd);
    }

    // gsALMCrisfield::predictor (gsALMCrisfield.hpp:111-164) with computeLambdaMU (:384-398);
    // g = {Ut.Ut, DUold.Ut, U.U}
    void crisfield_predictor(Inst &I, const double *g, VecDesc &d)
    {
        const double uu = g[0], ou = g[1], UU = g[2];
        if (I.prev_zero) {  // no previous step: the origin
            I.dl = I.arc / std::sqrt(2.0 * uu);
            if (!I.phi_user) I.phi = std::sqrt(uu / I.FF);
        } else {
            if (!I.phi_user) I.phi = std::sqrt(UU / (I.L * I.L * I.FF));
            const double A0 = I.phi * I.phi * I.FF;
            const int dir = sgn(ou + A0 * I.DLold);
            const double den = std::sqrt(uu + A0);
            const double mu = den == 0 ? I.arc : I.arc / den;
            I.dl = dir * mu;
        }
        d.add(I.v[V_DU_STEP], I.dl, I.v[V_UT]);
        d.add(I.v[V_DU], 1.0, I.v[V_DU_STEP]);
        d.add(I.v[V_X], 1.0, I.v[V_U], 1.0, I.v[V_DU]);
        I.DL = I.dl;
        if (m_o.angle_method == 1 || m_o.angle_method == 2) {
            d.add(I.v[V_DUOLD], 1.0, I.v[V_DU]);
            I.DLold = I.DL;
        }
    }

    // gsALMRiks::predictor (gsALMRiks.hpp:98-134); g = {Ut.Ut} on the first point of a path
    void riks_predictor(Inst &I, const double *g, VecDesc &d)
    {
        I.phi = I.n == 1 ? 0.99999999999 : 1.0 / I.n;
        if (I.riks_first) {
            const double den = std::sqrt(g[0] + I.DL * 1.0);
            I.dl = I.arc * (1.0 / den);
            d.add(I.v[V_DU_STEP], I.arc / den, I.v[V_UT]);
        } else {
            const double f = I.arc * (1.0 / I.arc_prev);
            I.dl = I.arc * ((I.L - I.Lprev) / I.arc_prev);
            d.add(I.v[V_DU_STEP], f, I.v[V_U], -f, I.v[V_UPREV]);
        }
        d.add(I.v[V_DU], 1.0, I.v[V_DU_STEP]);
        d.add(I.v[V_X], 1.0, I.v[V_U], 1.0, I.v[V_DU]);
        I.DL = I.dl;
    }

    // gsALMCrisfield::iteration with computeLambdas (gsALMCrisfield.hpp:74-100, 213-363, 401-472):
    // the quadratic constraint in the load increment, its roots chosen by the DOT rule; negative
    // discriminant: Lam's modified eta, else the simplified complex-root step (returns false: the
    // caller evaluates it).  g = {Ut.Ut, Ut.DU, Ub.Ut, DU.DU, DU.Ub, Ub.Ub, DUold.Ut}
    bool crisfield_corrector(Inst &I, const double *g, VecDesc &d)
    {
        const double uu = g[0], uD = g[1], bu = g[2], DD = g[3], Db = g[4], bb = g[5], ou = g[6];
        const double A0 = I.phi * I.phi * I.FF;
        const double a0 = uu + A0;
        const double b0 = 2.0 * (uD + I.DL * A0);
        const double b1 = 2.0 * bu;
        const double c0 = DD + I.DL * I.DL * A0 - I.arc * I.arc;
        const double c1 = 2.0 * Db;
        const double c2 = bb;
        const double lamold = I.dl;
        I.eta = 1.0;
        double al1 = a0, al2 = b0 + I.eta * b1, al3 = c0 + I.eta * c1 + I.eta * I.eta * c2;
        double disc = al2 * al2 - 4.0 * al1 * al3;
        double roots[2];
        if (disc >= 0) {
            roots[0] = (-al2 + std::sqrt(disc)) / (2.0 * al1);
            roots[1] = (-al2 - std::sqrt(disc)) / (2.0 * al1);
        } else {
            // Lam 1992: the eta that makes the discriminant vanish, shifted by 5% of the root gap
            const double m1 = b1 * b1 - 4.0 * a0 * c2;
            const double m2 = 2.0 * b0 * b1 - 4.0 * a0 * c1;
            const double m3 = b0 * b0 - 4.0 * a0 * c0;
            const double disc2 = m2 * m2 - 4.0 * m1 * m3;
            if (disc2 >= 0) {
                const double e0 = (-m2 + std::sqrt(disc2)) / (2.0 * m1), e1 = (-m2 - std::sqrt(disc2)) / (2.0 * m1);
                const double eta1 = std::min(e0, e1), eta2 = std::max(e0, e1);
                const double xi = 0.05 * std::abs(eta2 - eta1);
                if (eta2 < 1.0) I.eta = eta2 - xi;
                else if (eta2 > 1.0 && -m2 / m1 < 1.0) I.eta = eta2 + xi;
                else if (eta1 < 1.0 && -m2 / m1 > 1.0) I.eta = eta1 - xi;
                else if (eta1 > 1.0) I.eta = eta1 + xi;
            }
            if (disc2 >= 0 && I.eta > 0.05) {
                al2 = b0 + I.eta * b1;
                roots[0] = roots[1] = -al2 / (2.0 * a0);  // Zhou 1995: the double root
            } else {
                I.eta = 1.0;
                I.complex_root = true;
                return false;
            }
        }
        // DOT rule of Ritto-Correa et al. 2008 (as the reference writes it: phi^2 DLold)
        const double base = ou + I.phi * I.phi * I.DLold;
        const double d1 = roots[0] * base, d2 = roots[1] * base;
        const int pick = d1 < d2 ? 1 : 0;
        I.dl = roots[pick];
        double cu = I.eta, ct = I.dl;
        if (((lamold * I.dl < 0) && (std::abs(I.dl) <= std::abs(lamold))) && m_o.relaxation != 1.0) {
            // relaxation against an oscillating load factor (gsALMCrisfield.hpp:84-89)
            cu = m_o.relaxation * I.eta;
            ct = m_o.relaxation * I.dl;
            I.dl = m_o.relaxation * I.dl;
        }
        d.add(I.v[V_DU_STEP], cu, I.v[V_UB], ct, I.v[V_UT]);
        finish_correction(I, d);
        return true;
    }

    void finish_correction(Inst &I, VecDesc &d)
    {
        d.add(I.v[V_DU], 1.0, I.v[V_DU], 1.0, I.v[V_DU_STEP]);
        d.add(I.v[V_X], 1.0, I.v[V_U], 1.0, I.v[V_DU]);
        I.DL += I.dl;
        if (m_o.method == 0 && m_o.angle_method == 1) {
            d.add(I.v[V_DUOLD], 1.0, I.v[V_DU]);
            I.DLold = I.DL;
        }
    }

    // gsALMCrisfield::computeLambdasComplex (gsALMCrisfield.hpp:306-325): the load factor that
    // best matches the internal forces K (U + DU), and the step scaled back onto the sphere
    int complex_roots(std::vector<Inst *> &cx, std::vector<VecDesc> &vd)
    {
        int rc;
        std::vector<VecDesc> pre;
        for (Inst *I : cx) {
            VecDesc d;
            d.n = I->n;
            d.add(I->v[V_T2], 1.0, I->v[V_DU], 1.0, I->v[V_UB]);
            pre.push_back(d);
        }
        if ((rc = m_dev.vecops(pre)) || (rc = m_dev.synchronize())) return rc;
        for (Inst *I : cx) {
            if ((rc = band_symv(*I->K, I->v[V_X], I->v[V_T1]))) return rc;  // K (U + DU) on the problem stream
            KSH_HIP(hipStreamSynchronize(I->pb->stream));
        }
        std::vector<DotDesc> dd;
        for (Inst *I : cx) {
            DotDesc d = dot1(*I, I->v[V_T1], I->pb->F);
            add_pair(d, I->v[V_T2], I->v[V_T2]);
            dd.push_back(d);
        }
        std::vector<double> dv;
        if ((rc = m_dev.dots(dd, dv))) return rc;
        for (size_t i = 0; i < cx.size(); ++i) {
            Inst &I = *cx[i];
            const double A0 = I.phi * I.phi * I.FF;
            const double Lcr = dv[2 * i] / I.FF;
            const double DLcr = Lcr - I.L;
            const double arc_cr = std::sqrt(dv[2 * i + 1] + A0 * DLcr * DLcr);
            const double mu = I.arc / arc_cr;
            I.dl = mu * DLcr - I.DL;
            VecDesc d;
            d.n = I.n;
            d.add(I.v[V_DU_STEP], mu, I.v[V_T2], -1.0, I.v[V_DU]);
            finish_correction(I, d);
            vd.push_back(d);
            prof_count("crisfield_complex_roots", 1);
        }
        return KSH_SUCCESS;
    }

    // gsALMRiks::iteration (gsALMRiks.hpp:57-71); g = {DU.DU, DU.Ub, DU.Ut}
    void riks_corrector(Inst &I, const double *g, VecDesc &d)
    {
        const double DD = g[0], Db = g[1], Dt = g[2];
        const double r = I.phi * DD + (1.0 - I.phi) * I.DL * I.DL - I.arc * I.arc;
        I.dl = -(r + 2.0 * I.phi * Db) / (2.0 * (1.0 - I.phi) * I.DL + 2.0 * I.phi * Dt);
        d.add(I.v[V_DU_STEP], 1.0, I.v[V_UB], I.dl, I.v[V_UT]);
        finish_correction(I, d);
    }

    // ---- E -------------------------------------------------------------------------------------
    // Residual at U + DU, L + DL; computeResidualNorms (gsALMBase.hpp:151-183); convergence test
    // residueF < TolF and residueU < TolU (:394-405).  `predictor_only`: instances that just took a
    // learned proposal (norm bases of iteration 0, then a corrector iteration in this round).
    int residual_and_check(std::vector<Inst *> &act, bool predictor_only)
    {
        int rc;
        std::vector<ResidualJob> rj;
        for (Inst *I : act) {
            rj.push_back(ResidualJob{I->pb, I->v[V_X], I->L + I->DL, KSH_RES_ALM, I->v[V_R]});
            ++I->resid;
        }
        if ((rc = m_dev.residual(rj))) return rc;
        std::vector<DotDesc> dd;
        for (Inst *I : act) {
            DotDesc d = dot1(*I, I->v[V_R], I->v[V_R]);
            add_pair(d, I->v[V_DU_STEP], I->v[V_DU_STEP]);
            add_pair(d, I->v[V_DU], I->v[V_DU]);
            add_pair(d, I->v[V_X], I->v[V_X]);
            dd.push_back(d);
        }
        std::vector<double> dv;
        if ((rc = m_dev.dots(dd, dv))) return rc;
        std::vector<Inst *> conv;
        for (size_t i = 0; i < act.size(); ++i) {
            Inst &I = *act[i];
            const double rn = std::sqrt(dv[4 * i]), dn = std::sqrt(dv[4 * i + 1]), Dn = std::sqrt(dv[4 * i + 2]);
            I.unorm_new = std::sqrt(dv[4 * i + 3]);
            I.DUDU = dv[4 * i + 2];
            if (!std::isfinite(rn)) {  // the residual callback fails: AssemblyError, handled as a failed step
                fail_step(I, KSH_ASSEMBLY_ERROR);
                continue;
            }
            if (I.ph == Phase::Start || predictor_only) {
                I.basisF = std::abs(I.L + I.DL) * std::sqrt(I.FF);
                I.basisU = Dn;
                I.residueF = rn / I.basisF;
                I.residueU = dn / Dn;
                I.ph = Phase::Correct;
                I.it = 1;
                I.stability_prev = I.stability;
                continue;
            }
            I.residueF = rn / I.basisF;
            I.residueU = dn / I.basisU;
            if (I.residueF < m_o.tol_f && I.residueU < m_o.tol_u) {
                conv.push_back(&I);
            } else if (I.it == m_o.max_iter - 1) {
                fail_step(I, KSH_NOT_CONVERGED);
            } else {
                ++I.it;
            }
        }
        if (!conv.empty()) return complete_steps(conv);
        return KSH_SUCCESS;
    }

    // a step that did not converge: the learned proposal's step is redone from the tangent
    // predictor; otherwise the driver halves the arc length and retries from the last converged
    // point (benchmarks/benchmark_Roof.cpp:423-433)
    void fail_step(Inst &I, int code)
    {
        if (I.sp == Sp::BisectStep) {  // a bisection step that failed: gsALMBase.hpp:783-795
            I.ph = Phase::Singular;
            I.sp = Sp::BisectEval;
            I.sp_step_ok = false;
            return;
        }
        if (I.learned_step) {
            ++I.res->fallbacks;
            begin_step(I);
            I.learned_off = true;
            I.ph = Phase::Start;
            return;
        }
        I.learned_off = false;
        if (I.halvings >= m_o.max_halvings) {  // the path ends here (status of the last attempt kept)
            I.status = KSH_NOT_CONVERGED;
            I.last_fail = code;
            I.ph = Phase::Done;
            return;
        }
        ++I.halvings;
        I.arc = I.arc_prev = I.arc / 2.0;  // setLength(dL / 2)
        I.bisected = true;
        I.ph = Phase::Start;
    }

    // converged: gsALMCrisfield / gsALMRiks::iterationFinish, the stability of the counted
    // tangent, and the driver's bookkeeping (or the singular-point procedure when the stability
    // changed, or the next stage of a bisection)
    int complete_steps(std::vector<Inst *> &conv)
    {
        int rc;
        // FP32 pivot counts without a certificate are taken from an FP64 factor of the same tangent
        // (certificate: the estimate of rho(M^-1 E) below the guard, see ksh_batch.cpp)
        std::vector<Matrix *> cert, recount;
        for (Inst *I : conv) {
            Matrix &K = *I->K;
            if (m_o.bifurcation_method < 0 || K.fprec != KSH_FP32 || K.use64 || !(K.guard > 0)) continue;
            cert.push_back(&K);
        }
        if ((rc = m_dev.power_contraction(cert))) return rc;
        for (Matrix *K : cert)
            if (std::max(K->rho, K->rho_pw) > K->guard) recount.push_back(K);
        if (!recount.empty()) {
            prof_count("fp64_refactor_count", (long long)recount.size());
            if (!m_o.singular_point) {
                // the count is only recorded (no singular-point procedure branches on it): its FP64
                // factor runs beside the next rounds, and the record is completed when it is done
                std::vector<long long> tags;
                for (Matrix *K : recount)
                    for (Inst *I : conv)
                        if (I->K == K) {
                            tags.push_back((long long)m_pending.size());
                            m_pending.push_back(PendingCount{I, I->k});
                            break;
                        }
                std::vector<BatchDevice::Count> now;
                if ((rc = m_dev.recount_async(recount, tags, now))) return rc;
                for (const BatchDevice::Count &c : now) apply_count(c, true);
            } else {
                if ((rc = m_dev.refactor64(recount))) return rc;
                for (Inst *I : conv)
                    if (std::find(recount.begin(), recount.end(), I->K) != recount.end()) {
                        if (I->K->info) {
                            I->status = fail(KSH_SOLVER_ERROR, "FP64 refactorization for the pivot count failed");
                            I->ph = Phase::Done;
                            continue;
                        }
                        count_pivots(*I);
                    }
            }
        }
        std::vector<VecDesc> vd;
        std::vector<Inst *> fin;
        for (Inst *I : conv) {
            if (I->ph == Phase::Done) continue;
            VecDesc d;
            d.n = I->n;
            if (m_o.singular_point && I->sp == Sp::None) d.add(I->v[V_SPUOLD], 1.0, I->v[V_U]);  // the driver's Uold
            if (m_o.learned) {
                d.add(I->v[V_UH2], 1.0, I->v[V_UH1]);
                d.add(I->v[V_UH1], 1.0, I->v[V_U]);
            }
            if (m_o.method == 1) d.add(I->v[V_UPREV], 1.0, I->v[V_U]);
            d.add(I->v[V_U], 1.0, I->v[V_X]);
            if (m_o.method == 0 && m_o.angle_method == 0) d.add(I->v[V_DUOLD], 1.0, I->v[V_DU]);
            vd.push_back(d);
            fin.push_back(I);
        }
        if ((rc = m_dev.vecops(vd))) return rc;
        std::vector<Inst *> rec;
        for (Inst *I : fin) {
            const double arc_used = I->arc;
            I->last_arc = arc_used;
            if (m_o.method == 1) {
                I->Lprev = I->L;
                I->prev_dnorm = std::sqrt(I->DUDU);
            }
            I->Lh2 = I->Lh1;
            I->Lh1 = I->L;
            if (I->sp == Sp::None) I->sp_Lold = I->L;  // the driver's Lold
            I->L += I->DL;
            if (m_o.method == 0 && m_o.angle_method == 0) {
                I->DLold = I->DL;
                I->prev_zero = I->DUDU == 0 && I->DL == 0;
            } else if (m_o.method == 0) {
                I->prev_zero = false;
            }
            I->stability = I->indicator < 0 ? 1 : -1;
            // gsALMBase::_step after convergence (gsALMBase.hpp:396-403): adaptive length (off
            // inside a bisection, gsALMBase.hpp:763-764)
            if (m_o.adaptive && I->sp != Sp::BisectStep) {
                double fac = (double)m_o.adaptive_iterations / I->it;
                fac = std::min(2.0, std::max(0.5, fac));
                I->arc_prev = I->arc;
                I->arc *= fac;
            } else {
                I->arc_prev = I->arc;
            }
            if (I->sp == Sp::BisectStep) {  // gsALMBase.hpp:786-800: the bisection evaluates the point
                I->ph = Phase::Singular;
                I->sp = Sp::BisectEval;
                I->sp_step_ok = true;
                continue;
            }
            I->rec_it = I->it;
            I->rec_resF = I->residueF;
            I->rec_arc = arc_used;
            // the drivers: computeStability(false) (the same counted tangent), stabilityChange()
            if (m_o.singular_point && I->stability * I->stability_prev < 0) {
                if (m_o.verbose)
                    std::printf("[ksh path] step %d: stability changes at lambda %.10e, singular point\n", I->k, I->L);
                I->ph = Phase::Singular;
                I->sp = Sp::Test;
                continue;
            }
            rec.push_back(I);
        }
        return record_steps(rec);
    }

    // A deferred FP64 pivot count of step p.k: before that step is recorded (an early factor's
    // statistics), it replaces the instance's count as the synchronous count would; afterwards it
    // replaces the recorded count.  A count whose factor failed ends the path at that step, where
    // the synchronous count ends it.
    void apply_count(const BatchDevice::Count &c, bool before_record)
    {
        const PendingCount &p = m_pending[(size_t)c.tag];
        Inst &I = *p.I;
        if (c.info) {
            if (before_record || p.k < I.k) {
                I.status = fail(KSH_SOLVER_ERROR, "FP64 refactorization for the pivot count failed");
                I.k = std::min(I.k, p.k);
                I.ph = Phase::Done;
            }
            return;
        }
        if (before_record) {
            I.negatives = (int64_t)c.stats[0];
            I.indicator = c.stats[1];
            return;
        }
        if (p.k >= I.k) return;  // the path was ended before this step
        ksh_alm_result &r = *I.res;
        if (r.negatives) r.negatives[p.k] = (int64_t)c.stats[0];
        if (r.indicator) r.indicator[p.k] = c.stats[1];
    }

    int take_counts(bool wait)
    {
        if (m_pending.empty()) return KSH_SUCCESS;
        std::vector<BatchDevice::Count> done;
        int rc;
        if ((rc = m_dev.recount_take(wait, done))) return rc;
        for (const BatchDevice::Count &c : done) apply_count(c, false);
        return KSH_SUCCESS;
    }

    // the driver's record of converged points (Uold, Lold) and its arc-length reset
    int record_steps(std::vector<Inst *> &rec)
    {
        int rc;
        if (rec.empty()) return KSH_SUCCESS;
        // features of the new states (learned history, or requested by the caller)
        std::vector<FeatDesc> fd;
        std::vector<Inst *> fi;
        for (Inst *I : rec)
            if (m_o.learned || I->res->features) {
                FeatDesc f;
                f.ncp = I->pb->ncp;
                f.nfree = I->n;
                f.P = I->pb->modal_P;
                f.zfree = I->pb->modal_zfree;
                f.U = I->v[V_U];
                f.lambda = I->L;
                fd.push_back(f);
                fi.push_back(I);
            }
        std::vector<double> feats(fd.size() * kFeatures);
        if (!fd.empty()) {
            if ((rc = m_feat.ensure(feats.size() * sizeof(double)))) return rc;
            for (size_t i = 0; i < fd.size(); ++i) fd[i].out = (double *)m_feat.p + i * kFeatures;
            if ((rc = batch_features(m_dev.stream(), m_featdesc, fd))) return rc;
            KSH_HIP(hipMemcpyAsync(feats.data(), m_feat.p, feats.size() * sizeof(double), hipMemcpyDeviceToHost, m_dev.stream()));
        }
        if ((rc = m_dev.synchronize())) return rc;
        for (Inst *I : rec) {
            ksh_alm_result &r = *I->res;
            const int k = I->k;
            if (r.lambda) r.lambda[k] = I->L;
            if (r.unorm) r.unorm[k] = I->unorm_new;
            if (r.indicator) r.indicator[k] = I->indicator;
            if (r.negatives) r.negatives[k] = I->negatives;
            if (r.iterations) r.iterations[k] = I->rec_it;
            if (r.residue_f) r.residue_f[k] = I->rec_resF;
            if (r.length) r.length[k] = I->rec_arc;
            if (r.singular) r.singular[k] = I->sp_kind;
            if (r.U) KSH_HIP(hipMemcpy(r.U + (size_t)k * I->n, I->v[V_U], sizeof(double) * I->n, hipMemcpyDeviceToHost));
            if (r.w && r.probe_dof >= 0 && r.probe_dof < I->n)
                KSH_HIP(hipMemcpy(&r.w[k], I->v[V_U] + r.probe_dof, sizeof(double) * 1, hipMemcpyDeviceToHost));
            if (m_o.verbose)
                std::printf("[ksh path] step %d: lambda %.10e |U| %.6e iterations %d negatives %lld length %.3e%s\n", k,
                            I->L, I->unorm_new, I->rec_it, (long long)I->negatives, I->rec_arc,
                            I->sp_kind == 1 ? " (limit point)" : I->sp_kind >= 2 ? " (bifurcation point)" : "");
            // the driver: arc length back to its start value after a step that was not halved
            if (!I->bisected) I->arc = I->arc_prev = I->arc0;
            I->bisected = false;
            I->learned_off = false;
            if (I->sp_kind) {  // the path has a kink here: the learned history starts again
                I->hist.clear();
                I->lens.clear();
            }
            I->k = k + 1;
            I->ph = I->k >= m_o.nsteps ? Phase::Done : Phase::Start;
        }
        for (size_t i = 0; i < fi.size(); ++i) {
            Inst &I = *fi[i];
            std::array<double, kFeatures> f;
            std::copy(feats.begin() + i * kFeatures, feats.begin() + (i + 1) * kFeatures, f.begin());
            if (I.res->features) std::copy(f.begin(), f.end(), I.res->features + (size_t)(I.k - 1) * kFeatures);
            if (!I.lens.empty() && I.lens.back() != I.rec_arc) {  // uniform steps only
                I.hist.clear();
                I.lens.clear();
            }
            I.hist.push_back(f);
            I.lens.push_back(I.rec_arc);
            while ((int)I.hist.size() > kHist) I.hist.pop_front();
            while ((int)I.lens.size() > kHist) I.lens.pop_front();
        }
        for (Inst *I : rec) I->sp_kind = 0;
        return KSH_SUCCESS;
    }

    // ---- singular points -------------------------------------------------------------------------
    // computeSingularPoint(Uold, Lold) of the drivers (gsALMBase.hpp:422-446), then switchBranch
    // (gsALMBase.hpp:843-856) with singular_point = 2.  One stage per round for each instance.
    int singular_round(std::vector<Inst *> &sing)
    {
        int rc;
        auto stage = [&](Sp s) {
            std::vector<Inst *> g;
            for (Inst *I : sing)
                if (I->ph == Phase::Singular && I->sp == s) g.push_back(I);
            return g;
        };
        std::vector<Inst *> g;
        if (!(g = stage(Sp::Test)).empty() && (rc = sp_test(g))) return rc;
        if (!(g = stage(Sp::BisectStart)).empty() && (rc = sp_bisect_start(g))) return rc;
        if (!(g = stage(Sp::BisectEval)).empty() && (rc = sp_bisect_eval(g))) return rc;
        if (!(g = stage(Sp::BisectVector)).empty() && (rc = sp_bisect_vector(g))) return rc;
        if (!(g = stage(Sp::ExtStart)).empty() && (rc = sp_ext_start(g))) return rc;
        if (!(g = stage(Sp::ExtIter)).empty() && (rc = sp_ext_iter(g))) return rc;
        if (!(g = stage(Sp::Finish)).empty() && (rc = sp_finish(g))) return rc;
        return KSH_SUCCESS;
    }

    // solves of every job; a failed one ends its instance's path
    int sp_solve(std::vector<SolveJob> &sj)
    {
        int rc;
        if ((rc = m_dev.solve(sj))) return rc;
        for (const SolveJob &q : sj) {
            Inst *I = static_cast<Inst *>(q.tag);
            ++I->solves;
            if (q.rc != KSH_SUCCESS && I->ph != Phase::Done) {
                I->status = q.rc;
                I->ph = Phase::Done;
            }
        }
        return KSH_SUCCESS;
    }

    // tangent + factor of each instance's K at the given vector; a zero or non-finite pivot ends
    // the path (factorizeMatrix throws 3)
    int sp_tangent(std::vector<Inst *> &g, VecId at)
    {
        std::vector<TangentJob> tj;
        for (Inst *I : g) {
            tj.push_back(TangentJob{I->pb, I->K, I->v[at]});
            ++I->jac;
        }
        int rc;
        if ((rc = m_dev.tangent(tj))) return rc;
        for (Inst *I : g)
            if (I->K->info) {
                I->status = fail(KSH_SOLVER_ERROR, "LDL^T at the singular point: zero or non-finite pivot");
                I->ph = Phase::Done;
            }
        return KSH_SUCCESS;
    }

    static void keep_live(std::vector<Inst *> &g)
    {
        std::vector<Inst *> ok;
        for (Inst *I : g)
            if (I->ph != Phase::Done) ok.push_back(I);
        g.swap(ok);
    }

    // V <- K^-1 V / ||K^-1 V|| with each instance's current factor, from the normalised ones
    // vector; `stop`: the bisection's early exit ||V - Vold|| < Tol (gsALMBase.hpp:822-837)
    int sp_power(std::vector<Inst *> &g, int its, bool stop)
    {
        int rc;
        std::vector<VecDesc> vd;
        for (Inst *I : g) {
            VecDesc d;
            d.n = I->n;
            d.add(I->v[V_SPV], 1.0 / std::sqrt((double)I->n), m_ones);
            if (stop) d.add(I->v[V_SPQ], 0.0, m_ones);
            vd.push_back(d);
        }
        if ((rc = m_dev.vecops(vd))) return rc;
        std::vector<Inst *> act = g;
        for (int it = 0; it < its && !act.empty(); ++it) {
            std::vector<SolveJob> sj;
            for (Inst *I : act) sj.push_back(SolveJob{I->K, I->v[V_SPV], I->v[V_T1], m_o.max_refine, m_o.refine_tol, KSH_SUCCESS, I});
            if ((rc = sp_solve(sj))) return rc;
            keep_live(act);
            std::vector<DotDesc> dd;
            for (Inst *I : act) dd.push_back(dot1(*I, I->v[V_T1], I->v[V_T1]));
            std::vector<double> dv;
            if ((rc = m_dev.dots(dd, dv))) return rc;
            vd.clear();
            for (size_t i = 0; i < act.size(); ++i) {
                Inst &I = *act[i];
                VecDesc d;
                d.n = I.n;
                d.add(I.v[V_SPV], 1.0 / std::sqrt(dv[i]), I.v[V_T1]);
                if (stop) d.add(I.v[V_SPP], 1.0, I.v[V_SPV], -1.0, I.v[V_SPQ]);
                vd.push_back(d);
            }
            if ((rc = m_dev.vecops(vd))) return rc;
            if (!stop) continue;
            dd.clear();
            for (Inst *I : act) dd.push_back(dot1(*I, I->v[V_SPP], I->v[V_SPP]));
            if ((rc = m_dev.dots(dd, dv))) return rc;
            std::vector<Inst *> next;
            vd.clear();
            for (size_t i = 0; i < act.size(); ++i) {
                Inst &I = *act[i];
                if (std::sqrt(dv[i]) < m_o.tol) continue;
                VecDesc d;
                d.n = I.n;
                d.add(I.v[V_SPQ], 1.0, I.v[V_SPV]);
                vd.push_back(d);
                next.push_back(&I);
            }
            if ((rc = m_dev.vecops(vd))) return rc;
            act.swap(next);
        }
        return KSH_SUCCESS;
    }

    // _testSingularPoint(jacobian = false) (gsALMBase.hpp:475-511): power iterations with the
    // factor of the step's counted tangent; |V . F| < SingularPointTestTol: a bifurcation point.
    // Then the solution goes back to the driver's previous point (gsALMBase.hpp:427-428).
    int sp_test(std::vector<Inst *> &g)
    {
        int rc;
        if ((rc = sp_power(g, m_sp_test_it, false))) return rc;
        keep_live(g);
        std::vector<DotDesc> dd;
        for (Inst *I : g) dd.push_back(dot1(*I, I->v[V_SPV], I->pb->F));
        std::vector<double> dv;
        if ((rc = m_dev.dots(dd, dv))) return rc;
        std::vector<VecDesc> vd;
        for (size_t i = 0; i < g.size(); ++i) {
            Inst &I = *g[i];
            I.sp_test = std::abs(dv[i]) < m_sp_test_tol;
            VecDesc d;
            d.n = I.n;
            d.add(I.v[V_U], 1.0, I.v[V_SPUOLD]);
            vd.push_back(d);
            I.L = I.sp_Lold;
            I.sp_conv = false;
            I.sp = !I.sp_test ? Sp::Finish : m_o.sp_tol_b != 0 ? Sp::BisectStart : Sp::ExtStart;
            if (m_o.verbose)
                std::printf("[ksh path] singular point test: |V.F| = %.3e -> %s\n", std::abs(dv[i]),
                            I.sp_test ? "bifurcation point" : "limit point");
        }
        return m_dev.vecops(vd);
    }

    // _bisectionSolve (gsALMBase.hpp:754-840): the reference value of the termination function
    // and the pivot count at the start point, then the first step
    int sp_bisect_start(std::vector<Inst *> &g)
    {
        int rc;
        if ((rc = sp_tangent(g, V_U))) return rc;
        keep_live(g);
        std::vector<VecDesc> vd;
        for (Inst *I : g) {
            count_pivots(*I);
            I->stability = I->indicator < 0 ? 1 : -1;
            I->sp_ref = I->indicator;
            I->sp_fa = I->negatives;
            I->sp_Lbu = I->L;
            I->sp_arc = I->arc;
            I->sp_k = 1;
            VecDesc d;
            d.n = I->n;
            d.add(I->v[V_SPBU], 1.0, I->v[V_U]);
            vd.push_back(d);
            I->sp = Sp::BisectStep;
            I->ph = Phase::Start;
        }
        return m_dev.vecops(vd);
    }

    // after a bisection step: the pivot count at the new point (or at the start point when the
    // step failed), halve the arc length when it changed, stop when |min D / min D_0| < TolB
    int sp_bisect_eval(std::vector<Inst *> &g)
    {
        int rc;
        if ((rc = sp_tangent(g, V_U))) return rc;
        keep_live(g);
        std::vector<VecDesc> vd;
        for (Inst *I : g) {
            count_pivots(*I);
            I->stability = I->indicator < 0 ? 1 : -1;
            VecDesc d;
            d.n = I->n;
            if (I->sp_fa == I->negatives && I->sp_step_ok) {
                d.add(I->v[V_SPBU], 1.0, I->v[V_U]);
                I->sp_Lbu = I->L;
            } else {
                I->arc = I->arc / 2.0;
            }
            const double term = I->indicator / I->sp_ref;
            if (std::abs(term) < m_o.sp_tol_b || I->sp_k >= m_o.max_iter - 1) {
                I->sp = Sp::BisectVector;
            } else {
                ++I->sp_k;
                d.add(I->v[V_U], 1.0, I->v[V_SPBU]);
                I->L = I->sp_Lbu;
                I->sp = Sp::BisectStep;
                I->ph = Phase::Start;
            }
            vd.push_back(d);
        }
        return m_dev.vecops(vd);
    }

    // the eigenvector at the bisection's last point, the arc length restored
    int sp_bisect_vector(std::vector<Inst *> &g)
    {
        int rc;
        if ((rc = sp_power(g, 5, true))) return rc;
        for (Inst *I : g) {
            if (I->ph == Phase::Done) continue;
            I->arc = I->sp_arc;
            I->sp = Sp::ExtStart;
        }
        return KSH_SUCCESS;
    }

    // _extendedSystemSolve (gsALMBase.hpp:631-650): K and the residual at the start point, zero
    // increments
    int sp_ext_start(std::vector<Inst *> &g)
    {
        int rc;
        for (Inst *I : g)
            if (!I->Keps) {
                ksh_matrix *K = nullptr;
                if ((rc = ksh_matrix_create(static_cast<ksh_problem *>(I->pb), KSH_FP64, &K))) return rc;
                m_mats.ms.push_back(K);
                I->Keps = K;
            }
        if ((rc = sp_tangent(g, V_U))) return rc;
        keep_live(g);
        std::vector<ResidualJob> rj;
        std::vector<VecDesc> vd;
        for (Inst *I : g) {
            rj.push_back(ResidualJob{I->pb, I->v[V_U], I->L, KSH_RES_ALM, I->v[V_R]});
            ++I->resid;
            VecDesc d;
            d.n = I->n;
            d.add(I->v[V_DU], 0.0, I->v[V_DU]);
            d.add(I->v[V_SPDV], 0.0, I->v[V_SPDV]);
            d.add(I->v[V_X], 1.0, I->v[V_U]);
            d.add(I->v[V_SPW], 1.0, I->v[V_SPV]);
            vd.push_back(d);
            I->sp_DL = 0.0;
            I->sp_k = 0;
            I->sp = Sp::ExtIter;
        }
        if ((rc = m_dev.residual(rj))) return rc;
        return m_dev.vecops(vd);
    }

    // one iteration of _extendedSystemSolve / _extendedSystemIteration (gsALMBase.hpp:651-735):
    // Newton on R(U, L) = 0, K(U) W = 0, ||W|| = 1 with the derivative of K along W by a finite
    // difference, K(X + eps W) - K(X) over eps (eps = 1e-8):
    //   K dUt = F, K dUb = -R,  h1 = (K_eps dUt - F) / eps,  h2 = K W + (K_eps dUb + R) / eps,
    //   K dVt = -h1, K dVb = -h2,  dL = -(e . dVb + |W| - 1) / (e . dVt) with e = W / |W|,
    //   DU += dL dUt + dUb, DV += dL dVt + dVb, DL += dL;
    // converged when ||K(U + DU) (V + DV)|| < SingularPointComputeTolE
    int sp_ext_iter(std::vector<Inst *> &g)
    {
        int rc;
        const double eps = 1e-8, ie = 1.0 / eps;
        std::vector<VecDesc> vd;
        std::vector<SolveJob> sj;
        for (Inst *I : g) {
            VecDesc d;
            d.n = I->n;
            d.add(I->v[V_NR], -1.0, I->v[V_R]);
            d.add(I->v[V_SPY], 1.0, I->v[V_X], eps, I->v[V_SPW]);
            vd.push_back(d);
            sj.push_back(SolveJob{I->K, I->pb->F, I->v[V_UT], m_o.max_refine, m_o.refine_tol, KSH_SUCCESS, I});
            sj.push_back(SolveJob{I->K, I->v[V_NR], I->v[V_UB], m_o.max_refine, m_o.refine_tol, KSH_SUCCESS, I});
        }
        if ((rc = m_dev.vecops(vd)) || (rc = sp_solve(sj))) return rc;
        keep_live(g);
        std::vector<TangentJob> tj;
        for (Inst *I : g) {
            tj.push_back(TangentJob{I->pb, I->Keps, I->v[V_SPY]});
            ++I->jac;
        }
        if ((rc = m_dev.assemble(tj))) return rc;
        std::vector<ProductJob> pj;
        for (Inst *I : g) {  // -K_eps dUt, -K_eps dUb, -K W
            pj.push_back(ProductJob{I->Keps, I->v[V_UT], nullptr, I->v[V_SPP]});
            pj.push_back(ProductJob{I->Keps, I->v[V_UB], nullptr, I->v[V_SPQ]});
            pj.push_back(ProductJob{I->K, I->v[V_SPW], nullptr, I->v[V_T1]});
        }
        std::vector<double> rr;
        if ((rc = m_dev.products(pj, rr))) return rc;
        vd.clear();
        sj.clear();
        for (Inst *I : g) {  // -h1 and -h2
            VecDesc d;
            d.n = I->n;
            d.add(I->v[V_T2], ie, I->pb->F, ie, I->v[V_SPP]);
            d.add(I->v[V_SPY], 1.0, I->v[V_T1], ie, I->v[V_SPQ], -ie, I->v[V_R]);
            vd.push_back(d);
            sj.push_back(SolveJob{I->K, I->v[V_T2], I->v[V_SPVT], m_o.max_refine, m_o.refine_tol, KSH_SUCCESS, I});
            sj.push_back(SolveJob{I->K, I->v[V_SPY], I->v[V_SPVB], m_o.max_refine, m_o.refine_tol, KSH_SUCCESS, I});
        }
        if ((rc = m_dev.vecops(vd)) || (rc = sp_solve(sj))) return rc;
        keep_live(g);
        std::vector<DotDesc> dd;
        for (Inst *I : g) {
            DotDesc d = dot1(*I, I->v[V_SPW], I->v[V_SPW]);
            add_pair(d, I->v[V_SPW], I->v[V_SPVB]);
            add_pair(d, I->v[V_SPW], I->v[V_SPVT]);
            dd.push_back(d);
        }
        std::vector<double> dv;
        if ((rc = m_dev.dots(dd, dv))) return rc;
        vd.clear();
        for (size_t i = 0; i < g.size(); ++i) {
            Inst &I = *g[i];
            const double nw = std::sqrt(dv[3 * i]);
            const double dl = -(dv[3 * i + 1] / nw + nw - 1.0) / (dv[3 * i + 2] / nw);
            I.sp_DL += dl;
            VecDesc d;
            d.n = I.n;
            d.add(I.v[V_DU], 1.0, I.v[V_DU], dl, I.v[V_UT], 1.0, I.v[V_UB]);
            d.add(I.v[V_SPDV], 1.0, I.v[V_SPDV], dl, I.v[V_SPVT], 1.0, I.v[V_SPVB]);
            d.add(I.v[V_X], 1.0, I.v[V_U], 1.0, I.v[V_DU]);
            d.add(I.v[V_SPW], 1.0, I.v[V_SPV], 1.0, I.v[V_SPDV]);
            vd.push_back(d);
        }
        if ((rc = m_dev.vecops(vd))) return rc;
        // K and the residual at the new point (the next iteration's), ||K (V + DV)||
        if ((rc = sp_tangent(g, V_X))) return rc;
        keep_live(g);
        std::vector<ResidualJob> rj;
        pj.clear();
        for (Inst *I : g) {
            rj.push_back(ResidualJob{I->pb, I->v[V_X], I->L + I->sp_DL, KSH_RES_ALM, I->v[V_R]});
            ++I->resid;
            pj.push_back(ProductJob{I->K, I->v[V_SPW], nullptr, I->v[V_T1]});
        }
        if ((rc = m_dev.residual(rj)) || (rc = m_dev.products(pj, rr))) return rc;
        vd.clear();
        for (size_t i = 0; i < g.size(); ++i) {
            Inst &I = *g[i];
            const double res = std::sqrt(rr[i]);
            if (m_o.verbose)
                std::printf("[ksh path] extended iteration %d: ||K phi|| = %.3e, lambda %.10e\n", I.sp_k, res,
                            I.L + I.sp_DL);
            const bool conv = res < m_sp_tol_e;
            const bool last = I.sp_k == m_o.max_iter - 1;
            if (!conv && !last) {
                ++I.sp_k;
                continue;
            }
            I.sp_conv = conv;
            if (conv || m_o.sp_fail == 1) {  // gsALMBase.hpp:670-686
                VecDesc d;
                d.n = I.n;
                d.add(I.v[V_U], 1.0, I.v[V_X]);
                vd.push_back(d);
                I.L += I.sp_DL;
            }
            I.sp = Sp::Finish;
        }
        return m_dev.vecops(vd);
    }

    // the end of computeSingularPoint (m_stability = m_stabilityPrev after a bifurcation point,
    // gsALMBase.hpp:441-444), switchBranch with singular_point = 2 (gsALMBase.hpp:843-856:
    // U += (||V|| / tau) V with V normalised, DUold = 0, DLold = 0), the drivers' new arc length,
    // then the step's record
    int sp_finish(std::vector<Inst *> &g)
    {
        int rc;
        std::vector<DotDesc> dd;
        std::vector<double> dv;
        if (m_o.singular_point == 2) {
            for (Inst *I : g) dd.push_back(dot1(*I, I->v[V_SPV], I->v[V_SPV]));
            if ((rc = m_dev.dots(dd, dv))) return rc;
            std::vector<VecDesc> vd;
            for (size_t i = 0; i < g.size(); ++i) {
                Inst &I = *g[i];
                const double xi = 1.0 / m_sp_tau;  // ||V|| = 1 after the normalisation
                VecDesc d;
                d.n = I.n;
                d.add(I.v[V_SPV], 1.0 / std::sqrt(dv[i]), I.v[V_SPV]);
                d.add(I.v[V_U], 1.0, I.v[V_U], xi, I.v[V_SPV]);
                d.add(I.v[V_DUOLD], 0.0, I.v[V_DUOLD]);
                vd.push_back(d);
                I.DLold = 0.0;
                I.prev_zero = true;
            }
            if ((rc = m_dev.vecops(vd))) return rc;
        }
        std::vector<VecDesc> vd;
        dd.clear();
        for (Inst *I : g) {
            DotDesc d = dot1(*I, I->v[V_U], I->v[V_U]);
            if (m_o.method == 1) {
                VecDesc e;
                e.n = I->n;
                e.add(I->v[V_T1], 1.0, I->v[V_U], -1.0, I->v[V_UPREV]);
                vd.push_back(e);
                add_pair(d, I->v[V_T1], I->v[V_T1]);
            }
            dd.push_back(d);
        }
        if ((rc = m_dev.vecops(vd)) || (rc = m_dev.dots(dd, dv))) return rc;
        size_t o = 0;
        for (Inst *I : g) {
            I->unorm_new = std::sqrt(dv[o++]);
            if (m_o.method == 1) I->prev_dnorm = std::sqrt(dv[o++]);
            if (I->sp_test) I->stability = I->stability_prev;
            I->sp_kind = !I->sp_test ? 1 : I->sp_conv ? 2 : 3;
            if (m_o.length_after_sp > 0) I->arc0 = m_o.length_after_sp;
            I->arc = I->arc_prev = I->arc0;
            I->sp = Sp::None;
        }
        return record_steps(g);
    }

    ksh_alm_options m_o;
    BatchDevice m_dev;
    std::vector<Inst> m_in;
    struct PendingCount {  // a deferred pivot count (the tag is its index)
        Inst *I;
        int k;             // the step it belongs to
    };
    std::vector<PendingCount> m_pending;
    Slab m_slab;
    MatrixOwner m_mats;
    double *m_coef = nullptr;
    double *m_ones = nullptr;  // device ones (singular points)
    DevScratch m_decdesc, m_featdesc, m_feat;
    double m_sp_test_tol = 1e-6, m_sp_tol_e = 1e-10, m_sp_tau = 1e3;
    int m_sp_test_it = 5;
};

}  // namespace
}  // namespace ksh

namespace ksh {
namespace {

// gsStaticNewton::_solveNonlinear (src/gsStaticSolvers/gsStaticNewton.hpp:142-193, _start :274-316)
// under gsControlDisplacement::step (src/gsStaticSolvers/gsControlDisplacement.h:46-60): each load
// step lambda += dlambda starts from the residual at the converged point, then Newton iterations
// K(U + DU) dU = R, DU += dU until ||dU|| / ||DU|| < tol and ||R|| / ||R_0|| < tol.
struct StaticInst {
    Problem *pb = nullptr;
    Matrix *K = nullptr;
    int n = 0;
    ksh_static_result *res = nullptr;
    double *U = nullptr, *DU = nullptr, *dU = nullptr, *R = nullptr, *X = nullptr;
    double L = 0, lam = 0, res_ini = 1, residual = 0;
    int it = 0, k = 0, status = KSH_SUCCESS;
    Phase ph = Phase::Start;
    int64_t jac = 0;
};

class StaticEngine {
public:
    explicit StaticEngine(const ksh_static_options &o) : m_o(o) {}

    int run(ksh_problem *const *pbs, int count, ksh_static_result *res)
    {
        const ksh_static_options &o = m_o;
        if (o.precision != KSH_FP64 && o.precision != KSH_FP32_IR && o.precision != KSH_FP32)
            return fail(KSH_OTHER_ERROR, "unknown precision");
        if (o.nsteps < 0 || o.max_iter < 1) return fail(KSH_OTHER_ERROR, "bad step options");
        int rc;
        if ((rc = m_dev.init())) return rc;
        const double t0 = now_s();
        m_in.resize(count);
        size_t tot = 0;
        for (int i = 0; i < count; ++i) tot += (size_t)pbs[i]->nfree;
        const size_t bytes = tot * 5 * sizeof(double);
        KSH_HIP(hipMalloc(&m_slab.base, bytes));
        KSH_HIP(hipMemsetAsync(m_slab.base, 0, bytes, m_dev.stream()));
        double *p = m_slab.base;
        for (int i = 0; i < count; ++i) {
            StaticInst &I = m_in[i];
            I.pb = pbs[i];
            I.n = pbs[i]->nfree;
            I.res = &res[i];
            double **vs[5] = {&I.U, &I.DU, &I.dU, &I.R, &I.X};
            for (double **v : vs) {
                *v = p;
                p += I.n;
            }
            if (o.inertia_guard != 0) I.pb->inertia_guard = o.inertia_guard > 0 ? o.inertia_guard : 0.0;
            if (o.refine_accept > 0) I.pb->refine_accept = o.refine_accept;
            ksh_matrix *K = nullptr;
            if ((rc = ksh_matrix_create(static_cast<ksh_problem *>(I.pb), o.precision, &K))) return rc;
            m_mats.ms.push_back(K);
            I.K = K;
            if (o.nsteps == 0) I.ph = Phase::Done;
        }
        for (;;) {
            std::vector<StaticInst *> live;
            for (StaticInst &I : m_in)
                if (I.ph != Phase::Done) live.push_back(&I);
            if (live.empty()) break;
            if ((rc = round(live))) return rc;
        }
        if ((rc = m_dev.synchronize())) return rc;
        const double t1 = now_s();
        int worst = KSH_SUCCESS;
        for (StaticInst &I : m_in) {
            I.res->steps_done = I.k;
            I.res->final_status = I.status;
            I.res->wall_seconds = t1 - t0;
            I.res->jacobians = I.jac;
            if (I.status != KSH_SUCCESS && I.status != KSH_NOT_CONVERGED && worst == KSH_SUCCESS) worst = I.status;
        }
        if (worst != KSH_SUCCESS) return worst;
        for (StaticInst &I : m_in)
            if (I.k != m_o.nsteps) return KSH_NOT_CONVERGED;
        return KSH_SUCCESS;
    }

private:
    int round(std::vector<StaticInst *> &live)
    {
        int rc;
        // start of a load step: residual at the converged point (_start without a head start)
        std::vector<StaticInst *> st;
        for (StaticInst *I : live)
            if (I->ph == Phase::Start) st.push_back(I);
        if (!st.empty()) {
            std::vector<VecDesc> vd;
            std::vector<ResidualJob> rj;
            for (StaticInst *I : st) {
                I->lam = I->L + m_o.dlambda;
                I->it = 0;
                VecDesc d;
                d.n = I->n;
                d.add(I->DU, 0.0, nullptr);
                d.add(I->X, 1.0, I->U);
                vd.push_back(d);
                rj.push_back(ResidualJob{I->pb, I->U, I->lam, KSH_RES_STATIC, I->R});
            }
            if ((rc = m_dev.vecops(vd)) || (rc = m_dev.residual(rj))) return rc;
            std::vector<DotDesc> dd;
            for (StaticInst *I : st) {
                DotDesc d;
                d.n = I->n;
                d.npairs = 1;
                d.a[0] = d.b[0] = I->R;
                dd.push_back(d);
            }
            std::vector<double> dv;
            if ((rc = m_dev.dots(dd, dv))) return rc;
            for (size_t i = 0; i < st.size(); ++i) {
                double r = std::sqrt(dv[i]);
                if (!std::isfinite(r)) {
                    st[i]->status = fail(KSH_ASSEMBLY_ERROR, "non-finite residual");
                    st[i]->ph = Phase::Done;
                    continue;
                }
                st[i]->res_ini = r == 0 ? 1.0 : r;  // purely displacement-driven steps divide by 1
                st[i]->ph = Phase::Correct;
            }
        }
        std::vector<StaticInst *> act;
        std::vector<TangentJob> tj;
        for (StaticInst *I : live)
            if (I->ph == Phase::Correct) {
                act.push_back(I);
                tj.push_back(TangentJob{I->pb, I->K, I->X});
                ++I->jac;
            }
        if (act.empty()) return KSH_SUCCESS;
        if ((rc = m_dev.tangent(tj))) return rc;
        std::vector<SolveJob> sj;
        std::vector<StaticInst *> ok;
        for (StaticInst *I : act) {
            if (I->K->info) {
                I->status = fail(KSH_SOLVER_ERROR, "LDL^T: zero or non-finite pivot");
                I->ph = Phase::Done;
                continue;
            }
            ok.push_back(I);
            sj.push_back(SolveJob{I->K, I->R, I->dU, m_o.max_refine, m_o.refine_tol, KSH_SUCCESS, I});
        }
        if ((rc = m_dev.solve(sj))) return rc;
        act.clear();
        for (const SolveJob &q : sj) {
            StaticInst *I = static_cast<StaticInst *>(q.tag);
            if (q.rc != KSH_SUCCESS) {
                I->status = q.rc;
                I->ph = Phase::Done;
            } else {
                act.push_back(I);
            }
        }
        std::vector<VecDesc> vd;
        std::vector<ResidualJob> rj;
        for (StaticInst *I : act) {
            VecDesc d;
            d.n = I->n;
            d.add(I->DU, 1.0, I->DU, 1.0, I->dU);  // Relaxation 1
            d.add(I->X, 1.0, I->U, 1.0, I->DU);
            vd.push_back(d);
            rj.push_back(ResidualJob{I->pb, I->X, I->lam, KSH_RES_STATIC, I->R});
        }
        if ((rc = m_dev.vecops(vd)) || (rc = m_dev.residual(rj))) return rc;
        std::vector<DotDesc> dd;
        for (StaticInst *I : act) {
            DotDesc d;
            d.n = I->n;
            d.npairs = 3;
            d.a[0] = d.b[0] = I->R;
            d.a[1] = d.b[1] = I->dU;
            d.a[2] = d.b[2] = I->DU;
            dd.push_back(d);
        }
        std::vector<double> dv;
        if ((rc = m_dev.dots(dd, dv))) return rc;
        std::vector<VecDesc> fin;
        std::vector<StaticInst *> done;
        for (size_t i = 0; i < act.size(); ++i) {
            StaticInst &I = *act[i];
            const double rn = std::sqrt(dv[3 * i]), dn = std::sqrt(dv[3 * i + 1]), Dn = std::sqrt(dv[3 * i + 2]);
            I.residual = rn;
            if (!std::isfinite(rn)) {
                I.status = fail(KSH_ASSEMBLY_ERROR, "non-finite residual");
                I.ph = Phase::Done;
                continue;
            }
            if (dn / Dn < m_o.tol && rn / I.res_ini < m_o.tol) {
                VecDesc d;
                d.n = I.n;
                d.add(I.U, 1.0, I.X);
                fin.push_back(d);
                done.push_back(&I);
            } else if (I.it + 1 == m_o.max_iter) {  // maximum iterations: NotConverged ends the loading
                I.status = KSH_NOT_CONVERGED;
                I.ph = Phase::Done;
            } else {
                ++I.it;
            }
        }
        if ((rc = m_dev.vecops(fin)) || (rc = m_dev.synchronize())) return rc;
        {  // the recorded pivot count: FP32 counts only under the certificate, else an FP64 factor
            std::vector<Matrix *> cert, recount;
            for (StaticInst *I : done)
                if (I->K->fprec == KSH_FP32 && !I->K->use64 && I->K->guard > 0) cert.push_back(I->K);
            if ((rc = m_dev.power_contraction(cert))) return rc;
            for (Matrix *K : cert)
                if (std::max(K->rho, K->rho_pw) > K->guard) recount.push_back(K);
            if (!recount.empty() && (rc = m_dev.refactor64(recount))) return rc;
        }
        for (StaticInst *I : done) {
            ksh_static_result &r = *I->res;
            const int k = I->k;
            I->L = I->lam;
            if (r.lambda) r.lambda[k] = I->L;
            if (r.iterations) r.iterations[k] = I->it + 1;
            if (r.negatives) r.negatives[k] = (int64_t)I->K->fstats[0];
            if (r.residual) r.residual[k] = I->residual;
            if (r.unorm || r.U) {
                std::vector<double> u(I->n);
                KSH_HIP(hipMemcpy(u.data(), I->U, sizeof(double) * I->n, hipMemcpyDeviceToHost));
                double s = 0;
                for (double x : u) s += x * x;
                if (r.unorm) r.unorm[k] = std::sqrt(s);
                if (r.U) std::memcpy(r.U + (size_t)k * I->n, u.data(), sizeof(double) * I->n);
            }
            if (m_o.verbose)
                std::printf("[ksh static] step %d: lambda %.6f iterations %d residual %.3e\n", k, I->L, I->it + 1, I->residual);
            I->k = k + 1;
            I->ph = I->k >= m_o.nsteps ? Phase::Done : Phase::Start;
        }
        return KSH_SUCCESS;
    }

    ksh_static_options m_o;
    BatchDevice m_dev;
    std::vector<StaticInst> m_in;
    Slab m_slab;
    MatrixOwner m_mats;
};

template <class F>
int guarded(F &&f)
{
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return fail(KSH_OTHER_ERROR, "host memory exhausted");
    } catch (const std::exception &e) {
        return fail(KSH_OTHER_ERROR, e.what());
    } catch (...) {
        return fail(KSH_OTHER_ERROR, "unknown exception in the path engine");
    }
}

}  // namespace
}  // namespace ksh

// benchmarks/benchmark_Roof.cpp:320-512: the arc-length solver's options, initialize(), and the step
// loop with arc-length halving, for one problem (a batch of one)
extern "C" int ksh_run_alm(ksh_problem *pb, const ksh_alm_options *o, ksh_alm_result *res)
{
    if (!pb || !o || !res) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    return ksh::guarded([&] {
        ksh_alm_options one = *o;
        if (one.lengths) {
            one.length = one.lengths[0];
            one.lengths = nullptr;
        }
        ksh::AlmEngine e(one);
        ksh_problem *arr[1] = {pb};
        double wall = 0;
        return e.run(arr, 1, res, &wall);
    });
}

extern "C" int ksh_batch_run_alm(ksh_problem *const *pbs, int count, const ksh_alm_options *o, ksh_alm_result *res,
                                 double *wall_seconds)
{
    if (!pbs || count <= 0 || !o || !res) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    for (int i = 0; i < count; ++i)
        if (!pbs[i]) return ksh::fail(KSH_OTHER_ERROR, "null problem");
    return ksh::guarded([&] {
        ksh::AlmEngine e(*o);
        return e.run(pbs, count, res, wall_seconds);
    });
}

extern "C" int ksh_run_static(ksh_problem *pb, const ksh_static_options *o, ksh_static_result *res)
{
    if (!pb || !o || !res) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    return ksh::guarded([&] {
        ksh::StaticEngine e(*o);
        ksh_problem *arr[1] = {pb};
        return e.run(arr, 1, res);
    });
}
