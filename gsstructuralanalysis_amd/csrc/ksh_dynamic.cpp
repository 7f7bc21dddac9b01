// ksh_dynamic.cpp -- the residual-only and mass consumers of the shell operators: dynamic
// relaxation (explicit, residual evaluations only) and nonlinear Newmark time integration
// (consistent mass, tangent, band factor), both driven from one host thread over any number of
// problems with every vector in device memory.
//
// Dynamic relaxation (gsStaticDR, src/gsStaticSolvers/gsStaticDR.hpp:93-266).  A round is one
// explicit step of every live instance: the static residual F - F_int(U + DU) of all of them in
// one batched assembly, then one fused update kernel (damping term, velocities, increments and
// the four norms the stopping test and the kinetic-damping peak test read); instances at a
// kinetic-energy peak take a second residual and the peak update in the same round.
//
// Newmark (gsDynamicNewmark<T, true>, src/gsDynamicSolvers/gsDynamicNewmark.hpp:69-142).  A round
// is one Newton iteration of every live instance: K(U) assembled in place, the iteration matrix
// beta dt^2 K + (1 + gamma dt c) M formed over the band storage, factored, and solved for the
// acceleration increment; the residual F - F_int(U) - M (A + c V) with its norm closes the round.
// Instances that converge record their step and start the next one with the predictor.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ksh_batch.hpp"
#include "ksh_internal.hpp"
#include "ksh_profile.hpp"

namespace ksh {
namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct DevBlock {  // one device allocation, freed with the engine
    double *p = nullptr;
    ~DevBlock()
    {
        if (p) (void)hipFree(p);
    }
};

struct Matrices {
    std::vector<ksh_matrix *> ms;
    ~Matrices()
    {
        for (ksh_matrix *K : ms)
            if (K) ksh_matrix_destroy(K);
    }
    int create(Problem *pb, int prec, Matrix **out)
    {
        ksh_matrix *K = nullptr;
        int rc;
        if ((rc = ksh_matrix_create(static_cast<ksh_problem *>(pb), prec, &K))) return rc;
        ms.push_back(K);
        *out = K;
        return KSH_SUCCESS;
    }
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
        return fail(KSH_OTHER_ERROR, "unknown exception in the dynamic solvers");
    }
}

// consistent mass of every problem (assembled on its own stream) and, with `lumped`, its row
// sums m = M 1 (assembleMass(true)) into m_out[i] on the engine stream
int assemble_masses(BatchDevice &dev, Matrices &mats, std::vector<Problem *> &pbs, std::vector<Matrix *> &M,
                    const std::vector<double *> *lumped, double *ones)
{
    int rc;
    M.resize(pbs.size());
    for (size_t i = 0; i < pbs.size(); ++i) {
        if ((rc = mats.create(pbs[i], KSH_FP64, &M[i]))) return rc;
        if ((rc = ksh_assemble_mass(static_cast<ksh_problem *>(pbs[i]), static_cast<ksh_matrix *>(M[i])))) return rc;
        KSH_HIP(hipStreamSynchronize(pbs[i]->stream));
    }
    if (!lumped) return KSH_SUCCESS;
    std::vector<ProductJob> pj;
    for (size_t i = 0; i < pbs.size(); ++i) pj.push_back(ProductJob{M[i], ones, nullptr, (*lumped)[i]});
    std::vector<double> rr;
    if ((rc = dev.products(pj, rr))) return rc;  // -M 1
    std::vector<VecDesc> vd;
    for (size_t i = 0; i < pbs.size(); ++i) {
        VecDesc d;
        d.n = pbs[i]->nfree;
        d.add((*lumped)[i], -1.0, (*lumped)[i]);
        vd.push_back(d);
    }
    return dev.vecops(vd);
}

// ---------------------------------------------------------------------------------------------
class DrEngine {
public:
    explicit DrEngine(const ksh_dr_options &o) : m_o(o) {}

    int run(ksh_problem *const *pbs, int count, ksh_dr_result *res, double *wall)
    {
        const ksh_dr_options &o = m_o;
        if (o.max_iter < 2 || !(o.alpha > 0) || o.damping < 0 || !(o.tol > 0))
            return fail(KSH_OTHER_ERROR, "bad dynamic relaxation options");
        m_tolF = o.tol_f != -1 ? o.tol_f : o.tol;
        m_tolU = o.tol_u != -1 ? o.tol_u : o.tol;
        int rc;
        if ((rc = m_dev.init())) return rc;
        const double t0 = now_s();
        // per instance: U, DU, dU, v, R, X, m, minv, damp
        constexpr int kVec = 9;
        size_t tot = 0, nmax = 0;
        for (int i = 0; i < count; ++i) {
            tot += (size_t)pbs[i]->nfree;
            nmax = std::max(nmax, (size_t)pbs[i]->nfree);
        }
        const size_t bytes = (tot * kVec + nmax + 4 * (size_t)count) * sizeof(double);
        KSH_HIP(hipMalloc(&m_mem.p, bytes));
        KSH_HIP(hipMemsetAsync(m_mem.p, 0, bytes, m_dev.stream()));
        double *p = m_mem.p;
        m_in.resize(count);
        std::vector<Problem *> probs;
        std::vector<double *> lumped;
        for (int i = 0; i < count; ++i) {
            Inst &I = m_in[i];
            I.pb = pbs[i];
            I.n = pbs[i]->nfree;
            I.res = &res[i];
            double **vs[kVec] = {&I.U, &I.DU, &I.dU, &I.v, &I.R, &I.X, &I.m, &I.minv, &I.damp};
            for (double **v : vs) {
                *v = p;
                p += I.n;
            }
            probs.push_back(I.pb);
            lumped.push_back(I.m);
        }
        double *ones = p;
        p += nmax;
        m_sums = p;
        {
            std::vector<double> one(nmax, 1.0);
            KSH_HIP(hipMemcpyAsync(ones, one.data(), nmax * sizeof(double), hipMemcpyHostToDevice, m_dev.stream()));
            KSH_HIP(hipStreamSynchronize(m_dev.stream()));
        }
        // gsStaticDR::initialize (gsStaticDR.hpp:139-145): massInv = 1 / (alpha m), damp = c m
        std::vector<Matrix *> M;
        if ((rc = assemble_masses(m_dev, m_mats, probs, M, &lumped, ones))) return rc;
        if ((rc = m_dev.synchronize())) return rc;
        for (Inst &I : m_in) {
            std::vector<double> m(I.n), minv(I.n), damp(I.n);
            KSH_HIP(hipMemcpy(m.data(), I.m, sizeof(double) * I.n, hipMemcpyDeviceToHost));
            for (int k = 0; k < I.n; ++k) {
                if (!(m[k] > 0)) return fail(KSH_OTHER_ERROR, "lumped mass must be positive (density, thickness)");
                minv[k] = 1.0 / m[k] / o.alpha;
                damp[k] = o.damping * m[k];
            }
            KSH_HIP(hipMemcpy(I.minv, minv.data(), sizeof(double) * I.n, hipMemcpyHostToDevice));
            KSH_HIP(hipMemcpy(I.damp, damp.data(), sizeof(double) * I.n, hipMemcpyHostToDevice));
        }
        // ||F|| (the reference's initial residual norm, gsStaticDR.hpp:229-236)
        {
            std::vector<DotDesc> dd;
            for (size_t i = 0; i < m_in.size(); ++i) {
                DotDesc d;
                d.n = m_in[i].n;
                d.npairs = 1;
                d.a[0] = d.b[0] = m_in[i].pb->F;
                dd.push_back(d);
            }
            std::vector<double> ff;
            if ((rc = m_dev.dots(dd, ff))) return rc;
            for (size_t i = 0; i < m_in.size(); ++i) {
                const double f = std::sqrt(ff[i]);
                m_in[i].residual = m_in[i].residual_ini = f != 0 ? f : 1.0;
            }
        }
        // _start: residual at U, the half step
        std::vector<Inst *> all;
        for (Inst &I : m_in) all.push_back(&I);
        if ((rc = residuals(all)) || (rc = update(all, 0))) return rc;
        for (Inst *I : all) {
            I->Ek = I->Ek0 = I->s[1];
            if (I->res->ek_history) I->res->ek_history[0] = I->Ek;
            I->it = 1;
            I->reset = 0;
        }
        for (;;) {
            std::vector<Inst *> live;
            for (Inst &I : m_in)
                if (!I.done) live.push_back(&I);
            if (live.empty()) break;
            if ((rc = round(live))) return rc;
        }
        if ((rc = m_dev.synchronize())) return rc;
        const double t1 = now_s();
        if (wall) *wall = t1 - t0;
        int worst = KSH_SUCCESS;
        for (Inst &I : m_in) {
            ksh_dr_result &r = *I.res;
            r.iterations = I.it;
            r.final_status = I.status;
            r.peaks = I.peaks;
            r.residual = I.residual / I.residual_ini;
            r.ek = I.Ek;
            r.ek0 = I.Ek0;
            r.wall_seconds = t1 - t0;
            r.residuals = I.nres;
            if (r.U) KSH_HIP(hipMemcpy(r.U, I.U, sizeof(double) * I.n, hipMemcpyDeviceToHost));
            if (I.status != KSH_SUCCESS && worst == KSH_SUCCESS) worst = I.status;
        }
        return worst;
    }

private:
    struct Inst {
        Problem *pb = nullptr;
        int n = 0;
        ksh_dr_result *res = nullptr;
        double *U = nullptr, *DU = nullptr, *dU = nullptr, *v = nullptr, *R = nullptr, *X = nullptr;
        double *m = nullptr, *minv = nullptr, *damp = nullptr;
        double s[4] = {0, 0, 0, 0};
        double Ek = 0, Ek0 = 0, residual = 0, residual_ini = 0;
        int it = 0, reset = 0, peaks = 0, status = KSH_SUCCESS;
        int64_t nres = 0;
        bool done = false;
    };

    // static residual F - F_int(X) of every instance, X = U + DU
    int residuals(std::vector<Inst *> &g)
    {
        std::vector<ResidualJob> rj;
        for (Inst *I : g) {
            rj.push_back(ResidualJob{I->pb, I->X, 1.0, KSH_RES_STATIC, I->R});
            ++I->nres;
        }
        return m_dev.residual(rj);
    }

    // the fused update of every instance in `mode`, norms to the host
    int update(std::vector<Inst *> &g, int mode)
    {
        std::vector<DrDesc> h;
        for (size_t i = 0; i < g.size(); ++i) {
            Inst &I = *g[i];
            DrDesc d;
            d.n = I.n;
            d.mode = mode;
            d.dt = 1.0;
            d.R = I.R;
            d.m = I.m;
            d.minv = I.minv;
            d.damp = I.damp;
            d.U = I.U;
            d.v = I.v;
            d.dU = I.dU;
            d.DU = I.DU;
            d.X = I.X;
            d.out = m_sums + 4 * i;
            h.push_back(d);
        }
        int rc;
        if ((rc = batch_dr(m_dev.stream(), m_desc, h))) return rc;
        std::vector<double> s(4 * g.size());
        KSH_HIP(hipMemcpyAsync(s.data(), m_sums, s.size() * sizeof(double), hipMemcpyDeviceToHost, m_dev.stream()));
        KSH_HIP(hipStreamSynchronize(m_dev.stream()));
        for (size_t i = 0; i < g.size(); ++i) std::copy(s.begin() + 4 * i, s.begin() + 4 * i + 4, g[i]->s);
        return KSH_SUCCESS;
    }

    // gsStaticDR::_solve loop body (gsStaticDR.hpp:111-136)
    int round(std::vector<Inst *> &live)
    {
        int rc;
        std::vector<double> ek_prev(live.size());
        for (size_t i = 0; i < live.size(); ++i) ek_prev[i] = live[i]->Ek;
        if ((rc = residuals(live)) || (rc = update(live, 1))) return rc;
        std::vector<Inst *> peak;
        for (size_t i = 0; i < live.size(); ++i) {
            Inst &I = *live[i];
            I.residual = std::sqrt(I.s[0]);
            if (I.residual_ini == 0) I.residual_ini = I.residual;
            I.Ek = I.s[1];
            if ((m_o.damping == 0 && ek_prev[i] > I.Ek) || I.reset == m_o.reset_it) peak.push_back(&I);
        }
        if (!peak.empty()) {  // _peak (gsStaticDR.hpp:204-214)
            if ((rc = residuals(peak)) || (rc = update(peak, 2))) return rc;
            for (Inst *I : peak) {
                I->reset = 0;
                I->Ek = 0.0;
                ++I->peaks;
            }
        }
        std::vector<VecDesc> fin;
        for (Inst *I : live) {
            if (I->res->ek_history) I->res->ek_history[I->it] = I->Ek;
            const double dUn = std::sqrt(I->s[2]), DUn = std::sqrt(I->s[3]);
            const bool conv = I->residual / I->residual_ini < m_tolF && I->Ek / I->Ek0 < m_o.tol_e && dUn / DUn < m_tolU;
            if (m_o.verbose > 0 && (I->it % m_o.verbose == 0 || conv))
                std::printf("[ksh dr] iteration %d: |R|/|R0| %.3e Ek/Ek0 %.3e |dU|/|DU| %.3e peaks %d\n", I->it,
                            I->residual / I->residual_ini, I->Ek / I->Ek0, dUn / DUn, I->peaks);
            if (conv || I->it == m_o.max_iter - 1) {
                VecDesc d;
                d.n = I->n;
                d.add(I->U, 1.0, I->U, 1.0, I->DU);
                fin.push_back(d);
                I->status = conv ? KSH_SUCCESS : KSH_NOT_CONVERGED;
                I->done = true;
                continue;
            }
            ++I->it;
            ++I->reset;
        }
        return m_dev.vecops(fin);
    }

    ksh_dr_options m_o;
    double m_tolF = 0, m_tolU = 0;
    BatchDevice m_dev;
    std::vector<Inst> m_in;
    DevBlock m_mem;
    Matrices m_mats;
    double *m_sums = nullptr;
    DevScratch m_desc;
};

// ---------------------------------------------------------------------------------------------
class NewmarkEngine {
public:
    explicit NewmarkEngine(const ksh_newmark_options &o) : m_o(o) {}

    int run(ksh_problem *const *pbs, int count, const double *const *U0, const double *const *V0,
            const double *const *A0, ksh_newmark_result *res, double *wall)
    {
        const ksh_newmark_options &o = m_o;
        if (o.precision != KSH_FP64 && o.precision != KSH_FP32_IR)
            return fail(KSH_OTHER_ERROR, "Newmark: precision KSH_FP64 or KSH_FP32_IR");
        if (o.nsteps < 0 || !(o.dt > 0) || o.max_iter < 1 || !(o.beta > 0) || o.gamma < 0 || o.damping_mass < 0)
            return fail(KSH_OTHER_ERROR, "bad Newmark options");
        int rc;
        if ((rc = m_dev.init())) return rc;
        const double t0 = now_s();
        constexpr int kVec = 7;  // U, V, A, dA, R, rhs, W
        size_t tot = 0;
        for (int i = 0; i < count; ++i) tot += (size_t)pbs[i]->nfree;
        KSH_HIP(hipMalloc(&m_mem.p, tot * kVec * sizeof(double)));
        KSH_HIP(hipMemsetAsync(m_mem.p, 0, tot * kVec * sizeof(double), m_dev.stream()));
        double *p = m_mem.p;
        m_in.resize(count);
        std::vector<Problem *> probs;
        for (int i = 0; i < count; ++i) {
            Inst &I = m_in[i];
            I.pb = pbs[i];
            I.n = pbs[i]->nfree;
            I.res = &res[i];
            double **vs[kVec] = {&I.U, &I.V, &I.A, &I.dA, &I.R, &I.rhs, &I.W};
            for (double **v : vs) {
                *v = p;
                p += I.n;
            }
            if ((rc = m_mats.create(I.pb, o.precision, &I.K))) return rc;
            if (I.K->sprec != KSH_FP64) return fail(KSH_OTHER_ERROR, "Newmark needs FP64 storage of K");
            probs.push_back(I.pb);
            if (o.nsteps == 0) I.done = true;
        }
        if ((rc = m_dev.synchronize())) return rc;
        for (int i = 0; i < count; ++i) {
            Inst &I = m_in[i];
            const double *src[3] = {U0 ? U0[i] : nullptr, V0 ? V0[i] : nullptr, A0 ? A0[i] : nullptr};
            double *dst[3] = {I.U, I.V, I.A};
            for (int q = 0; q < 3; ++q)
                if (src[q]) KSH_HIP(hipMemcpy(dst[q], src[q], sizeof(double) * I.n, hipMemcpyHostToDevice));
        }
        std::vector<Matrix *> M;
        if ((rc = assemble_masses(m_dev, m_mats, probs, M, nullptr, nullptr))) return rc;
        for (int i = 0; i < count; ++i) m_in[i].M = M[i];
        for (;;) {
            std::vector<Inst *> live;
            for (Inst &I : m_in)
                if (!I.done) live.push_back(&I);
            if (live.empty()) break;
            if ((rc = round(live))) return rc;
        }
        if ((rc = m_dev.synchronize())) return rc;
        const double t1 = now_s();
        if (wall) *wall = t1 - t0;
        int worst = KSH_SUCCESS;
        for (Inst &I : m_in) {
            ksh_newmark_result &r = *I.res;
            r.steps_done = I.k;
            r.final_status = I.status;
            r.wall_seconds = t1 - t0;
            r.jacobians = I.jac;
            if (I.status != KSH_SUCCESS && worst == KSH_SUCCESS) worst = I.status;
        }
        return worst;
    }

private:
    struct Inst {
        Problem *pb = nullptr;
        int n = 0;
        ksh_newmark_result *res = nullptr;
        Matrix *K = nullptr, *M = nullptr;
        double *U = nullptr, *V = nullptr, *A = nullptr, *dA = nullptr, *R = nullptr, *rhs = nullptr, *W = nullptr;
        double r0 = 1.0, t = 0.0;
        int it = 0, k = 0, status = KSH_SUCCESS;
        bool started = false, done = false;
        int64_t jac = 0;
    };

    // R = F - F_int(U), rhs = R - M (A + c V) and ||rhs||^2 for every instance (W = A + c V set)
    int out_of_balance(std::vector<Inst *> &g, std::vector<double> &rr)
    {
        std::vector<ResidualJob> rj;
        std::vector<ProductJob> pj;
        for (Inst *I : g) {
            rj.push_back(ResidualJob{I->pb, I->U, 1.0, KSH_RES_STATIC, I->R});
            pj.push_back(ProductJob{I->M, I->W, I->R, I->rhs});
        }
        int rc;
        if ((rc = m_dev.residual(rj))) return rc;
        return m_dev.products(pj, rr);
    }

    int round(std::vector<Inst *> &live)
    {
        const ksh_newmark_options &o = m_o;
        const double dt = o.dt, b = o.beta, gm = o.gamma, c = o.damping_mass;
        int rc;
        // step starts: the predictor with A = 0 (gsDynamicNewmark.hpp:81-89) and its residual
        std::vector<Inst *> st;
        std::vector<VecDesc> vd;
        for (Inst *I : live)
            if (!I->started) {
                VecDesc d;
                d.n = I->n;
                d.add(I->U, 1.0, I->U, dt, I->V, (0.5 - b) * dt * dt, I->A);
                d.add(I->V, 1.0, I->V, (1.0 - gm) * dt, I->A);
                d.add(I->A, 0.0, I->A);
                d.add(I->W, c, I->V);
                vd.push_back(d);
                st.push_back(I);
            }
        if (!st.empty()) {
            std::vector<double> rr;
            if ((rc = m_dev.vecops(vd)) || (rc = out_of_balance(st, rr))) return rc;
            for (size_t i = 0; i < st.size(); ++i) {
                const double r = std::sqrt(rr[i]);
                st[i]->r0 = r != 0 ? r : 1.0;
                st[i]->it = 0;
                st[i]->started = true;
            }
        }
        // Newton iteration (gsDynamicNewmark.hpp:104-138): K(U), the iteration matrix, its factor
        std::vector<TangentJob> tj;
        for (Inst *I : live) {
            tj.push_back(TangentJob{I->pb, I->K, I->U});
            ++I->jac;
        }
        if ((rc = m_dev.assemble(tj))) return rc;
        vd.clear();
        std::vector<Matrix *> ms;
        for (Inst *I : live) {
            VecDesc d;
            const size_t cnt = (size_t)I->K->n * I->K->ld;
            if (cnt > (size_t)0x7fffffff) return fail(KSH_OTHER_ERROR, "band too large for the combination kernel");
            d.n = (int)cnt;
            d.add((double *)I->K->ab, b * dt * dt, (const double *)I->K->ab, 1.0 + gm * dt * c, (const double *)I->M->ab);
            vd.push_back(d);
            ms.push_back(I->K);
        }
        if ((rc = m_dev.vecops(vd)) || (rc = m_dev.factor(ms))) return rc;
        std::vector<Inst *> act;
        for (Inst *I : live) {
            if (I->K->info) {
                I->status = fail(KSH_SOLVER_ERROR, "LDL^T of the Newmark iteration matrix: zero or non-finite pivot");
                I->done = true;
            } else {
                act.push_back(I);
            }
        }
        std::vector<SolveJob> sj;
        for (Inst *I : act) sj.push_back(SolveJob{I->K, I->rhs, I->dA, o.max_refine, o.refine_tol, KSH_SUCCESS, I});
        if ((rc = m_dev.solve(sj))) return rc;
        for (const SolveJob &q : sj)
            if (q.rc != KSH_SUCCESS) {
                Inst *I = static_cast<Inst *>(q.tag);
                I->status = q.rc;
                I->done = true;
            }
        {
            std::vector<Inst *> ok;
            for (Inst *I : act)
                if (!I->done) ok.push_back(I);
            act.swap(ok);
        }
        vd.clear();
        std::vector<DotDesc> dd;
        for (Inst *I : act) {
            VecDesc d;
            d.n = I->n;
            d.add(I->A, 1.0, I->A, 1.0, I->dA);
            d.add(I->V, 1.0, I->V, gm * dt, I->dA);
            d.add(I->U, 1.0, I->U, b * dt * dt, I->dA);
            d.add(I->W, 1.0, I->A, c, I->V);
            vd.push_back(d);
            DotDesc e;
            e.n = I->n;
            e.npairs = 2;
            e.a[0] = e.b[0] = I->A;
            e.a[1] = e.b[1] = I->dA;
            dd.push_back(e);
        }
        std::vector<double> dv, rr;
        if ((rc = m_dev.vecops(vd)) || (rc = m_dev.dots(dd, dv)) || (rc = out_of_balance(act, rr))) return rc;
        for (size_t i = 0; i < act.size(); ++i) {
            Inst &I = *act[i];
            const double an = std::sqrt(dv[2 * i]), dan = std::sqrt(dv[2 * i + 1]);
            const double upd = an != 0 ? dan / an : dan;
            const double rn = std::sqrt(rr[i]) / I.r0;
            if (o.verbose)
                std::printf("[ksh newmark] step %d iteration %d: |R|/|R0| %.3e |dA|/|A| %.3e\n", I.k, I.it, rn, upd);
            if (upd < o.tol_u && rn < o.tol_f) {
                if ((rc = record(I))) return rc;
            } else if (I.it == o.max_iter - 1) {
                I.status = KSH_NOT_CONVERGED;
                I.done = true;
            } else {
                ++I.it;
            }
        }
        return KSH_SUCCESS;
    }

    int record(Inst &I)
    {
        ksh_newmark_result &r = *I.res;
        const int k = I.k;
        I.t += m_o.dt;
        const size_t n = (size_t)I.n;
        int rc;
        if ((rc = m_dev.synchronize())) return rc;
        if (r.time) r.time[k] = I.t;
        if (r.iterations) r.iterations[k] = I.it + 1;
        if (r.U) KSH_HIP(hipMemcpy(r.U + k * n, I.U, n * sizeof(double), hipMemcpyDeviceToHost));
        if (r.V) KSH_HIP(hipMemcpy(r.V + k * n, I.V, n * sizeof(double), hipMemcpyDeviceToHost));
        if (r.A) KSH_HIP(hipMemcpy(r.A + k * n, I.A, n * sizeof(double), hipMemcpyDeviceToHost));
        if (r.w && r.probe_dof >= 0 && r.probe_dof < I.n)
            KSH_HIP(hipMemcpy(&r.w[k], I.U + r.probe_dof, sizeof(double), hipMemcpyDeviceToHost));
        if (r.unorm) {
            DotDesc d;
            d.n = I.n;
            d.npairs = 1;
            d.a[0] = d.b[0] = I.U;
            std::vector<double> uu;
            if ((rc = m_dev.dots({d}, uu))) return rc;
            r.unorm[k] = std::sqrt(uu[0]);
        }
        I.k = k + 1;
        I.started = false;
        if (I.k >= m_o.nsteps) I.done = true;
        return KSH_SUCCESS;
    }

    ksh_newmark_options m_o;
    BatchDevice m_dev;
    std::vector<Inst> m_in;
    DevBlock m_mem;
    Matrices m_mats;
};

}  // namespace
}  // namespace ksh

extern "C" int ksh_batch_run_dr(ksh_problem *const *pbs, int count, const ksh_dr_options *o, ksh_dr_result *res,
                                double *wall_seconds)
{
    if (!pbs || count <= 0 || !o || !res) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    for (int i = 0; i < count; ++i)
        if (!pbs[i]) return ksh::fail(KSH_OTHER_ERROR, "null problem");
    return ksh::guarded([&] {
        ksh::DrEngine e(*o);
        return e.run(pbs, count, res, wall_seconds);
    });
}

// solvers/static_shell_XML.cpp:243-275: the lumped mass, gsStaticDR(M, F, Residual), solve()
extern "C" int ksh_run_dr(ksh_problem *pb, const ksh_dr_options *o, ksh_dr_result *res)
{
    if (!pb) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    ksh_problem *arr[1] = {pb};
    double wall = 0;
    return ksh_batch_run_dr(arr, 1, o, res, &wall);
}

extern "C" int ksh_run_newmark(ksh_problem *pb, const ksh_newmark_options *o, const double *U0, const double *V0,
                               const double *A0, ksh_newmark_result *res)
{
    if (!pb || !o || !res) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    return ksh::guarded([&] {
        ksh::NewmarkEngine e(*o);
        ksh_problem *arr[1] = {pb};
        const double *u[1] = {U0}, *v[1] = {V0}, *a[1] = {A0};
        double wall = 0;
        return e.run(arr, 1, u, v, a, res, &wall);
    });
}

extern "C" int ksh_batch_run_newmark(ksh_problem *const *pbs, int count, const ksh_newmark_options *o,
                                     ksh_newmark_result *res, double *wall_seconds)
{
    if (!pbs || count <= 0 || !o || !res) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    for (int i = 0; i < count; ++i)
        if (!pbs[i]) return ksh::fail(KSH_OTHER_ERROR, "null problem");
    return ksh::guarded([&] {
        ksh::NewmarkEngine e(*o);
        return e.run(pbs, count, nullptr, nullptr, nullptr, res, wall_seconds);
    });
}
