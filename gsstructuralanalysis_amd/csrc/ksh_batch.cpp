// ksh_batch.cpp -- batch engine: many independent shell problems run through the unchanged solver
// classes (one host thread per problem) while their device calls are merged into batched launches.
//
// Each problem's gsALMCrisfield / gsALMRiks runs in its own host thread on a gsShellOps whose
// executor is an engine proxy: every operator-surface call (tangent, residual, factor, solve,
// inertia) posts a request and blocks.  When every live thread is blocked, the engine thread
// takes the pending requests, groups them by kind and runs each group as one launch over all the
// problems in it (element assembly over the union of all elements, one factorization workgroup
// per band, one solve workgroup per band, batched refinement), then releases the threads.  The
// solver logic per instance is exactly the single-problem logic; only the device work is merged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "gs/gsALMCrisfield.h"
#include "gs/gsALMRiks.h"
#include "gs/ksh_ops.h"
#include "ksh_internal.hpp"
#include "ksh_profile.hpp"

namespace ksh {

int alm_driver(ksh_problem *pb, gismo::gsExec *exec, const ksh_alm_options *o, ksh_alm_result *res);

namespace {

enum ReqKind { RQ_TANGENT = 0, RQ_RESIDUAL, RQ_FACTOR, RQ_SOLVE, RQ_INERTIA, RQ_PREDICT, RQ_KINDS };

struct Req {
    ReqKind kind;
    ksh_problem *pb = nullptr;
    ksh_matrix *K = nullptr;
    const double *in = nullptr;
    double *out = nullptr;
    double lam = 0.0;
    int conv = 0;
    int max_refine = 0;
    double tol = 0.0;
    int64_t *neg = nullptr;
    double *dmin = nullptr;
    ksh_mlp *net = nullptr;
    int rc = 0;
    bool done = false;
    int waited = 0;  // engine rounds this request has been left pending
    // the instance waits on its own request: the engine wakes each one separately, instead of
    // every instance thread queueing for one mutex after a broadcast
    std::mutex mu;
    std::condition_variable cv;
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b)
    {
        if (b <= bytes) return KSH_SUCCESS;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        KSH_HIP(hipMalloc(&p, b));
        bytes = b;
        return KSH_SUCCESS;
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b)
    {
        if (b <= bytes) return KSH_SUCCESS;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        KSH_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
        return KSH_SUCCESS;
    }
    ~HostBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};

class Engine {
public:
    explicit Engine(int nthreads) : m_active(nthreads) {}
    ~Engine()
    {
        if (m_early_stream) {
            (void)hipStreamSynchronize(m_early_stream);
            (void)hipStreamDestroy(m_early_stream);
        }
        if (m_ev_copied) (void)hipEventDestroy(m_ev_copied);
        if (m_ev_early) (void)hipEventDestroy(m_ev_early);
        if (m_stream) (void)hipStreamDestroy(m_stream);
    }
    int init()
    {
        KSH_HIP(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking));
        const char *e = std::getenv("KSH_EARLY64");
        m_early_on = !(e && std::strcmp(e, "0") == 0);
        if (m_early_on) {
            KSH_HIP(hipStreamCreateWithFlags(&m_early_stream, hipStreamNonBlocking));
            KSH_HIP(hipEventCreateWithFlags(&m_ev_copied, hipEventDisableTiming));
            KSH_HIP(hipEventCreateWithFlags(&m_ev_early, hipEventDisableTiming));
        }
        return KSH_SUCCESS;
    }

    // instance side --------------------------------------------------------------------------
    int submit(Req &r)
    {
        {
            std::lock_guard<std::mutex> lk(m_mu);
            m_pending.push_back(&r);
            ++m_blocked;
            if (m_blocked == m_active) m_cv_main.notify_one();  // the engine waits for all of them
        }
        std::unique_lock<std::mutex> lk(r.mu);
        r.cv.wait(lk, [&] { return r.done; });
        return r.rc;
    }
    void finished()
    {
        std::lock_guard<std::mutex> lk(m_mu);
        --m_active;
        if (m_blocked == m_active) m_cv_main.notify_one();
    }

    // engine side ------------------------------------------------------------------------------
    int run()
    {
        int first_error = KSH_SUCCESS;
        for (;;) {
            std::vector<Req *> batch;
            const auto t0 = std::chrono::steady_clock::now();
            {
                std::unique_lock<std::mutex> lk(m_mu);
                m_cv_main.wait(lk, [&] { return m_active == 0 || (m_blocked == m_active && !m_pending.empty()); });
                if (m_active == 0 && m_pending.empty()) break;
                select(batch);
            }
            const auto t1 = std::chrono::steady_clock::now();
            int rc = execute(batch);
            const auto t2 = std::chrono::steady_clock::now();
            // host time between rounds (instances running their solver code and submitting) and
            // time of the round itself, in microseconds
            prof_count("engine_rounds", 1);
            prof_count("engine_wait_us", (long long)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
            prof_count("engine_round_us", (long long)std::chrono::duration_cast<std::chrono::microseconds>(t2 - t1).count());
            if (rc && !first_error) first_error = rc;
            {
                std::lock_guard<std::mutex> lk(m_mu);
                m_blocked -= (int)batch.size();
                ++m_rounds;
            }
            for (Req *r : batch) {
                // notified under the request's lock: the instance cannot return (and release
                // its request) before the engine is done with it
                std::lock_guard<std::mutex> lk(r->mu);
                if (rc && r->rc == KSH_SUCCESS) r->rc = rc;
                r->done = true;
                r->cv.notify_one();
            }
        }
        return first_error;
    }
    long long rounds() const { return m_rounds; }

private:
    // Requests of one kind per round: the kind most instances are waiting for (inertia queries,
    // answered from the factor's pivot statistics, always go along).  Instances that want another
    // kind wait, so the instances fall into step and each launch carries most of the batch; a
    // request left waiting kMaxWait rounds has its kind taken next.
    void select(std::vector<Req *> &batch)
    {
        constexpr int kMaxWait = 16;
        int cnt[RQ_KINDS] = {};
        int forced = -1, oldest = -1;
        for (Req *r : m_pending) {
            ++cnt[r->kind];
            if (r->waited >= kMaxWait && r->waited > oldest) {
                oldest = r->waited;
                forced = r->kind;
            }
        }
        int kind = forced;
        if (kind < 0) {
            kind = RQ_TANGENT;
            for (int k = 0; k < RQ_KINDS; ++k)
                if (k != RQ_INERTIA && cnt[k] > cnt[kind]) kind = k;
            if (cnt[kind] == 0) kind = RQ_INERTIA;
        }
        std::vector<Req *> keep;
        for (Req *r : m_pending) {
            if (r->kind == kind || r->kind == RQ_INERTIA) {
                batch.push_back(r);
            } else {
                ++r->waited;
                keep.push_back(r);
            }
        }
        m_pending.swap(keep);
    }

    int execute(std::vector<Req *> &batch)
    {
        std::vector<Req *> by[RQ_KINDS];
        for (Req *r : batch) by[r->kind].push_back(r);
        int rc;
        if (!by[RQ_TANGENT].empty() && (rc = do_tangent(by[RQ_TANGENT]))) return rc;
        if (!by[RQ_RESIDUAL].empty() && (rc = do_residual(by[RQ_RESIDUAL]))) return rc;
        if (!by[RQ_FACTOR].empty() && (rc = do_factor(by[RQ_FACTOR]))) return rc;
        if (!by[RQ_SOLVE].empty() && (rc = do_solve(by[RQ_SOLVE]))) return rc;
        for (Req *r : by[RQ_INERTIA]) r->rc = ksh_inertia(r->K, r->neg, r->dmin);
        if (!by[RQ_PREDICT].empty() && (rc = do_predict(by[RQ_PREDICT]))) return rc;
        KSH_HIP(hipStreamSynchronize(m_stream));
        return KSH_SUCCESS;
    }

    // Stage the host input vectors of the group into one pinned buffer and one device buffer.
    // Each request kind has its own pair: a kind's host-to-device copy may still be queued when
    // the next kind of the same round stages its inputs (a round ends with a stream sync).
    int stage_in(std::vector<Req *> &g, std::vector<size_t> &off, bool need_out, int slot)
    {
        size_t tot = 0;
        off.resize(g.size());
        for (size_t i = 0; i < g.size(); ++i) {
            off[i] = tot;
            tot += (size_t)g[i]->pb->nfree;
        }
        int rc;
        if ((rc = m_hin[slot].ensure(tot * sizeof(double)))) return rc;
        if ((rc = m_din[slot].ensure(tot * sizeof(double)))) return rc;
        if (need_out) {
            if ((rc = m_hout.ensure(tot * sizeof(double)))) return rc;
            if ((rc = m_dout.ensure(tot * sizeof(double)))) return rc;
        }
        double *h = (double *)m_hin[slot].p;
        for (size_t i = 0; i < g.size(); ++i) std::memcpy(h + off[i], g[i]->in, sizeof(double) * g[i]->pb->nfree);
        KSH_HIP(hipMemcpyAsync(m_din[slot].p, m_hin[slot].p, tot * sizeof(double), hipMemcpyHostToDevice, m_stream));
        m_tot = tot;
        return KSH_SUCCESS;
    }

    int ensure_desc(size_t n)
    {
        int rc;
        if ((rc = m_desc.ensure(n * sizeof(InstDesc)))) return rc;
        if ((rc = m_bdesc.ensure(n * kBandDescBytes))) return rc;
        if ((rc = m_rdesc.ensure(n * sizeof(RefDesc)))) return rc;
        if ((rc = m_stats.ensure(n * 4 * sizeof(double)))) return rc;
        if ((rc = m_info.ensure(n * sizeof(int)))) return rc;
        if ((rc = m_cdesc.ensure(n * sizeof(CopyDesc)))) return rc;
        if ((rc = m_norms.ensure(n * 2 * sizeof(double)))) return rc;
        return KSH_SUCCESS;
    }

    int do_tangent(std::vector<Req *> &g)
    {
        std::vector<size_t> off;
        int rc;
        if ((rc = ensure_desc(g.size())) || (rc = stage_in(g, off, false, 0))) return rc;
        const double *din = (const double *)m_din[0].p;
        // one launch per storage precision (normally a single group)
        for (int sp : {KSH_FP64, KSH_FP32}) {
            std::vector<InstDesc> h;
            for (size_t i = 0; i < g.size(); ++i) {
                Matrix &K = *g[i]->K;
                if (K.sprec != sp) continue;
                InstDesc d = make_desc(*g[i]->pb);
                d.u_free = din + off[i];
                d.ab = K.ab;
                d.sprec = K.sprec;
                d.skip_zero = K.ab_clean ? 1 : 0;
                h.push_back(d);
            }
            if (!h.empty() && (rc = batch_tangent(m_stream, (InstDesc *)m_desc.p, h))) return rc;
        }
        for (Req *r : g) r->K->ab_clean = true;
        std::vector<Matrix *> ms;
        for (size_t i = 0; i < g.size(); ++i) {
            Matrix &K = *g[i]->K;
            if (K.sprec == KSH_FP32) {
                KSH_HIP(hipMemcpyAsync(K.u_at, din + off[i], sizeof(double) * K.n, hipMemcpyDeviceToDevice,
                                       m_stream));
                K.has_u = true;
            }
            ++K.version;
            g[i]->rc = KSH_SUCCESS;
            ms.push_back(&K);
        }
        // The solvers factor every tangent they assemble right after assembling it
        // (gsALMBase::computeJacobian, src/gsALMSolvers/gsALMBase.hpp:186-211), so the factor
        // goes into this round; the instance's factor request then finds the matrix factored
        // and returns without a round of its own.
        return factor_matrices(ms);
    }

    int do_residual(std::vector<Req *> &g)
    {
        std::vector<size_t> off;
        int rc;
        if ((rc = ensure_desc(g.size())) || (rc = stage_in(g, off, true, 1))) return rc;
        std::vector<InstDesc> h;
        for (size_t i = 0; i < g.size(); ++i) {
            InstDesc d = make_desc(*g[i]->pb);
            d.u_free = (const double *)m_din[1].p + off[i];
            d.lambda = g[i]->lam;
            d.mode = (g[i]->conv == KSH_RES_ALM) ? 1 : 2;
            d.R = (double *)m_dout.p + off[i];
            h.push_back(d);
        }
        if ((rc = batch_residual(m_stream, (InstDesc *)m_desc.p, h))) return rc;
        KSH_HIP(hipMemcpyAsync(m_hout.p, m_dout.p, m_tot * sizeof(double), hipMemcpyDeviceToHost, m_stream));
        KSH_HIP(hipStreamSynchronize(m_stream));
        const double *ho = (const double *)m_hout.p;
        for (size_t i = 0; i < g.size(); ++i) {
            int n = g[i]->pb->nfree;
            std::memcpy(g[i]->out, ho + off[i], sizeof(double) * n);
            g[i]->rc = KSH_SUCCESS;
            for (int k = 0; k < n; ++k)
                if (!std::isfinite(ho[off[i] + k])) {
                    g[i]->rc = fail(KSH_ASSEMBLY_ERROR, "non-finite residual");
                    break;
                }
        }
        return KSH_SUCCESS;
    }

    // learned-predictor inference: one batched forward per network over the waiting instances
    int do_predict(std::vector<Req *> &g)
    {
        constexpr int kIn = 264, kOut = 65;
        std::vector<Req *> rest = g;
        while (!rest.empty()) {
            ksh_mlp *net = rest[0]->net;
            std::vector<Req *> grp, other;
            for (Req *r : rest) (r->net == net ? grp : other).push_back(r);
            std::vector<double> X((size_t)grp.size() * kIn), Y((size_t)grp.size() * kOut);
            for (size_t i = 0; i < grp.size(); ++i) std::memcpy(X.data() + i * kIn, grp[i]->in, sizeof(double) * kIn);
            const int rc = ksh_mlp_predict(net, X.data(), (int)grp.size(), Y.data());
            for (size_t i = 0; i < grp.size(); ++i) {
                if (rc == KSH_SUCCESS) std::memcpy(grp[i]->out, Y.data() + i * kOut, sizeof(double) * kOut);
                grp[i]->rc = rc;
            }
            rest.swap(other);
        }
        return KSH_SUCCESS;
    }

    int do_factor(std::vector<Req *> &g)
    {
        std::vector<Matrix *> todo;
        for (Req *r : g) {
            Matrix &K = *r->K;
            if (!(K.factored && K.factored_version == K.version)) todo.push_back(&K);
        }
        int rc;
        if ((rc = factor_matrices(todo))) return rc;
        for (Req *r : g)
            r->rc = r->K->info ? fail(KSH_SOLVER_ERROR, "LDL^T: zero or non-finite pivot") : KSH_SUCCESS;
        return KSH_SUCCESS;
    }

    // LDL^T of the current values of several matrices: one batched copy into the factor storage,
    // one batched factor launch per factor precision, and the FP64 refactorization of the FP32
    // factors whose smallest pivot is too small to trust (one more batched launch).  The
    // outcome is in each matrix (factored_version, info, pivot statistics, use64).
    int factor_matrices(const std::vector<Matrix *> &ms)
    {
        if (ms.empty()) return KSH_SUCCESS;
        int rc;
        if ((rc = retire_early())) return rc;
        if ((rc = ensure_desc(ms.size()))) return rc;
        for (int fp : {KSH_FP64, KSH_FP32}) {
            std::vector<BandDescHost> h;
            std::vector<CopyDesc> cp;
            std::vector<Matrix *> grp;
            for (Matrix *K : ms) {
                if (K->fprec != fp) continue;
                cp.push_back(CopyDesc{K->ab, K->fac, (size_t)K->n * K->ld, K->sprec, fp, 0});
                const size_t k = grp.size();
                h.push_back(BandDescHost{K->fac, K->n, K->bw, (int *)m_info.p + k, (double *)m_stats.p + 4 * k, nullptr, 1});
                grp.push_back(K);
            }
            if (grp.empty()) continue;
            if (fp == KSH_FP32) {
                // a solve of this matrix's previous FP32 factor fell back to FP64: so will one of
                // this factor's, most likely; its FP64 factor runs beside the FP32 one
                std::vector<Matrix *> early;
                for (Matrix *K : grp)
                    if (K->stalled && K->sprec == KSH_FP64) early.push_back(K);
                if ((rc = launch_early(early))) return rc;
            }
            if ((rc = band_copy_batch(m_stream, m_cdesc.p, cp))) return rc;
            if ((rc = band_factor_batch(m_stream, m_bdesc.p, h, fp))) return rc;
            std::vector<int> info(grp.size());
            std::vector<double> st(4 * grp.size());
            KSH_HIP(hipMemcpyAsync(info.data(), m_info.p, sizeof(int) * grp.size(), hipMemcpyDeviceToHost, m_stream));
            KSH_HIP(hipMemcpyAsync(st.data(), m_stats.p, sizeof(double) * 4 * grp.size(), hipMemcpyDeviceToHost, m_stream));
            KSH_HIP(hipStreamSynchronize(m_stream));
            std::vector<Matrix *> re;
            for (size_t k = 0; k < grp.size(); ++k) {
                Matrix &K = *grp[k];
                K.stalled = false;
                K.factored = true;
                K.factored_version = K.version;
                K.use64 = false;
                K.info = info[k];
                for (int q = 0; q < 4; ++q) K.fstats[q] = st[4 * k + q];
                K.has_stats = info[k] == 0;
                // FP32 factors with a pivot too small to trust (or a failed one) are redone in FP64
                if (fp == KSH_FP32 && (info[k] != 0 || !(K.fstats[3] >= K.guard * K.fstats[2]))) re.push_back(&K);
            }
            prof_count("fp64_refactor_guard", (long long)re.size());
            if ((rc = refactor64(re))) return rc;
        }
        return KSH_SUCCESS;
    }

    // FP64 factors of the current values of several matrices (each at most once): the early
    // factor where one was started for this version, else one batched refactorization.
    int refactor64(const std::vector<Matrix *> &ms)
    {
        std::vector<Matrix *> re;
        std::vector<std::pair<Matrix *, long>> taken;
        for (Matrix *K : ms) {
            if (std::find(re.begin(), re.end(), K) != re.end() ||
                std::find_if(taken.begin(), taken.end(), [&](const auto &t) { return t.first == K; }) != taken.end())
                continue;
            const long k = early_index(K);
            if (k >= 0) taken.emplace_back(K, k);
            else re.push_back(K);
        }
        if (!taken.empty()) {
            KSH_HIP(hipEventSynchronize(m_ev_early));
            KSH_HIP(hipStreamWaitEvent(m_stream, m_ev_early, 0));
            const size_t n = m_early.size();
            const int *hi = (const int *)m_hearly.p;
            const double *hs = (const double *)((const char *)m_hearly.p + n * sizeof(int));
            for (const auto &t : taken) {  // as refactor64_batch leaves a matrix
                Matrix &K = *t.first;
                const long k = t.second;
                K.use64 = true;
                K.info = hi[k] == 0 ? 0 : 1;
                for (int q = 0; q < 4; ++q) K.fstats[q] = hs[4 * k + q];
                K.has_stats = hi[k] == 0;
            }
            prof_count("fp64_refactor_early_used", (long long)taken.size());
        }
        return refactor64_batch(re);
    }

    // Early FP64 factors.  A matrix whose previous FP32 factor needed the FP64 fallback in a
    // solve is likely to need it again (the stalls follow the conditioning, which changes
    // slowly along the path), so its FP64 factor is started together with the FP32 one, on a
    // second stream: the two launch chains, each bound by its panel latency, run at once.  A
    // solve that stalls (or a guard refactorization) takes that factor instead of refactoring:
    // the same kernels on the same values, so the same bits, pivot statistics and decisions as
    // the refactorization it replaces; unused early factors are dropped.
    // Only FP64 storage (the copy of ab); FP32 storage re-assembles on the problem's element
    // tiles, which the main stream may be using.  fac64 of these matrices belongs to the second
    // stream until retire_early (start of the next factor call) or until a stall takes it.
    int launch_early(const std::vector<Matrix *> &ms)
    {
        if (!m_early_on || ms.empty()) return KSH_SUCCESS;
        int rc;
        const size_t n = ms.size();
        const size_t hb = n * (sizeof(int) + 4 * sizeof(double));
        if (m_ebdesc.bytes < n * kBandDescBytes || m_einfo.bytes < n * sizeof(int) ||
            m_estats.bytes < n * 4 * sizeof(double) || m_hearly.bytes < hb)
            KSH_HIP(hipStreamSynchronize(m_early_stream));  // buffers about to be replaced
        if ((rc = m_ebdesc.ensure(n * kBandDescBytes)) || (rc = m_einfo.ensure(n * sizeof(int))) ||
            (rc = m_estats.ensure(n * 4 * sizeof(double))) || (rc = m_hearly.ensure(hb)))
            return rc;
        std::vector<CopyDesc> cp;
        std::vector<BandDescHost> h;
        std::vector<Early> entries;  // published only once everything is queued
        for (size_t k = 0; k < n; ++k) {
            Matrix *K = ms[k];
            if ((rc = ensure_fac64(*K))) return rc;
            cp.push_back(CopyDesc{K->ab, K->fac64, (size_t)K->n * K->ld, K->sprec, KSH_FP64, 0});
            h.push_back(BandDescHost{K->fac64, K->n, K->bw, (int *)m_einfo.p + k, (double *)m_estats.p + 4 * k, nullptr, 1});
            entries.push_back(Early{K, K->version});  // the version about to be factored
        }
        if ((rc = band_copy_batch(m_stream, m_cdesc.p, cp))) return rc;
        KSH_HIP(hipEventRecord(m_ev_copied, m_stream));
        KSH_HIP(hipStreamWaitEvent(m_early_stream, m_ev_copied, 0));
        if ((rc = band_factor_batch(m_early_stream, m_ebdesc.p, h, KSH_FP64, "band_ldlt_early_fp64"))) {
            (void)hipStreamSynchronize(m_early_stream);  // nothing of a failed launch left in flight
            return rc;
        }
        char *hp = (char *)m_hearly.p;
        KSH_HIP(hipMemcpyAsync(hp, m_einfo.p, n * sizeof(int), hipMemcpyDeviceToHost, m_early_stream));
        KSH_HIP(hipMemcpyAsync(hp + n * sizeof(int), m_estats.p, n * 4 * sizeof(double), hipMemcpyDeviceToHost,
                               m_early_stream));
        KSH_HIP(hipEventRecord(m_ev_early, m_early_stream));
        for (Matrix *K : ms) {  // the matrix's own copy of the event, for whoever frees or rewrites fac64
            if (!K->early_done) KSH_HIP(hipEventCreateWithFlags(&K->early_done, hipEventDisableTiming));
            KSH_HIP(hipEventRecord(K->early_done, m_early_stream));
        }
        m_early.swap(entries);
        m_early_pending = true;
        prof_count("fp64_refactor_early", (long long)n);
        return KSH_SUCCESS;
    }

    // The main stream waits for the early factors (their fac64 may be written again) and they
    // are forgotten.
    int retire_early()
    {
        if (m_early_pending) KSH_HIP(hipStreamWaitEvent(m_stream, m_ev_early, 0));
        m_early_pending = false;
        m_early.clear();
        return KSH_SUCCESS;
    }

    // Position of K's early factor of its current version, or -1.  Compares addresses only:
    // the entries are dropped at every factor call, and a matrix can only be freed and another
    // created at its address after that.
    long early_index(const Matrix *K) const
    {
        for (size_t k = 0; k < m_early.size(); ++k)
            if (m_early[k].K == K && m_early[k].version == K->factored_version) return (long)k;
        return -1;
    }

    // FP64 refactorization of several matrices whose FP32 factor cannot be trusted, in one
    // batched launch; per matrix the same result as refactor_fp64 (an FP64 factor of the
    // assembled values, re-assembled in FP64 from the stored point for FP32 storage), and the
    // solves and pivot count of the current version use it.
    int refactor64_batch(const std::vector<Matrix *> &ms)
    {
        if (ms.empty()) return KSH_SUCCESS;
        int rc;
        if ((rc = ensure_desc(ms.size()))) return rc;
        std::vector<InstDesc> tang;
        std::vector<CopyDesc> cp;
        for (Matrix *K : ms) {
            if ((rc = ensure_fac64(*K))) return rc;
            if (K->sprec == KSH_FP32 && K->has_u) {
                InstDesc d = make_desc(*K->pb);
                d.u_free = K->u_at;
                d.ab = K->fac64;
                d.sprec = KSH_FP64;
                tang.push_back(d);
            } else {
                cp.push_back(CopyDesc{K->ab, K->fac64, (size_t)K->n * K->ld, K->sprec, KSH_FP64, 0});
            }
        }
        if ((rc = band_copy_batch(m_stream, m_cdesc.p, cp))) return rc;
        if (!tang.empty() && (rc = batch_tangent(m_stream, (InstDesc *)m_desc.p, tang))) return rc;
        std::vector<BandDescHost> h;
        for (size_t k = 0; k < ms.size(); ++k)
            h.push_back(BandDescHost{ms[k]->fac64, ms[k]->n, ms[k]->bw, (int *)m_info.p + k, (double *)m_stats.p + 4 * k, nullptr, 1});
        if ((rc = band_factor_batch(m_stream, m_bdesc.p, h, KSH_FP64))) return rc;
        std::vector<int> info(ms.size());
        std::vector<double> st(4 * ms.size());
        KSH_HIP(hipMemcpyAsync(info.data(), m_info.p, sizeof(int) * ms.size(), hipMemcpyDeviceToHost, m_stream));
        KSH_HIP(hipMemcpyAsync(st.data(), m_stats.p, sizeof(double) * 4 * ms.size(), hipMemcpyDeviceToHost, m_stream));
        KSH_HIP(hipStreamSynchronize(m_stream));
        for (size_t k = 0; k < ms.size(); ++k) {
            Matrix &K = *ms[k];
            K.use64 = true;
            K.info = info[k] == 0 ? 0 : 1;
            for (int q = 0; q < 4; ++q) K.fstats[q] = st[4 * k + q];
            K.has_stats = info[k] == 0;
        }
        ++m_refactor_launches;
        m_refactors += (long long)ms.size();
        return KSH_SUCCESS;
    }

    int do_solve(std::vector<Req *> &g)
    {
        std::vector<size_t> off;
        int rc;
        if ((rc = ensure_desc(g.size())) || (rc = stage_in(g, off, true, 2))) return rc;
        if ((rc = m_dr.ensure(m_tot * sizeof(double)))) return rc;
        double *b = (double *)m_din[2].p, *x = (double *)m_dout.p, *r = (double *)m_dr.p;
        KSH_HIP(hipMemcpyAsync(x, b, m_tot * sizeof(double), hipMemcpyDeviceToDevice, m_stream));
        for (Req *q : g)
            if (!q->K->factored || q->K->factored_version != q->K->version || q->K->info)
                return fail(KSH_SOLVER_ERROR, "matrix is not factored");
        // solve groups by the precision of the factor in use
        auto solve_all = [&](double *rhs_base, const std::vector<int> &active) -> int {
            for (int fp : {KSH_FP64, KSH_FP32}) {
                std::vector<BandDescHost> h;
                for (size_t i = 0; i < g.size(); ++i) {
                    Matrix &K = *g[i]->K;
                    int efp = K.use64 ? KSH_FP64 : K.fprec;
                    if (efp != fp) continue;
                    void *F = K.use64 ? (void *)K.fac64 : K.fac;
                    h.push_back(BandDescHost{F, K.n, K.bw, nullptr, nullptr, rhs_base + off[i], active[i]});
                }
                int rc2;
                if (!h.empty() && (rc2 = band_solve_batch(m_stream, m_bdesc.p, h, fp))) return rc2;
            }
            return KSH_SUCCESS;
        };
        std::vector<int> active(g.size(), 1);
        if ((rc = solve_all(x, active))) return rc;
        int max_refine = 0;
        double tol = 0.0;
        for (Req *q : g) {
            max_refine = std::max(max_refine, q->max_refine);
            tol = std::max(tol, q->tol);
        }
        std::vector<size_t> poff(g.size());
        size_t ptot = 0;
        for (size_t i = 0; i < g.size(); ++i) {
            poff[i] = ptot;
            ptot += 2 * (((size_t)g[i]->K->n + 63) / 64);
        }
        if (max_refine > 0 && (rc = m_part.ensure(ptot * sizeof(double)))) return rc;
        // Refinement, instance by instance the same decisions as the single-problem solve
        // (ksh_solve_dev): stop at ||r|| <= tol ||b||; on the last sweep of an FP32 factor whose
        // residual is still above accept ||b|| (K.accept, default 1e-6), replace the factor by an
        // FP64 one and solve again.
        std::vector<int> efp(g.size()), fallback(g.size(), 0), err(g.size(), 0);
        for (size_t i = 0; i < g.size(); ++i) {
            efp[i] = g[i]->K->use64 ? KSH_FP64 : g[i]->K->fprec;
            if (g[i]->max_refine <= 0) active[i] = 0;
        }
        for (int it = 0; it < max_refine; ++it) {
            std::vector<RefDesc> h(g.size());
            for (size_t i = 0; i < g.size(); ++i) {
                Matrix &K = *g[i]->K;
                h[i] = RefDesc{K.ab, K.sprec, K.n, K.bw, active[i], b + off[i], x + off[i], r + off[i],
                               (double *)m_norms.p + 2 * i, (double *)m_part.p + poff[i], 0, 0};
            }
            if ((rc = refine_residual_batch(m_stream, (RefDesc *)m_rdesc.p, h))) return rc;
            std::vector<double> nr(2 * g.size());
            KSH_HIP(hipMemcpyAsync(nr.data(), m_norms.p, sizeof(double) * nr.size(), hipMemcpyDeviceToHost, m_stream));
            KSH_HIP(hipStreamSynchronize(m_stream));
            std::vector<int> last(g.size(), 0);
            bool any = false;
            for (size_t i = 0; i < g.size(); ++i) {
                if (!active[i]) continue;
                const double rn = std::sqrt(nr[2 * i]), bn = std::sqrt(nr[2 * i + 1]);
                const bool conv = rn <= g[i]->tol * bn;
                if (!conv && !std::isfinite(rn)) {
                    err[i] = 1;
                    active[i] = 0;
                } else if (conv) {
                    active[i] = 0;
                } else if (it == g[i]->max_refine - 1 && efp[i] == KSH_FP32 && rn > g[i]->K->accept * bn) {
                    fallback[i] = 1;
                    active[i] = 0;
                } else {
                    last[i] = it == g[i]->max_refine - 1;
                    any = true;
                }
            }
            if (!any) break;
            if ((rc = solve_all(r, active))) return rc;
            for (size_t i = 0; i < g.size(); ++i) h[i].active = active[i];
            if ((rc = refine_update_batch(m_stream, (RefDesc *)m_rdesc.p, h))) return rc;
            for (size_t i = 0; i < g.size(); ++i)
                if (last[i]) active[i] = 0;
        }
        // FP64 fallback of stalled refinements: x = K64^-1 b, from the early FP64 factor where
        // one was started, else from one batched refactorization
        {
            std::vector<Matrix *> re;
            for (size_t i = 0; i < g.size(); ++i) {
                if (!fallback[i]) continue;
                g[i]->K->stalled = true;
                re.push_back(g[i]->K);
            }
            prof_count("fp64_refactor_stall", (long long)re.size());
            if ((rc = refactor64(re))) return rc;
            std::vector<BandDescHost> h64;
            for (size_t i = 0; i < g.size(); ++i) {
                if (!fallback[i]) continue;
                Matrix &K = *g[i]->K;
                if (K.info) {
                    err[i] = 1;
                    continue;
                }
                KSH_HIP(hipMemcpyAsync(x + off[i], b + off[i], sizeof(double) * K.n, hipMemcpyDeviceToDevice, m_stream));
                h64.push_back(BandDescHost{K.fac64, K.n, K.bw, nullptr, nullptr, x + off[i], 1});
            }
            if (!h64.empty() && (rc = band_solve_batch(m_stream, m_bdesc.p, h64, KSH_FP64))) return rc;
        }
        KSH_HIP(hipMemcpyAsync(m_hout.p, x, m_tot * sizeof(double), hipMemcpyDeviceToHost, m_stream));
        KSH_HIP(hipStreamSynchronize(m_stream));
        const double *ho = (const double *)m_hout.p;
        for (size_t i = 0; i < g.size(); ++i) {
            int n = g[i]->pb->nfree;
            std::memcpy(g[i]->out, ho + off[i], sizeof(double) * n);
            g[i]->rc = err[i] ? fail(KSH_SOLVER_ERROR, "refinement diverged") : KSH_SUCCESS;
            for (int k = 0; k < n; ++k)
                if (!std::isfinite(ho[off[i] + k])) {
                    g[i]->rc = fail(KSH_SOLVER_ERROR, "non-finite solution");
                    break;
                }
        }
        return KSH_SUCCESS;
    }

    std::mutex m_mu;
    std::condition_variable m_cv_main;
    int m_active = 0, m_blocked = 0;
    long long m_rounds = 0;
    long long m_refactors = 0, m_refactor_launches = 0;  // batched FP64 refactorizations (matrices, launches)
    std::vector<Req *> m_pending;
    hipStream_t m_stream = nullptr;
    // early FP64 factors (launch_early)
    struct Early {
        const Matrix *K;
        uint64_t version;
    };
    bool m_early_on = false, m_early_pending = false;
    hipStream_t m_early_stream = nullptr;
    hipEvent_t m_ev_copied = nullptr, m_ev_early = nullptr;
    std::vector<Early> m_early;
    DevBuf m_ebdesc, m_einfo, m_estats;
    HostBuf m_hearly;
    DevBuf m_din[3], m_dout, m_dr, m_desc, m_bdesc, m_cdesc, m_rdesc, m_stats, m_info, m_norms, m_part;
    HostBuf m_hin[3], m_hout;
    size_t m_tot = 0;
};

// executor of one instance: posts requests to the engine
class EngineExec : public gismo::gsExec {
public:
    explicit EngineExec(Engine *e) : m_e(e) {}
    int tangent(ksh_problem *pb, const double *u, ksh_matrix *K) override
    {
        Req r{RQ_TANGENT};
        r.pb = pb;
        r.K = K;
        r.in = u;
        return m_e->submit(r);
    }
    int residual(ksh_problem *pb, const double *u, double lam, int conv, double *R) override
    {
        Req r{RQ_RESIDUAL};
        r.pb = pb;
        r.in = u;
        r.out = R;
        r.lam = lam;
        r.conv = conv;
        return m_e->submit(r);
    }
    int factor(ksh_matrix *K) override
    {
        // factored by the engine in the round that assembled it (or by an earlier request)
        if (K->factored && K->factored_version == K->version) return K->info ? KSH_SOLVER_ERROR : KSH_SUCCESS;
        Req r{RQ_FACTOR};
        r.pb = K->pb == nullptr ? nullptr : static_cast<ksh_problem *>(K->pb);
        r.K = K;
        return m_e->submit(r);
    }
    int solve(ksh_matrix *K, const double *b, double *x, int max_refine, double tol) override
    {
        Req r{RQ_SOLVE};
        r.pb = static_cast<ksh_problem *>(K->pb);
        r.K = K;
        r.in = b;
        r.out = x;
        r.max_refine = max_refine;
        r.tol = tol;
        return m_e->submit(r);
    }
    int inertia(ksh_matrix *K, int64_t *neg, double *dmin) override
    {
        Req r{RQ_INERTIA};
        r.K = K;
        r.neg = neg;
        r.dmin = dmin;
        return m_e->submit(r);
    }
    int predict(ksh_mlp *m, const double *x, double *y) override
    {
        Req r{RQ_PREDICT};
        r.net = m;
        r.in = x;
        r.out = y;
        return m_e->submit(r);
    }

private:
    Engine *m_e;
};

}  // namespace
}  // namespace ksh

extern "C" int ksh_batch_run_alm(ksh_problem *const *pbs, int count, const ksh_alm_options *opt, ksh_alm_result *res,
                                 double *wall_seconds)
{
    using namespace ksh;
    if (!pbs || count <= 0 || !opt || !res) return fail(KSH_OTHER_ERROR, "null argument");
    for (int i = 0; i < count; ++i)
        if (!pbs[i]) return fail(KSH_OTHER_ERROR, "null problem");
    Engine eng(count);
    int rc;
    if ((rc = eng.init())) return rc;
    std::vector<int> codes(count, KSH_SUCCESS);
    std::vector<std::thread> threads;
    threads.reserve(count);
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < count; ++i) {
        threads.emplace_back([&, i] {
            EngineExec ex(&eng);
            try {
                codes[i] = alm_driver(pbs[i], &ex, opt, &res[i]);
            } catch (...) {
                codes[i] = KSH_OTHER_ERROR;
            }
            eng.finished();  // always: the engine waits for every live instance
        });
    }
    int erc = eng.run();
    for (auto &t : threads) t.join();
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (wall_seconds) *wall_seconds = wall;
    if (erc) return erc;
    for (int i = 0; i < count; ++i)
        if (codes[i] != KSH_SUCCESS && codes[i] != KSH_NOT_CONVERGED) return codes[i];
    for (int i = 0; i < count; ++i)
        if (codes[i] != KSH_SUCCESS) return KSH_NOT_CONVERGED;
    return KSH_SUCCESS;
}
