// ksh_batch.cpp -- batched device operations of the path engine (see ksh_batch.hpp).
//
// Precision handling of the mixed form lives here:
//   - FP32 factors of the assembled FP64 (or FP32) band; a zero or non-finite FP32 pivot gets an
//     FP64 factor of the same values;
//   - solves with an FP32 factor run FGMRES in FP64 against the assembled K, preconditioned by the
//     factor (GMRES-IR); a solve still above its tolerance after its Krylov steps is redone with an
//     FP64 factor (one batched refactorization per round, or the early FP64 factor started
//     beside the FP32 one);
//   - each FP32 factor M = K + E records the largest residual contraction its solves showed, an
//     estimate of rho(M^-1 E).  While rho(M^-1 E) < 1 the path K(t) = M - t E,
//     t in [0, 1], stays nonsingular (K(t) = M (I - t M^-1 E)), so no eigenvalue crosses zero and
//     M has the inertia of K: the FP32 pivot count is exact.  The engine trusts the count when the
//     estimate is below the matrix's guard (default 0.5, a factor 2 of margin on the estimate) and
//     otherwise refactors in FP64 before counting.
#include "ksh_batch.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "ksh_profile.hpp"

namespace ksh {

BatchDevice::~BatchDevice()
{
    if (m_early_stream) {
        (void)hipStreamSynchronize(m_early_stream);
        (void)hipStreamDestroy(m_early_stream);
    }
    if (m_rc_stream) {
        (void)hipStreamSynchronize(m_rc_stream);
        (void)hipStreamDestroy(m_rc_stream);
    }
    for (auto &b : m_rcb) {
        if (b->done) (void)hipEventDestroy(b->done);
        if (b->host) (void)hipHostFree(b->host);
    }
    if (m_ev_rc_ready) (void)hipEventDestroy(m_ev_rc_ready);
    if (m_stream) (void)hipStreamSynchronize(m_stream);
    if (m_ev_copied) (void)hipEventDestroy(m_ev_copied);
    if (m_ev_early) (void)hipEventDestroy(m_ev_early);
    if (m_hearly) (void)hipHostFree(m_hearly);
    if (m_stream) (void)hipStreamDestroy(m_stream);
}

int BatchDevice::init()
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

int BatchDevice::synchronize()
{
    KSH_HIP(hipStreamSynchronize(m_stream));
    return KSH_SUCCESS;
}

int BatchDevice::ensure_desc(size_t n)
{
    int rc;
    if ((rc = m_desc.ensure(n * sizeof(InstDesc)))) return rc;
    if ((rc = m_bdesc.ensure(2 * n * kBandDescBytes))) return rc;
    if ((rc = m_rdesc.ensure(n * sizeof(RefDesc)))) return rc;
    if ((rc = m_stats.ensure(n * 4 * sizeof(double)))) return rc;
    if ((rc = m_info.ensure(n * sizeof(int)))) return rc;
    if ((rc = m_cdesc.ensure(n * sizeof(CopyDesc)))) return rc;
    if ((rc = m_norms.ensure(n * 2 * sizeof(double)))) return rc;
    return KSH_SUCCESS;
}

int BatchDevice::tangent(const std::vector<TangentJob> &jobs)
{
    if (jobs.empty()) return KSH_SUCCESS;
    int rc;
    if ((rc = assemble(jobs))) return rc;
    std::vector<Matrix *> ms;
    for (const TangentJob &j : jobs) ms.push_back(j.K);
    return factor_matrices(ms);
}

int BatchDevice::assemble(const std::vector<TangentJob> &jobs)
{
    if (jobs.empty()) return KSH_SUCCESS;
    int rc;
    if ((rc = ensure_desc(jobs.size()))) return rc;
    for (int sp : {KSH_FP64, KSH_FP32}) {  // one launch per storage precision (normally one group)
        std::vector<InstDesc> h;
        for (const TangentJob &j : jobs) {
            if (j.K->sprec != sp) continue;
            InstDesc d = make_desc(*j.pb);
            d.u_free = j.u;
            d.ab = j.K->ab;
            d.sprec = j.K->sprec;
            d.skip_zero = j.K->ab_clean ? 1 : 0;
            h.push_back(d);
        }
        if (!h.empty() && (rc = batch_tangent(m_stream, (InstDesc *)m_desc.p, h))) return rc;
    }
    for (const TangentJob &j : jobs) {
        Matrix &K = *j.K;
        K.ab_clean = true;
        if (K.sprec == KSH_FP32) {
            KSH_HIP(hipMemcpyAsync(K.u_at, j.u, sizeof(double) * K.n, hipMemcpyDeviceToDevice, m_stream));
            K.has_u = true;
        }
        ++K.version;
    }
    return KSH_SUCCESS;
}

int BatchDevice::power_contraction(const std::vector<Matrix *> &ms)
{
    std::vector<Matrix *> todo;
    for (Matrix *K : ms)
        if (K->fprec == KSH_FP32 && !K->use64 && K->rho_pw < 0) todo.push_back(K);
    if (todo.empty()) return KSH_SUCCESS;
    const size_t nj = todo.size();
    size_t tot = 0, nmax = 0;
    std::vector<size_t> off(nj);
    for (size_t i = 0; i < nj; ++i) {
        off[i] = tot;
        tot += (size_t)todo[i]->n;
        nmax = std::max(nmax, (size_t)todo[i]->n);
    }
    int rc;
    if (nmax > m_start_n) {
        if ((rc = m_start.ensure(nmax * sizeof(double)))) return rc;
        std::vector<double> h(nmax);
        power_start_vector(h.data(), nmax);
        KSH_HIP(hipMemcpyAsync(m_start.p, h.data(), nmax * sizeof(double), hipMemcpyHostToDevice, m_stream));
        KSH_HIP(hipStreamSynchronize(m_stream));
        m_start_n = nmax;
    }
    if ((rc = m_pw.ensure(3 * tot * sizeof(double)))) return rc;
    double *base = (double *)m_pw.p;
    auto Z = [&](size_t i) { return base + off[i]; };
    auto T = [&](size_t i) { return base + tot + off[i]; };
    auto W = [&](size_t i) { return base + 2 * tot + off[i]; };
    std::vector<VecDesc> vd(nj);
    std::vector<DotDesc> dd(nj);
    for (size_t i = 0; i < nj; ++i) {
        vd[i].n = todo[i]->n;
        vd[i].add(Z(i), 1.0, (const double *)m_start.p);
        dd[i].n = todo[i]->n;
        dd[i].npairs = 1;
        dd[i].a[0] = dd[i].b[0] = Z(i);
    }
    std::vector<double> prev, cur;
    if ((rc = vecops(vd)) || (rc = dots(dd, prev))) return rc;
    for (size_t i = 0; i < nj; ++i) todo[i]->rho_pw = 0.0;
    std::vector<int> live(nj, 1);
    for (int k = 0; k < kPowerSteps; ++k) {
        std::vector<ProductJob> pj;
        std::vector<SolveJob> sj;
        std::vector<size_t> idx;
        for (size_t i = 0; i < nj; ++i) {
            if (!live[i]) continue;
            pj.push_back(ProductJob{todo[i], Z(i), nullptr, T(i)});                            // t = -K z
            sj.push_back(SolveJob{todo[i], T(i), W(i), 0, 0.0, KSH_SUCCESS, nullptr});       // w = -M^-1 K z
            idx.push_back(i);
        }
        if (idx.empty()) break;
        std::vector<double> rr;
        if ((rc = products(pj, rr)) || (rc = solve(sj))) return rc;
        for (const SolveJob &q : sj)
            if (q.rc != KSH_SUCCESS) return q.rc;
        std::vector<VecDesc> up;
        std::vector<DotDesc> nd;
        for (size_t i : idx) {
            VecDesc d;
            d.n = todo[i]->n;
            d.add(Z(i), 1.0, Z(i), 1.0, W(i));  // z - M^-1 K z = M^-1 E z
            up.push_back(d);
            nd.push_back(dd[i]);
        }
        if ((rc = vecops(up)) || (rc = dots(nd, cur))) return rc;
        std::vector<VecDesc> sc;
        for (size_t q = 0; q < idx.size(); ++q) {
            const size_t i = idx[q];
            const double ratio = prev[i] > 0 ? std::sqrt(cur[q] / prev[i]) : 0.0;
            Matrix &K = *todo[i];
            if (!std::isfinite(ratio)) {
                K.rho_pw = INFINITY;
                live[i] = 0;
                continue;
            }
            K.rho_pw = ratio;  // the last step's contraction
            if (cur[q] == 0) {
                live[i] = 0;
                continue;
            }
            VecDesc d;  // rescale to unit length
            d.n = K.n;
            d.add(Z(i), 1.0 / std::sqrt(cur[q]), Z(i));
            sc.push_back(d);
            prev[i] = 1.0;
        }
        if ((rc = vecops(sc))) return rc;
    }
    prof_count("power_contraction", (long long)nj);
    return KSH_SUCCESS;
}

int BatchDevice::products(std::vector<ProductJob> &jobs, std::vector<double> &rr)
{
    rr.assign(jobs.size(), 0.0);
    if (jobs.empty()) return KSH_SUCCESS;
    const size_t nj = jobs.size();
    int rc;
    if ((rc = ensure_desc(nj))) return rc;
    size_t ptot = 0, nmax = 0;
    for (const ProductJob &q : jobs) {
        if (q.x == q.r) return fail(KSH_OTHER_ERROR, "product needs distinct x and r");
        ptot += 2 * (((size_t)q.K->n + 63) / 64);
        nmax = std::max(nmax, (size_t)q.K->n);
    }
    if ((rc = m_ppart.ensure(ptot * sizeof(double)))) return rc;
    if (nmax > m_zero_n) {
        if ((rc = m_zero.ensure(nmax * sizeof(double)))) return rc;
        KSH_HIP(hipMemsetAsync(m_zero.p, 0, nmax * sizeof(double), m_stream));
        m_zero_n = nmax;
    }
    std::vector<RefDesc> h(nj);
    size_t po = 0;
    for (size_t i = 0; i < nj; ++i) {
        const Matrix &K = *jobs[i].K;
        h[i] = RefDesc{K.ab, K.sprec, K.n, K.bw, 1, jobs[i].b ? jobs[i].b : (const double *)m_zero.p,
                       const_cast<double *>(jobs[i].x), jobs[i].r, (double *)m_norms.p + 2 * i, (double *)m_ppart.p + po, 0, 0};
        po += 2 * (((size_t)K.n + 63) / 64);
    }
    if ((rc = refine_residual_batch(m_stream, (RefDesc *)m_rdesc.p, h))) return rc;
    std::vector<double> nr(2 * nj);
    KSH_HIP(hipMemcpyAsync(nr.data(), m_norms.p, sizeof(double) * nr.size(), hipMemcpyDeviceToHost, m_stream));
    KSH_HIP(hipStreamSynchronize(m_stream));
    for (size_t i = 0; i < nj; ++i) rr[i] = nr[2 * i];
    return KSH_SUCCESS;
}

int BatchDevice::factor(const std::vector<Matrix *> &ms)
{
    std::vector<Matrix *> todo;
    for (Matrix *K : ms)
        if (!(K->factored && K->factored_version == K->version)) todo.push_back(K);
    return factor_matrices(todo);
}

// LDL^T of the current values of several matrices: one batched copy into the factor storage, one
// batched factor launch per factor precision, and the FP64 refactorization of FP32 factors that
// hit a zero or non-finite pivot.
int BatchDevice::factor_matrices(const std::vector<Matrix *> &ms)
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
            // a solve of this matrix's previous FP32 factor fell back to FP64: so will one of this
            // factor's, most likely; its FP64 factor runs beside the FP32 one
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
            K.rho = 0.0;
            K.rho_obs = 0;
            K.rho_pw = -1.0;
            for (int q = 0; q < 4; ++q) K.fstats[q] = st[4 * k + q];
            K.has_stats = info[k] == 0;
            if (fp == KSH_FP32 && info[k] != 0) re.push_back(&K);
        }
        prof_count("fp64_refactor_pivot", (long long)re.size());
        if ((rc = refactor64(re))) return rc;
    }
    return KSH_SUCCESS;
}

// FP64 factors of the current values of several matrices (each at most once): the early factor
// where one was started for this version, else one batched refactorization.
int BatchDevice::refactor64(const std::vector<Matrix *> &ms)
{
    std::vector<Matrix *> re;
    std::vector<std::pair<Matrix *, long>> taken;
    for (Matrix *K : ms) {
        if (K->use64 || K->fprec == KSH_FP64) continue;
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
        const int *hi = (const int *)m_hearly;
        const double *hs = (const double *)((const char *)m_hearly + n * sizeof(int));
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

int BatchDevice::recount_async(const std::vector<Matrix *> &ms, const std::vector<long long> &tags, std::vector<Count> &now)
{
    if (ms.size() != tags.size()) return fail(KSH_OTHER_ERROR, "recount_async: one tag per matrix");
    // matrices with an early FP64 factor of their current values: its statistics now
    std::vector<Matrix *> rest;
    std::vector<long long> rest_tags;
    bool early_waited = false;
    for (size_t i = 0; i < ms.size(); ++i) {
        const long k = early_index(ms[i]);
        if (k < 0) {
            rest.push_back(ms[i]);
            rest_tags.push_back(tags[i]);
            continue;
        }
        if (!early_waited) {
            KSH_HIP(hipEventSynchronize(m_ev_early));
            early_waited = true;
        }
        const size_t ne = m_early.size();
        const int *hi = (const int *)m_hearly;
        const double *hs = (const double *)((const char *)m_hearly + ne * sizeof(int));
        Count c;
        c.tag = tags[i];
        c.info = hi[k];
        for (int q = 0; q < 4; ++q) c.stats[q] = hs[4 * k + q];
        now.push_back(c);
    }
    if (!early_waited && rest.empty()) return KSH_SUCCESS;
    if (early_waited) prof_count("fp64_refactor_early_used", (long long)(ms.size() - rest.size()));
    if (rest.empty()) return KSH_SUCCESS;
    int rc;
    if (!m_rc_stream) {
        KSH_HIP(hipStreamCreateWithFlags(&m_rc_stream, hipStreamNonBlocking));
        KSH_HIP(hipEventCreateWithFlags(&m_ev_rc_ready, hipEventDisableTiming));
    }
    RecountBatch *B = nullptr;  // a record whose counts were taken, else a new one
    for (auto &b : m_rcb)
        if (!b->busy) {
            B = b.get();
            break;
        }
    if (!B) {
        m_rcb.push_back(std::make_unique<RecountBatch>());
        B = m_rcb.back().get();
        KSH_HIP(hipEventCreateWithFlags(&B->done, hipEventDisableTiming));
    }
    const size_t n = rest.size();
    const size_t hb = n * (sizeof(int) + 4 * sizeof(double));
    if ((rc = ensure_desc(n)) || (rc = B->desc.ensure(n * kBandDescBytes)) || (rc = B->info.ensure(n * sizeof(int))) ||
        (rc = B->stats.ensure(n * 4 * sizeof(double))))
        return rc;
    if (B->host_bytes < hb) {
        if (B->host) (void)hipHostFree(B->host);
        B->host = nullptr;
        B->host_bytes = 0;
        KSH_HIP(hipHostMalloc(&B->host, hb, hipHostMallocDefault));
        B->host_bytes = hb;
    }
    std::vector<CopyDesc> cp;
    std::vector<InstDesc> tang;
    std::vector<BandDescHost> h;
    for (size_t k = 0; k < n; ++k) {
        Matrix *K = rest[k];
        if ((rc = ensure_rc64(*K))) return rc;
        if (K->rc_done) KSH_HIP(hipStreamWaitEvent(m_stream, K->rc_done, 0));  // rc64 free again
        if (K->sprec == KSH_FP32 && K->has_u) {
            InstDesc d = make_desc(*K->pb);
            d.u_free = K->u_at;
            d.ab = K->rc64;
            d.sprec = KSH_FP64;
            tang.push_back(d);
        } else {
            cp.push_back(CopyDesc{K->ab, K->rc64, (size_t)K->n * K->ld, K->sprec, KSH_FP64, 0});
        }
        h.push_back(BandDescHost{K->rc64, K->n, K->bw, (int *)B->info.p + k, (double *)B->stats.p + 4 * k, nullptr, 1});
    }
    if ((rc = band_copy_batch(m_stream, m_cdesc.p, cp))) return rc;
    if (!tang.empty() && (rc = batch_tangent(m_stream, (InstDesc *)m_desc.p, tang))) return rc;
    KSH_HIP(hipEventRecord(m_ev_rc_ready, m_stream));
    KSH_HIP(hipStreamWaitEvent(m_rc_stream, m_ev_rc_ready, 0));
    if ((rc = band_factor_batch(m_rc_stream, B->desc.p, h, KSH_FP64, "band_ldlt_recount_fp64"))) {
        (void)hipStreamSynchronize(m_rc_stream);  // nothing of a failed launch left in flight
        return rc;
    }
    char *hp = (char *)B->host;
    KSH_HIP(hipMemcpyAsync(hp, B->info.p, n * sizeof(int), hipMemcpyDeviceToHost, m_rc_stream));
    KSH_HIP(hipMemcpyAsync(hp + n * sizeof(int), B->stats.p, n * 4 * sizeof(double), hipMemcpyDeviceToHost, m_rc_stream));
    KSH_HIP(hipEventRecord(B->done, m_rc_stream));
    for (Matrix *K : rest) {
        if (!K->rc_done) KSH_HIP(hipEventCreateWithFlags(&K->rc_done, hipEventDisableTiming));
        KSH_HIP(hipEventRecord(K->rc_done, m_rc_stream));
    }
    B->tags = rest_tags;
    B->n = n;
    B->busy = true;
    m_refactors += (long long)n;
    prof_count("fp64_recount_deferred", (long long)n);
    return KSH_SUCCESS;
}

int BatchDevice::recount_take(bool wait, std::vector<Count> &out)
{
    for (auto &b : m_rcb) {
        if (!b->busy) continue;
        if (wait) {
            KSH_HIP(hipEventSynchronize(b->done));
        } else {
            const hipError_t e = hipEventQuery(b->done);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return fail(KSH_OTHER_ERROR, std::string("deferred pivot count: ") + hipGetErrorString(e));
        }
        const int *hi = (const int *)b->host;
        const double *hs = (const double *)((const char *)b->host + b->n * sizeof(int));
        for (size_t k = 0; k < b->n; ++k) {
            Count c;
            c.tag = b->tags[k];
            c.info = hi[k];
            for (int q = 0; q < 4; ++q) c.stats[q] = hs[4 * k + q];
            out.push_back(c);
        }
        b->busy = false;
    }
    return KSH_SUCCESS;
}

// Early FP64 factors: a matrix whose previous FP32 factor needed the FP64 fallback in a solve is
// likely to need it again (the stalls follow the conditioning, which changes slowly along the
// path), so its FP64 factor is started together with the FP32 one, on a second stream of default
// priority: the two launch chains, each bound by its panel latency, run at once.  A solve that
// stalls (or a count that is not certified) takes that factor instead of refactoring: the same
// kernels on the same values, so the same bits.  Only FP64 storage (a copy of ab); FP32 storage
// re-assembles on the problem's element tiles, which the main stream may be using.  fac64 of
// these matrices belongs to the second stream until retire_early or until a fallback takes it.
int BatchDevice::launch_early(const std::vector<Matrix *> &ms)
{
    if (!m_early_on || ms.empty()) return KSH_SUCCESS;
    int rc;
    const size_t n = ms.size();
    const size_t hb = n * (sizeof(int) + 4 * sizeof(double));
    if (m_ebdesc.bytes < n * kBandDescBytes || m_einfo.bytes < n * sizeof(int) || m_estats.bytes < n * 4 * sizeof(double) ||
        m_hearly_bytes < hb)
        KSH_HIP(hipStreamSynchronize(m_early_stream));  // buffers about to be replaced
    if ((rc = m_ebdesc.ensure(n * kBandDescBytes)) || (rc = m_einfo.ensure(n * sizeof(int))) ||
        (rc = m_estats.ensure(n * 4 * sizeof(double))))
        return rc;
    if (m_hearly_bytes < hb) {
        if (m_hearly) (void)hipHostFree(m_hearly);
        m_hearly = nullptr;
        m_hearly_bytes = 0;
        KSH_HIP(hipHostMalloc(&m_hearly, hb, hipHostMallocDefault));
        m_hearly_bytes = hb;
    }
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
    char *hp = (char *)m_hearly;
    KSH_HIP(hipMemcpyAsync(hp, m_einfo.p, n * sizeof(int), hipMemcpyDeviceToHost, m_early_stream));
    KSH_HIP(hipMemcpyAsync(hp + n * sizeof(int), m_estats.p, n * 4 * sizeof(double), hipMemcpyDeviceToHost, m_early_stream));
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

// The main stream waits for the early factors (their fac64 may be written again); they are
// forgotten.
int BatchDevice::retire_early()
{
    if (m_early_pending) KSH_HIP(hipStreamWaitEvent(m_stream, m_ev_early, 0));
    m_early_pending = false;
    m_early.clear();
    return KSH_SUCCESS;
}

// Position of K's early factor of its current version, or -1.  Compares addresses only: the
// entries are dropped at every factor call, and a matrix can only be freed and another created at
// its address after that.
long BatchDevice::early_index(const Matrix *K) const
{
    for (size_t k = 0; k < m_early.size(); ++k)
        if (m_early[k].K == K && m_early[k].version == K->factored_version) return (long)k;
    return -1;
}

// FP64 refactorization of several matrices in one batched launch: the assembled values copied (or
// re-assembled in FP64 from the stored point, for FP32 storage) into fac64 and factored there.
int BatchDevice::refactor64_batch(const std::vector<Matrix *> &ms)
{
    if (ms.empty()) return KSH_SUCCESS;
    int rc;
    if ((rc = ensure_desc(ms.size()))) return rc;
    std::vector<InstDesc> tang;
    std::vector<CopyDesc> cp;
    for (Matrix *K : ms) {
        if ((rc = ensure_fac64(*K))) return rc;
        if (K->early_done) KSH_HIP(hipStreamWaitEvent(m_stream, K->early_done, 0));
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
    m_refactors += (long long)ms.size();
    return KSH_SUCCESS;
}

int BatchDevice::residual(const std::vector<ResidualJob> &jobs)
{
    if (jobs.empty()) return KSH_SUCCESS;
    int rc;
    if ((rc = ensure_desc(jobs.size()))) return rc;
    std::vector<InstDesc> h;
    for (const ResidualJob &j : jobs) {
        if (j.conv != KSH_RES_ALM && j.conv != KSH_RES_STATIC) return fail(KSH_OTHER_ERROR, "unknown residual convention");
        InstDesc d = make_desc(*j.pb);
        d.u_free = j.u;
        d.lambda = j.lambda;
        d.mode = (j.conv == KSH_RES_ALM) ? 1 : 2;
        d.R = j.R;
        h.push_back(d);
    }
    return batch_residual(m_stream, (InstDesc *)m_desc.p, h);
}

// Solves K x = b.  FP64 factors: direct solve (plus classical refinement sweeps if asked).  FP32
// factors M = K + E: FGMRES in FP64 on K with M^-1 as the preconditioner (GMRES-IR): x0 = M^-1 b,
// r0 = b - K x0, then at most max_refine Arnoldi steps z_j = M^-1 v_j, u = K z_j, orthogonalised
// against v_0..v_j by classical Gram-Schmidt with one reorthogonalisation, Givens rotations for
// the least-squares residual |g_j+1| = ||b - K x_j||; x = x0 + Z y.  Refinement sweeps stall once
// rho(M^-1 E) approaches 1 (thin shells near limit points); the Krylov space takes the few
// outlying eigenvalues of K M^-1 out in as many steps.  A solve still above `tol` after its steps
// is redone with an FP64 factor.  The first two residual contractions (||r0|| / ||b|| and
// ||(I - K M^-1) v_0||, one power step of E M^-1 each) go into the factor's rho estimate.
int BatchDevice::solve(std::vector<SolveJob> &jobs)
{
    if (jobs.empty()) return KSH_SUCCESS;
    const size_t nj = jobs.size();
    int rc;
    if ((rc = ensure_desc(nj))) return rc;
    int nmax = 0, mmax = 0;
    for (SolveJob &q : jobs) {
        Matrix &K = *q.K;
        if (!K.factored || K.factored_version != K.version || K.info) return fail(KSH_SOLVER_ERROR, "matrix is not factored");
        if (q.b == q.x) return fail(KSH_OTHER_ERROR, "solve needs distinct b and x");
        if (q.max_refine > kMaxKrylov) return fail(KSH_OTHER_ERROR, "at most 8 refinement steps per solve");
        q.rc = KSH_SUCCESS;
        nmax = std::max(nmax, K.n);
        mmax = std::max(mmax, q.max_refine);
    }
    std::vector<size_t> off(nj), poff(nj);
    size_t tot = 0, ptot = 0;
    for (size_t i = 0; i < nj; ++i) {
        off[i] = tot;
        tot += (size_t)jobs[i].K->n;
        poff[i] = ptot;
        ptot += 2 * (((size_t)jobs[i].K->n + 63) / 64);
    }
    // per job: r, then the Krylov basis V (mmax + 1 vectors) and Z (mmax vectors)
    const size_t per = 1 + (mmax > 0 ? 2 * (size_t)mmax + 1 : 0);
    if ((rc = m_r.ensure(tot * per * sizeof(double)))) return rc;
    if (mmax > 0 && (rc = m_part.ensure(ptot * sizeof(double)))) return rc;
    if (mmax > 0 && (size_t)nmax > m_zero_n) {
        if ((rc = m_zero.ensure((size_t)nmax * sizeof(double)))) return rc;
        KSH_HIP(hipMemsetAsync(m_zero.p, 0, (size_t)nmax * sizeof(double), m_stream));
        m_zero_n = (size_t)nmax;
    }
    double *rbase = (double *)m_r.p;
    auto R = [&](size_t i) { return rbase + off[i]; };
    auto V = [&](size_t i, int k) { return rbase + tot + (size_t)off[i] * (2 * mmax + 1) + (size_t)k * jobs[i].K->n; };
    auto Z = [&](size_t i, int k) { return V(i, mmax + 1 + k); };
    {  // x = b
        std::vector<VecDesc> cp(nj);
        for (size_t i = 0; i < nj; ++i) {
            cp[i].n = jobs[i].K->n;
            cp[i].add(jobs[i].x, 1.0, jobs[i].b);
        }
        if ((rc = batch_vecops(m_stream, m_vecdesc, cp))) return rc;
    }
    std::vector<int> efp(nj);
    for (size_t i = 0; i < nj; ++i) efp[i] = jobs[i].K->use64 ? KSH_FP64 : jobs[i].K->fprec;
    // in-place solves of the given right-hand sides with each job's factor in use
    auto solve_all = [&](const std::vector<double *> &rhs, const std::vector<int> &active) -> int {
        for (int fp : {KSH_FP64, KSH_FP32}) {
            std::vector<BandDescHost> h;
            for (size_t i = 0; i < nj; ++i) {
                if (efp[i] != fp || !active[i]) continue;
                Matrix &K = *jobs[i].K;
                void *F = K.use64 ? (void *)K.fac64 : K.fac;
                h.push_back(BandDescHost{F, K.n, K.bw, nullptr, nullptr, rhs[i], 1});
            }
            int rc2;
            if (!h.empty() && (rc2 = band_solve_batch(m_stream, m_bdesc.p, h, fp))) return rc2;
        }
        return KSH_SUCCESS;
    };
    // r = b - K x for the active jobs (b = nullptr: the zero vector), norms to the host
    std::vector<double> nr(2 * nj);
    auto residuals = [&](const std::vector<const double *> &b, const std::vector<const double *> &x,
                         const std::vector<double *> &r, const std::vector<int> &active) -> int {
        std::vector<RefDesc> h(nj);
        for (size_t i = 0; i < nj; ++i) {
            Matrix &K = *jobs[i].K;
            h[i] = RefDesc{K.ab, K.sprec, K.n, K.bw, active[i], b[i] ? b[i] : (const double *)m_zero.p,
                           const_cast<double *>(x[i]), r[i], (double *)m_norms.p + 2 * i, (double *)m_part.p + poff[i], 0, 0};
        }
        int rc2;
        if ((rc2 = refine_residual_batch(m_stream, (RefDesc *)m_rdesc.p, h))) return rc2;
        KSH_HIP(hipMemcpyAsync(nr.data(), m_norms.p, sizeof(double) * nr.size(), hipMemcpyDeviceToHost, m_stream));
        KSH_HIP(hipStreamSynchronize(m_stream));
        return KSH_SUCCESS;
    };
    std::vector<double *> xs(nj);
    for (size_t i = 0; i < nj; ++i) xs[i] = jobs[i].x;
    std::vector<int> active(nj, 1);
    if ((rc = solve_all(xs, active))) return rc;
    if (mmax == 0) return KSH_SUCCESS;

    // ---- residual of the first solve ---------------------------------------------------------
    std::vector<const double *> bs(nj), xc(nj);
    std::vector<double *> rs(nj);
    for (size_t i = 0; i < nj; ++i) {
        bs[i] = jobs[i].b;
        xc[i] = jobs[i].x;
        rs[i] = R(i);
        active[i] = jobs[i].max_refine > 0;
    }
    if ((rc = residuals(bs, xc, rs, active))) return rc;
    std::vector<double> bn(nj, 0.0), beta(nj, 0.0);
    std::vector<int> fallback(nj, 0), gm(nj, 0), ir(nj, 0);
    for (size_t i = 0; i < nj; ++i) {
        if (!active[i]) continue;
        const double rn = std::sqrt(nr[2 * i]);
        bn[i] = std::sqrt(nr[2 * i + 1]);
        beta[i] = rn;
        if (!std::isfinite(rn)) {
            jobs[i].rc = fail(KSH_SOLVER_ERROR, "non-finite residual in the refined solve");
            active[i] = 0;
            continue;
        }
        if (efp[i] == KSH_FP32 && bn[i] > 0) {
            Matrix &K = *jobs[i].K;
            K.rho = std::max(K.rho, rn / bn[i]);
            ++K.rho_obs;
        }
        if (rn <= jobs[i].tol * bn[i]) active[i] = 0;
        else if (efp[i] == KSH_FP32) gm[i] = 1;
        else ir[i] = 1;
    }
    // ---- FP64 factors: classical refinement sweeps ---------------------------------------------
    for (int it = 0; it < mmax; ++it) {
        std::vector<int> act(nj, 0);
        bool any = false;
        for (size_t i = 0; i < nj; ++i)
            if (ir[i] && it < jobs[i].max_refine) {
                act[i] = 1;
                any = true;
            }
        if (!any) break;
        if ((rc = solve_all(rs, act))) return rc;
        std::vector<RefDesc> h(nj);
        for (size_t i = 0; i < nj; ++i) {
            Matrix &K = *jobs[i].K;
            h[i] = RefDesc{K.ab, K.sprec, K.n, K.bw, act[i], jobs[i].b, jobs[i].x, R(i), (double *)m_norms.p + 2 * i,
                           (double *)m_part.p + poff[i], 0, 0};
        }
        if ((rc = refine_update_batch(m_stream, (RefDesc *)m_rdesc.p, h))) return rc;
        if ((rc = residuals(bs, xc, rs, act))) return rc;
        for (size_t i = 0; i < nj; ++i) {
            if (!act[i]) continue;
            const double rn = std::sqrt(nr[2 * i]);
            if (!std::isfinite(rn)) {
                jobs[i].rc = fail(KSH_SOLVER_ERROR, "refinement diverged");
                ir[i] = 0;
            } else if (rn <= jobs[i].tol * bn[i]) {
                ir[i] = 0;
            }
        }
    }
    // ---- FP32 factors: FGMRES with the FP32 factor as preconditioner -----------------------------
    std::vector<std::vector<double>> H(nj), cs(nj), sn(nj), g(nj);
    std::vector<int> steps(nj, 0);
    {
        std::vector<VecDesc> vd;
        for (size_t i = 0; i < nj; ++i) {
            if (!gm[i]) continue;
            const int m = jobs[i].max_refine;
            H[i].assign((size_t)(m + 1) * m, 0.0);
            cs[i].assign(m, 0.0);
            sn[i].assign(m, 0.0);
            g[i].assign(m + 1, 0.0);
            g[i][0] = beta[i];
            VecDesc d;
            d.n = jobs[i].K->n;
            d.add(V(i, 0), 1.0 / beta[i], R(i));
            vd.push_back(d);
        }
        if ((rc = batch_vecops(m_stream, m_vecdesc, vd))) return rc;
    }
    for (int j = 0; j < mmax; ++j) {
        std::vector<int> act(nj, 0);
        bool any = false;
        for (size_t i = 0; i < nj; ++i)
            if (gm[i] == 1 && j < jobs[i].max_refine) {
                act[i] = 1;
                any = true;
            }
        if (!any) break;
        // z_j = M^-1 v_j ; u = K z_j (as -u = 0 - K z_j, in R)
        std::vector<VecDesc> cp;
        std::vector<double *> zs(nj, nullptr);
        for (size_t i = 0; i < nj; ++i) {
            if (!act[i]) continue;
            VecDesc d;
            d.n = jobs[i].K->n;
            d.add(Z(i, j), 1.0, V(i, j));
            cp.push_back(d);
            zs[i] = Z(i, j);
        }
        if ((rc = batch_vecops(m_stream, m_vecdesc, cp)) || (rc = solve_all(zs, act))) return rc;
        std::vector<const double *> zero(nj, nullptr), zc(nj);
        for (size_t i = 0; i < nj; ++i) zc[i] = zs[i] ? zs[i] : jobs[i].x;
        if ((rc = residuals(zero, zc, rs, act))) return rc;
        // two Gram-Schmidt passes against v_0..v_j (R holds -u)
        std::vector<double> h1(nj * (kMaxKrylov + 1), 0.0), uu(nj, 0.0);
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<DotDesc> dd;
            std::vector<size_t> who;
            for (size_t i = 0; i < nj; ++i) {
                if (!act[i]) continue;
                DotDesc d;
                d.n = jobs[i].K->n;
                for (int k = 0; k <= j; ++k) {
                    d.a[d.npairs] = R(i);
                    d.b[d.npairs] = V(i, k);
                    ++d.npairs;
                }
                d.a[d.npairs] = R(i);
                d.b[d.npairs] = R(i);
                ++d.npairs;
                dd.push_back(d);
                who.push_back(i);
            }
            std::vector<double> dv;
            if ((rc = dots(dd, dv))) return rc;
            std::vector<VecDesc> vd;
            size_t o = 0;
            for (size_t q = 0; q < who.size(); ++q) {
                const size_t i = who[q];
                const double sign = pass == 0 ? -1.0 : 1.0;  // pass 0 works on -u, pass 1 on the projected u
                VecDesc d;
                d.n = jobs[i].K->n;
                std::vector<double> c(j + 1);
                for (int k = 0; k <= j; ++k) c[k] = sign * dv[o + k];
                const double ww = dv[o + j + 1];
                o += (size_t)j + 2;
                for (int k = 0; k <= j; ++k) h1[i * (kMaxKrylov + 1) + k] += c[k];
                if (pass == 0) uu[i] = ww;
                else uu[i] = ww;  // ||u - V h||^2 before the second projection
                // R <- sign R - sum c_k v_k  (pass 0 turns -u into u on the way)
                int k = 0;
                d.add(R(i), sign, R(i), -(k <= j ? c[k] : 0.0), V(i, 0), k + 1 <= j ? -c[k + 1] : 0.0,
                      k + 1 <= j ? V(i, k + 1) : nullptr);
                for (k = 2; k <= j; k += 2)
                    d.add(R(i), 1.0, R(i), -c[k], V(i, k), k + 1 <= j ? -c[k + 1] : 0.0, k + 1 <= j ? V(i, k + 1) : nullptr);
                vd.push_back(d);
                if (pass == 1) {
                    double proj = 0.0;
                    for (int kk = 0; kk <= j; ++kk) proj += c[kk] * c[kk];
                    uu[i] = std::max(0.0, ww - proj);  // ||w||^2 after the second projection
                }
            }
            if ((rc = batch_vecops(m_stream, m_vecdesc, vd))) return rc;
            if (pass == 0) {
                // first power step of E M^-1 on v_0: ||v_0 - u||^2 = 1 - 2 v_0.u + u.u
                for (size_t i = 0; i < nj; ++i)
                    if (act[i] && j == 0) {
                        Matrix &K = *jobs[i].K;
                        const double hh = h1[i * (kMaxKrylov + 1)];
                        const double r2 = 1.0 - 2.0 * hh + uu[i];
                        K.rho = std::max(K.rho, std::sqrt(std::max(0.0, r2)));
                        ++K.rho_obs;
                    }
            }
        }
        std::vector<VecDesc> vd;
        for (size_t i = 0; i < nj; ++i) {
            if (!act[i]) continue;
            const int m = jobs[i].max_refine;
            double *Hc = H[i].data() + (size_t)j * (m + 1);  // column j
            for (int k = 0; k <= j; ++k) Hc[k] = h1[i * (kMaxKrylov + 1) + k];
            const double hn = std::sqrt(uu[i]);
            Hc[j + 1] = hn;
            for (int k = 0; k < j; ++k) {  // previous rotations
                const double a = Hc[k], b = Hc[k + 1];
                Hc[k] = cs[i][k] * a + sn[i][k] * b;
                Hc[k + 1] = -sn[i][k] * a + cs[i][k] * b;
            }
            const double den = std::hypot(Hc[j], Hc[j + 1]);
            cs[i][j] = den > 0 ? Hc[j] / den : 1.0;
            sn[i][j] = den > 0 ? Hc[j + 1] / den : 0.0;
            Hc[j] = den;
            Hc[j + 1] = 0.0;
            g[i][j + 1] = -sn[i][j] * g[i][j];
            g[i][j] = cs[i][j] * g[i][j];
            steps[i] = j + 1;
            const double res = std::abs(g[i][j + 1]);
            if (!std::isfinite(res)) {
                gm[i] = 0;
                fallback[i] = 1;
                continue;
            }
            if (res <= jobs[i].tol * bn[i] || hn == 0.0) {
                gm[i] = 2;  // converged: solution update below
            } else if (j + 1 < m) {
                VecDesc d;
                d.n = jobs[i].K->n;
                d.add(V(i, j + 1), 1.0 / hn, R(i));
                vd.push_back(d);
            }
        }
        if ((rc = batch_vecops(m_stream, m_vecdesc, vd))) return rc;
        for (size_t i = 0; i < nj; ++i)
            if (gm[i] == 2) gm[i] = 3;  // stop iterating, keep for the update
    }
    // x = x0 + Z y, R y = g (back substitution on the rotated Hessenberg)
    {
        std::vector<VecDesc> vd;
        for (size_t i = 0; i < nj; ++i) {
            if (!gm[i]) continue;
            const int m = jobs[i].max_refine, k = steps[i];
            if (gm[i] == 1 && std::abs(g[i][k]) > jobs[i].tol * bn[i]) {
                fallback[i] = 1;  // not converged within the steps: FP64 factor below
                continue;
            }
            std::vector<double> y(k, 0.0);
            for (int r = k - 1; r >= 0; --r) {
                double sum = g[i][r];
                for (int c = r + 1; c < k; ++c) sum -= H[i][(size_t)c * (m + 1) + r] * y[c];
                y[r] = sum / H[i][(size_t)r * (m + 1) + r];
            }
            VecDesc d;
            d.n = jobs[i].K->n;
            for (int c = 0; c < k; c += 2)
                d.add(jobs[i].x, 1.0, jobs[i].x, y[c], Z(i, c), c + 1 < k ? y[c + 1] : 0.0, c + 1 < k ? Z(i, c + 1) : nullptr);
            vd.push_back(d);
            prof_count("gmres_steps", k);
        }
        if ((rc = batch_vecops(m_stream, m_vecdesc, vd))) return rc;
    }
    // FP64 fallback of unconverged solves: x = K64^-1 b, from the early FP64 factor where one was
    // started, else from one batched refactorization
    std::vector<Matrix *> re;
    for (size_t i = 0; i < nj; ++i)
        if (fallback[i]) {
            jobs[i].K->stalled = true;
            re.push_back(jobs[i].K);
        }
    if (!re.empty()) {
        prof_count("fp64_refactor_stall", (long long)re.size());
        if ((rc = refactor64(re))) return rc;
        std::vector<VecDesc> cp;
        std::vector<BandDescHost> h64;
        for (size_t i = 0; i < nj; ++i) {
            if (!fallback[i]) continue;
            Matrix &K = *jobs[i].K;
            if (K.info) {
                jobs[i].rc = fail(KSH_SOLVER_ERROR, "FP64 refactorization failed");
                continue;
            }
            VecDesc v;
            v.n = K.n;
            v.add(jobs[i].x, 1.0, jobs[i].b);
            cp.push_back(v);
            h64.push_back(BandDescHost{K.fac64, K.n, K.bw, nullptr, nullptr, jobs[i].x, 1});
        }
        if ((rc = batch_vecops(m_stream, m_vecdesc, cp))) return rc;
        if (!h64.empty() && (rc = band_solve_batch(m_stream, m_bdesc.p, h64, KSH_FP64))) return rc;
    }
    return KSH_SUCCESS;
}

int BatchDevice::dots(const std::vector<DotDesc> &d, std::vector<double> &host_out)
{
    size_t tot = 0;
    for (const DotDesc &x : d) tot += (size_t)x.npairs;
    host_out.assign(tot, 0.0);
    if (tot == 0) return KSH_SUCCESS;
    int rc;
    if ((rc = m_dotout.ensure(tot * sizeof(double)))) return rc;
    std::vector<DotDesc> dd = d;
    size_t o = 0;
    for (DotDesc &x : dd) {
        x.out = (double *)m_dotout.p + o;
        o += (size_t)x.npairs;
    }
    if ((rc = batch_dots(m_stream, m_dotdesc, dd))) return rc;
    KSH_HIP(hipMemcpyAsync(host_out.data(), m_dotout.p, tot * sizeof(double), hipMemcpyDeviceToHost, m_stream));
    KSH_HIP(hipStreamSynchronize(m_stream));
    return KSH_SUCCESS;
}

int BatchDevice::vecops(const std::vector<VecDesc> &d) { return batch_vecops(m_stream, m_vecdesc, d); }

}  // namespace ksh
