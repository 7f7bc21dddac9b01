// ksh_batch.cpp -- batched device operations of the path engine (see ksh_batch.hpp).
//
// Precision handling of the mixed form lives here:
//   - FP32 factors of the assembled FP64 (or FP32) band; a zero or non-finite FP32 pivot gets an
//     FP64 factor of the same values;
//   - refined solves against the assembled K with FP64 residuals; a solve still above
//     accept ||b|| after its sweeps is redone with an FP64 factor (one batched refactorization
//     per round, or the early FP64 factor started beside the FP32 one);
//   - each FP32 factor M = K + E records the largest residual contraction its refinement sweeps
//     showed, an estimate of rho(M^-1 E).  While rho(M^-1 E) < 1 the path K(t) = M - t E,
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
    std::vector<Matrix *> ms;
    for (const TangentJob &j : jobs) {
        Matrix &K = *j.K;
        K.ab_clean = true;
        if (K.sprec == KSH_FP32) {
            KSH_HIP(hipMemcpyAsync(K.u_at, j.u, sizeof(double) * K.n, hipMemcpyDeviceToDevice, m_stream));
            K.has_u = true;
        }
        ++K.version;
        ms.push_back(&K);
    }
    return factor_matrices(ms);
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

int BatchDevice::solve(std::vector<SolveJob> &jobs)
{
    if (jobs.empty()) return KSH_SUCCESS;
    const size_t nj = jobs.size();
    int rc;
    if ((rc = ensure_desc(nj))) return rc;
    for (SolveJob &q : jobs) {
        Matrix &K = *q.K;
        if (!K.factored || K.factored_version != K.version || K.info) return fail(KSH_SOLVER_ERROR, "matrix is not factored");
        if (q.b == q.x) return fail(KSH_OTHER_ERROR, "solve needs distinct b and x");
        q.rc = KSH_SUCCESS;
    }
    std::vector<size_t> off(nj), poff(nj);
    size_t tot = 0, ptot = 0;
    for (size_t i = 0; i < nj; ++i) {
        off[i] = tot;
        tot += (size_t)jobs[i].K->n;
        poff[i] = ptot;
        ptot += 2 * (((size_t)jobs[i].K->n + 63) / 64);
    }
    if ((rc = m_r.ensure(tot * sizeof(double)))) return rc;
    double *rbase = (double *)m_r.p;
    {  // x = b
        std::vector<VecDesc> cp(nj);
        for (size_t i = 0; i < nj; ++i) {
            cp[i].n = jobs[i].K->n;
            cp[i].add(jobs[i].x, 1.0, jobs[i].b);
        }
        if ((rc = batch_vecops(m_stream, m_vecdesc, cp))) return rc;
    }
    // solves of the given right-hand sides (in place), grouped by the precision of the factor in use
    auto solve_all = [&](const std::vector<double *> &rhs, const std::vector<int> &active) -> int {
        for (int fp : {KSH_FP64, KSH_FP32}) {
            std::vector<BandDescHost> h;
            for (size_t i = 0; i < nj; ++i) {
                Matrix &K = *jobs[i].K;
                const int efp = K.use64 ? KSH_FP64 : K.fprec;
                if (efp != fp) continue;
                void *F = K.use64 ? (void *)K.fac64 : K.fac;
                h.push_back(BandDescHost{F, K.n, K.bw, nullptr, nullptr, rhs[i], active[i]});
            }
            int rc2;
            if (!h.empty() && (rc2 = band_solve_batch(m_stream, m_bdesc.p, h, fp))) return rc2;
        }
        return KSH_SUCCESS;
    };
    std::vector<double *> xs(nj), rs(nj);
    for (size_t i = 0; i < nj; ++i) {
        xs[i] = jobs[i].x;
        rs[i] = rbase + off[i];
    }
    std::vector<int> active(nj, 1);
    if ((rc = solve_all(xs, active))) return rc;
    int max_refine = 0;
    for (const SolveJob &q : jobs) max_refine = std::max(max_refine, q.max_refine);
    if (max_refine > 0 && (rc = m_part.ensure(ptot * sizeof(double)))) return rc;
    // Refinement, per job the decisions of the single-problem solve: stop at ||r|| <= tol ||b||;
    // on the last sweep of an FP32 factor whose residual is still above accept ||b||, replace the
    // factor by an FP64 one and solve again.  Residual ratios feed the FP32 count certificate.
    std::vector<int> efp(nj), fallback(nj, 0);
    std::vector<double> rprev(nj, 0.0);
    for (size_t i = 0; i < nj; ++i) {
        efp[i] = jobs[i].K->use64 ? KSH_FP64 : jobs[i].K->fprec;
        if (jobs[i].max_refine <= 0) active[i] = 0;
    }
    for (int it = 0; it < max_refine; ++it) {
        std::vector<RefDesc> h(nj);
        for (size_t i = 0; i < nj; ++i) {
            Matrix &K = *jobs[i].K;
            h[i] = RefDesc{K.ab, K.sprec, K.n, K.bw, active[i], jobs[i].b, jobs[i].x, rs[i], (double *)m_norms.p + 2 * i,
                           (double *)m_part.p + poff[i], 0, 0};
        }
        if ((rc = refine_residual_batch(m_stream, (RefDesc *)m_rdesc.p, h))) return rc;
        std::vector<double> nr(2 * nj);
        KSH_HIP(hipMemcpyAsync(nr.data(), m_norms.p, sizeof(double) * nr.size(), hipMemcpyDeviceToHost, m_stream));
        KSH_HIP(hipStreamSynchronize(m_stream));
        std::vector<int> last(nj, 0);
        bool any = false;
        for (size_t i = 0; i < nj; ++i) {
            if (!active[i]) continue;
            Matrix &K = *jobs[i].K;
            const double rn = std::sqrt(nr[2 * i]), bn = std::sqrt(nr[2 * i + 1]);
            if (efp[i] == KSH_FP32 && bn > 0 && std::isfinite(rn)) {
                // contraction of this sweep: ||r_0|| / ||b||, then ||r_k|| / ||r_k-1||
                const double ratio = it == 0 ? rn / bn : (rprev[i] > 0 ? rn / rprev[i] : 0.0);
                K.rho = std::max(K.rho, ratio);
                ++K.rho_obs;
            }
            rprev[i] = rn;
            const bool conv = rn <= jobs[i].tol * bn;
            if (!conv && !std::isfinite(rn)) {
                jobs[i].rc = fail(KSH_SOLVER_ERROR, "refinement diverged");
                active[i] = 0;
            } else if (conv) {
                active[i] = 0;
            } else if (it == jobs[i].max_refine - 1 && efp[i] == KSH_FP32 && rn > K.accept * bn) {
                fallback[i] = 1;
                active[i] = 0;
            } else {
                last[i] = it == jobs[i].max_refine - 1;
                any = true;
            }
        }
        if (!any) break;
        if ((rc = solve_all(rs, active))) return rc;
        for (size_t i = 0; i < nj; ++i) h[i].active = active[i];
        if ((rc = refine_update_batch(m_stream, (RefDesc *)m_rdesc.p, h))) return rc;
        for (size_t i = 0; i < nj; ++i)
            if (last[i]) active[i] = 0;
    }
    // FP64 fallback of stalled refinements: x = K64^-1 b, from the early FP64 factor where one was
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
