// ksh_batch.hpp -- batched device operations over many shell problems (one engine stream).
//
// Every operation takes the device vectors of all instances of a round and runs as a few
// launches over all of them: assembly over the union of all elements, one factor / solve
// workgroup (or tile set) per band, batched refinement residuals.  Inputs and outputs stay in
// device memory; only per-instance scalars (pivot statistics, norms) come back to the host.
#pragma once

#include <memory>
#include <vector>

#include "ksh_internal.hpp"
#include "ksh_vec.hpp"

namespace ksh {

constexpr int kMaxKrylov = 8;  // refinement (Krylov) steps per solve at most

struct TangentJob {
    Problem *pb;
    Matrix *K;
    const double *u;    // free vector (device)
};

struct ResidualJob {
    Problem *pb;
    const double *u;    // free vector (device)
    double lambda;
    int conv;           // KSH_RES_STATIC / KSH_RES_ALM
    double *R;          // result (device)
};

// r = b - K x with the assembled values of K (b = nullptr: the zero vector, r = -K x)
struct ProductJob {
    const Matrix *K;
    const double *x;    // device
    const double *b;    // device or nullptr
    double *r;          // device, distinct from x
};

struct SolveJob {
    Matrix *K;
    const double *b;    // right-hand side (device, not overwritten)
    double *x;          // solution (device, distinct from b)
    int max_refine;
    double tol;
    int rc = KSH_SUCCESS;   // per-job outcome
    void *tag = nullptr;    // caller's cookie
};

class BatchDevice {
public:
    BatchDevice() = default;
    BatchDevice(const BatchDevice &) = delete;
    BatchDevice &operator=(const BatchDevice &) = delete;
    ~BatchDevice();
    int init();
    hipStream_t stream() const { return m_stream; }

    // K(u) of every job, then its LDL^T (factor precision of the matrix; FP32 factors with a
    // zero / non-finite pivot are redone in FP64).  Afterwards every K holds factored_version ==
    // version, info, fstats and the certificate state of its pivot count.
    int tangent(const std::vector<TangentJob> &jobs);
    // K(u) of every job without a factorization (the factor, if any, becomes stale)
    int assemble(const std::vector<TangentJob> &jobs);
    // r = b - K x of every job; rr[i] = ||r||^2
    int products(std::vector<ProductJob> &jobs, std::vector<double> &rr);
    int factor(const std::vector<Matrix *> &ms);
    int residual(const std::vector<ResidualJob> &jobs);
    // x = K^-1 b with refinement against the assembled K; FP32 solves whose refinement stalls are
    // redone with an FP64 factor.  Each FP32 factor keeps the largest residual contraction seen
    // over its solves (Matrix::rho), the estimate its pivot-count certificate uses.
    int solve(std::vector<SolveJob> &jobs);
    // FP64 refactorization of the given FP32-factored matrices (their pivot count and solves then
    // use it); matrices already on FP64 are skipped
    int refactor64(const std::vector<Matrix *> &ms);
    // Matrix::rho_pw of the given FP32-factored matrices that lack it: kPowerSteps batched power
    // steps z <- z - M^-1 K z from the fixed pseudo-random start
    int power_contraction(const std::vector<Matrix *> &ms);

    // Deferred FP64 pivot counts, for counts that are only recorded (the path does not branch on
    // them).  Each matrix's current values are copied (FP32 storage: re-assembled in FP64 at the
    // stored point) on the engine stream into a buffer of the matrix's own, factored in FP64 on a
    // third stream of default priority, and their statistics copied to pinned host memory, while
    // the engine stream goes on with the next round.  A matrix with an early FP64 factor of its
    // current values takes that factor's statistics at once instead (appended to `now`).
    // recount_take returns the deferred counts whose stream work has completed (all of them with
    // wait), each with the caller's tag.
    struct Count {
        long long tag;
        int info;          // 0, or the factor hit a zero / non-finite pivot
        double stats[4];   // negatives, min d, max |d|, min |d|
    };
    int recount_async(const std::vector<Matrix *> &ms, const std::vector<long long> &tags, std::vector<Count> &now);
    int recount_take(bool wait, std::vector<Count> &out);

    int dots(const std::vector<DotDesc> &d, std::vector<double> &host_out);  // npairs doubles each, in order
    int vecops(const std::vector<VecDesc> &d);
    int synchronize();

    long long refactors() const { return m_refactors; }

private:
    int ensure_desc(size_t n);
    int factor_matrices(const std::vector<Matrix *> &ms);
    int refactor64_batch(const std::vector<Matrix *> &ms);
    int launch_early(const std::vector<Matrix *> &ms);
    int retire_early();
    long early_index(const Matrix *K) const;

    hipStream_t m_stream = nullptr;
    struct Early {
        const Matrix *K;
        uint64_t version;
    };
    bool m_early_on = false, m_early_pending = false;
    hipStream_t m_early_stream = nullptr;
    hipEvent_t m_ev_copied = nullptr, m_ev_early = nullptr;
    std::vector<Early> m_early;
    DevScratch m_ebdesc, m_einfo, m_estats;
    void *m_hearly = nullptr;  // pinned: info + stats of the early factors
    size_t m_hearly_bytes = 0;
    struct RecountBatch {
        hipEvent_t done = nullptr;
        DevScratch desc, info, stats;
        void *host = nullptr;  // pinned: info + stats
        size_t host_bytes = 0, n = 0;
        std::vector<long long> tags;
        bool busy = false;
    };
    std::vector<std::unique_ptr<RecountBatch>> m_rcb;
    hipStream_t m_rc_stream = nullptr;
    hipEvent_t m_ev_rc_ready = nullptr;
    DevScratch m_desc, m_bdesc, m_cdesc, m_rdesc, m_stats, m_info, m_norms, m_part, m_r, m_dotdesc, m_vecdesc, m_dotout;
    DevScratch m_zero, m_ppart, m_pw, m_start;
    size_t m_start_n = 0;
    size_t m_zero_n = 0;
    long long m_refactors = 0;
};

}  // namespace ksh
