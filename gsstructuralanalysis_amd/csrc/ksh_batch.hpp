// ksh_batch.hpp -- batched device operations over many shell problems (one engine stream).
//
// Every operation takes the device vectors of all instances of a round and runs as a few
// launches over all of them: assembly over the union of all elements, one factor / solve
// workgroup (or tile set) per band, batched refinement residuals.  Inputs and outputs stay in
// device memory; only per-instance scalars (pivot statistics, norms) come back to the host.
#pragma once

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
    DevScratch m_desc, m_bdesc, m_cdesc, m_rdesc, m_stats, m_info, m_norms, m_part, m_r, m_dotdesc, m_vecdesc, m_dotout;
    DevScratch m_zero, m_ppart, m_pw, m_start;
    size_t m_start_n = 0;
    size_t m_zero_n = 0;
    long long m_refactors = 0;
};

}  // namespace ksh
