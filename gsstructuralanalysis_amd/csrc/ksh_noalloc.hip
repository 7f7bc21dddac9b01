// ksh_noalloc.hip -- ksh_assemble_residual_dev_noalloc: the residual of ksh_assemble_residual_dev
// for callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops).  Its partial results -- the full control-net displacement, the per-Gauss-point stress
// resultants and the element force vectors -- go to a fixed device array of this library, one per
// GPU, instead of the problem's own workspace; nothing is allocated or freed per call.
//
// `scratch` is a file-scope device array: its host-side name is only a stand-in, so its address on
// each device is taken once with hipGetSymbolAddress (with that device current) and that pointer is
// what the kernels receive.  Every call on one GPU shares that GPU's array and leaves it holding its
// partial results; nothing in it needs to start at zero (each call writes every entry it reads).
// Calls that overlap on one GPU race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ksh_internal.hpp"

namespace ksh {

int run_residual_into(Problem &pb, const double *u_free_dev, double lambda, int conv, double *R_dev, double *u_full,
                      double *Hg, double *Fe);

namespace {

constexpr size_t kScratchBytes = 48ull << 20;  // 48 MiB: about 20k degree-3 elements
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// the device address of `scratch` on `device`, looked up once per device with it current
int scratch_ptr(int device, double **out)
{
    if (device < 0 || device >= kMaxDevices) return fail(KSH_OTHER_ERROR, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        int cur = 0;
        KSH_HIP(hipGetDevice(&cur));
        if (cur != device) KSH_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) KSH_HIP(hipSetDevice(cur));
        if (e != hipSuccess || !p) return fail(KSH_OTHER_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[device] = (double *)p;
    }
    *out = g_ptr[device];
    return KSH_SUCCESS;
}

}  // namespace

// resultants per Gauss point of one residual evaluation (three more with a follower pressure)
size_t noalloc_fq(const Problem &pb) { return pb.pressure != 0.0 ? (size_t)kFqP : (size_t)kFq; }

// scratch doubles one residual evaluation of `pb` needs: u_full, the resultants per Gauss point,
// element force vectors
size_t noalloc_residual_doubles(const Problem &pb)
{
    return (size_t)pb.ndof + noalloc_fq(pb) * (size_t)pb.ngp + 3ull * (size_t)pb.nb * (size_t)pb.nel + 3 * 32;
}

}  // namespace ksh

extern "C" int ksh_assemble_residual_dev_noalloc(ksh_problem *pb, const double *u_free_dev, double lambda, int conv,
                                                 double *R_out_dev)
{
    using namespace ksh;
    if (!pb || !u_free_dev || !R_out_dev) return fail(KSH_OTHER_ERROR, "null argument");
    if (conv != KSH_RES_STATIC && conv != KSH_RES_ALM) return fail(KSH_OTHER_ERROR, "unknown residual convention");
    const size_t need = noalloc_residual_doubles(*pb);
    if (need * sizeof(double) > kScratchBytes)
        return fail(KSH_OTHER_ERROR, "problem needs " + std::to_string(need * sizeof(double)) +
                                         " bytes of residual scratch; the allocation-free entry holds " +
                                         std::to_string(kScratchBytes) + " (use ksh_assemble_residual_dev)");
    double *base = nullptr;
    int rc;
    if ((rc = scratch_ptr(pb->device, &base))) return rc;
    // 256-byte aligned sub-arrays
    auto align = [](size_t n) { return (n + 31) & ~size_t(31); };
    double *u_full = base;
    double *Hg = u_full + align((size_t)pb->ndof);
    double *Fe = Hg + align(noalloc_fq(*pb) * (size_t)pb->ngp);
    if ((rc = run_residual_into(*pb, u_free_dev, lambda, conv, R_out_dev, u_full, Hg, Fe))) return fail(KSH_ASSEMBLY_ERROR, ksh_last_error());
    return KSH_SUCCESS;
}
