// ksh_mlp.hip -- the learned arc-length predictor's MLP: forward, backward and Adam on gfx950.
//
// Network (fixed by the predictor's feature layout): 264 -> 256 -> 256 -> 65, GELU after the two
// hidden layers, linear output.  Weights are kept in FP32 (Adam's master copy) and enter every
// product rounded to BF16 (round to nearest even), as do the activations; products accumulate in
// FP32 on v_mfma_f32_16x16x16bf16_1k.  Loss: mean squared error over the batch and the 65
// outputs (the caller normalizes inputs and targets).
//
// One GEMM kernel serves the forward layer (Y = act(X W + b)), the input gradient
// (dX = dY W^T) and the weight gradient (dW = X^T dY): operands are addressed through strides,
// so a transposed operand is just other strides.  Workgroup tile 64 x 64 (four waves of 32 x 32
// = 2 x 2 MFMA tiles), K in steps of 16 staged in LDS as BF16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "ksh_internal.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float x)
{
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

constexpr int kTM = 64, kTN = 64, kTK = 16;

// C(m, n) = act(sum_k A(m, k) B(k, n) + bias[n]), A(m, k) = A[m sam + k sak], B(k, n) = B[k sbk + n sbn]
// act: 0 identity, 1 GELU (the pre-activation goes to Z when Z != null)
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float *__restrict__ A, int sam, int sak,
                                                   const float *__restrict__ B, int sbk, int sbn,
                                                   const float *__restrict__ bias, int act, float *__restrict__ C, int ldc,
                                                   float *__restrict__ Z)
{
    __shared__ unsigned short sA[kTM][kTK + 4];  // [m][k]
    __shared__ unsigned short sB[kTN][kTK + 4];  // [n][k]
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += kTK) {
        // stage 64 x 16 of A and 16 x 64 of B (4 entries per thread each), rounded to BF16
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = tid + 256 * q, r = idx >> 4, k = idx & 15;
            const int m = m0 + r, n = n0 + r, kk = k0 + k;
            sA[r][k] = bf16_rne((m < M && kk < K) ? A[(size_t)m * sam + (size_t)kk * sak] : 0.f);
            sB[r][k] = bf16_rne((n < N && kk < K) ? B[(size_t)kk * sbk + (size_t)n * sbn] : 0.f);
        }
        __syncthreads();
        // lane supplies A[row = lane % 16][k = 4 (lane / 16) .. +3] and B[k = same][col = lane % 16]
        const int r = lane & 15, kq = 4 * (lane >> 4);
        v4s fa[2], fb[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int j = 0; j < 4; ++j) fa[a][j] = (short)sA[wm + 16 * a + r][kq + j];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[b][j] = (short)sB[wn + 16 * b + r][kq + j];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // result: lane holds D[4 (lane / 16) + v][lane % 16]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + 16 * a + 4 * (lane >> 4) + v, n = n0 + wn + 16 * b + (lane & 15);
                if (m >= M || n >= N) continue;
                float z = acc[a][b][v] + (bias ? bias[n] : 0.f);
                if (Z) Z[(size_t)m * ldc + n] = z;
                if (act == 1) z = 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
                C[(size_t)m * ldc + n] = z;
            }
}

// dZ = dH * GELU'(Z), GELU'(z) = Phi(z) + z phi(z)
__global__ void k_gelu_back(int count, const float *Z, const float *dH, float *dZ)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float z = Z[i];
    const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    dZ[i] = dH[i] * (cdf + z * pdf);
}

// dY = 2 (Y - T) / (rows cols); loss partial sums per block into part[]
__global__ __launch_bounds__(256) void k_mse(int count, float scale, const float *Y, const float *T, float *dY, float *part)
{
    __shared__ float s[256];
    float acc = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) {
        const float d = Y[i] - T[i];
        dY[i] = 2.f * scale * d;
        acc += d * d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0] * scale;
}

// db[n] = sum_m dZ[m][n]: a workgroup of 4 waves per 64 columns, lanes along the columns (coalesced
// rows), wave w summing rows w, w + 4, ... in order, then the four partial sums in wave order
// (a fixed order, so the result is reproducible)
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float *dZ, float *db)
{
    __shared__ float part[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + lane;
    float s = 0.f;
    if (n < N)
        for (int m = wave; m < M; m += 4) s += dZ[(size_t)m * N + n];
    part[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && n < N) db[n] = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

__global__ void k_adam(int count, float *w, const float *g, float *mo, float *ve, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i] + wd * w[i];
    const float m = b1 * mo[i] + (1.f - b1) * gi;
    const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    w[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// rows of a device table picked by an index list
__global__ void k_gather_rows(int rows, int cols, const int *idx, const float *src, float *dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, c = i % cols;
    dst[i] = src[(size_t)idx[r] * cols + c];
}

inline unsigned nb(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct Mlp {
    static constexpr int kLayers = 3;
    int width[kLayers + 1];
    int batch_cap = 0;
    hipStream_t s = nullptr;
    float *param = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;  // all weights then biases
    size_t nparam = 0;
    size_t woff[kLayers], boff[kLayers];
    // activations: H[0] = input, H[l] layer outputs; Z[l] pre-activations; dH / dZ gradients
    float *H[kLayers + 1] = {}, *Zp[kLayers] = {}, *dH[kLayers + 1] = {}, *dZ[kLayers] = {};
    float *target = nullptr, *loss_part = nullptr;
    int *idx = nullptr;
    float *data_x = nullptr, *data_y = nullptr;
    int ndata = 0;
    long long adam_t = 0;
    // normalization of the predictor's features / targets (identity until set)
    std::vector<double> in_mean, in_std, out_mean, out_std;

    float *W(int l) { return param + woff[l]; }
    float *b(int l) { return param + boff[l]; }
    float *gW(int l) { return grad + woff[l]; }
    float *gb(int l) { return grad + boff[l]; }

    int ensure_batch(int B)
    {
        if (B <= batch_cap) return KSH_SUCCESS;
        for (int l = 0; l <= kLayers; ++l) {
            (void)hipFree(H[l]);
            (void)hipFree(dH[l]);
            KSH_HIP(hipMalloc(&H[l], sizeof(float) * (size_t)B * width[l]));
            KSH_HIP(hipMalloc(&dH[l], sizeof(float) * (size_t)B * width[l]));
        }
        for (int l = 0; l < kLayers; ++l) {
            (void)hipFree(Zp[l]);
            (void)hipFree(dZ[l]);
            KSH_HIP(hipMalloc(&Zp[l], sizeof(float) * (size_t)B * width[l + 1]));
            KSH_HIP(hipMalloc(&dZ[l], sizeof(float) * (size_t)B * width[l + 1]));
        }
        (void)hipFree(target);
        (void)hipFree(idx);
        KSH_HIP(hipMalloc(&target, sizeof(float) * (size_t)B * width[kLayers]));
        KSH_HIP(hipMalloc(&idx, sizeof(int) * (size_t)B));
        batch_cap = B;
        return KSH_SUCCESS;
    }

    int gemm(int M, int N, int K, const float *A, int sam, int sak, const float *Bm, int sbk, int sbn, const float *bias,
             int act, float *C, int ldc, float *Z)
    {
        dim3 g(nb(N, kTN), nb(M, kTM));
        hipLaunchKernelGGL(k_gemm_bf16, g, dim3(256), 0, s, M, N, K, A, sam, sak, Bm, sbk, sbn, bias, act, C, ldc, Z);
        KSH_HIP(hipGetLastError());
        return KSH_SUCCESS;
    }

    // H[0] holds the batch (B x width[0]); H[kLayers] gets the output
    int forward(int B)
    {
        ProfScope ps("mlp_forward", s);
        int rc;
        for (int l = 0; l < kLayers; ++l) {
            const int K = width[l], N = width[l + 1];
            const int act = l + 1 < kLayers ? 1 : 0;
            if ((rc = gemm(B, N, K, H[l], K, 1, W(l), N, 1, b(l), act, H[l + 1], N, Zp[l]))) return rc;
        }
        return KSH_SUCCESS;
    }

    // gradients of the mean squared error against `target`; loss (mean) on the host
    int backward(int B, double *loss)
    {
        ProfScope ps("mlp_backward", s);
        const int out = width[kLayers];
        const int count = B * out;
        const unsigned nblk = std::min(64u, nb(count, 256));
        hipLaunchKernelGGL(k_mse, dim3(nblk), dim3(256), 0, s, count, 1.f / (float)count, H[kLayers], target, dZ[kLayers - 1],
                           loss_part);
        KSH_HIP(hipGetLastError());
        int rc;
        for (int l = kLayers - 1; l >= 0; --l) {
            const int K = width[l], N = width[l + 1];
            // dW = H[l]^T dZ[l]   (K x N, reduction over the batch)
            if ((rc = gemm(K, N, B, H[l], 1, K, dZ[l], N, 1, nullptr, 0, gW(l), N, nullptr))) return rc;
            hipLaunchKernelGGL(k_colsum, dim3(nb(N, 64)), dim3(256), 0, s, B, N, dZ[l], gb(l));
            KSH_HIP(hipGetLastError());
            if (l > 0) {
                // dH[l] = dZ[l] W^T, then through GELU of layer l-1
                if ((rc = gemm(B, K, N, dZ[l], N, 1, W(l), 1, N, nullptr, 0, dH[l], K, nullptr))) return rc;
                hipLaunchKernelGGL(k_gelu_back, dim3(nb((size_t)B * K, 256)), dim3(256), 0, s, B * K, Zp[l - 1], dH[l],
                                   dZ[l - 1]);
                KSH_HIP(hipGetLastError());
            }
        }
        if (loss) {
            std::vector<float> part(nblk);
            KSH_HIP(hipMemcpyAsync(part.data(), loss_part, sizeof(float) * nblk, hipMemcpyDeviceToHost, s));
            KSH_HIP(hipStreamSynchronize(s));
            double l = 0.0;
            for (float p : part) l += p;
            *loss = l;
        }
        return KSH_SUCCESS;
    }

    int adam(float lr, float wd)
    {
        ++adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nb(nparam, 256)), dim3(256), 0, s, (int)nparam, param, grad, m1, m2, lr, b1, b2,
                           1e-8f, c1, c2, wd);
        KSH_HIP(hipGetLastError());
        return KSH_SUCCESS;
    }
};

}  // namespace ksh

struct ksh_mlp : ksh::Mlp {};

using namespace ksh;

extern "C" {

int ksh_mlp_create(const int *widths, uint64_t seed, ksh_mlp **out)
{
    if (!widths || !out) return fail(KSH_OTHER_ERROR, "null argument");
    ksh_mlp *m = new ksh_mlp();
    for (int l = 0; l <= Mlp::kLayers; ++l) {
        if (widths[l] <= 0) {
            delete m;
            return fail(KSH_OTHER_ERROR, "layer width must be positive");
        }
        m->width[l] = widths[l];
    }
    size_t off = 0;
    for (int l = 0; l < Mlp::kLayers; ++l) {
        m->woff[l] = off;
        off += (size_t)m->width[l] * m->width[l + 1];
    }
    for (int l = 0; l < Mlp::kLayers; ++l) {
        m->boff[l] = off;
        off += (size_t)m->width[l + 1];
    }
    m->nparam = off;
    auto bail = [&](int rc) {
        ksh_mlp_destroy(m);
        return rc;
    };
    if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess) return bail(fail(KSH_OTHER_ERROR, "stream"));
    for (float **p : {&m->param, &m->grad, &m->m1, &m->m2})
        if (hipMalloc(p, sizeof(float) * off) != hipSuccess) return bail(fail(KSH_OTHER_ERROR, "hipMalloc"));
    if (hipMalloc(&m->loss_part, sizeof(float) * 64) != hipSuccess) return bail(fail(KSH_OTHER_ERROR, "hipMalloc"));
    // He-style init for the GELU layers, zero biases
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(seed ? seed : 20261202ull);
    for (int l = 0; l < Mlp::kLayers; ++l) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)m->width[l]) * (l + 1 < Mlp::kLayers ? 1.f : 0.5f));
        for (size_t i = 0; i < (size_t)m->width[l] * m->width[l + 1]; ++i) h[m->woff[l] + i] = nd(rng);
    }
    if (hipMemcpy(m->param, h.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m1, 0, sizeof(float) * off) != hipSuccess || hipMemset(m->m2, 0, sizeof(float) * off) != hipSuccess)
        return bail(fail(KSH_OTHER_ERROR, "init"));
    *out = m;
    return KSH_SUCCESS;
}

int ksh_mlp_destroy(ksh_mlp *m)
{
    if (!m) return KSH_SUCCESS;
    if (m->s) (void)hipStreamSynchronize(m->s);
    for (float *p : {m->param, m->grad, m->m1, m->m2, m->target, m->loss_part, m->data_x, m->data_y}) (void)hipFree(p);
    for (int l = 0; l <= Mlp::kLayers; ++l) {
        (void)hipFree(m->H[l]);
        (void)hipFree(m->dH[l]);
    }
    for (int l = 0; l < Mlp::kLayers; ++l) {
        (void)hipFree(m->Zp[l]);
        (void)hipFree(m->dZ[l]);
    }
    (void)hipFree(m->idx);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return KSH_SUCCESS;
}

int ksh_mlp_num_params(const ksh_mlp *m, int64_t *n)
{
    if (!m || !n) return fail(KSH_OTHER_ERROR, "null argument");
    *n = (int64_t)m->nparam;
    return KSH_SUCCESS;
}

int ksh_mlp_get_params(const ksh_mlp *m, float *out)
{
    if (!m || !out) return fail(KSH_OTHER_ERROR, "null argument");
    KSH_HIP(hipMemcpy(out, m->param, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
    return KSH_SUCCESS;
}

int ksh_mlp_set_params(ksh_mlp *m, const float *in)
{
    if (!m || !in) return fail(KSH_OTHER_ERROR, "null argument");
    KSH_HIP(hipMemcpy(m->param, in, sizeof(float) * m->nparam, hipMemcpyHostToDevice));
    return KSH_SUCCESS;
}

int ksh_mlp_forward(ksh_mlp *m, const float *x, int batch, float *y)
{
    if (!m || !x || !y || batch <= 0) return fail(KSH_OTHER_ERROR, "bad argument");
    int rc;
    if ((rc = m->ensure_batch(batch))) return rc;
    KSH_HIP(hipMemcpyAsync(m->H[0], x, sizeof(float) * (size_t)batch * m->width[0], hipMemcpyHostToDevice, m->s));
    if ((rc = m->forward(batch))) return rc;
    KSH_HIP(hipMemcpyAsync(y, m->H[Mlp::kLayers], sizeof(float) * (size_t)batch * m->width[Mlp::kLayers],
                           hipMemcpyDeviceToHost, m->s));
    KSH_HIP(hipStreamSynchronize(m->s));
    return KSH_SUCCESS;
}

int ksh_mlp_set_normalization(ksh_mlp *m, const double *in_mean, const double *in_std, const double *out_mean,
                              const double *out_std)
{
    if (!m || !in_mean || !in_std || !out_mean || !out_std) return fail(KSH_OTHER_ERROR, "null argument");
    const int in = m->width[0], out = m->width[Mlp::kLayers];
    m->in_mean.assign(in_mean, in_mean + in);
    m->in_std.assign(in_std, in_std + in);
    m->out_mean.assign(out_mean, out_mean + out);
    m->out_std.assign(out_std, out_std + out);
    return KSH_SUCCESS;
}

int ksh_mlp_predict(ksh_mlp *m, const double *x, int batch, double *y)
{
    if (!m || !x || !y || batch <= 0) return fail(KSH_OTHER_ERROR, "bad argument");
    const int in = m->width[0], out = m->width[Mlp::kLayers];
    const bool norm = !m->in_mean.empty();
    std::vector<float> xf((size_t)batch * in), yf((size_t)batch * out);
    for (int b = 0; b < batch; ++b)
        for (int i = 0; i < in; ++i) {
            const double v = x[(size_t)b * in + i];
            xf[(size_t)b * in + i] = (float)(norm ? (v - m->in_mean[i]) / m->in_std[i] : v);
        }
    int rc;
    if ((rc = ksh_mlp_forward(m, xf.data(), batch, yf.data()))) return rc;
    for (int b = 0; b < batch; ++b)
        for (int o = 0; o < out; ++o) {
            const double v = yf[(size_t)b * out + o];
            y[(size_t)b * out + o] = norm ? v * m->out_std[o] + m->out_mean[o] : v;
        }
    return KSH_SUCCESS;
}

int ksh_mlp_gradients(ksh_mlp *m, const float *x, const float *t, int batch, float *grads, double *loss)
{
    if (!m || !x || !t || batch <= 0) return fail(KSH_OTHER_ERROR, "bad argument");
    int rc;
    if ((rc = m->ensure_batch(batch))) return rc;
    KSH_HIP(hipMemcpyAsync(m->H[0], x, sizeof(float) * (size_t)batch * m->width[0], hipMemcpyHostToDevice, m->s));
    KSH_HIP(hipMemcpyAsync(m->target, t, sizeof(float) * (size_t)batch * m->width[Mlp::kLayers], hipMemcpyHostToDevice, m->s));
    if ((rc = m->forward(batch)) || (rc = m->backward(batch, loss))) return rc;
    if (grads) KSH_HIP(hipMemcpyAsync(grads, m->grad, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->s));
    KSH_HIP(hipStreamSynchronize(m->s));
    return KSH_SUCCESS;
}

int ksh_mlp_train(ksh_mlp *m, const float *x, const float *t, int nsamples, int epochs, int batch, double lr,
                  double weight_decay, uint64_t seed, double *loss_per_epoch)
{
    if (!m || !x || !t || nsamples <= 0 || epochs <= 0 || batch <= 0) return fail(KSH_OTHER_ERROR, "bad argument");
    int rc;
    if ((rc = m->ensure_batch(batch))) return rc;
    const int in = m->width[0], out = m->width[Mlp::kLayers];
    if (nsamples > m->ndata) {
        (void)hipFree(m->data_x);
        (void)hipFree(m->data_y);
        m->data_x = m->data_y = nullptr;
        KSH_HIP(hipMalloc(&m->data_x, sizeof(float) * (size_t)nsamples * in));
        KSH_HIP(hipMalloc(&m->data_y, sizeof(float) * (size_t)nsamples * out));
        m->ndata = nsamples;
    }
    KSH_HIP(hipMemcpyAsync(m->data_x, x, sizeof(float) * (size_t)nsamples * in, hipMemcpyHostToDevice, m->s));
    KSH_HIP(hipMemcpyAsync(m->data_y, t, sizeof(float) * (size_t)nsamples * out, hipMemcpyHostToDevice, m->s));
    std::vector<int> order(nsamples);
    for (int i = 0; i < nsamples; ++i) order[i] = i;
    std::mt19937_64 rng(seed ? seed : 20261202ull);
    std::vector<int> hidx(batch);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(order.begin(), order.end(), rng);
        double acc = 0.0;
        int nb_ = 0;
        for (int b0 = 0; b0 + batch <= nsamples; b0 += batch) {  // full batches only
            KSH_HIP(hipMemcpyAsync(m->idx, order.data() + b0, sizeof(int) * batch, hipMemcpyHostToDevice, m->s));
            hipLaunchKernelGGL(k_gather_rows, dim3(nb((size_t)batch * in, 256)), dim3(256), 0, m->s, batch, in, m->idx,
                               m->data_x, m->H[0]);
            hipLaunchKernelGGL(k_gather_rows, dim3(nb((size_t)batch * out, 256)), dim3(256), 0, m->s, batch, out, m->idx,
                               m->data_y, m->target);
            KSH_HIP(hipGetLastError());
            double l = 0.0;
            if ((rc = m->forward(batch)) || (rc = m->backward(batch, &l)) || (rc = m->adam((float)lr, (float)weight_decay)))
                return rc;
            acc += l;
            ++nb_;
        }
        if (loss_per_epoch) loss_per_epoch[e] = nb_ ? acc / nb_ : 0.0;
    }
    KSH_HIP(hipStreamSynchronize(m->s));
    return KSH_SUCCESS;
}

}  // extern "C"

namespace ksh {

// input and output widths of a network (the path engine checks them against its feature layout)
int mlp_io_widths(const ksh_mlp *m, int *in, int *out)
{
    if (!m) return fail(KSH_OTHER_ERROR, "null network");
    *in = m->width[0];
    *out = m->width[Mlp::kLayers];
    return KSH_SUCCESS;
}

}  // namespace ksh
