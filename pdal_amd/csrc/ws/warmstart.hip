// Learned warm start: MLP training / inference and POD expansion on gfx950.
//
// GEMMs use the CDNA matrix cores: v_mfma_f32_16x16x16_bf16 through
// __builtin_amdgcn_mfma_f32_16x16x16bf16_1k.  A 256-thread workgroup owns a
// 64x64 tile of C as a 2x2 arrangement of 64-lane wavefronts, each holding a
// 2x2 block of 16x16 accumulators (4 FP32 per lane per block).  Each K step of
// 16 stages a 64x16 slice of A and a 16x64 slice of B in LDS as BF16 (FP32
// sources rounded to nearest even on the way in), then every wavefront feeds
// its fragments to the MFMA unit.  Fragment layout of the 16x16x16 BF16 MFMA
// (wave64): A[i][k] lives in lane i + 16*(k/4), element k%4; B[k][j] in lane
// j + 16*(k/4), element k%4; D[i][j] in lane j + 16*(i/4), element i%4.
//
// The MLP is small (64 -> 512 -> 512 -> 512 -> 256 at batch 256 is 0.35
// GFLOP forward), so the whole training step stays on the device: gather of
// the minibatch, forward GEMMs with fused bias/ReLU, loss and its gradient,
// backward GEMMs with the ReLU mask fused, column-sum bias gradients and the
// Adam update over the FP32 master weights.  Reductions are fixed-order.
#include "pdal_warmstart.h"

#include "../common/hip_util.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <mutex>
#include <numeric>
#include <random>
#include <vector>

#define PDAL_API extern "C" __attribute__((visibility("default")))

namespace pdal {
namespace ws {

using bf16x4 = short __attribute__((ext_vector_type(4)));
using f32x4 = float __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;
constexpr int kK = 16;

__device__ __forceinline__ unsigned short to_bf16(float f)
{   // round to nearest even (finite inputs)
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

enum Epilogue { EPI_STORE = 0, EPI_BIAS_RELU = 1, EPI_BIAS = 2, EPI_MASK = 3 };

// C (M x N, ldc) = op(A) (M x K) . op(B) (K x N) with the epilogue:
//   EPI_BIAS_RELU: max(acc + bias[n], 0); EPI_BIAS: acc + bias[n];
//   EPI_MASK: acc * (mask[m*ldm + n] > 0).
// op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k]; op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n].
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256)
    gemm_bf16_kernel(const float *__restrict__ A, int lda, const float *__restrict__ B, int ldb,
                     float *__restrict__ C, int ldc, int M, int N, int K, const float *__restrict__ bias,
                     const float *__restrict__ mask, int ldm)
{
    __shared__ unsigned short As[kTile][kK + 2];
    __shared__ unsigned short Bs[kK][kTile + 2];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += kK) {
        for (int e = tid; e < kTile * kK; e += 256) {
            const int r = e / kK, kk = e % kK;
            const int gm = m0 + r, gk = k0 + kk;
            float v = 0.f;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[r][kk] = to_bf16(v);
        }
        for (int e = tid; e < kTile * kK; e += 256) {
            const int kk = e / kTile, c = e % kTile;
            const int gk = k0 + kk, gn = n0 + c;
            float v = 0.f;
            if (gk < K && gn < N) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            Bs[kk][c] = to_bf16(v);
        }
        __syncthreads();
        const int kb = (lane >> 4) * 4;
        for (int i = 0; i < 2; ++i) {
            const int ra = wm * 32 + i * 16 + (lane & 15);
            bf16x4 a;
            a[0] = (short)As[ra][kb + 0];
            a[1] = (short)As[ra][kb + 1];
            a[2] = (short)As[ra][kb + 2];
            a[3] = (short)As[ra][kb + 3];
            for (int j = 0; j < 2; ++j) {
                const int cb = wn * 32 + j * 16 + (lane & 15);
                bf16x4 b;
                b[0] = (short)Bs[kb + 0][cb];
                b[1] = (short)Bs[kb + 1][cb];
                b[2] = (short)Bs[kb + 2][cb];
                b[3] = (short)Bs[kb + 3][cb];
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + r;
                const int gn = n0 + wn * 32 + j * 16 + (lane & 15);
                if (gm >= M || gn >= N) continue;
                float v = acc[i][j][r];
                if (EPI == EPI_BIAS_RELU) v = fmaxf(v + bias[gn], 0.f);
                else if (EPI == EPI_BIAS) v = v + bias[gn];
                else if (EPI == EPI_MASK) v = mask[(size_t)gm * ldm + gn] > 0.f ? v : 0.f;
                C[(size_t)gm * ldc + gn] = v;
            }
}

// db[n] = sum over rows of G (rows x n), one thread per column, rows in order.
__global__ void __launch_bounds__(256) colsum_kernel(const float *__restrict__ G, int rows, int cols,
                                                     float *__restrict__ db)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += G[(size_t)r * cols + c];
    db[c] = s;
}

// dP = 2 (P - Y) / (rows*cols); loss partial sums per block (fixed order).
__global__ void __launch_bounds__(256) mse_kernel(const float *__restrict__ P, const float *__restrict__ Y, long len,
                                                  float scale, float *__restrict__ dP, float *__restrict__ parts)
{
    __shared__ float red[256];
    float s = 0.f;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long)gridDim.x * 256) {
        const float d = P[i] - Y[i];
        dP[i] = 2.f * d * scale;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) sum_parts_kernel(const float *__restrict__ parts, int n, float scale,
                                                        float *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += parts[i];
    *out = s * scale;
}

__global__ void __launch_bounds__(256)
    adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void __launch_bounds__(256) gather_rows_kernel(const float *__restrict__ src, const int *__restrict__ idx,
                                                          int rows, int cols, float *__restrict__ dst)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)rows * cols) return;
    const int r = (int)(e / cols), c = (int)(e % cols);
    dst[e] = src[(size_t)idx[r] * cols + c];
}

// out[b][d] = mean[d] + sum_k coef[b][k] basis[k][d]; a workgroup covers 256
// columns for 16 rows of the batch, coefficients staged in LDS.
constexpr int kExpandRows = 16;
__global__ void __launch_bounds__(256)
    pod_expand_kernel(const float *__restrict__ basis, const float *__restrict__ mean, int dim, int k,
                      const float *__restrict__ coef, int batch, double *__restrict__ out,
                      const double *__restrict__ lo = nullptr, const double *__restrict__ hi = nullptr)
{
    extern __shared__ float cs[];  // kExpandRows x k
    const int b0 = blockIdx.y * kExpandRows;
    for (int e = threadIdx.x; e < kExpandRows * k; e += 256) {
        const int r = e / k, j = e % k;
        cs[e] = (b0 + r < batch) ? coef[(size_t)(b0 + r) * k + j] : 0.f;
    }
    __syncthreads();
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= dim) return;
    float acc[kExpandRows];
    for (int r = 0; r < kExpandRows; ++r) acc[r] = 0.f;
    for (int j = 0; j < k; ++j) {
        const float w = basis[(size_t)j * dim + d];
        for (int r = 0; r < kExpandRows; ++r) acc[r] += cs[r * k + j] * w;
    }
    const float mu = mean[d];
    const double l = lo ? lo[d] : -INFINITY, h = hi ? hi[d] : INFINITY;
    for (int r = 0; r < kExpandRows; ++r)
        if (b0 + r < batch) out[(size_t)(b0 + r) * dim + d] = fmin(fmax((double)(acc[r] + mu), l), h);
}

// ---------------------------------------------------------------------------

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t s, const float *A, int lda, const float *B, int ldb, float *C, int ldc, int M, int N, int K,
          const float *bias = nullptr, const float *mask = nullptr, int ldm = 0)
{
    const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
    gemm_bf16_kernel<TA, TB, EPI><<<grid, 256, 0, s>>>(A, lda, B, ldb, C, ldc, M, N, K, bias, mask, ldm);
    PDAL_LAUNCHED();
}

inline int grid1(long n) { return (int)std::max<long>(1, (n + 255) / 256); }

struct Mlp {
    std::vector<int> w;  // widths
    int L = 0;
    std::vector<long> woff, boff;  // offsets of W_l and b_l in the parameter vector
    long nparam = 0;
    DBuf<float> params, grads, adam_m, adam_v;
    int cap = 0;  // batch capacity of the activation buffers
    std::vector<DBuf<float>> H, G;  // H[0..L] activations, G[0..L] gradients
    DBuf<float> Y, parts, loss;
    DBuf<int> idx;
    long step = 0;
    hipStream_t s = nullptr;

    float *W(int l) { return params.get() + woff[l]; }
    float *b(int l) { return params.get() + boff[l]; }
    float *dW(int l) { return grads.get() + woff[l]; }
    float *db(int l) { return grads.get() + boff[l]; }

    void ensure(int batch)
    {
        if (batch <= cap) return;
        cap = batch;
        H.clear();
        G.clear();
        H.resize(L + 1);
        G.resize(L + 1);
        for (int l = 0; l <= L; ++l) {
            H[l].alloc((size_t)batch * w[l]);
            G[l].alloc((size_t)batch * w[l]);
        }
        Y.alloc((size_t)batch * w[L]);
        idx.alloc(batch);
    }

    void forward(int B)
    {
        for (int l = 0; l < L; ++l) {
            if (l + 1 < L)
                gemm<false, true, EPI_BIAS_RELU>(s, H[l].get(), w[l], W(l), w[l], H[l + 1].get(), w[l + 1], B,
                                                 w[l + 1], w[l], b(l));
            else
                gemm<false, true, EPI_BIAS>(s, H[l].get(), w[l], W(l), w[l], H[l + 1].get(), w[l + 1], B, w[l + 1],
                                            w[l], b(l));
        }
    }

    // loss of H[L] against Y, gradients into grads
    void backward(int B)
    {
        const long len = (long)B * w[L];
        const int nb = std::min(256, grid1(len));
        mse_kernel<<<nb, 256, 0, s>>>(H[L].get(), Y.get(), len, 1.f / (float)len, G[L].get(), parts.get());
        sum_parts_kernel<<<1, 256, 0, s>>>(parts.get(), nb, 1.f / (float)len, loss.get());
        PDAL_LAUNCHED();
        for (int l = L - 1; l >= 0; --l) {
            // dW_l = G[l+1]^T H[l]  (out x in), db_l = column sums of G[l+1]
            gemm<true, false, EPI_STORE>(s, G[l + 1].get(), w[l + 1], H[l].get(), w[l], dW(l), w[l], w[l + 1], w[l],
                                         B);
            colsum_kernel<<<grid1(w[l + 1]), 256, 0, s>>>(G[l + 1].get(), B, w[l + 1], db(l));
            PDAL_LAUNCHED();
            if (l > 0)  // G[l] = (G[l+1] W_l) * (H[l] > 0)
                gemm<false, false, EPI_MASK>(s, G[l + 1].get(), w[l + 1], W(l), w[l], G[l].get(), w[l], B, w[l],
                                             w[l + 1], nullptr, H[l].get(), w[l]);
        }
    }

    void adam(float lr)
    {
        ++step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
        adam_kernel<<<grid1(nparam), 256, 0, s>>>(params.get(), grads.get(), adam_m.get(), adam_v.get(), nparam, lr,
                                                  b1, b2, 1e-8f, c1, c2);
        PDAL_LAUNCHED();
    }
};

Mlp *create(int n_layers, const int *widths, unsigned long long seed)
{
    auto *m = new Mlp;
    try {
        m->L = n_layers;
        m->w.assign(widths, widths + n_layers + 1);
        for (int l = 0; l < n_layers; ++l) {
            m->woff.push_back(m->nparam);
            m->nparam += (long)m->w[l + 1] * m->w[l];
            m->boff.push_back(m->nparam);
            m->nparam += m->w[l + 1];
        }
        PDAL_HIP(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
        std::vector<float> p(m->nparam, 0.f);
        std::mt19937_64 gen(seed);
        for (int l = 0; l < n_layers; ++l) {
            const float a = std::sqrt(6.f / (float)m->w[l]);
            std::uniform_real_distribution<float> u(-a, a);
            for (long i = 0; i < (long)m->w[l + 1] * m->w[l]; ++i) p[m->woff[l] + i] = u(gen);
        }
        m->params.alloc(m->nparam);
        m->grads.alloc(m->nparam);
        m->adam_m.alloc(m->nparam);
        m->adam_v.alloc(m->nparam);
        m->params.upload(p.data(), m->nparam, m->s);
        PDAL_HIP(hipMemsetAsync(m->adam_m.get(), 0, m->nparam * sizeof(float), m->s));
        PDAL_HIP(hipMemsetAsync(m->adam_v.get(), 0, m->nparam * sizeof(float), m->s));
        PDAL_HIP(hipMemsetAsync(m->grads.get(), 0, m->nparam * sizeof(float), m->s));
        m->parts.alloc(256);
        m->loss.alloc(1);
        PDAL_HIP(hipStreamSynchronize(m->s));
    } catch (...) {
        if (m->s) (void)hipStreamDestroy(m->s);
        delete m;
        throw;
    }
    return m;
}

}  // namespace ws
}  // namespace pdal

using pdal::ws::Mlp;

struct pdal_mlp {
    Mlp *m;
};

template <class F>
static int guarded(const char *what, F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "[pdal-ws] %s failed: %s\n", what, e.what());
        return -2;
    }
}

PDAL_API pdal_mlp_t *pdal_mlp_create(int n_layers, const int *widths, unsigned long long seed)
{
    if (n_layers <= 0 || !widths) return nullptr;
    for (int l = 0; l <= n_layers; ++l)
        if (widths[l] <= 0) return nullptr;
    try {
        return new pdal_mlp{pdal::ws::create(n_layers, widths, seed)};
    } catch (const std::exception &e) {
        fprintf(stderr, "[pdal-ws] mlp create failed: %s\n", e.what());
        return nullptr;
    }
}

PDAL_API void pdal_mlp_free(pdal_mlp_t *h)
{
    if (!h) return;
    if (h->m->s) (void)hipStreamSynchronize(h->m->s);
    hipStream_t s = h->m->s;
    delete h->m;
    if (s) (void)hipStreamDestroy(s);
    delete h;
}

PDAL_API long long pdal_mlp_param_count(const pdal_mlp_t *h) { return h ? h->m->nparam : -1; }

PDAL_API int pdal_mlp_get_params(const pdal_mlp_t *h, float *out)
{
    if (!h || !out) return -1;
    return guarded("get_params", [&] {
        h->m->params.download(out, h->m->nparam, h->m->s);
        PDAL_HIP(hipStreamSynchronize(h->m->s));
    });
}

PDAL_API int pdal_mlp_set_params(pdal_mlp_t *h, const float *in)
{
    if (!h || !in) return -1;
    return guarded("set_params", [&] {
        h->m->params.upload(in, h->m->nparam, h->m->s);
        PDAL_HIP(hipStreamSynchronize(h->m->s));
    });
}

PDAL_API int pdal_mlp_forward(pdal_mlp_t *h, const float *x, int batch, float *y)
{
    if (!h || !x || !y || batch <= 0) return -1;
    return guarded("forward", [&] {
        Mlp &m = *h->m;
        m.ensure(batch);
        m.H[0].upload(x, (size_t)batch * m.w[0], m.s);
        m.forward(batch);
        m.H[m.L].download(y, (size_t)batch * m.w[m.L], m.s);
        PDAL_HIP(hipStreamSynchronize(m.s));
    });
}

PDAL_API int pdal_mlp_loss_grad(pdal_mlp_t *h, const float *x, const float *y, int batch, float *loss, float *grad)
{
    if (!h || !x || !y || batch <= 0) return -1;
    return guarded("loss_grad", [&] {
        Mlp &m = *h->m;
        m.ensure(batch);
        m.H[0].upload(x, (size_t)batch * m.w[0], m.s);
        m.Y.upload(y, (size_t)batch * m.w[m.L], m.s);
        m.forward(batch);
        m.backward(batch);
        if (loss) m.loss.download(loss, 1, m.s);
        if (grad) m.grads.download(grad, m.nparam, m.s);
        PDAL_HIP(hipStreamSynchronize(m.s));
    });
}

PDAL_API int pdal_mlp_train(pdal_mlp_t *h, const float *x, const float *y, int n, int batch, int epochs, float lr,
                            unsigned long long seed, float *loss_hist)
{
    if (!h || !x || !y || n <= 0 || batch <= 0 || batch > n || epochs < 0) return -1;
    return guarded("train", [&] {
        Mlp &m = *h->m;
        m.ensure(batch);
        const int din = m.w[0], dout = m.w[m.L];
        pdal::DBuf<float> X((size_t)n * din), Y((size_t)n * dout);
        X.upload(x, (size_t)n * din, m.s);
        Y.upload(y, (size_t)n * dout, m.s);
        std::vector<int> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::mt19937_64 gen(seed);
        const int steps = n / batch;
        pdal::DBuf<float> losses(std::max(steps, 1));
        std::vector<int> order((size_t)steps * batch);
        pdal::DBuf<int> d_order(std::max<size_t>(order.size(), 1));
        std::vector<float> hl(std::max(steps, 1));
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(perm.begin(), perm.end(), gen);
            std::copy(perm.begin(), perm.begin() + order.size(), order.begin());
            d_order.upload(order.data(), order.size(), m.s);
            for (int st = 0; st < steps; ++st) {
                const int *ix = d_order.get() + (size_t)st * batch;
                pdal::ws::gather_rows_kernel<<<pdal::ws::grid1((long)batch * din), 256, 0, m.s>>>(
                    X.get(), ix, batch, din, m.H[0].get());
                pdal::ws::gather_rows_kernel<<<pdal::ws::grid1((long)batch * dout), 256, 0, m.s>>>(
                    Y.get(), ix, batch, dout, m.Y.get());
                PDAL_LAUNCHED();
                m.forward(batch);
                m.backward(batch);
                PDAL_HIP(hipMemcpyAsync(losses.get() + st, m.loss.get(), sizeof(float), hipMemcpyDeviceToDevice,
                                        m.s));
                m.adam(lr);
            }
            if (loss_hist) {
                losses.download(hl.data(), std::max(steps, 1), m.s);
                PDAL_HIP(hipStreamSynchronize(m.s));
                double s = 0.0;
                for (int st = 0; st < steps; ++st) s += hl[st];
                loss_hist[ep] = steps ? (float)(s / steps) : 0.f;
            }
        }
        PDAL_HIP(hipStreamSynchronize(m.s));
    });
}

PDAL_API int pdal_pod_expand(const float *basis, const float *mean, int dim, int k, const float *coef, int batch,
                             double *out)
{
    if (!basis || !mean || !coef || !out || dim <= 0 || k <= 0 || batch <= 0) return -1;
    if ((size_t)pdal::ws::kExpandRows * k * sizeof(float) > 64 * 1024) return -1;  // LDS staging limit
    return guarded("pod_expand", [&] {
        hipStream_t s;
        PDAL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        pdal::DBuf<float> B((size_t)k * dim), mu(dim), C((size_t)batch * k);
        pdal::DBuf<double> O((size_t)batch * dim);
        B.upload(basis, (size_t)k * dim, s);
        mu.upload(mean, dim, s);
        C.upload(coef, (size_t)batch * k, s);
        const dim3 grid((dim + 255) / 256, (batch + pdal::ws::kExpandRows - 1) / pdal::ws::kExpandRows);
        pdal::ws::pod_expand_kernel<<<grid, 256, pdal::ws::kExpandRows * k * sizeof(float), s>>>(
            B.get(), mu.get(), dim, k, C.get(), batch, O.get());
        PDAL_LAUNCHED();
        O.download(out, (size_t)batch * dim, s);
        PDAL_HIP(hipStreamSynchronize(s));
        (void)hipStreamDestroy(s);
    });
}

// ---------------------------------------------------------------------------
// POD model resident on the device: basis, mean and bounds uploaded once; the
// expansion writes into a device buffer of the caller (a batch's start buffer).

struct pdal_pod {
    hipStream_t s = nullptr;
    int dev = 0;  // device current at create: where B, mu, lo, hi live
    int dim = 0, k = 0;
    pdal::DBuf<float> B, mu;
    pdal::DBuf<double> lo, hi;
    pdal::DBuf<float> C;
    bool has_lo = false, has_hi = false;
};

PDAL_API pdal_pod_t *pdal_pod_create(const float *basis, const float *mean, int dim, int k, const double *lo,
                                     const double *hi)
{
    if (!basis || !mean || dim <= 0 || k <= 0) return nullptr;
    if ((size_t)pdal::ws::kExpandRows * k * sizeof(float) > 64 * 1024) return nullptr;  // LDS staging limit
    pdal_pod *p = new pdal_pod;
    const int rc = guarded("pod_create", [&] {
        PDAL_HIP(hipGetDevice(&p->dev));
        PDAL_HIP(hipStreamCreateWithFlags(&p->s, hipStreamNonBlocking));
        p->dim = dim;
        p->k = k;
        p->B.alloc((size_t)k * dim);
        p->mu.alloc(dim);
        p->B.upload(basis, (size_t)k * dim, p->s);
        p->mu.upload(mean, dim, p->s);
        if (lo) {
            p->lo.alloc(dim);
            p->lo.upload(lo, dim, p->s);
            p->has_lo = true;
        }
        if (hi) {
            p->hi.alloc(dim);
            p->hi.upload(hi, dim, p->s);
            p->has_hi = true;
        }
        PDAL_HIP(hipStreamSynchronize(p->s));
    });
    if (rc != 0) {
        if (p->s) (void)hipStreamDestroy(p->s);
        delete p;
        return nullptr;
    }
    return p;
}

PDAL_API void pdal_pod_free(pdal_pod_t *p)
{
    if (!p) return;
    if (p->s) {
        (void)hipStreamSynchronize(p->s);
        (void)hipStreamDestroy(p->s);
    }
    delete p;
}

PDAL_API int pdal_pod_expand_device(pdal_pod_t *p, const float *coef, int batch, double *out,
                                    long long out_capacity)
{
    if (!p || !coef || !out || batch <= 0) return -1;
    if ((long long)batch * p->dim > out_capacity) return -1;  // the expansion would overrun `out`
    return guarded("pod_expand_device", [&] {
        if (p->C.size() < (size_t)batch * p->k) p->C.alloc((size_t)batch * p->k);
        p->C.upload(coef, (size_t)batch * p->k, p->s);
        const dim3 grid((p->dim + 255) / 256, (batch + pdal::ws::kExpandRows - 1) / pdal::ws::kExpandRows);
        pdal::ws::pod_expand_kernel<<<grid, 256, pdal::ws::kExpandRows * p->k * sizeof(float), p->s>>>(
            p->B.get(), p->mu.get(), p->dim, p->k, p->C.get(), batch, out, p->has_lo ? p->lo.get() : nullptr,
            p->has_hi ? p->hi.get() : nullptr);
        PDAL_LAUNCHED();
        PDAL_HIP(hipStreamSynchronize(p->s));
    });
}

// ---------------------------------------------------------------------------
// Expansion without allocation.  The coefficients of a call are staged in a
// file-scope device array instead of a buffer the pod grows with hipMalloc:
// every GPU the library runs on holds one copy of the array (512 KiB) until
// the process ends.  Host code never uses the array's name as an address: its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and kept.  Calls on one GPU share the array, so calls that
// overlap on one GPU race; nothing in it needs clearing (a call overwrites
// exactly the coefficients it reads).

namespace {

constexpr size_t kStageFloats = 128 * 1024;  // batch x k coefficients of one call, at most
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kStageFloats];

std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

float *scratch_on(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) throw pdal::HipError("device index beyond the staging table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return g_scratch_ptr[dev];
    int cur = 0;
    PDAL_HIP(hipGetDevice(&cur));
    if (cur != dev) PDAL_HIP(hipSetDevice(dev));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) PDAL_HIP(hipSetDevice(cur));
    PDAL_HIP(e);
    if (!ptr) throw pdal::HipError("hipGetSymbolAddress returned no address for the staging array");
    g_scratch_ptr[dev] = static_cast<float *>(ptr);
    return g_scratch_ptr[dev];
}

}  // namespace

PDAL_API int pdal_pod_expand_device_noalloc(pdal_pod_t *p, const float *coef, int batch, double *out,
                                            long long out_capacity)
{
    if (!p || !coef || !out || batch <= 0) return -1;
    if ((long long)batch * p->dim > out_capacity) return -1;  // the expansion would overrun `out`
    if ((size_t)batch * (size_t)p->k > kStageFloats) {
        fprintf(stderr,
                "[pdal-ws] pod_expand_device_noalloc: %d x %d coefficients exceed the %zu-float staging area "
                "(use pdal_pod_expand_device)\n",
                batch, p->k, kStageFloats);
        return -3;
    }
    return guarded("pod_expand_device_noalloc", [&] {
        float *C = scratch_on(p->dev);
        PDAL_HIP(hipMemcpyAsync(C, coef, (size_t)batch * p->k * sizeof(float), hipMemcpyHostToDevice, p->s));
        const dim3 grid((p->dim + 255) / 256, (batch + pdal::ws::kExpandRows - 1) / pdal::ws::kExpandRows);
        pdal::ws::pod_expand_kernel<<<grid, 256, pdal::ws::kExpandRows * p->k * sizeof(float), p->s>>>(
            p->B.get(), p->mu.get(), p->dim, p->k, C, batch, out, p->has_lo ? p->lo.get() : nullptr,
            p->has_hi ? p->hi.get() : nullptr);
        PDAL_LAUNCHED();
        PDAL_HIP(hipStreamSynchronize(p->s));
    });
}
