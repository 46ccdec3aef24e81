// Surrogate network of the hip platform: an ensemble of multilayer
// perceptrons (inputs -> hidden GELU layers -> outputs) trained with Adam,
// every member in the same launches (blockIdx.z = member).
//
// Arithmetic: weights, activations and backward signals are stored in bf16;
// every product runs on the matrix cores (v_mfma_f32_16x16x32_bf16, with one 16x16x16 step for a K tail of 16) with
// fp32 accumulation; biases, pre-activations, outputs, gradients and the
// Adam state are fp32.  Feature dimensions are padded to multiples of 16
// with zeros (the padded weights get zero gradients and stay zero).
//
// All products are written C[m][n] = sum_k A[m][k] * B[n][k] with both
// operands k-contiguous, so activations are kept in both orientations:
//   forward   Z_l  = X_l  . W_l^T + b_l,   X_{l+1} = gelu(Z_l)
//   backward  dW_l = dZ_l^T . X_l           (operands dZ_l^T, X_l^T)
//             dX_l = dZ_l . W_l             (operands dZ_l, W_l^T)
//             dZ_{l-1} = dX_l * gelu'(Z_{l-1})
// No reference counterpart: the reference has no learned model; this is the
// surrogate of its Monte Carlo runs (alsuq/src/run/Runner.cpp:39-90).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "avh_internal.hpp"

using namespace avh;

namespace {

constexpr int MAX_LAYERS = 8;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even; NaN stays NaN
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// exact GELU, x * Phi(x), and its derivative Phi(x) + x * phi(x)
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

enum GemmEpilogue { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_DX = 2, EPI_DW = 3 };

struct GemmArgs {
    const uint16_t* A;  // [M][K] per member
    const uint16_t* B;  // [N][K] per member
    long long a_ms, b_ms;
    int lda, ldb;
    int M, N, K;
    const float* bias;  // [N] per member (EPI_HIDDEN / EPI_OUT)
    long long bias_ms;
    float* C32;         // [M][N] fp32 out: Z, outputs or dW
    long long c32_ms;
    uint16_t* Cb;       // [M][N] bf16 out: activations or dZ_prev
    uint16_t* CbT;      // [N][M] bf16 out, transposed
    long long cb_ms;
    const float* Zp;    // [M][N] pre-activations of the previous layer (EPI_DX)
    long long zp_ms;
};

// One wavefront per 16x16 tile of C, four per workgroup (32x32).  The whole
// wavefront skips a tile outside C, so the matrix instruction always runs
// with every lane active.
template <int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(GemmArgs g) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tm = blockIdx.y * 2 + (w >> 1), tn = blockIdx.x * 2 + (w & 1);
    if (tm * 16 >= g.M || tn * 16 >= g.N) return;
    const int mem = blockIdx.z;
    const uint16_t* A = g.A + mem * g.a_ms + (long long)(tm * 16 + (lane & 15)) * g.lda;
    const uint16_t* B = g.B + mem * g.b_ms + (long long)(tn * 16 + (lane & 15)) * g.ldb;
    const int kq = lane >> 4;  // the lane's quarter of each K step
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    // K steps of 32 on gfx950's v_mfma_f32_16x16x32_bf16 (8 consecutive k
    // per lane, 16-byte loads), then a step of 16 for K = 16 mod 32 (the
    // padded widths 48 and 80)
    int k0 = 0;
    for (; k0 + 32 <= g.K; k0 += 32) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + k0 + 8 * kq);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(B + k0 + 8 * kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    if (k0 < g.K) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(A + k0 + 4 * kq);
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(B + k0 + 4 * kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int n = tn * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = tm * 16 + 4 * (lane >> 4) + r;
        const long long o = (long long)m * g.N + n;
        const long long ot = (long long)n * g.M + m;
        float v = acc[r];
        if (EPI == EPI_HIDDEN) {
            v += g.bias[mem * g.bias_ms + n];
            g.C32[mem * g.c32_ms + o] = v;
            const uint16_t a = f2bf(gelu(v));
            g.Cb[mem * g.cb_ms + o] = a;
            g.CbT[mem * g.cb_ms + ot] = a;
        } else if (EPI == EPI_OUT) {
            g.C32[mem * g.c32_ms + o] = v + g.bias[mem * g.bias_ms + n];
        } else if (EPI == EPI_DX) {
            const uint16_t d = f2bf(v * gelu_grad(g.Zp[mem * g.zp_ms + o]));
            g.Cb[mem * g.cb_ms + o] = d;
            g.CbT[mem * g.cb_ms + ot] = d;
        } else {
            g.C32[mem * g.c32_ms + o] = v;
        }
    }
}

// Minibatch rows: X0 = bf16((x - mean) / std) (both orientations), targets
// (y - mean) / std; rows past `rows` and padded features are zero.
__global__ __launch_bounds__(256) void mlp_gather_kernel(const float* X, const float* Y, const int* idx, long long idx_ms,
                                                         int rows, int nin, int pin, int nout, int pout, int B,
                                                         const float* in_mu, const float* in_sd, const float* out_mu,
                                                         const float* out_sd, uint16_t* X0, uint16_t* X0T, float* Yn) {
    const int mem = blockIdx.z;
    const int r = blockIdx.x;  // row of the minibatch
    const bool live = r < rows;
    const long long src = live ? (long long)idx[mem * idx_ms + r] : 0;
    for (int k = threadIdx.x; k < pin; k += blockDim.x) {
        float v = 0.f;
        if (live && k < nin) v = (X[src * nin + k] - in_mu[k]) / in_sd[k];
        const uint16_t b = f2bf(v);
        X0[((long long)mem * B + r) * pin + k] = b;
        X0T[((long long)mem * pin + k) * B + r] = b;
    }
    if (Y) {
        for (int o = threadIdx.x; o < pout; o += blockDim.x) {
            float v = 0.f;
            if (live && o < nout) v = (Y[src * nout + o] - out_mu[o]) / out_sd[o];
            Yn[((long long)mem * B + r) * pout + o] = v;
        }
    }
}

// Mean-squared error over the real outputs and its gradient
// 2 (yhat - y) / (B * nout) in both orientations; the loss of each member is
// added to loss[mem] (fixed-order tree, so it is reproducible).
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* Yhat, const float* Yn, int B, int nout, int pout,
                                                       uint16_t* dZ, uint16_t* dZT, long long dz_ms, float* loss) {
    __shared__ float red[256];
    const int mem = blockIdx.x;
    const float scale = 2.0f / (float)(B * nout);
    float acc = 0.f;
    for (int t = threadIdx.x; t < B * pout; t += 256) {
        const int r = t / pout, o = t % pout;
        const long long i = (long long)mem * B * pout + t;
        const float d = o < nout ? Yhat[i] - Yn[i] : 0.f;
        acc += d * d;
        const uint16_t g = f2bf(d * scale);
        dZ[mem * dz_ms + (long long)r * pout + o] = g;
        dZT[mem * dz_ms + (long long)o * B + r] = g;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[mem] += red[0] / (float)(B * nout);
}

// db[o] = sum over the minibatch rows of dZ[r][o], in row order.
__global__ __launch_bounds__(256) void mlp_bias_grad_kernel(const uint16_t* dZT, long long dz_ms, int nout, int B,
                                                            float* grad, long long p_ms, long long b_off) {
    const int mem = blockIdx.y;
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= nout) return;
    const uint16_t* row = dZT + mem * dz_ms + (long long)o * B;
    float acc = 0.f;
    for (int r = 0; r < B; ++r) acc += bf2f(row[r]);
    grad[mem * p_ms + b_off + o] = acc;
}

struct LayerDesc {
    int nin, nout;
    long long w_off, b_off;  // in the fp32 parameter block of a member
    long long wb_off;        // in the bf16 weight blocks of a member
};

struct AdamArgs {
    float *p, *m, *v;
    const float* g;
    uint16_t *wb, *wtb;
    long long p_ms, wb_ms;
    LayerDesc L[MAX_LAYERS];
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
    float wd;                       // decoupled weight decay of the weights (not the biases)
};

// Adam (bias-corrected) on the fp32 master parameters of layer blockIdx.y,
// with decoupled weight decay on the weights (AdamW, Loshchilov and Hutter,
// "Decoupled Weight Decay Regularization", 2019: p -= lr * wd * p beside the
// Adam step), then the bf16 copies of the weights in both orientations.
__global__ __launch_bounds__(256) void mlp_adam_kernel(AdamArgs a) {
    const LayerDesc& L = a.L[blockIdx.y];
    const int mem = blockIdx.z;
    const long long nw = (long long)L.nin * L.nout;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= nw + L.nout) return;
    const long long i = mem * a.p_ms + (e < nw ? L.w_off + e : L.b_off + (e - nw));
    const float g = a.g[i];
    const float m = a.b1 * a.m[i] + (1.f - a.b1) * g;
    const float v = a.b2 * a.v[i] + (1.f - a.b2) * (g * g);
    a.m[i] = m;
    a.v[i] = v;
    const float p0 = a.p[i];
    float p = p0 - a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
    if (e < nw && a.wd != 0.f) p = p - a.lr * a.wd * p0;
    a.p[i] = p;
    if (e < nw) {
        const int o = (int)(e / L.nin), k = (int)(e % L.nin);
        const uint16_t b = f2bf(p);
        a.wb[mem * a.wb_ms + L.wb_off + e] = b;
        a.wtb[mem * a.wb_ms + L.wb_off + (long long)k * L.nout + o] = b;
    }
}

// bf16 copies of every weight after the host set the fp32 parameters.
__global__ __launch_bounds__(256) void mlp_cast_kernel(AdamArgs a) {
    const LayerDesc& L = a.L[blockIdx.y];
    const int mem = blockIdx.z;
    const long long nw = (long long)L.nin * L.nout;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= nw) return;
    const int o = (int)(e / L.nin), k = (int)(e % L.nin);
    const uint16_t b = f2bf(a.p[mem * a.p_ms + L.w_off + e]);
    a.wb[mem * a.wb_ms + L.wb_off + e] = b;
    a.wtb[mem * a.wb_ms + L.wb_off + (long long)k * L.nout + o] = b;
}

inline int pad16(int x) { return (x + 15) / 16 * 16; }

}  // namespace

struct avh_mlp {
    avh_ctx* ctx = nullptr;
    avh_mlp_config cfg{};
    int nl = 0, pin = 0, pout = 0, B = 0, M = 0, maxw = 0;
    LayerDesc L[MAX_LAYERS];
    long long P = 0, WB = 0;
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *wb = nullptr, *wtb = nullptr;
    uint16_t *x0 = nullptr, *x0t = nullptr;
    float* z = nullptr;        // [hidden][M][B][width]
    uint16_t *a = nullptr, *at = nullptr;
    float *yhat = nullptr, *yn = nullptr;
    uint16_t *dz[2] = {nullptr, nullptr}, *dzt[2] = {nullptr, nullptr};
    float *mu_in = nullptr, *sd_in = nullptr, *mu_out = nullptr, *sd_out = nullptr;
    float* loss = nullptr;
    float *dx = nullptr, *dy = nullptr;
    long long dcap = 0;
    int* idx = nullptr;
    long long icap = 0;
    long long t = 0;  // Adam steps taken
    long long epochs_done = 0;  // epochs of avh_mlp_train so far (seeds the shuffles)
    float wd = 0.f;   // decoupled weight decay
    hipStream_t stream = nullptr;
};

namespace {

void free_mlp(avh_mlp* n) {
    if (!n) return;
    (void)hipSetDevice(n->ctx->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    void* ptrs[] = {n->params, n->m, n->v, n->grad, n->wb, n->wtb, n->x0, n->x0t, n->z, n->a, n->at, n->yhat,
                    n->yn, n->dz[0], n->dz[1], n->dzt[0], n->dzt[1], n->mu_in, n->sd_in, n->mu_out, n->sd_out,
                    n->loss, n->dx, n->dy, n->idx};
    for (void* p : ptrs) (void)hipFree(p);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

AdamArgs adam_args(const avh_mlp* n) {
    AdamArgs a;
    std::memset(&a, 0, sizeof a);
    a.p = n->params;
    a.m = n->m;
    a.v = n->v;
    a.g = n->grad;
    a.wb = n->wb;
    a.wtb = n->wtb;
    a.p_ms = n->P;
    a.wb_ms = n->WB;
    for (int l = 0; l < n->nl; ++l) a.L[l] = n->L[l];
    return a;
}

unsigned int max_layer_blocks(const avh_mlp* n) {
    long long mx = 0;
    for (int l = 0; l < n->nl; ++l) mx = std::max(mx, (long long)n->L[l].nin * n->L[l].nout + n->L[l].nout);
    return (unsigned int)((mx + 255) / 256);
}

avh_status cast_weights(avh_mlp* n) {
    AdamArgs a = adam_args(n);
    hipLaunchKernelGGL(mlp_cast_kernel, dim3(max_layer_blocks(n), n->nl, n->M), dim3(256), 0, n->stream, a);
    AVH_HIP(hipGetLastError());
    return AVH_OK;
}

avh_status ensure_data(avh_mlp* n, long long rows) {
    if (rows <= n->dcap) return AVH_OK;
    (void)hipFree(n->dx);
    (void)hipFree(n->dy);
    n->dx = n->dy = nullptr;
    n->dcap = 0;
    AVH_HIP(hipMalloc(&n->dx, sizeof(float) * rows * n->cfg.n_in));
    AVH_HIP(hipMalloc(&n->dy, sizeof(float) * rows * n->cfg.n_out));
    n->dcap = rows;
    return AVH_OK;
}

avh_status ensure_idx(avh_mlp* n, long long count) {
    if (count <= n->icap) return AVH_OK;
    (void)hipFree(n->idx);
    n->idx = nullptr;
    n->icap = 0;
    AVH_HIP(hipMalloc(&n->idx, sizeof(int) * count));
    n->icap = count;
    return AVH_OK;
}

template <int EPI>
avh_status gemm(avh_mlp* n, const GemmArgs& g) {
    const dim3 grid((g.N + 31) / 32, (g.M + 31) / 32, n->M);
    hipLaunchKernelGGL(mlp_gemm_kernel<EPI>, grid, dim3(256), 0, n->stream, g);
    AVH_HIP(hipGetLastError());
    return AVH_OK;
}

// rows of the minibatch from the device data through idx[mem][row0 ...]
avh_status gather(avh_mlp* n, long long row0, long long idx_ms, int rows, bool targets) {
    hipLaunchKernelGGL(mlp_gather_kernel, dim3(n->B, 1, n->M), dim3(64), 0, n->stream, n->dx,
                       targets ? n->dy : nullptr, n->idx + row0, idx_ms, rows, n->cfg.n_in, n->pin,
                       n->cfg.n_out, n->pout, n->B, n->mu_in, n->sd_in, n->mu_out, n->sd_out, n->x0, n->x0t,
                       n->yn);
    AVH_HIP(hipGetLastError());
    return AVH_OK;
}

avh_status forward(avh_mlp* n) {
    const int H = n->nl - 1, W = n->maxw, B = n->B;
    for (int l = 0; l < n->nl; ++l) {
        const LayerDesc& L = n->L[l];
        GemmArgs g;
        std::memset(&g, 0, sizeof g);
        if (l == 0) {
            g.A = n->x0;
            g.a_ms = (long long)B * n->pin;
        } else {
            g.A = n->a + (long long)(l - 1) * n->M * B * W;
            g.a_ms = (long long)B * W;
        }
        g.lda = L.nin;
        g.B = n->wb + L.wb_off;
        g.b_ms = n->WB;
        g.ldb = L.nin;
        g.M = B;
        g.N = L.nout;
        g.K = L.nin;
        g.bias = n->params + L.b_off;
        g.bias_ms = n->P;
        if (l < H) {
            g.C32 = n->z + (long long)l * n->M * B * W;
            g.c32_ms = (long long)B * W;
            g.Cb = n->a + (long long)l * n->M * B * W;
            g.CbT = n->at + (long long)l * n->M * B * W;
            g.cb_ms = (long long)B * W;
            avh_status r = gemm<EPI_HIDDEN>(n, g);
            if (r) return r;
        } else {
            g.C32 = n->yhat;
            g.c32_ms = (long long)B * n->pout;
            avh_status r = gemm<EPI_OUT>(n, g);
            if (r) return r;
        }
    }
    return AVH_OK;
}

// Loss, gradients of every parameter into n->grad (no update).
avh_status backward(avh_mlp* n) {
    const int H = n->nl - 1, W = n->maxw, B = n->B;
    const long long dz_ms = (long long)B * std::max(W, n->pout);
    hipLaunchKernelGGL(mlp_loss_kernel, dim3(n->M), dim3(256), 0, n->stream, n->yhat, n->yn, B, n->cfg.n_out,
                       n->pout, n->dz[0], n->dzt[0], dz_ms, n->loss);
    AVH_HIP(hipGetLastError());
    int cur = 0;
    for (int l = n->nl - 1; l >= 0; --l) {
        const LayerDesc& L = n->L[l];
        // dW_l = dZ_l^T . X_l
        GemmArgs g;
        std::memset(&g, 0, sizeof g);
        g.A = n->dzt[cur];
        g.a_ms = dz_ms;
        g.lda = B;
        if (l == 0) {
            g.B = n->x0t;
            g.b_ms = (long long)n->pin * B;
        } else {
            g.B = n->at + (long long)(l - 1) * n->M * B * W;
            g.b_ms = (long long)W * B;
        }
        g.ldb = B;
        g.M = L.nout;
        g.N = L.nin;
        g.K = B;
        g.C32 = n->grad + L.w_off;
        g.c32_ms = n->P;
        avh_status r = gemm<EPI_DW>(n, g);
        if (r) return r;
        hipLaunchKernelGGL(mlp_bias_grad_kernel, dim3((L.nout + 255) / 256, n->M), dim3(256), 0, n->stream,
                           n->dzt[cur], dz_ms, L.nout, B, n->grad, n->P, L.b_off);
        AVH_HIP(hipGetLastError());
        if (l == 0) break;
        // dZ_{l-1} = (dZ_l . W_l) * gelu'(Z_{l-1})
        GemmArgs d;
        std::memset(&d, 0, sizeof d);
        d.A = n->dz[cur];
        d.a_ms = dz_ms;
        d.lda = L.nout;
        d.B = n->wtb + L.wb_off;
        d.b_ms = n->WB;
        d.ldb = L.nout;
        d.M = B;
        d.N = L.nin;
        d.K = L.nout;
        d.Zp = n->z + (long long)(l - 1) * n->M * B * W;
        d.zp_ms = (long long)B * W;
        d.Cb = n->dz[1 - cur];
        d.CbT = n->dzt[1 - cur];
        d.cb_ms = dz_ms;
        if ((r = gemm<EPI_DX>(n, d))) return r;
        cur = 1 - cur;
    }
    (void)H;
    return AVH_OK;
}

avh_status adam_step(avh_mlp* n, float lr) {
    n->t += 1;
    AdamArgs a = adam_args(n);
    a.lr = lr;
    a.wd = n->wd;
    a.b1 = n->cfg.beta1;
    a.b2 = n->cfg.beta2;
    a.eps = n->cfg.eps;
    a.c1 = (float)(1.0 - std::pow((double)n->cfg.beta1, (double)n->t));
    a.c2 = (float)(1.0 - std::pow((double)n->cfg.beta2, (double)n->t));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(max_layer_blocks(n), n->nl, n->M), dim3(256), 0, n->stream, a);
    AVH_HIP(hipGetLastError());
    return AVH_OK;
}

avh_status upload_data(avh_mlp* n, const float* X, const float* Y, long long rows) {
    avh_status r = ensure_data(n, rows);
    if (r) return r;
    AVH_HIP(hipMemcpyAsync(n->dx, X, sizeof(float) * rows * n->cfg.n_in, hipMemcpyHostToDevice, n->stream));
    if (Y) AVH_HIP(hipMemcpyAsync(n->dy, Y, sizeof(float) * rows * n->cfg.n_out, hipMemcpyHostToDevice, n->stream));
    return AVH_OK;
}

}  // namespace

extern "C" {

avh_status avh_mlp_create(avh_ctx* ctx, const avh_mlp_config* cfg, avh_mlp** out) {
    if (!ctx || !cfg || !out) return fail(AVH_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->n_in < 1 || cfg->n_out < 1 || cfg->width < 16 || cfg->width % 16 || cfg->hidden < 1 ||
        cfg->hidden > MAX_LAYERS - 1 || cfg->members < 1 || cfg->members > 64 || cfg->batch < 16 ||
        cfg->batch % 16 || cfg->batch > 4096 || cfg->n_in > 4096 || cfg->n_out > 4096 || cfg->width > 4096)
        return fail(AVH_EINVAL,
                    "bad network shape (width and batch multiples of 16, 1..7 hidden layers, 1..64 members)");
    AVH_HIP(hipSetDevice(ctx->device));
    auto n = new (std::nothrow) avh_mlp;
    if (!n) return fail(AVH_ENOMEM, "host allocation failed");
    n->ctx = ctx;
    n->cfg = *cfg;
    n->nl = cfg->hidden + 1;
    n->pin = pad16(cfg->n_in);
    n->pout = pad16(cfg->n_out);
    n->B = cfg->batch;
    n->M = cfg->members;
    n->maxw = cfg->width;
    long long P = 0, WB = 0;
    for (int l = 0; l < n->nl; ++l) {
        LayerDesc& L = n->L[l];
        L.nin = l == 0 ? n->pin : cfg->width;
        L.nout = l == n->nl - 1 ? n->pout : cfg->width;
        L.w_off = P;
        P += (long long)L.nin * L.nout;
        L.b_off = P;
        P += L.nout;
        L.wb_off = WB;
        WB += (long long)L.nin * L.nout;
    }
    n->P = P;
    n->WB = WB;
    const long long M = n->M, B = n->B, W = cfg->width, H = cfg->hidden;
    const long long dzw = B * std::max<long long>(W, n->pout);
    bool ok = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) == hipSuccess;
    if (!ok) n->stream = nullptr;
    auto alloc = [&](void** p, size_t bytes) { ok = ok && hipMalloc(p, bytes) == hipSuccess; };
    alloc((void**)&n->params, sizeof(float) * M * P);
    alloc((void**)&n->m, sizeof(float) * M * P);
    alloc((void**)&n->v, sizeof(float) * M * P);
    alloc((void**)&n->grad, sizeof(float) * M * P);
    alloc((void**)&n->wb, 2 * M * WB);
    alloc((void**)&n->wtb, 2 * M * WB);
    alloc((void**)&n->x0, 2 * M * B * n->pin);
    alloc((void**)&n->x0t, 2 * M * B * n->pin);
    alloc((void**)&n->z, sizeof(float) * H * M * B * W);
    alloc((void**)&n->a, 2 * H * M * B * W);
    alloc((void**)&n->at, 2 * H * M * B * W);
    alloc((void**)&n->yhat, sizeof(float) * M * B * n->pout);
    alloc((void**)&n->yn, sizeof(float) * M * B * n->pout);
    for (int i = 0; i < 2; ++i) {
        alloc((void**)&n->dz[i], 2 * M * dzw);
        alloc((void**)&n->dzt[i], 2 * M * dzw);
    }
    alloc((void**)&n->mu_in, sizeof(float) * cfg->n_in);
    alloc((void**)&n->sd_in, sizeof(float) * cfg->n_in);
    alloc((void**)&n->mu_out, sizeof(float) * cfg->n_out);
    alloc((void**)&n->sd_out, sizeof(float) * cfg->n_out);
    alloc((void**)&n->loss, sizeof(float) * M);
    if (!ok) {
        free_mlp(n);
        return fail(AVH_ENOMEM, "device allocation failed");
    }
    ok = hipMemset(n->m, 0, sizeof(float) * M * P) == hipSuccess &&
         hipMemset(n->v, 0, sizeof(float) * M * P) == hipSuccess &&
         hipMemset(n->grad, 0, sizeof(float) * M * P) == hipSuccess;
    std::vector<float> zeros(std::max(cfg->n_in, cfg->n_out), 0.f), ones(std::max(cfg->n_in, cfg->n_out), 1.f);
    ok = ok && hipMemcpy(n->mu_in, zeros.data(), sizeof(float) * cfg->n_in, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(n->sd_in, ones.data(), sizeof(float) * cfg->n_in, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(n->mu_out, zeros.data(), sizeof(float) * cfg->n_out, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(n->sd_out, ones.data(), sizeof(float) * cfg->n_out, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        free_mlp(n);
        return fail(AVH_EHIP, "device initialisation failed");
    }
    // initial weights and biases ~ U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from
    // std::mt19937_64(seed + member), padded entries zero
    std::vector<float> h((size_t)(M * P), 0.f);
    for (long long mem = 0; mem < M; ++mem) {
        std::mt19937_64 gen(cfg->seed + (uint64_t)mem);
        std::uniform_real_distribution<float> U(-1.f, 1.f);
        for (int l = 0; l < n->nl; ++l) {
            const LayerDesc& L = n->L[l];
            const int fin = l == 0 ? cfg->n_in : cfg->width;
            const int fout = l == n->nl - 1 ? cfg->n_out : cfg->width;
            const float bound = 1.f / std::sqrt((float)fin);
            for (int o = 0; o < fout; ++o)
                for (int k = 0; k < fin; ++k) h[mem * P + L.w_off + (long long)o * L.nin + k] = U(gen) * bound;
            for (int o = 0; o < fout; ++o) h[mem * P + L.b_off + o] = U(gen) * bound;
        }
    }
    if (hipMemcpy(n->params, h.data(), sizeof(float) * M * P, hipMemcpyHostToDevice) != hipSuccess) {
        free_mlp(n);
        return fail(AVH_EHIP, "device initialisation failed");
    }
    avh_status r = cast_weights(n);
    if (!r && hipStreamSynchronize(n->stream) != hipSuccess) r = fail(AVH_EHIP, "hipStreamSynchronize failed");
    if (r) {
        free_mlp(n);
        return r;
    }
    *out = n;
    return AVH_OK;
}

avh_status avh_mlp_destroy(avh_mlp* n) {
    free_mlp(n);
    return AVH_OK;
}

avh_status avh_mlp_num_params(avh_mlp* n, int64_t* count) {
    if (!n || !count) return fail(AVH_EINVAL, "null argument");
    int64_t c = 0;
    for (int l = 0; l < n->nl; ++l) {
        const int fin = l == 0 ? n->cfg.n_in : n->cfg.width;
        const int fout = l == n->nl - 1 ? n->cfg.n_out : n->cfg.width;
        c += (int64_t)fin * fout + fout;
    }
    *count = c;
    return AVH_OK;
}

avh_status avh_mlp_set_normalization(avh_mlp* n, const float* in_mean, const float* in_std, const float* out_mean,
                                     const float* out_std) {
    if (!n || !in_mean || !in_std || !out_mean || !out_std) return fail(AVH_EINVAL, "null argument");
    for (int i = 0; i < n->cfg.n_in; ++i)
        if (!(in_std[i] > 0)) return fail(AVH_EINVAL, "input scales must be positive");
    for (int i = 0; i < n->cfg.n_out; ++i)
        if (!(out_std[i] > 0)) return fail(AVH_EINVAL, "output scales must be positive");
    AVH_HIP(hipSetDevice(n->ctx->device));
    AVH_HIP(hipStreamSynchronize(n->stream));
    AVH_HIP(hipMemcpy(n->mu_in, in_mean, sizeof(float) * n->cfg.n_in, hipMemcpyHostToDevice));
    AVH_HIP(hipMemcpy(n->sd_in, in_std, sizeof(float) * n->cfg.n_in, hipMemcpyHostToDevice));
    AVH_HIP(hipMemcpy(n->mu_out, out_mean, sizeof(float) * n->cfg.n_out, hipMemcpyHostToDevice));
    AVH_HIP(hipMemcpy(n->sd_out, out_std, sizeof(float) * n->cfg.n_out, hipMemcpyHostToDevice));
    return AVH_OK;
}

// Unpadded parameter order per member: for each layer, W[out][in] row-major
// then b[out].
avh_status avh_mlp_get_params(avh_mlp* n, int member, float* out) {
    if (!n || !out || member < 0 || member >= n->M) return fail(AVH_EINVAL, "bad argument");
    AVH_HIP(hipSetDevice(n->ctx->device));
    AVH_HIP(hipStreamSynchronize(n->stream));
    std::vector<float> h(n->P);
    AVH_HIP(hipMemcpy(h.data(), n->params + member * n->P, sizeof(float) * n->P, hipMemcpyDeviceToHost));
    size_t o = 0;
    for (int l = 0; l < n->nl; ++l) {
        const LayerDesc& L = n->L[l];
        const int fin = l == 0 ? n->cfg.n_in : n->cfg.width;
        const int fout = l == n->nl - 1 ? n->cfg.n_out : n->cfg.width;
        for (int r = 0; r < fout; ++r)
            for (int k = 0; k < fin; ++k) out[o++] = h[L.w_off + (long long)r * L.nin + k];
        for (int r = 0; r < fout; ++r) out[o++] = h[L.b_off + r];
    }
    return AVH_OK;
}

avh_status avh_mlp_set_params(avh_mlp* n, int member, const float* in) {
    if (!n || !in || member < 0 || member >= n->M) return fail(AVH_EINVAL, "bad argument");
    AVH_HIP(hipSetDevice(n->ctx->device));
    AVH_HIP(hipStreamSynchronize(n->stream));
    std::vector<float> h(n->P, 0.f);
    size_t o = 0;
    for (int l = 0; l < n->nl; ++l) {
        const LayerDesc& L = n->L[l];
        const int fin = l == 0 ? n->cfg.n_in : n->cfg.width;
        const int fout = l == n->nl - 1 ? n->cfg.n_out : n->cfg.width;
        for (int r = 0; r < fout; ++r)
            for (int k = 0; k < fin; ++k) h[L.w_off + (long long)r * L.nin + k] = in[o++];
        for (int r = 0; r < fout; ++r) h[L.b_off + r] = in[o++];
    }
    AVH_HIP(hipMemcpy(n->params + member * n->P, h.data(), sizeof(float) * n->P, hipMemcpyHostToDevice));
    avh_status r = cast_weights(n);
    if (r) return r;
    AVH_HIP(hipStreamSynchronize(n->stream));
    return AVH_OK;
}

// Loss and gradients of every member on the same `batch` rows (no update):
// grads[member][num_params] in the get_params order, loss[member].
avh_status avh_mlp_gradients(avh_mlp* n, const float* X, const float* Y, int64_t rows, float* grads, float* loss) {
    if (!n || !X || !Y || !grads || !loss) return fail(AVH_EINVAL, "null argument");
    if (rows != n->B) return fail(AVH_EINVAL, "rows must equal the minibatch size");
    AVH_HIP(hipSetDevice(n->ctx->device));
    avh_status r = upload_data(n, X, Y, rows);
    if (r) return r;
    if ((r = ensure_idx(n, rows))) return r;
    std::vector<int> id(rows);
    for (int i = 0; i < rows; ++i) id[i] = i;
    AVH_HIP(hipMemcpyAsync(n->idx, id.data(), sizeof(int) * rows, hipMemcpyHostToDevice, n->stream));
    AVH_HIP(hipMemsetAsync(n->loss, 0, sizeof(float) * n->M, n->stream));
    if ((r = gather(n, 0, 0, (int)rows, true))) return r;
    if ((r = forward(n))) return r;
    if ((r = backward(n))) return r;
    std::vector<float> h((size_t)(n->M * n->P));
    AVH_HIP(hipMemcpyAsync(h.data(), n->grad, sizeof(float) * n->M * n->P, hipMemcpyDeviceToHost, n->stream));
    AVH_HIP(hipMemcpyAsync(loss, n->loss, sizeof(float) * n->M, hipMemcpyDeviceToHost, n->stream));
    AVH_HIP(hipStreamSynchronize(n->stream));
    int64_t np = 0;
    avh_mlp_num_params(n, &np);
    for (int mem = 0; mem < n->M; ++mem) {
        size_t o = 0;
        const float* src = h.data() + (size_t)mem * n->P;
        float* dst = grads + (size_t)mem * np;
        for (int l = 0; l < n->nl; ++l) {
            const LayerDesc& L = n->L[l];
            const int fin = l == 0 ? n->cfg.n_in : n->cfg.width;
            const int fout = l == n->nl - 1 ? n->cfg.n_out : n->cfg.width;
            for (int rr = 0; rr < fout; ++rr)
                for (int k = 0; k < fin; ++k) dst[o++] = src[L.w_off + (long long)rr * L.nin + k];
            for (int rr = 0; rr < fout; ++rr) dst[o++] = src[L.b_off + rr];
        }
    }
    return AVH_OK;
}

// One Adam step of every member from the gradients of the last
// avh_mlp_gradients / training step (for step-by-step checks).
avh_status avh_mlp_adam_step(avh_mlp* n, float lr) {
    if (!n) return fail(AVH_EINVAL, "null argument");
    AVH_HIP(hipSetDevice(n->ctx->device));
    avh_status r = adam_step(n, lr);
    if (r) return r;
    AVH_HIP(hipStreamSynchronize(n->stream));
    return AVH_OK;
}

// Minibatch Adam over `epochs` passes of the rows; each member visits the
// rows in its own seeded order (Fisher-Yates from std::mt19937_64(seed +
// 1000003 * (epoch + 1) + member), epochs counted over every call on this
// network), floor(rows / batch) steps per epoch;
// learning rate lr * (final_lr_fraction + (1 - final_lr_fraction) * (1 +
// cos(pi * step / steps)) / 2).  epoch_loss[e][member] = mean minibatch loss.
avh_status avh_mlp_train(avh_mlp* n, const float* X, const float* Y, int64_t rows, int epochs, float lr,
                         float final_lr_fraction, float* epoch_loss) {
    if (!n || !X || !Y) return fail(AVH_EINVAL, "null argument");
    if (rows < n->B || epochs < 1) return fail(AVH_EINVAL, "need at least one minibatch and one epoch");
    if (rows > (1ll << 30)) return fail(AVH_EINVAL, "too many rows");
    AVH_HIP(hipSetDevice(n->ctx->device));
    avh_status r = upload_data(n, X, Y, rows);
    if (r) return r;
    if ((r = ensure_idx(n, rows * n->M))) return r;
    const long long steps_per_epoch = rows / n->B;
    const long long total = steps_per_epoch * epochs;
    std::vector<int> perm((size_t)(rows * n->M));
    std::vector<float> lh(n->M);
    long long k = 0;
    for (int e = 0; e < epochs; ++e) {
        for (int mem = 0; mem < n->M; ++mem) {
            std::mt19937_64 gen(n->cfg.seed + 1000003ull * (uint64_t)(n->epochs_done + e + 1) + (uint64_t)mem);
            int* p = perm.data() + (size_t)mem * rows;
            for (long long i = 0; i < rows; ++i) p[i] = (int)i;
            for (long long i = rows - 1; i > 0; --i) {
                const long long j = (long long)(gen() % (uint64_t)(i + 1));
                std::swap(p[i], p[j]);
            }
        }
        AVH_HIP(hipMemcpyAsync(n->idx, perm.data(), sizeof(int) * rows * n->M, hipMemcpyHostToDevice, n->stream));
        AVH_HIP(hipMemsetAsync(n->loss, 0, sizeof(float) * n->M, n->stream));
        for (long long s = 0; s < steps_per_epoch; ++s, ++k) {
            const double c = 0.5 * (1.0 + std::cos(3.14159265358979323846 * (double)k / (double)total));
            const float lr_k = lr * (final_lr_fraction + (1.f - final_lr_fraction) * (float)c);
            if ((r = gather(n, s * n->B, rows, n->B, true))) return r;
            if ((r = forward(n))) return r;
            if ((r = backward(n))) return r;
            if ((r = adam_step(n, lr_k))) return r;
        }
        AVH_HIP(hipMemcpyAsync(lh.data(), n->loss, sizeof(float) * n->M, hipMemcpyDeviceToHost, n->stream));
        AVH_HIP(hipStreamSynchronize(n->stream));
        if (epoch_loss)
            for (int mem = 0; mem < n->M; ++mem) epoch_loss[(size_t)e * n->M + mem] = lh[mem] / (float)steps_per_epoch;
    }
    n->epochs_done += epochs;
    return AVH_OK;
}

avh_status avh_mlp_set_weight_decay(avh_mlp* n, float weight_decay) {
    if (!n) return fail(AVH_EINVAL, "null argument");
    if (!(weight_decay >= 0.f) || weight_decay > 1.f) return fail(AVH_EINVAL, "weight decay must be in [0, 1]");
    n->wd = weight_decay;
    return AVH_OK;
}

// Outputs of every member for `rows` inputs, in data units:
// out[member][row][n_out].
avh_status avh_mlp_predict(avh_mlp* n, const float* X, int64_t rows, float* out) {
    if (!n || !X || !out) return fail(AVH_EINVAL, "null argument");
    if (rows < 1 || rows > (1ll << 30)) return fail(AVH_EINVAL, "bad row count");
    AVH_HIP(hipSetDevice(n->ctx->device));
    avh_status r = upload_data(n, X, nullptr, rows);
    if (r) return r;
    const long long padded = (rows + n->B - 1) / n->B * n->B;
    if ((r = ensure_idx(n, padded))) return r;
    std::vector<int> id(padded);
    for (long long i = 0; i < padded; ++i) id[i] = (int)std::min(i, (long long)rows - 1);
    AVH_HIP(hipMemcpyAsync(n->idx, id.data(), sizeof(int) * padded, hipMemcpyHostToDevice, n->stream));
    std::vector<float> mu(n->cfg.n_out), sd(n->cfg.n_out);
    AVH_HIP(hipMemcpyAsync(mu.data(), n->mu_out, sizeof(float) * n->cfg.n_out, hipMemcpyDeviceToHost, n->stream));
    AVH_HIP(hipMemcpyAsync(sd.data(), n->sd_out, sizeof(float) * n->cfg.n_out, hipMemcpyDeviceToHost, n->stream));
    std::vector<float> h((size_t)(n->M * n->B * n->pout));
    for (long long c0 = 0; c0 < rows; c0 += n->B) {
        const int cnt = (int)std::min<long long>(n->B, rows - c0);
        if ((r = gather(n, c0, 0, cnt, false))) return r;
        if ((r = forward(n))) return r;
        AVH_HIP(hipMemcpyAsync(h.data(), n->yhat, sizeof(float) * h.size(), hipMemcpyDeviceToHost, n->stream));
        AVH_HIP(hipStreamSynchronize(n->stream));
        for (int mem = 0; mem < n->M; ++mem)
            for (int i = 0; i < cnt; ++i)
                for (int o = 0; o < n->cfg.n_out; ++o)
                    out[((size_t)mem * rows + c0 + i) * n->cfg.n_out + o] =
                        h[((size_t)mem * n->B + i) * n->pout + o] * sd[o] + mu[o];
    }
    return AVH_OK;
}

}  // extern "C"
