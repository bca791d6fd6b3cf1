// mlp.hip — learned phase-count / log-K initialiser of the flash (no counterpart
// in the reference; it only proposes a starting point, the flash's own stability
// tests accept or overrule it).
//
// Network: features (scaled ln P, scaled T, z[nc]) -> 3 hidden layers of 64
// SiLU units -> 3 phase-count logits (1, 2, 3 phases) + nc values of ln K_go.
// Training is FP32 with this file's own kernels: one fused kernel per mini-batch
// tile runs forward, loss gradient and backward for 64 states with every
// activation in LDS and writes the tile's partial weight gradients; a
// fixed-order reduction over tiles (deterministic) feeds Adam.  Inference rounds
// weights and activations to BF16 (FP32 accumulation).
//
// Work mapping of the fused kernel: 256 threads = 4 wavefronts own a tile of
// 64 states; thread t computes a 4 x 4 sub-tile (states 4*(t>>4)+i, units
// 4*(t&15)+q) of every 64 x 64 layer product, so the 64x64 weight matrix and
// the 64 x 65 activation rows are read from LDS with at most 4 distinct
// addresses per operand per wavefront.  A layer's pre-activations stay in the
// owning thread's registers for the backward pass (the backward sub-tile of
// dA_l uses the same (state, unit) ownership).
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>
#include <cstdio>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "../../include/nonenone_amd.h"

namespace
{

constexpr int kH = 64;        // hidden width
constexpr int kTile = 64;     // states per block
constexpr int kThreads = 256;
constexpr int kLd = kH + 1;   // padded LDS row (bank spread)
constexpr int kClasses = 3;
constexpr int64_t kMlpMaxBatch = NN_MLP_MAX_BATCH;

thread_local char g_mlpError[256] = "";

int mlpFail(int code, const char *msg)
{
    std::snprintf(g_mlpError, sizeof(g_mlpError), "%s", msg);
    return code;
}

template <int NC>
struct Dims
{
    static constexpr int NIN = NC + 2;
    static constexpr int NOUT = kClasses + NC;
    // flat parameter layout: per layer W (out x in, row-major) then b (out)
    static constexpr int W1 = 0, B1 = W1 + kH * NIN;
    static constexpr int W2 = B1 + kH, B2 = W2 + kH * kH;
    static constexpr int W3 = B2 + kH, B3 = W3 + kH * kH;
    static constexpr int W4 = B3 + kH, B4 = W4 + NOUT * kH;
    static constexpr int NPARAM = B4 + NOUT;
};

__device__ inline float sigmoidf(float a) { return 1.0f / (1.0f + ::expf(-a)); }
__device__ inline float siluf(float a) { return a * sigmoidf(a); }
__device__ inline float siluGrad(float a)
{
    const float s = sigmoidf(a);
    return s * (1.0f + a * (1.0f - s));
}

// Round to the nearest BF16 (ties to even), returned as the FP32 it denotes.
__host__ __device__ inline float bf16Round(float f)
{
    unsigned int u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    __builtin_memcpy(&r, &u, 4);
    return r;
}

// acc[i][q] = b[j] + sum_k Hin[s_i][k] W[j_q][k], s_i = 4gs+i, j_q = 4gn+q.
template <int IN, int OUT, int LDI>
__device__ inline void layerForward(const float *W, const float *b, const float *Hin, int gs, int gn,
                                    float acc[4][4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        const int j = gn * 4 + q;
        const float bj = j < OUT ? b[j] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i][q] = bj;
    }
#pragma unroll 8
    for (int k = 0; k < IN; ++k)
    {
        float a[4], w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Hin[(gs * 4 + i) * LDI + k];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const int j = gn * 4 + q;
            w[q] = j < OUT ? W[j * IN + k] : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[i][q] = ::fmaf(a[i], w[q], acc[i][q]);
    }
}

template <int IN, int LDI>
__device__ inline void storeTile(float *H, int gs, int gn, const float v[4][4], bool round)
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const int j = gn * 4 + q;
            if (j < IN)
                H[(gs * 4 + i) * LDI + j] = round ? bf16Round(v[i][q]) : v[i][q];
        }
}

template <int NC>
struct Smem
{
    using D = Dims<NC>;
    float W[D::NPARAM];
    float X[kTile][D::NIN + 1];
    float H[3][kTile * kLd];
    float G[2][kTile * kLd];
    float O[kTile][D::NOUT + 1];
    float red[kThreads / 64];
};

// Loads the tile's features (zero rows past n) and the parameters into LDS.
template <int NC>
__device__ inline void loadTile(Smem<NC> &sm, const float *__restrict__ params,
                                const float *__restrict__ X, int64_t n, int64_t base, bool round)
{
    using D = Dims<NC>;
    for (int p = threadIdx.x; p < D::NPARAM; p += kThreads)
        sm.W[p] = round ? bf16Round(params[p]) : params[p];
    for (int e = threadIdx.x; e < kTile * D::NIN; e += kThreads)
    {
        const int s = e / D::NIN, k = e % D::NIN;
        const int64_t g = base + s;
        const float v = g < n ? X[g * D::NIN + k] : 0.0f;
        sm.X[s][k] = round ? bf16Round(v) : v;
    }
    __syncthreads();
}

// Forward through the three hidden layers and the output layer.  Keeps the
// thread's pre-activation sub-tiles in A (for the backward pass).
template <int NC>
__device__ inline void forwardTile(Smem<NC> &sm, int gs, int gn, float A[3][4][4], bool round)
{
    using D = Dims<NC>;
    float acc[4][4], h[4][4];
    layerForward<D::NIN, kH, D::NIN + 1>(sm.W + D::W1, sm.W + D::B1, &sm.X[0][0], gs, gn, acc);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) { A[0][i][q] = acc[i][q]; h[i][q] = siluf(acc[i][q]); }
    storeTile<kH, kLd>(sm.H[0], gs, gn, h, round);
    __syncthreads();
    layerForward<kH, kH, kLd>(sm.W + D::W2, sm.W + D::B2, sm.H[0], gs, gn, acc);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) { A[1][i][q] = acc[i][q]; h[i][q] = siluf(acc[i][q]); }
    storeTile<kH, kLd>(sm.H[1], gs, gn, h, round);
    __syncthreads();
    layerForward<kH, kH, kLd>(sm.W + D::W3, sm.W + D::B3, sm.H[1], gs, gn, acc);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) { A[2][i][q] = acc[i][q]; h[i][q] = siluf(acc[i][q]); }
    storeTile<kH, kLd>(sm.H[2], gs, gn, h, round);
    __syncthreads();
    layerForward<kH, D::NOUT, kLd>(sm.W + D::W4, sm.W + D::B4, sm.H[2], gs, gn, acc);
    storeTile<D::NOUT, D::NOUT + 1>(&sm.O[0][0], gs, gn, acc, false);
    __syncthreads();
}

// Partial weight/bias gradient of one layer over the tile:
// dW[j][k] = sum_s dA[s][j] Hin[s][k], db[j] = sum_s dA[s][j].
template <int IN, int OUT, int LDI, int LDG>
__device__ inline void layerWeightGrad(const float *dA, const float *Hin, float *wsW, float *wsB)
{
    const int t = threadIdx.x;
    const int gj = t >> 4, gk = t & 15;
    if (gj * 4 < OUT && gk * 4 < IN)
    {
        float acc[4][4] = {};
#pragma unroll 8
        for (int s = 0; s < kTile; ++s)
        {
            float d[4], h[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) { const int j = gj * 4 + i; d[i] = j < OUT ? dA[s * LDG + j] : 0.0f; }
#pragma unroll
            for (int q = 0; q < 4; ++q) { const int k = gk * 4 + q; h[q] = k < IN ? Hin[s * LDI + k] : 0.0f; }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[i][q] = ::fmaf(d[i], h[q], acc[i][q]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const int j = gj * 4 + i, k = gk * 4 + q;
                if (j < OUT && k < IN)
                    wsW[j * IN + k] = acc[i][q];
            }
    }
    if (t < OUT)
    {
        float s0 = 0.0f;
        for (int s = 0; s < kTile; ++s) s0 += dA[s * LDG + t];
        wsB[t] = s0;
    }
}

// dA_prev[s][k] = (sum_j dA[s][j] W[j][k]) * silu'(A_prev[s][k]) on the thread's
// (state, unit) sub-tile.
template <int IN, int OUT, int LDG>
__device__ inline void layerInputGrad(const float *W, const float *dA, const float Aprev[4][4],
                                      float *dAprev, int gs, int gn)
{
    float acc[4][4] = {};
#pragma unroll 4
    for (int j = 0; j < OUT; ++j)
    {
        float d[4], w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = dA[(gs * 4 + i) * LDG + j];
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = W[j * IN + gn * 4 + q];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[i][q] = ::fmaf(d[i], w[q], acc[i][q]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            dAprev[(gs * 4 + i) * kLd + gn * 4 + q] = acc[i][q] * siluGrad(Aprev[i][q]);
}

// Fused forward + loss + backward over one tile of 64 states.  Loss per state:
//   -ln softmax(o[0..2])[label] + w * mask * mean_k (o[3+k] - logk[k])^2,
// scaled by `scale` (1/batch).  Writes the tile's partial gradient (NPARAM
// floats) to ws + blockIdx.x * NPARAM and its partial loss to lossPart.
template <int NC>
__global__ void __launch_bounds__(kThreads) mlpTrainTileKernel(
    int64_t n, const float *__restrict__ params, const float *__restrict__ X,
    const int32_t *__restrict__ label, const float *__restrict__ logk,
    const uint8_t *__restrict__ mask, float logkWeight, float scale, float *__restrict__ ws,
    float *__restrict__ lossPart)
{
    using D = Dims<NC>;
    __shared__ Smem<NC> sm;
    const int64_t base = int64_t(blockIdx.x) * kTile;
    const int gs = threadIdx.x >> 4, gn = threadIdx.x & 15;
    loadTile<NC>(sm, params, X, n, base, false);
    float A[3][4][4];
    forwardTile<NC>(sm, gs, gn, A, false);

    // loss gradient at the outputs (one thread per state, wavefront 0)
    float loss = 0.0f;
    float *dO = sm.G[0];  // row stride kLd
    if (threadIdx.x < kTile)
    {
        const int s = threadIdx.x;
        const int64_t g = base + s;
        float d[D::NOUT];
        for (int j = 0; j < D::NOUT; ++j) d[j] = 0.0f;
        if (g < n)
        {
            const float *o = sm.O[s];
            const float m0 = fmaxf(o[0], fmaxf(o[1], o[2]));
            float e[3], se = 0.0f;
            for (int c = 0; c < 3; ++c) { e[c] = ::expf(o[c] - m0); se += e[c]; }
            const int lab = label[g];
            loss = -(o[lab] - m0 - ::logf(se));
            for (int c = 0; c < 3; ++c) d[c] = (e[c] / se - (c == lab ? 1.0f : 0.0f)) * scale;
            if (mask[g])
            {
                float sq = 0.0f;
                for (int k = 0; k < NC; ++k)
                {
                    const float r = o[3 + k] - logk[g * NC + k];
                    sq += r * r;
                    d[3 + k] = logkWeight * (2.0f / NC) * r * scale;
                }
                loss += logkWeight * sq / NC;
            }
            loss *= scale;
        }
        for (int j = 0; j < D::NOUT; ++j) dO[s * kLd + j] = d[j];
        // wavefront sum of the tile's loss (fixed order)
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off);
        if (s == 0)
            lossPart[blockIdx.x] = loss;
    }
    __syncthreads();

    float *wsb = ws + int64_t(blockIdx.x) * D::NPARAM;
    // output layer
    layerWeightGrad<kH, D::NOUT, kLd, kLd>(sm.G[0], sm.H[2], wsb + D::W4, wsb + D::B4);
    layerInputGrad<kH, D::NOUT, kLd>(sm.W + D::W4, sm.G[0], A[2], sm.G[1], gs, gn);
    __syncthreads();
    // hidden layer 3
    layerWeightGrad<kH, kH, kLd, kLd>(sm.G[1], sm.H[1], wsb + D::W3, wsb + D::B3);
    layerInputGrad<kH, kH, kLd>(sm.W + D::W3, sm.G[1], A[1], sm.G[0], gs, gn);
    __syncthreads();
    // hidden layer 2
    layerWeightGrad<kH, kH, kLd, kLd>(sm.G[0], sm.H[0], wsb + D::W2, wsb + D::B2);
    layerInputGrad<kH, kH, kLd>(sm.W + D::W2, sm.G[0], A[0], sm.G[1], gs, gn);
    __syncthreads();
    // hidden layer 1 (input features)
    layerWeightGrad<D::NIN, kH, D::NIN + 1, kLd>(sm.G[1], &sm.X[0][0], wsb + D::W1, wsb + D::B1);
}

// grad[p] = sum over tiles of ws[tile][p] (fixed order); with `adam`, one Adam
// step on params (m, v moments; bias corrections bc1 = 1 - b1^t, bc2 = 1 - b2^t).
__global__ void mlpReduceAdamKernel(int64_t nparam, int64_t ntile, const float *__restrict__ ws,
                                    float *__restrict__ grad, int adam, float *__restrict__ params,
                                    float *__restrict__ m, float *__restrict__ v, float lr,
                                    float b1, float b2, float eps, float bc1, float bc2)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= nparam)
        return;
    float g = 0.0f;
    for (int64_t t = 0; t < ntile; ++t) g += ws[t * nparam + p];
    if (grad)
        grad[p] = g;
    if (adam)
    {
        const float mp = b1 * m[p] + (1.0f - b1) * g;
        const float vp = b2 * v[p] + (1.0f - b2) * g * g;
        m[p] = mp;
        v[p] = vp;
        params[p] -= lr * (mp / bc1) / (::sqrtf(vp / bc2) + eps);
    }
}

__global__ void mlpLossReduceKernel(int64_t ntile, const float *__restrict__ part,
                                    float *__restrict__ loss)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int64_t t = threadIdx.x; t < ntile; t += 256) s += part[t];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        loss[0] = red[0];
}

// BF16 inference: class probabilities (softmax of the 3 logits) and log-K.
template <int NC>
__global__ void __launch_bounds__(kThreads) mlpPredictKernel(
    int64_t n, const float *__restrict__ params, const float *__restrict__ X,
    float *__restrict__ prob, float *__restrict__ logk)
{
    using D = Dims<NC>;
    __shared__ Smem<NC> sm;
    const int64_t base = int64_t(blockIdx.x) * kTile;
    const int gs = threadIdx.x >> 4, gn = threadIdx.x & 15;
    loadTile<NC>(sm, params, X, n, base, true);
    float A[3][4][4];
    forwardTile<NC>(sm, gs, gn, A, true);
    if (threadIdx.x < kTile)
    {
        const int s = threadIdx.x;
        const int64_t g = base + s;
        if (g < n)
        {
            const float *o = sm.O[s];
            const float m0 = fmaxf(o[0], fmaxf(o[1], o[2]));
            float e[3], se = 0.0f;
            for (int c = 0; c < 3; ++c) { e[c] = ::expf(o[c] - m0); se += e[c]; }
            for (int c = 0; c < 3; ++c) prob[g * 3 + c] = e[c] / se;
            for (int k = 0; k < NC; ++k) logk[g * NC + k] = o[3 + k];
        }
    }
}

// Features of a flash input: (ln P - 15.5) / 1.5, (T - 450) / 150, z[0..nc-1].
template <int NC>
__global__ void mlpFeatureKernel(int64_t n, const double *__restrict__ P,
                                 const double *__restrict__ T, const double *__restrict__ z,
                                 float *__restrict__ X)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    constexpr int NIN = NC + 2;
    const double p = P[s] > 1.0 ? P[s] : 1.0;
    X[s * NIN + 0] = float((::log(p) - 15.5) / 1.5);
    X[s * NIN + 1] = float((T[s] - 450.0) / 150.0);
    for (int c = 0; c < NC; ++c) X[s * NIN + 2 + c] = float(z[s * NC + c]);
}

// Flash hints from predictions: 1 = confident single phase, 2 = confident
// two or three phases (log-K seed), 0 = no hint (the plain flash); bits 2-3 =
// most probable phase count - 1 (ordering only).
template <int NC>
__global__ void mlpHintKernel(int64_t n, const float *__restrict__ prob,
                              const float *__restrict__ logk, float threshold, int32_t modes,
                              uint8_t *__restrict__ hint, double *__restrict__ seedLogK)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const float p0 = prob[s * 3], p1 = prob[s * 3 + 1], p2 = prob[s * 3 + 2];
    uint8_t h = 0;
    if (p0 >= threshold && (modes & NN_HINT_SINGLE_PHASE))
        h = NN_HINT_SINGLE_PHASE;
    else if ((p1 >= threshold || p2 >= threshold) && (modes & NN_HINT_LOG_K))
        h = NN_HINT_LOG_K;
    if (modes & NN_HINT_ORDER)
    {
        const int cls = (p1 > p0 && p1 >= p2) ? 1 : ((p2 > p0 && p2 > p1) ? 2 : 0);
        h = uint8_t(h | (cls << 2));
    }
    hint[s] = h;
    for (int c = 0; c < NC; ++c) seedLogK[s * NC + c] = double(logk[s * NC + c]);
}

template <class F>
int dispatchMlpNc(int nc, F &&f)
{
    switch (nc)
    {
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 6: return f(std::integral_constant<int, 6>{});
    default: return mlpFail(NN_E_UNSUPPORTED, "MLP compiled for nc = 3, 4, 6");
    }
}

int paramCount(int nc)
{
    switch (nc)
    {
    case 3: return Dims<3>::NPARAM;
    case 4: return Dims<4>::NPARAM;
    case 6: return Dims<6>::NPARAM;
    default: return 0;
    }
}

#define MLP_HIP(call)                                                              \
    do                                                                             \
    {                                                                              \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
        {                                                                          \
            std::snprintf(g_mlpError, sizeof(g_mlpError), "%s: %s", #call,        \
                          hipGetErrorString(e_));                                  \
            return NN_E_HIP;                                                       \
        }                                                                          \
    } while (0)

} // namespace

struct nn_mlp
{
    int nc = 0;
    int64_t nparam = 0;
    int64_t step = 0;
    float *dParams = nullptr, *dM = nullptr, *dV = nullptr;
    float *dWs = nullptr, *dLossPart = nullptr;
    int64_t wsTiles = 0;
};

namespace
{
int ensureWorkspace(nn_mlp *m, int64_t tiles)
{
    if (tiles <= m->wsTiles)
        return NN_OK;
    if (m->dWs) (void)hipFree(m->dWs);
    if (m->dLossPart) (void)hipFree(m->dLossPart);
    m->dWs = nullptr;
    m->dLossPart = nullptr;
    m->wsTiles = 0;
    if (hipMalloc(&m->dWs, size_t(tiles) * m->nparam * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->dLossPart, size_t(tiles) * sizeof(float)) != hipSuccess)
        return mlpFail(NN_E_OUT_OF_MEMORY, "MLP gradient workspace allocation failed");
    m->wsTiles = tiles;
    return NN_OK;
}

int lossGrad(nn_mlp *m, int64_t n, const float *X, const int32_t *label, const float *logk,
             const uint8_t *mask, float logkWeight, float *loss, float *grad, bool adam, float lr,
             hipStream_t st)
{
    if (!m || n <= 0 || !X || !label || !logk || !mask || !loss)
        return mlpFail(NN_E_INVALID_ARGUMENT, "invalid MLP argument");
    if (n > kMlpMaxBatch)
        return mlpFail(NN_E_INVALID_ARGUMENT, "MLP batch above NN_MLP_MAX_BATCH");
    const int64_t tiles = (n + kTile - 1) / kTile;
    if (const int rc = ensureWorkspace(m, tiles); rc != NN_OK)
        return rc;
    const float scale = 1.0f / float(n);
    const int rc = dispatchMlpNc(m->nc, [&](auto tag) {
        constexpr int NC = decltype(tag)::value;
        mlpTrainTileKernel<NC><<<unsigned(tiles), kThreads, 0, st>>>(
            n, m->dParams, X, label, logk, mask, logkWeight, scale, m->dWs, m->dLossPart);
        MLP_HIP(hipGetLastError());
        return NN_OK;
    });
    if (rc != NN_OK)
        return rc;
    float bc1 = 1.0f, bc2 = 1.0f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam)
    {
        m->step += 1;
        bc1 = 1.0f - std::pow(b1, float(m->step));
        bc2 = 1.0f - std::pow(b2, float(m->step));
    }
    const unsigned g = unsigned((m->nparam + 255) / 256);
    mlpReduceAdamKernel<<<g, 256, 0, st>>>(m->nparam, tiles, m->dWs, grad, adam ? 1 : 0,
                                           m->dParams, m->dM, m->dV, lr, b1, b2, 1.0e-8f, bc1,
                                           bc2);
    MLP_HIP(hipGetLastError());
    mlpLossReduceKernel<<<1, 256, 0, st>>>(tiles, m->dLossPart, loss);
    MLP_HIP(hipGetLastError());
    return NN_OK;
}
} // namespace

extern "C" {

const char *nn_mlp_last_error(void) { return g_mlpError; }

int nn_mlp_create(int32_t nc, uint64_t seed, nn_mlp **out)
{
    if (!out)
        return mlpFail(NN_E_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    const int np = paramCount(nc);
    if (np == 0)
        return mlpFail(NN_E_UNSUPPORTED, "MLP compiled for nc = 3, 4, 6");
    auto *m = new (std::nothrow) nn_mlp;
    if (!m)
        return mlpFail(NN_E_OUT_OF_MEMORY, "host allocation failed");
    m->nc = nc;
    m->nparam = np;
    // PyTorch-style Linear initialisation: U(-1/sqrt(in), 1/sqrt(in)) for W and b.
    std::vector<float> h(size_t(np), 0.0f);
    std::mt19937_64 rng(seed);
    const int nin = nc + 2, nout = kClasses + nc;
    const int ins[4] = {nin, kH, kH, kH}, outs[4] = {kH, kH, kH, nout};
    size_t off = 0;
    for (int l = 0; l < 4; ++l)
    {
        std::uniform_real_distribution<float> u(-1.0f / std::sqrt(float(ins[l])),
                                                1.0f / std::sqrt(float(ins[l])));
        for (int e = 0; e < ins[l] * outs[l] + outs[l]; ++e) h[off++] = u(rng);
    }
    const size_t bytes = size_t(np) * sizeof(float);
    if (hipMalloc(&m->dParams, bytes) != hipSuccess || hipMalloc(&m->dM, bytes) != hipSuccess ||
        hipMalloc(&m->dV, bytes) != hipSuccess)
    {
        nn_mlp_destroy(m);
        return mlpFail(NN_E_OUT_OF_MEMORY, "MLP parameter allocation failed");
    }
    if (hipMemcpy(m->dParams, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->dM, 0, bytes) != hipSuccess || hipMemset(m->dV, 0, bytes) != hipSuccess)
    {
        nn_mlp_destroy(m);
        return mlpFail(NN_E_HIP, "MLP parameter upload failed");
    }
    *out = m;
    return NN_OK;
}

int nn_mlp_destroy(nn_mlp *m)
{
    if (!m)
        return NN_OK;
    if (m->dParams) (void)hipFree(m->dParams);
    if (m->dM) (void)hipFree(m->dM);
    if (m->dV) (void)hipFree(m->dV);
    if (m->dWs) (void)hipFree(m->dWs);
    if (m->dLossPart) (void)hipFree(m->dLossPart);
    delete m;
    return NN_OK;
}

int64_t nn_mlp_param_count(const nn_mlp *m) { return m ? m->nparam : 0; }

float *nn_mlp_params(nn_mlp *m) { return m ? m->dParams : nullptr; }

int nn_mlp_get_params(nn_mlp *m, float *d_out, void *stream)
{
    if (!m || !d_out)
        return mlpFail(NN_E_INVALID_ARGUMENT, "null argument");
    MLP_HIP(hipMemcpyAsync(d_out, m->dParams, size_t(m->nparam) * sizeof(float),
                           hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
    return NN_OK;
}

int nn_mlp_set_params(nn_mlp *m, const float *d_in, void *stream)
{
    if (!m || !d_in)
        return mlpFail(NN_E_INVALID_ARGUMENT, "null argument");
    MLP_HIP(hipMemcpyAsync(m->dParams, d_in, size_t(m->nparam) * sizeof(float),
                           hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
    return NN_OK;
}

int nn_mlp_reset_optimizer(nn_mlp *m)
{
    if (!m)
        return mlpFail(NN_E_INVALID_ARGUMENT, "null MLP");
    m->step = 0;
    MLP_HIP(hipMemset(m->dM, 0, size_t(m->nparam) * sizeof(float)));
    MLP_HIP(hipMemset(m->dV, 0, size_t(m->nparam) * sizeof(float)));
    return NN_OK;
}

int nn_mlp_features(nn_mlp *m, int64_t n, const double *d_P, const double *d_T, const double *d_z,
                    float *d_X, void *stream)
{
    if (!m || n < 0 || (n > 0 && (!d_P || !d_T || !d_z || !d_X)))
        return mlpFail(NN_E_INVALID_ARGUMENT, "invalid MLP argument");
    if (n == 0)
        return NN_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    return dispatchMlpNc(m->nc, [&](auto tag) {
        constexpr int NC = decltype(tag)::value;
        mlpFeatureKernel<NC><<<unsigned((n + 255) / 256), 256, 0, st>>>(n, d_P, d_T, d_z, d_X);
        MLP_HIP(hipGetLastError());
        return NN_OK;
    });
}

int nn_mlp_loss_grad(nn_mlp *m, int64_t n, const float *d_X, const int32_t *d_label,
                     const float *d_logk, const uint8_t *d_mask, float logk_weight, float *d_loss,
                     float *d_grad, void *stream)
{
    if (!d_grad)
        return mlpFail(NN_E_INVALID_ARGUMENT, "null gradient output");
    return lossGrad(m, n, d_X, d_label, d_logk, d_mask, logk_weight, d_loss, d_grad, false, 0.0f,
                    static_cast<hipStream_t>(stream));
}

int nn_mlp_train_step(nn_mlp *m, int64_t n, const float *d_X, const int32_t *d_label,
                      const float *d_logk, const uint8_t *d_mask, float logk_weight, float lr,
                      float *d_loss, void *stream)
{
    return lossGrad(m, n, d_X, d_label, d_logk, d_mask, logk_weight, d_loss, nullptr, true, lr,
                    static_cast<hipStream_t>(stream));
}

int nn_mlp_predict(nn_mlp *m, int64_t n, const float *d_X, float *d_prob, float *d_logk,
                   void *stream)
{
    if (!m || n < 0 || (n > 0 && (!d_X || !d_prob || !d_logk)))
        return mlpFail(NN_E_INVALID_ARGUMENT, "invalid MLP argument");
    if (n == 0)
        return NN_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t tiles = (n + kTile - 1) / kTile;
    if (tiles > 0x7FFFFFFF)
        return mlpFail(NN_E_INVALID_ARGUMENT, "batch too large");
    return dispatchMlpNc(m->nc, [&](auto tag) {
        constexpr int NC = decltype(tag)::value;
        mlpPredictKernel<NC><<<unsigned(tiles), kThreads, 0, st>>>(n, m->dParams, d_X, d_prob,
                                                                  d_logk);
        MLP_HIP(hipGetLastError());
        return NN_OK;
    });
}

int nn_mlp_hints(nn_mlp *m, int64_t n, const float *d_prob, const float *d_logk, float threshold,
                 int32_t modes, uint8_t *d_hint, double *d_seed_log_k, void *stream)
{
    if (!m || n < 0 || (n > 0 && (!d_prob || !d_logk || !d_hint || !d_seed_log_k)))
        return mlpFail(NN_E_INVALID_ARGUMENT, "invalid MLP argument");
    if (n == 0)
        return NN_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    return dispatchMlpNc(m->nc, [&](auto tag) {
        constexpr int NC = decltype(tag)::value;
        mlpHintKernel<NC><<<unsigned((n + 255) / 256), 256, 0, st>>>(
            n, d_prob, d_logk, threshold, modes, d_hint, d_seed_log_k);
        MLP_HIP(hipGetLastError());
        return NN_OK;
    });
}

} // extern "C"
