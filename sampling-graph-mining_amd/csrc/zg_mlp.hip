// zg_mlp.hip -- the learned error-latency profile's network: a 16-64-64-64-1
// ReLU MLP with BF16 operands on the MFMA units and FP32 accumulation.
// No reference counterpart: it stands in for the regression fit of
// ErrorProfile::get_num_estimators (src/ErrorLatencyProfile.cpp:127-200).
//
// One wavefront owns a 16-row tile.  Products with K = 64 (the hidden layers
// forward, dH = dZ W^T backward) use gfx950's v_mfma_f32_16x16x32_bf16 (A
// 16x32, B 32x16), two per 16x16 output tile; products with K = 16 (the
// first layer, the per-tile weight gradients) use v_mfma_f32_16x16x16_bf16:
//   16x16x16 A operand: lane l holds A[l % 16][4 (l / 16) + j], j = 0..3
//            B operand: lane l holds B[4 (l / 16) + j][l % 16]
//   16x16x32 A operand: lane l holds A[l % 16][8 (l / 16) + j], j = 0..7
//            B operand: lane l holds B[8 (l / 16) + j][l % 16]
//   D result (both)   : lane l holds D[4 (l / 16) + j][l % 16]
// Forward: X(16x16) W1 -> relu -> (16x64) W2 -> relu -> W3 -> relu -> . w4.
// Backward (MSE loss mean over the batch) per tile: dZ3 = dy w4^T * relu',
// gW3 = H2^T dZ3, dH2 = dZ3 W3^T, ... gW1 = X^T dZ1; each workgroup writes
// its gradient to its own partial row; a second kernel sums the rows in a
// fixed order (deterministic) and applies Adam to the fp32 master weights.
#include <math.h>
#include <string.h>

#include <memory>

#include <vector>

#include "zg_internal.h"

namespace zg {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int IN = 16, H = 64;
constexpr int OFF_W1 = 0;                 // [IN][H] row-major (in, out)
constexpr int OFF_B1 = OFF_W1 + IN * H;
constexpr int OFF_W2 = OFF_B1 + H;        // [H][H]
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;        // [H][H]
constexpr int OFF_B3 = OFF_W3 + H * H;
constexpr int OFF_W4 = OFF_B3 + H;        // [H]
constexpr int OFF_B4 = OFF_W4 + H;
constexpr int NPARAM = OFF_B4 + 1;
constexpr int PSTRIDE = NPARAM + 1;       // partial row: gradient + loss

// fp32 -> bf16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct Smem {
    alignas(16) uint16_t w1t[H][IN];   // W1^T (n-major) for the forward B operand
    alignas(16) uint16_t w2t[H][H];    // W2^T
    alignas(16) uint16_t w3t[H][H];    // W3^T
    alignas(16) uint16_t w2[H][H];     // W2 (row-major) for dH1 = dZ2 W2^T
    alignas(16) uint16_t w3[H][H];     // W3
    alignas(16) float b1[H], b2[H], b3[H], w4[H];
    alignas(16) uint16_t x[16][IN];
    alignas(16) uint16_t h1[16][H];
    alignas(16) uint16_t h2[16][H];
    alignas(16) uint16_t h3[16][H];
    alignas(16) uint16_t dz[16][H];
    alignas(16) float y[16];
    alignas(16) float dy[16];
    float b4;
};

__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) { return *reinterpret_cast<const bf16x4 *>(p); }

// A operand from row-major S[16][ld], K chunk kc
__device__ __forceinline__ bf16x4 opA(const uint16_t *S, int ld, int kc, int l)
{
    return ld4(S + (l & 15) * ld + kc * 16 + 4 * (l >> 4));
}
// B operand from n-major T[n][ldk] (T[n][k] = B[k][n]), K chunk kc, N tile nt
__device__ __forceinline__ bf16x4 opB(const uint16_t *T, int ldk, int kc, int nt, int l)
{
    return ld4(T + (nt * 16 + (l & 15)) * ldk + kc * 16 + 4 * (l >> 4));
}
// A operand = S^T for row-major S[16 rows = K][ld]: A[mm][kk] = S[kk][mt*16 + mm]
__device__ __forceinline__ bf16x4 opAT(const uint16_t *S, int ld, int mt, int l)
{
    bf16x4 r;
    const int col = mt * 16 + (l & 15), row = 4 * (l >> 4);
    r[0] = (short)S[(row + 0) * ld + col];
    r[1] = (short)S[(row + 1) * ld + col];
    r[2] = (short)S[(row + 2) * ld + col];
    r[3] = (short)S[(row + 3) * ld + col];
    return r;
}
// B operand = row-major S[16 rows = K][ld] itself: B[kk][nn] = S[kk][nt*16 + nn]
__device__ __forceinline__ bf16x4 opBrow(const uint16_t *S, int ld, int nt, int l)
{
    return opAT(S, ld, nt, l);
}

__device__ __forceinline__ f32x4 mma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// K = 32 operands: 8 consecutive bf16 of a row (A) or of an n-major column (B)
__device__ __forceinline__ bf16x8 ld8(const uint16_t *p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8 *>(p));
}
// A operand from row-major S[16][ld], K chunk kc of 32
__device__ __forceinline__ bf16x8 opA32(const uint16_t *S, int ld, int kc, int l)
{
    return ld8(S + (l & 15) * ld + kc * 32 + 8 * (l >> 4));
}
// B operand from n-major T[n][ldk], K chunk kc of 32, N tile nt
__device__ __forceinline__ bf16x8 opB32(const uint16_t *T, int ldk, int kc, int nt, int l)
{
    return ld8(T + (nt * 16 + (l & 15)) * ldk + kc * 32 + 8 * (l >> 4));
}
__device__ __forceinline__ f32x4 mma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ void load_weights(Smem &s, const float *__restrict__ P, int l)
{
    for (int i = l; i < IN * H; i += 64) {
        const int k = i / H, n = i % H;
        s.w1t[n][k] = f2bf(P[OFF_W1 + i]);
    }
    for (int i = l; i < H * H; i += 64) {
        const int k = i / H, n = i % H;
        const uint16_t a = f2bf(P[OFF_W2 + i]), b = f2bf(P[OFF_W3 + i]);
        s.w2t[n][k] = a; s.w2[k][n] = a;
        s.w3t[n][k] = b; s.w3[k][n] = b;
    }
    s.b1[l] = P[OFF_B1 + l];
    s.b2[l] = P[OFF_B2 + l];
    s.b3[l] = P[OFF_B3 + l];
    s.w4[l] = P[OFF_W4 + l];
    if (l == 0) s.b4 = P[OFF_B4];
}

// hidden layer: out[16][64] = relu(in[16][K] W + b) as bf16 into `out`
template <int K>
__device__ void dense_relu(const uint16_t *in, const uint16_t *wt, const float *b, uint16_t *out, int l)
{
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        if constexpr (K % 32 == 0) {
#pragma unroll
            for (int kc = 0; kc < K / 32; ++kc) acc = mma32(opA32(in, K, kc, l), opB32(wt, K, kc, nt, l), acc);
        } else {
#pragma unroll
            for (int kc = 0; kc < K / 16; ++kc) acc = mma(opA(in, K, kc, l), opB(wt, K, kc, nt, l), acc);
        }
        const int n = nt * 16 + (l & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = 4 * (l >> 4) + j;
            const float z = acc[j] + b[n];
            out[m * H + n] = f2bf(z > 0.f ? z : 0.f);
        }
    }
}

__device__ void forward_tile(Smem &s, const float *__restrict__ X, int row0, int B, int l)
{
    for (int i = l; i < 16 * IN; i += 64) {
        const int m = i / IN, k = i % IN;
        s.x[m][k] = (row0 + m < B) ? f2bf(X[(size_t)(row0 + m) * IN + k]) : 0;
    }
    __syncthreads();
    dense_relu<IN>(&s.x[0][0], &s.w1t[0][0], s.b1, &s.h1[0][0], l);
    __syncthreads();
    dense_relu<H>(&s.h1[0][0], &s.w2t[0][0], s.b2, &s.h2[0][0], l);
    __syncthreads();
    dense_relu<H>(&s.h2[0][0], &s.w3t[0][0], s.b3, &s.h3[0][0], l);
    __syncthreads();
    if (l < 16) {
        float acc = 0.f;
        for (int n = 0; n < H; ++n) acc += bf2f(s.h3[l][n]) * s.w4[n];
        s.y[l] = acc + s.b4;
    }
    __syncthreads();
}

__global__ __launch_bounds__(64) void k_forward(const float *__restrict__ P, const float *__restrict__ X, int B,
                                                float *__restrict__ Y)
{
    __shared__ Smem s;
    const int l = threadIdx.x;
    load_weights(s, P, l);
    __syncthreads();
    const int row0 = blockIdx.x * 16;
    forward_tile(s, X, row0, B, l);
    if (l < 16 && row0 + l < B) Y[row0 + l] = s.y[l];
}

// gradient tile store: D layout -> partial[off + (mt*16 + 4(l/16)+j) * ldn + nt*16 + l%16]
__device__ __forceinline__ void store_grad(float *dst, int ldn, int mt, int nt, f32x4 d, int l)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) dst[(mt * 16 + 4 * (l >> 4) + j) * ldn + nt * 16 + (l & 15)] = d[j];
}

// dZ (bf16, LDS) of a hidden layer from dH in D layout, with relu' from the
// layer's own activations; returns nothing, writes the bias gradient
__device__ void backprop_hidden(Smem &s, const uint16_t *hb /* this layer's activations */,
                                const uint16_t *wt_next_row /* W_{l+1} row-major */, float *gbias, int l)
{
    // dH = dZ_{l+1} W_{l+1}^T : A = dz (row-major [m][n_out]), B[kk = n_out][nn = k_in] = W[k_in][n_out]
    f32x4 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        acc[nt] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < H / 32; ++kc)
            acc[nt] = mma32(opA32(&s.dz[0][0], H, kc, l), opB32(wt_next_row, H, kc, nt, l), acc[nt]);
    }
    __syncthreads();   // everyone has read dz before it is overwritten
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int n = nt * 16 + (l & 15);
        float colsum = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = 4 * (l >> 4) + j;
            const float d = hb[m * H + n] != 0 ? acc[nt][j] : 0.f;   // relu'(z) = (h > 0); h >= 0
            colsum += d;
            s.dz[m][n] = f2bf(d);
        }
        colsum += __shfl_xor(colsum, 16, 64);
        colsum += __shfl_xor(colsum, 32, 64);
        if (l < 16) gbias[n] = colsum;
    }
    __syncthreads();
}

// weight gradient: G[k][n] = sum_m Hb[m][k] dz[m][n]; K_IN rows of G
template <int K_IN>
__device__ void weight_grad(const uint16_t *hb, int ld_h, const uint16_t *dz, float *G, int l)
{
#pragma unroll
    for (int mt = 0; mt < K_IN / 16; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mma(opAT(hb, ld_h, mt, l), opBrow(dz, H, nt, l), acc);
            store_grad(G, H, mt, nt, acc, l);
        }
}

__global__ __launch_bounds__(64) void k_grad(const float *__restrict__ P, const float *__restrict__ X,
                                             const float *__restrict__ T, int B, float *__restrict__ partials)
{
    __shared__ Smem s;
    const int l = threadIdx.x;
    load_weights(s, P, l);
    __syncthreads();
    const int row0 = blockIdx.x * 16;
    forward_tile(s, X, row0, B, l);
    float *G = partials + (size_t)blockIdx.x * PSTRIDE;
    // output layer: loss = mean (y - t)^2
    if (l < 16) {
        const bool ok = row0 + l < B;
        const float e = ok ? s.y[l] - T[row0 + l] : 0.f;
        s.dy[l] = 2.f * e / (float)B;
        float sq = e * e;
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 16);
        if (l == 0) G[NPARAM] = sq;
    }
    __syncthreads();
    {   // gw4, gb4, dZ3 (lanes over columns)
        const int n = l;
        float gw = 0.f, gb = 0.f;
        for (int m = 0; m < 16; ++m) {
            const float h = bf2f(s.h3[m][n]);
            gw += h * s.dy[m];
            const float d = s.h3[m][n] != 0 ? s.dy[m] * s.w4[n] : 0.f;
            gb += d;
            s.dz[m][n] = f2bf(d);
        }
        G[OFF_W4 + n] = gw;
        G[OFF_B3 + n] = gb;
        if (l == 0) {
            float t = 0.f;
            for (int m = 0; m < 16; ++m) t += s.dy[m];
            G[OFF_B4] = t;
        }
    }
    __syncthreads();
    weight_grad<H>(&s.h2[0][0], H, &s.dz[0][0], G + OFF_W3, l);      // gW3 = H2^T dZ3
    backprop_hidden(s, &s.h2[0][0], &s.w3[0][0], G + OFF_B2, l);      // dZ2
    weight_grad<H>(&s.h1[0][0], H, &s.dz[0][0], G + OFF_W2, l);      // gW2 = H1^T dZ2
    backprop_hidden(s, &s.h1[0][0], &s.w2[0][0], G + OFF_B1, l);      // dZ1
    weight_grad<IN>(&s.x[0][0], IN, &s.dz[0][0], G + OFF_W1, l);     // gW1 = X^T dZ1
}

// sums the partial rows in a fixed order; optionally applies Adam
__global__ void k_reduce_adam(const float *__restrict__ partials, int nrows, float *__restrict__ grad,
                              float *__restrict__ loss, float *__restrict__ P, float *__restrict__ M,
                              float *__restrict__ V, float lr, float b1t, float b2t, int apply)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NPARAM) return;
    float g = 0.f;
    for (int r = 0; r < nrows; ++r) g += partials[(size_t)r * PSTRIDE + p];
    if (p == NPARAM) { *loss = g; return; }
    grad[p] = g;
    if (!apply) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float m = beta1 * M[p] + (1.f - beta1) * g;
    const float v = beta2 * V[p] + (1.f - beta2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / (1.f - b1t)) / (sqrtf(v / (1.f - b2t)) + eps);
}

// minibatch gather: rows chosen by Philox (seed, TAG 5, step * batch + i)
__global__ void k_gather_batch(const float *__restrict__ X, const float *__restrict__ T, int N, int batch,
                               uint64_t seed, uint64_t step, float *__restrict__ xb, float *__restrict__ tb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= batch) return;
    Rng r;
    r.init(seed, 5u, step * (uint64_t)batch + i);
    const uint32_t row = (uint32_t)r.uniform((uint64_t)N);
    for (int k = 0; k < IN; ++k) xb[(size_t)i * IN + k] = X[(size_t)row * IN + k];
    tb[i] = T[row];
}

}  // namespace mlp
}  // namespace zg

struct zg_mlp {
    float *P = nullptr, *M = nullptr, *V = nullptr, *grad = nullptr, *loss = nullptr, *partials = nullptr;
    size_t partial_rows = 0;
    int t = 0;
    hipStream_t stream = nullptr;
    float *io = nullptr;         // inputs and outputs of small forward calls (kSmallRows rows)
};
constexpr int kSmallRows = 64;   // forward calls up to this many rows reuse zg_mlp::io

using namespace zg;
using namespace zg::mlp;

namespace {

int ensure_partials(zg_mlp *m, int rows)
{
    if ((size_t)rows <= m->partial_rows) return ZG_OK;
    if (m->partials) ZG_HIP(hipFree(m->partials));
    m->partials = nullptr;
    ZG_HIP(hipMalloc(&m->partials, sizeof(float) * PSTRIDE * (size_t)rows));
    m->partial_rows = rows;
    return ZG_OK;
}

struct DevF {
    float *p = nullptr;
    ~DevF() { if (p) (void)hipFree(p); }
};

int grad_step(zg_mlp *m, const float *dX, const float *dT, int B, float lr, int apply)
{
    const int tiles = (B + 15) / 16;
    int rc = ensure_partials(m, tiles);
    if (rc) return rc;
    hipLaunchKernelGGL(k_grad, dim3(tiles), dim3(64), 0, m->stream, m->P, dX, dT, B, m->partials);
    ZG_HIP(hipGetLastError());
    float b1t = 1.f, b2t = 1.f;
    if (apply) {
        m->t++;
        b1t = powf(0.9f, (float)m->t);
        b2t = powf(0.999f, (float)m->t);
    }
    hipLaunchKernelGGL(k_reduce_adam, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, m->stream, m->partials, tiles,
                       m->grad, m->loss, m->P, m->M, m->V, lr, b1t, b2t, apply);
    ZG_HIP(hipGetLastError());
    return ZG_OK;
}

}  // namespace

extern "C" int zg_mlp_num_params(void) { return NPARAM; }
extern "C" int zg_mlp_input_dim(void) { return IN; }

extern "C" int zg_mlp_create(uint64_t seed, zg_mlp **out)
{
    if (!out) return set_error(ZG_EINVAL, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    std::unique_ptr<zg_mlp> m(new zg_mlp());
    ZG_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    ZG_HIP(hipMalloc(&m->P, sizeof(float) * NPARAM));
    ZG_HIP(hipMalloc(&m->M, sizeof(float) * NPARAM));
    ZG_HIP(hipMalloc(&m->V, sizeof(float) * NPARAM));
    ZG_HIP(hipMalloc(&m->grad, sizeof(float) * NPARAM));
    ZG_HIP(hipMalloc(&m->loss, sizeof(float)));
    // He-uniform weights from Philox (seed, tag 6, index), zero biases
    std::vector<float> h(NPARAM, 0.f);
    // He-uniform hidden layers; the output layer starts at 1 % of that scale so
    // the untrained network predicts ~0 (the targets are small corrections)
    auto fill = [&](int off, int fan_in, int count, float scale = 1.f) {
        const float a = scale * sqrtf(6.f / (float)fan_in);
        for (int i = 0; i < count; ++i) {
            const uint4 r = philox10(make_uint4(0u, 6u, (uint32_t)(off + i), 0u), (uint32_t)seed, (uint32_t)(seed >> 32));
            const float u = (float)(r.x >> 8) * 0x1p-24f;
            h[off + i] = (2.f * u - 1.f) * a;
        }
    };
    fill(OFF_W1, IN, IN * H);
    fill(OFF_W2, H, H * H);
    fill(OFF_W3, H, H * H);
    fill(OFF_W4, H, H, 0.01f);
    ZG_HIP(hipMemcpy(m->P, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    ZG_HIP(hipMemset(m->M, 0, sizeof(float) * NPARAM));
    ZG_HIP(hipMemset(m->V, 0, sizeof(float) * NPARAM));
    *out = m.release();
    return ZG_OK;
}

extern "C" int zg_mlp_destroy(zg_mlp *m)
{
    if (!m) return ZG_OK;
    for (float *p : {m->P, m->M, m->V, m->grad, m->loss, m->partials, m->io})
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return ZG_OK;
}

extern "C" int zg_mlp_get_params(const zg_mlp *m, float *host)
{
    if (!m || !host) return set_error(ZG_EINVAL, "null argument");
    ZG_HIP(hipMemcpy(host, m->P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return ZG_OK;
}

extern "C" int zg_mlp_set_params(zg_mlp *m, const float *host)
{
    if (!m || !host) return set_error(ZG_EINVAL, "null argument");
    ZG_HIP(hipMemcpy(m->P, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    ZG_HIP(hipMemset(m->M, 0, sizeof(float) * NPARAM));
    ZG_HIP(hipMemset(m->V, 0, sizeof(float) * NPARAM));
    m->t = 0;
    return ZG_OK;
}

extern "C" int zg_mlp_forward(zg_mlp *m, const float *X, int B, float *Y)
{
    if (!m || !X || !Y || B <= 0) return set_error(ZG_EINVAL, "bad arguments");
    DevF dX, dY;
    float *x = nullptr, *y = nullptr;
    if (B <= kSmallRows) {   // the sizing path: one row per query, no allocation per call
        if (!m->io) ZG_HIP(hipMalloc(&m->io, sizeof(float) * (IN + 1) * kSmallRows));
        x = m->io;
        y = m->io + IN * kSmallRows;
    } else {
        ZG_HIP(hipMalloc(&dX.p, sizeof(float) * IN * (size_t)B));
        ZG_HIP(hipMalloc(&dY.p, sizeof(float) * (size_t)B));
        x = dX.p;
        y = dY.p;
    }
    ZG_HIP(hipMemcpyAsync(x, X, sizeof(float) * IN * (size_t)B, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_forward, dim3((B + 15) / 16), dim3(64), 0, m->stream, m->P, x, B, y);
    ZG_HIP(hipGetLastError());
    ZG_HIP(hipMemcpyAsync(Y, y, sizeof(float) * (size_t)B, hipMemcpyDeviceToHost, m->stream));
    ZG_HIP(hipStreamSynchronize(m->stream));
    return ZG_OK;
}

extern "C" int zg_mlp_gradient(zg_mlp *m, const float *X, const float *T, int B, float *grad, float *loss)
{
    if (!m || !X || !T || !grad || B <= 0) return set_error(ZG_EINVAL, "bad arguments");
    DevF dX, dT;
    ZG_HIP(hipMalloc(&dX.p, sizeof(float) * IN * (size_t)B));
    ZG_HIP(hipMalloc(&dT.p, sizeof(float) * (size_t)B));
    ZG_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * IN * (size_t)B, hipMemcpyHostToDevice, m->stream));
    ZG_HIP(hipMemcpyAsync(dT.p, T, sizeof(float) * (size_t)B, hipMemcpyHostToDevice, m->stream));
    int rc = grad_step(m, dX.p, dT.p, B, 0.f, 0);
    if (rc) return rc;
    ZG_HIP(hipMemcpyAsync(grad, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
    float l = 0.f;
    ZG_HIP(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    ZG_HIP(hipStreamSynchronize(m->stream));
    if (loss) *loss = l / (float)B;
    return ZG_OK;
}

extern "C" int zg_mlp_train(zg_mlp *m, const float *X, const float *T, int N, int steps, int batch, float lr,
                            uint64_t seed, float *final_loss)
{
    if (!m || !X || !T || N <= 0 || steps < 0 || batch <= 0) return set_error(ZG_EINVAL, "bad arguments");
    DevF dX, dT, xb, tb;
    ZG_HIP(hipMalloc(&dX.p, sizeof(float) * IN * (size_t)N));
    ZG_HIP(hipMalloc(&dT.p, sizeof(float) * (size_t)N));
    ZG_HIP(hipMalloc(&xb.p, sizeof(float) * IN * (size_t)batch));
    ZG_HIP(hipMalloc(&tb.p, sizeof(float) * (size_t)batch));
    ZG_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * IN * (size_t)N, hipMemcpyHostToDevice, m->stream));
    ZG_HIP(hipMemcpyAsync(dT.p, T, sizeof(float) * (size_t)N, hipMemcpyHostToDevice, m->stream));
    for (int s = 0; s < steps; ++s) {
        hipLaunchKernelGGL(k_gather_batch, dim3((batch + 255) / 256), dim3(256), 0, m->stream, dX.p, dT.p, N, batch,
                           seed, (uint64_t)s, xb.p, tb.p);
        int rc = grad_step(m, xb.p, tb.p, batch, lr, 1);
        if (rc) return rc;
    }
    float l = 0.f;
    ZG_HIP(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    ZG_HIP(hipStreamSynchronize(m->stream));
    if (final_loss) *final_loss = l / (float)batch;
    return ZG_OK;
}
