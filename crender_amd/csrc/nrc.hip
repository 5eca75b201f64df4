// nrc.hip -- neural radiance cache: MLP inference, training (forward + backward) and Adam,
// hand-written for gfx950 bf16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
//
// Network: 96 input features (12-frequency sin/cos of the position, direction, normal,
// albedo, roughness, a constant 1, zero padding) -> 5 hidden layers of 64 (ReLU) -> 32
// outputs of which 0..2 are RGB radiance.  No biases (the constant feature biases the
// first layer).  Weights: FP32 master copy updated by Adam, bf16 copies in row-major
// W[k][n] (B operand of the backward data product) and transposed WT[n][k] (B operand of
// the forward product).
//
// MFMA operand mapping used throughout (one 16x16 output tile, 32-deep reduction):
//   A: lane l supplies row (l % 16), reduction elements 8*(l/16) .. 8*(l/16)+7
//   B: lane l supplies column (l % 16), the same reduction elements
//   D: lane l holds rows 4*(l/16) .. 4*(l/16)+3 of column (l % 16)
// A and B use the same element mapping, so the reduction is exact whatever order the
// hardware sums the 32 products in.
//
// Inference: each wave owns 64 samples (4 row tiles of 16); activations between layers go
// through the wave's slice of LDS.  Training: a workgroup of four waves owns 64 samples and
// splits every layer's output columns over its waves; it stores every layer's input
// activations (sample-major, bf16) in a global scratch so the backward pass can form
// dW = X^T G, writes per-tile gradient partials (no atomics), and the Adam kernel sums the
// partials per parameter.
#include <hip/hip_runtime.h>
#include <math.h>
#include "crt_internal.h"
#include "nrc.h"

namespace crt {

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float  f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16 to_bf16(float x) { return (bf16)x; }   // round to nearest even
__device__ __forceinline__ float from_bf16(bf16 x) { return (float)x; }

constexpr int kXStride = kNrcIn + 8;    // LDS row stride (bf16) of the sample-major activations
constexpr int kGStride = kNrcHid + 8;

__host__ __device__ constexpr int layer_k(int l) { return l == 0 ? kNrcIn : kNrcHid; }
__host__ __device__ constexpr int layer_n(int l) { return l == kNrcLayers - 1 ? kNrcOut : kNrcHid; }
__host__ __device__ constexpr int act_off(int l) { return l == 0 ? 0 : kNrcIn + (l - 1) * kNrcHid; }

// features of one query: p[3] in [0,1], d[3], n[3], albedo[3], roughness
__device__ __forceinline__ void encode_row(const float *q, bf16 *row)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float f = 3.14159265358979323846f;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const float arg = q[c] * f;
            float s, co;
            sincosf(arg, &s, &co);
            row[c * 24 + 2 * k]     = to_bf16(s);
            row[c * 24 + 2 * k + 1] = to_bf16(co);
            f *= 2.0f;
        }
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) row[72 + k] = to_bf16(q[3 + k]);
    row[82] = to_bf16(1.0f);
#pragma unroll
    for (int k = 83; k < kNrcIn; ++k) row[k] = to_bf16(0.0f);
}

__device__ __forceinline__ bf16x8 lds8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ __forceinline__ bf16x8 glb8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// Forward product of one layer for the wave's 64 rows held in X (sample-major, LDS).
template <int L>
__device__ __forceinline__ void layer_fwd(const bf16 *X, const bf16 *WT, f32x4 (&acc)[4][4])
{
    constexpr int K = layer_k(L), N = layer_n(L), NT = N / 16;
    const int lane = threadIdx.x & 63, r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const bf16 *w = WT + nrc_w_off(L);
#pragma unroll
    for (int kk = 0; kk < K / 32; ++kk) {
        bf16x8 a[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) a[rt] = lds8(X + (rt * 16 + r16) * kXStride + kk * 32 + 8 * g);
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
            const bf16x8 b = glb8(w + (ct * 16 + r16) * K + kk * 32 + 8 * g);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) acc[rt][ct] = mma(a[rt], b, acc[rt][ct]);
        }
    }
}

// ReLU and write the layer output back into X (sample-major) as the next layer's input.
template <int N>
__device__ __forceinline__ void store_relu(bf16 *X, const f32x4 (&acc)[4][4])
{
    const int lane = threadIdx.x & 63, r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < N / 16; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                X[(rt * 16 + 4 * g + i) * kXStride + ct * 16 + r16] = to_bf16(fmaxf(acc[rt][ct][i], 0.0f));
}

template <int L>
__device__ __forceinline__ void hidden(bf16 *X, const bf16 *WT)
{
    f32x4 acc[4][4];
    layer_fwd<L>(X, WT, acc);
    __syncthreads();
    store_relu<kNrcHid>(X, acc);
    __syncthreads();
}

// n_dev (when set) is a device-side count, clamped to the buffers' capacity n_host.
__global__ __launch_bounds__(256) void k_nrc_infer(const float *__restrict__ q, float *__restrict__ out, uint32_t n_host,
                                                   const uint32_t *__restrict__ n_dev, const bf16 *__restrict__ WT)
{
    __shared__ bf16 xs[4 * 64 * kXStride];
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r16 = lane & 15, g = lane >> 4;
    bf16 *X = xs + wv * 64 * kXStride;
    const uint32_t ntiles = (n + 255) / 256;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t s0 = tile * 256 + wv * 64;
        const uint32_t s  = s0 + lane;
        float qq[kNrcQuery];
#pragma unroll
        for (int k = 0; k < kNrcQuery; ++k) qq[k] = s < n ? q[(size_t)s * kNrcQuery + k] : 0.0f;
        encode_row(qq, X + lane * kXStride);
        __syncthreads();
        hidden<0>(X, WT);
        hidden<1>(X, WT);
        hidden<2>(X, WT);
        hidden<3>(X, WT);
        hidden<4>(X, WT);
        f32x4 acc[4][4];
        layer_fwd<5>(X, WT, acc);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = s0 + rt * 16 + 4 * g + i;
                if (r16 < 3 && row < n) out[(size_t)row * 3 + r16] = acc[rt][0][i];
            }
        __syncthreads();
    }
}

// One part of a query's features, for the four lane groups of the training workgroup:
// part c < 3 = the 24 sin/cos features of position component c; part 3 = direction, normal,
// albedo, roughness, the constant 1 and the zero padding.  Same values as encode_row.
__device__ __forceinline__ void encode_part(const float *q, bf16 *row, int part)
{
    if (part < 3) {
        float f = 3.14159265358979323846f;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const float arg = q[part] * f;
            float s, co;
            sincosf(arg, &s, &co);
            row[part * 24 + 2 * k]     = to_bf16(s);
            row[part * 24 + 2 * k + 1] = to_bf16(co);
            f *= 2.0f;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 10; ++k) row[72 + k] = to_bf16(q[3 + k]);
        row[82] = to_bf16(1.0f);
#pragma unroll
        for (int k = 83; k < kNrcIn; ++k) row[k] = to_bf16(0.0f);
    }
}

// Copies the tile's 64 activation rows (columns [0, C)) from LDS to the global scratch.
template <int C>
__device__ __forceinline__ void save_acts(const bf16 *X, bf16 *A, int off)
{
    for (int c = threadIdx.x; c < 64 * (C / 8); c += 256) {
        const int row = c / (C / 8), col = (c % (C / 8)) * 8;
        *reinterpret_cast<bf16x8 *>(A + row * kNrcActs + off + col) = lds8(X + row * kXStride + col);
    }
}

// Training: one workgroup (4 waves) = one tile of 64 samples.  The output columns of every
// layer product are split over the waves (forward and dX: column tile = wave; dW: row tile
// kt = wave, wave + 4, ...), with the tile's activations and output gradients shared in LDS,
// so a step's serial chain of layer products is a quarter as long as with one wave per
// tile.  Writes the tile's gradient partial (every parameter by exactly one wave, no
// atomics) and its loss partial.
__global__ __launch_bounds__(256) void k_nrc_train(const float *__restrict__ q, const float *__restrict__ target,
                                                   uint32_t n, float inv_count, const bf16 *__restrict__ W,
                                                   const bf16 *__restrict__ WT, bf16 *__restrict__ act,
                                                   float *__restrict__ partial, float *__restrict__ partial_loss)
{
    __shared__ bf16 X[64 * kXStride];
    __shared__ bf16 Gs[64 * kGStride];    // G sample-major [s][n]
    __shared__ bf16 Gt[64 * kGStride];    // G feature-major [n][s]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r16 = lane & 15, g = lane >> 4;
    const uint32_t tile = blockIdx.x;
    const uint32_t s0 = tile * 64;
    bf16 *A = act + (size_t)tile * 64 * kNrcActs;    // [64][kNrcActs] sample-major
    {
        const uint32_t s = s0 + lane;
        float qq[kNrcQuery];
#pragma unroll
        for (int k = 0; k < kNrcQuery; ++k) qq[k] = s < n ? q[(size_t)s * kNrcQuery + k] : 0.0f;
        encode_part(qq, X + lane * kXStride, wv);
    }
    __syncthreads();
    save_acts<kNrcIn>(X, A, 0);
    // ---- forward: wave w computes output columns [16w, 16w + 16) of each hidden layer ----
#define FWD_HIDDEN(L)                                                                             \
    {                                                                                             \
        constexpr int K = layer_k(L);                                                             \
        f32x4 acc[4];                                                                             \
        for (int rt = 0; rt < 4; ++rt) acc[rt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};                   \
        const bf16 *w = WT + nrc_w_off(L);                                                        \
        for (int kk = 0; kk < K / 32; ++kk) {                                                     \
            const bf16x8 b = glb8(w + (wv * 16 + r16) * K + kk * 32 + 8 * g);                     \
            for (int rt = 0; rt < 4; ++rt)                                                        \
                acc[rt] = mma(lds8(X + (rt * 16 + r16) * kXStride + kk * 32 + 8 * g), b, acc[rt]); \
        }                                                                                         \
        __syncthreads();                                                                          \
        for (int rt = 0; rt < 4; ++rt)                                                            \
            for (int i = 0; i < 4; ++i)                                                           \
                X[(rt * 16 + 4 * g + i) * kXStride + wv * 16 + r16] = to_bf16(fmaxf(acc[rt][i], 0.0f)); \
        __syncthreads();                                                                          \
        save_acts<kNrcHid>(X, A, act_off(L + 1));                                                 \
    }
    FWD_HIDDEN(0)
    FWD_HIDDEN(1)
    FWD_HIDDEN(2)
    FWD_HIDDEN(3)
    FWD_HIDDEN(4)
#undef FWD_HIDDEN
    // ---- output layer (waves 0, 1: column tiles 0, 1), loss and dL/dy ----
    if (wv < kNrcOut / 16) {
        constexpr int K = layer_k(5);
        f32x4 acc[4];
        for (int rt = 0; rt < 4; ++rt) acc[rt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const bf16 *w = WT + nrc_w_off(5);
        for (int kk = 0; kk < K / 32; ++kk) {
            const bf16x8 b = glb8(w + (wv * 16 + r16) * K + kk * 32 + 8 * g);
            for (int rt = 0; rt < 4; ++rt) acc[rt] = mma(lds8(X + (rt * 16 + r16) * kXStride + kk * 32 + 8 * g), b, acc[rt]);
        }
        float lsum = 0.0f;
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = rt * 16 + 4 * g + i, col = wv * 16 + r16;
                const uint32_t srow = s0 + row;
                float gr = 0.0f;
                if (col < 3 && srow < n) {
                    const float y = acc[rt][i];
                    const float t = target[(size_t)srow * 3 + col];
                    const float e = y - t;
                    const float den = y * y + kNrcEpsRel;
                    lsum += e * e / den * inv_count;
                    gr = 2.0f * e / den * inv_count;
                }
                const bf16 gb = to_bf16(gr);
                Gs[row * kGStride + col] = gb;
                Gt[col * kGStride + row] = gb;
            }
        for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
        if (wv == 0 && lane == 0) partial_loss[tile] = lsum;
    }
    __syncthreads();
    float *P = partial + (size_t)tile * kNrcParams;
    // ---- backward ----
#define BWD(L)                                                                                              \
    {                                                                                                       \
        constexpr int K = layer_k(L), N = layer_n(L);                                                       \
        /* dW = X_L^T G: row tiles kt = wave, wave + 4, ...; all N / 16 column tiles */                     \
        for (int kt = wv; kt < K / 16; kt += 4) {                                                           \
            bf16x8 a[2];                                                                                    \
            for (int ss = 0; ss < 2; ++ss) {                                                                \
                bf16 tmp[8];                                                                                \
                for (int i = 0; i < 8; ++i) tmp[i] = A[(ss * 32 + 8 * g + i) * kNrcActs + act_off(L) + kt * 16 + r16]; \
                a[ss] = bf16x8{tmp[0], tmp[1], tmp[2], tmp[3], tmp[4], tmp[5], tmp[6], tmp[7]};            \
            }                                                                                               \
            for (int ct = 0; ct < N / 16; ++ct) {                                                           \
                f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};                                                    \
                for (int ss = 0; ss < 2; ++ss) d = mma(a[ss], lds8(Gt + (ct * 16 + r16) * kGStride + ss * 32 + 8 * g), d); \
                for (int i = 0; i < 4; ++i) P[nrc_w_off(L) + (kt * 16 + 4 * g + i) * N + ct * 16 + r16] = d[i]; \
            }                                                                                               \
        }                                                                                                   \
        if (L > 0) {                                                                                        \
            /* dX = (G W_L^T) * relu'(X_L): wave w computes columns [16w, 16w + 16); reduction over n */    \
            f32x4 dx[4];                                                                                    \
            for (int rt = 0; rt < 4; ++rt) dx[rt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};                         \
            for (int nn = 0; nn < N / 32; ++nn) {                                                           \
                const bf16x8 b = glb8(W + nrc_w_off(L) + (wv * 16 + r16) * N + nn * 32 + 8 * g);            \
                for (int rt = 0; rt < 4; ++rt)                                                              \
                    dx[rt] = mma(lds8(Gs + (rt * 16 + r16) * kGStride + nn * 32 + 8 * g), b, dx[rt]);       \
            }                                                                                               \
            __syncthreads();                                                                                \
            for (int rt = 0; rt < 4; ++rt)                                                                  \
                for (int i = 0; i < 4; ++i) {                                                               \
                    const int row = rt * 16 + 4 * g + i, col = wv * 16 + r16;                               \
                    const float xv = from_bf16(A[row * kNrcActs + act_off(L) + col]);                        \
                    const bf16 gb = to_bf16(xv > 0.0f ? dx[rt][i] : 0.0f);                                  \
                    Gs[row * kGStride + col] = gb;                                                          \
                    Gt[col * kGStride + row] = gb;                                                          \
                }                                                                                           \
            __syncthreads();                                                                                \
        }                                                                                                   \
    }
    BWD(5)
    BWD(4)
    BWD(3)
    BWD(2)
    BWD(1)
    BWD(0)
#undef BWD
}

// Adam on the FP32 master weights from the summed per-tile partials; refreshes the bf16
// copies.  Padded outputs (columns >= 3 of the last layer) and padded input features
// (rows >= 83 of the first layer) stay zero.
__global__ void k_nrc_adam(const float *__restrict__ partial, uint32_t nrows, float *params, float *m, float *v,
                           bf16 *W, bf16 *WT, float lr, float b1, float b2, float eps, float bc1, float bc2, int update)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kNrcParams) return;
    int l = 0;
    while (l + 1 < kNrcLayers && p >= (uint32_t)nrc_w_off(l + 1)) ++l;
    const uint32_t K = layer_k(l), N = layer_n(l);
    const uint32_t local = p - nrc_w_off(l), k = local / N, c = local % N;
    const bool frozen = (l == kNrcLayers - 1 && c >= 3) || (l == 0 && k >= 83);
    float w = params[p];
    if (update && !frozen) {
        float gsum = 0.0f;
        for (uint32_t t = 0; t < nrows; ++t) gsum += partial[(size_t)t * kNrcParams + p];
        const float mm = b1 * m[p] + (1.0f - b1) * gsum;
        const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
        m[p] = mm;
        v[p] = vv;
        w -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
        params[p] = w;
    }
    if (frozen) {
        w = 0.0f;
        params[p] = 0.0f;
    }
    W[p] = to_bf16(w);
    WT[nrc_w_off(l) + c * K + k] = to_bf16(w);
}

__global__ void k_nrc_sum_partials(const float *__restrict__ partial, uint32_t nrows, float *out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kNrcParams) return;
    float s = 0.0f;
    for (uint32_t t = 0; t < nrows; ++t) s += partial[(size_t)t * kNrcParams + p];
    out[p] = s;
}

}  // namespace

hipError_t nrc_launch_infer(const float *q, float *out, uint32_t n, const __bf16 *WT, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t tiles = (n + 255) / 256;
    k_nrc_infer<<<std::min<uint32_t>(tiles, 2048u), 256, 0, s>>>(q, out, n, nullptr, WT);
    return hipGetLastError();
}

hipError_t nrc_launch_infer_dev(const float *q, float *out, const uint32_t *n_dev, uint32_t n_max, const __bf16 *WT,
                                hipStream_t s)
{
    if (n_max == 0) return hipSuccess;
    const uint32_t tiles = (n_max + 255) / 256;
    k_nrc_infer<<<std::min<uint32_t>(tiles, 2048u), 256, 0, s>>>(q, out, n_max, n_dev, WT);
    return hipGetLastError();
}

hipError_t nrc_launch_train(const float *q, const float *target, uint32_t n, const __bf16 *W, const __bf16 *WT,
                            __bf16 *act, float *partial, float *partial_loss, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t tiles = (n + 63) / 64;
    k_nrc_train<<<tiles, 256, 0, s>>>(q, target, n, 1.0f / (3.0f * (float)n), W, WT, act, partial, partial_loss);
    return hipGetLastError();
}

hipError_t nrc_launch_adam(const float *partial, uint32_t nrows, float *params, float *m, float *v, __bf16 *W,
                           __bf16 *WT, float lr, float b1, float b2, float eps, uint64_t step, int update,
                           hipStream_t s)
{
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    k_nrc_adam<<<(kNrcParams + 255) / 256, 256, 0, s>>>(partial, nrows, params, m, v, W, WT, lr, b1, b2, eps,
                                                         update ? bc1 : 1.0f, update ? bc2 : 1.0f, update);
    return hipGetLastError();
}

hipError_t nrc_launch_sum_partials(const float *partial, uint32_t nrows, float *out, hipStream_t s)
{
    k_nrc_sum_partials<<<(kNrcParams + 255) / 256, 256, 0, s>>>(partial, nrows, out);
    return hipGetLastError();
}

}  // namespace crt
