// nrc.h -- layout constants and launchers of the neural radiance cache (nrc.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crt {

constexpr int   kNrcQuery  = 13;   // p[3] (scene-normalised), d[3], n[3], albedo[3], roughness
constexpr int   kNrcIn     = 96;   // encoded features (82 used + constant 1 + padding)
constexpr int   kNrcHid    = 64;
constexpr int   kNrcOut    = 32;   // outputs 0..2 = RGB
constexpr int   kNrcLayers = 6;    // 96->64, 4 x 64->64, 64->32
constexpr int   kNrcParams = kNrcIn * kNrcHid + 4 * kNrcHid * kNrcHid + kNrcHid * kNrcOut;   // 24576
constexpr int   kNrcActs   = kNrcIn + 5 * kNrcHid;   // saved inputs of the six layers per sample
constexpr float kNrcEpsRel = 0.01f;

__host__ __device__ constexpr int nrc_w_off(int l)
{
    return l == 0 ? 0 : kNrcIn * kNrcHid + (l - 1) * kNrcHid * kNrcHid;
}

// Gradient partial rows written by nrc_launch_train for n samples: one per 64-sample tile.
__host__ __device__ constexpr uint32_t nrc_partial_rows(uint32_t n) { return (n + 63u) / 64u; }

hipError_t nrc_launch_infer(const float *q, float *out, uint32_t n, const __bf16 *WT, hipStream_t s);
hipError_t nrc_launch_train(const float *q, const float *target, uint32_t n, const __bf16 *W, const __bf16 *WT,
                            __bf16 *act, float *partial, float *partial_loss, hipStream_t s);
hipError_t nrc_launch_adam(const float *partial, uint32_t nrows, float *params, float *m, float *v, __bf16 *W,
                           __bf16 *WT, float lr, float b1, float b2, float eps, uint64_t step, int update,
                           hipStream_t s);
hipError_t nrc_launch_sum_partials(const float *partial, uint32_t nrows, float *out, hipStream_t s);

}  // namespace crt
