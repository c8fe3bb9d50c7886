/*
 * tr_mlp.h - the radiance-cache network: layout of its parameters, the
 * frequency encoding, and the launchers of its HIP kernels (tr_mlp.hip).
 * The network has no reference counterpart.
 *
 *   input    10 scalars in [0, 1]: position (scene box), outgoing direction
 *            ((w + 1) / 2), shading normal ((n + 1) / 2), roughness
 *   encoding per scalar x, band k = 0..11: sin(2^k pi x), cos(2^k pi x),
 *            feature (i * 12 + k) * 2 + {sin, cos}: 240 features (padded
 *            to 256 with zeros for the matrix cores)
 *   network  240 -> 64 -> 64 -> 64 -> 64 -> 3, ReLU on hidden layers
 *   params   FP32 master vector W0 (64 x 240 row major), b0, W1, b1, W2,
 *            b2, W3, b3, W4 (3 x 64), b4: 28,099
 *
 * The matrix products run on v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
 * accumulation.  Weights live as BF16 copies packed per layer with padded
 * shapes (forward: [out_pad][in_pad]; backward: the transposes of layers
 * 1..4), biases as FP32.
 */
#ifndef TR_MLP_H
#define TR_MLP_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tr_fmath.h"

namespace trk {

constexpr int kMlpIn = 10;
constexpr int kMlpBands = 12;
constexpr int kMlpFeat = kMlpIn * kMlpBands * 2;  /* 240 */
constexpr int kMlpFeatPad = 256;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 3;
constexpr int kMlpParams = 28099;

/* offsets in the FP32 parameter vector */
constexpr int kMlpOffW[5] = {0, 15424, 19584, 23744, 27904};
constexpr int kMlpOffB[5] = {15360, 19520, 23680, 27840, 28096};
constexpr int kMlpFanIn[5] = {240, 64, 64, 64, 64};
constexpr int kMlpFanOut[5] = {64, 64, 64, 64, 3};

/* packed BF16 forward weights, row major [out_pad][in_pad] */
constexpr int kPackOff[5] = {0, 16384, 20480, 24576, 28672};
constexpr int kPackInPad[5] = {256, 64, 64, 64, 64};
constexpr int kPackTotal = 28672 + 16 * 64;       /* layer 4 padded to 16 rows */
/* packed BF16 transposes of layers 1..4, row major [in][out_pad] */
constexpr int kPackTOff[5] = {0, 0, 4096, 8192, 12288};
constexpr int kPackTOutPad[5] = {0, 64, 64, 64, 32};
constexpr int kPackTTotal = 12288 + 64 * 32;
/* FP32 biases: b0..b3 at 64 l, b4 at 256 (padded to 16) */
constexpr int kBiasTotal = 272;

/* Cache queries as structure of arrays (`stride` entries per field): the
 * network inputs, the path throughput that scales the prediction and the
 * path slot whose radiance receives it (thr and slot may be null when
 * only the outputs are wanted). */
struct MlpQueries {
    float* in;        /* [kMlpIn][stride] */
    float* thr;       /* [3][stride] */
    uint32_t* slot;   /* [stride] */
    uint32_t stride;
};

/* One training sample: inputs, target radiance, weight 1 (valid) or 0. */
struct MlpSampleSoA {
    float* in;        /* [kMlpIn][stride] */
    float* target;    /* [3][stride] */
    float* valid;     /* [stride] */
    uint32_t stride;
};

struct MlpWeights {
    const uint16_t* w;    /* [kPackTotal] */
    const uint16_t* wt;   /* [kPackTTotal] */
    const float* bias;    /* [kBiasTotal] */
};

/* sin and cos of 2^k pi x for the four bands k = k0 .. k0 + 3 of one
 * input scalar (k0 a multiple of 4): band k0 from the shared sin/cos in
 * turns (2^k0 pi x = 2 pi (2^(k0-1) x), the scaling exact in binary32),
 * each next band by angle doubling, sin 2a = (2 sin a) cos a and
 * cos 2a = 1 - (2 sin a) sin a.  Three doublings keep the error within a
 * few units in the last place of 1, far below the BF16 rounding of the
 * features, at a third of the cost of four sin/cos evaluations. */
__host__ __device__ inline void mlp_band_quad(float x, int k0, float s[4], float c[4])
{
    tr_sincos_turns(x * __builtin_ldexpf(1.0f, k0 - 1), &s[0], &c[0]);
    for (int b = 1; b < 4; ++b) {
        float t = 2.0f * s[b - 1];
        s[b] = t * c[b - 1];
        c[b] = 1.0f - t * s[b - 1];
    }
}

/* A training input as the network reads it: samples of weight 0 (records
 * of paths that never reached a query vertex, padding up to a whole
 * slice) may hold any bits, so a value outside [-1e30, 1e30] or NaN reads
 * as 0; their features then stay finite and their zero pre-activation
 * gradients give exactly zero weight gradients. */
__host__ __device__ inline float mlp_clean_input(float x) { return fabsf(x) <= 1e30f ? x : 0.0f; }

/* Feature of an input scalar: band k, sin (c = 0) or cos (c = 1), by the
 * same operations as mlp_band_quad (bit-identical to it). */
__host__ __device__ inline float mlp_feature(float x, int k, int c)
{
    float s, co;
    tr_sincos_turns(x * __builtin_ldexpf(1.0f, (k & ~3) - 1), &s, &co);
    for (int b = 0; b < (k & 3); ++b) {
        float t = 2.0f * s;
        float s2 = t * co;
        co = 1.0f - t * s;
        s = s2;
    }
    return c ? co : s;
}

/* Layout of one weight-gradient partial ("gradient slots"), the order in
 * which the gradient kernel's matrix-core tiles leave the registers: 16x16
 * tiles of 256 floats, lane l's four rows 4 (l / 16) + i of column l % 16
 * at 4 l + i, so a wave stores a tile as 1 KB of consecutive bytes.
 *   tiles 0..63     layer 0: wave block w = 0..3, tile 4 nt + mt; column
 *                   l % 16 of block w is feature group q = 16 w + l % 16
 *                   (input q / 6, bands 2 (q % 6) + nt / 2, sin for even
 *                   nt), feature 4 q + nt; groups 60..63 are padding
 *   tiles 64..111   layers 1..3, 16 tiles each, tile 4 nt + mt
 *   tiles 112..115  layer 4, tile nt (rows 0..2 of 16 used)
 *   then the biases b0..b3 (64 each) and b4 (3, padded to 16).
 * mlp_grad_param maps a slot to its parameter index, or -1 for padding. */
constexpr int kGradTiles = 116;
constexpr int kGradBiasOff = kGradTiles * 256;
constexpr int kGradSlots = kGradBiasOff + 4 * 64 + 16;

__host__ __device__ inline int mlp_grad_param(int q)
{
    if (q >= kGradBiasOff) {
        const int b = q - kGradBiasOff, L = b / 64, m = b % 64;
        if (L < 4) return kMlpOffB[L] + m;
        return m < kMlpOut ? kMlpOffB[4] + m : -1;
    }
    const int tile = q >> 8, l = (q >> 2) & 63, i = q & 3;
    const int row = 4 * (l >> 4) + i, col = l & 15;
    if (tile < 64) {
        const int w = tile >> 4, nt = (tile >> 2) & 3, mt = tile & 3;
        const int grp = 16 * w + col;
        if (grp >= 60) return -1;
        return kMlpOffW[0] + (16 * mt + row) * kMlpFeat + 4 * grp + nt;
    }
    if (tile < 112) {
        const int L = 1 + (tile - 64) / 16, t = (tile - 64) % 16, nt = t >> 2, mt = t & 3;
        return kMlpOffW[L] + (16 * mt + row) * kMlpHidden + 16 * nt + col;
    }
    const int nt = tile - 112;
    return row < kMlpOut ? kMlpOffW[4] + row * kMlpHidden + 16 * nt + col : -1;
}

/* Device buffers of one training step (sized for `capacity` samples). */
struct MlpTrainBufs {
    uint16_t* ht;       /* [4][64][capacity]: H1..H4, [neuron][sample] */
    uint16_t* dzt;      /* [5][64][capacity]: dZ of layers 0..4 (layer 4: 16 rows used) */
    float* partial;     /* [slices][kGradSlots] */
    float* loss_part;   /* [grid] */
    uint32_t capacity;
};

/* Outputs of queries 0..n-1 to out[3 i + c]. */
void launch_mlp_infer(hipStream_t s, const MlpWeights& w, const MlpQueries& q, uint32_t n, float* out,
                      uint32_t grid);
/* Queries index[i], i < *count (count read on the device): radiance of
 * their slots += thr * max(prediction, 0). */
void launch_mlp_infer_counted(hipStream_t s, const MlpWeights& w, const MlpQueries& q, const uint32_t* index,
                              const uint32_t* count, float4* radiance, uint32_t grid);
void launch_mlp_train(hipStream_t s, const MlpWeights& w, const MlpSampleSoA& smp, uint32_t n, float loss_scale,
                      float inv_valid, const MlpTrainBufs& tb, uint32_t grid);
/* Weight-gradient partials of n samples (a multiple of 128): one partial
 * per slice, mlp_wgrad_slices(n) of them. */
uint32_t mlp_wgrad_slices(uint32_t n);
void launch_mlp_wgrad(hipStream_t s, const MlpSampleSoA& smp, uint32_t n, const MlpTrainBufs& tb);
void launch_mlp_adam(hipStream_t s, float* master, float* m, float* v, const float* partial, uint32_t slices,
                     float grad_scale, float lr, float bc1, float bc2, uint16_t* w, uint16_t* wt, float* bias,
                     int update);

}  // namespace trk

#endif
