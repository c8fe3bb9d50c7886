/*
 * tr_mlp.hip - radiance-cache network kernels for gfx950 (no reference
 * counterpart).
 *
 * All products use v_mfma_f32_16x16x32_bf16 (D[16x16] += A[16x32] B[32x16],
 * BF16 in, FP32 accumulate).  Lane l of a wave supplies A[l % 16][8 (l/16)
 * + j] and B[8 (l/16) + j][l % 16] (j = 0..7) and receives D[4 (l/16) + i]
 * [l % 16] (i = 0..3).  Layers are evaluated transposed, Z^T = W H^T: the
 * weights are the A operand (rows = output neurons), a 16-sample tile is
 * the B operand (columns = samples), so a layer's output lands in D as
 * four consecutive neurons of one sample per lane and is written to LDS as
 * [sample][neuron], which is exactly the next layer's B layout.
 *
 *   k_mlp_infer  forward for cache queries, weights staged in LDS
 *   k_mlp_train  forward + relative-L2 loss gradient + backward data pass;
 *                writes activations and pre-activation gradients
 *                [neuron][sample] for the weight-gradient pass
 *   k_mlp_wgrad  dW = dZ H^T over sample slices (split K, deterministic)
 *   k_mlp_adam   sums the slices in order, Adam on the FP32 master, repacks
 *                the BF16 operand copies
 */
#include "tr_mlp.h"

#include <type_traits>

namespace trk {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

/* binary32 -> bfloat16, round to nearest even */
__device__ inline uint32_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ inline float bf2f(uint32_t h) { return __uint_as_float(h << 16); }

/* Two binary32 values to a packed BF16 pair (a in the low half), round to
 * nearest even, in one v_cvt_pk_bf16_f32. */
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t cvt2(float a, float b)
{
    bf16x2 r = __builtin_convertvector((f32x2){a, b}, bf16x2);
    uint32_t u;
    __builtin_memcpy(&u, &r, 4);
    return u;
}

__device__ inline bf16x8 ld8(const uint16_t* p)
{
    bf16x8 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}
__device__ inline bf16x8 pack8(const float f[8])
{
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) w[j] = cvt2(f[2 * j], f[2 * j + 1]);
    bf16x8 v;
    __builtin_memcpy(&v, w, 16);
    return v;
}
__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/* Eight consecutive features (chunk c = 0..31 of the 256) of one sample
 * whose input scalar c / 3 is x: bands 4 (c % 3) .. +3, (sin, cos) each.
 * Chunks 30 and 31 are the zero padding. */
__device__ inline bf16x8 feature_chunk(float x, int c)
{
    float f[8];
    if (c >= 30) {
#pragma unroll
        for (int j = 0; j < 8; ++j) f[j] = 0.0f;
    } else {
        float s[4], co[4];
        mlp_band_quad(x, (c % 3) * 4, s, co);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            f[2 * b] = s[b];
            f[2 * b + 1] = co[b];
        }
    }
    return pack8(f);
}

constexpr int kTile = 16;          /* samples per wave tile */

/* ---------------- LDS weight layout ---------------- */

/* The packed weights staged in LDS with padded rows: a 16-row MFMA
 * fragment read (16 bytes per lane) then spreads over all 64 banks instead
 * of hitting the same banks 16 times (unpadded rows are 128 or 512 bytes
 * apart).  Row strides in halves: layer 0 264, layers 1..4 72; transposes
 * of layers 1..3 72, of layer 4 (32 outputs padded) 40. */
constexpr int kLdsStride[5] = {264, 72, 72, 72, 72};
constexpr int kLdsOff[5] = {0, 64 * 264, 64 * 264 + 64 * 72, 64 * 264 + 2 * 64 * 72, 64 * 264 + 3 * 64 * 72};
constexpr int kLdsTotal = 64 * 264 + 3 * 64 * 72 + 16 * 72;
constexpr int kLdsTStride[5] = {0, 72, 72, 72, 40};
constexpr int kLdsTOff[5] = {0, 0, 64 * 72, 2 * 64 * 72, 3 * 64 * 72};
constexpr int kLdsTTotal = 3 * 64 * 72 + 64 * 40;

/* Copies `rows` rows of `row_len` halves (multiple of 8) from packed
 * global memory to LDS rows of `stride` halves. */
__device__ inline void stage_rows(uint16_t* dst, const uint16_t* src, int rows, int row_len, int stride)
{
    const int per_row = row_len / 8;
    for (int i = threadIdx.x; i < rows * per_row; i += blockDim.x) {
        const int r = i / per_row, c = i - r * per_row;
        *reinterpret_cast<uint4*>(dst + r * stride + 8 * c) = reinterpret_cast<const uint4*>(src + r * row_len)[c];
    }
}

__device__ inline void stage_forward(uint16_t* sw, const uint16_t* w)
{
#pragma unroll
    for (int L = 0; L < 5; ++L) stage_rows(sw + kLdsOff[L], w + kPackOff[L], L == 4 ? 16 : 64, kPackInPad[L], kLdsStride[L]);
}

__device__ inline void stage_transposed(uint16_t* swt, const uint16_t* wt)
{
#pragma unroll
    for (int L = 1; L < 5; ++L) stage_rows(swt + kLdsTOff[L], wt + kPackTOff[L], 64, kPackTOutPad[L], kLdsTStride[L]);
}

/* ---------------- inference ---------------- */

constexpr int kInferWaves = 8;                     /* waves per block (one weight copy) */
constexpr int kInferCols = 2;                      /* 16-query column tiles per wave */
constexpr int kInferWaveQ = 16 * kInferCols;       /* queries per wave iteration */
constexpr int kInferBlockQ = kInferWaves * kInferWaveQ;
constexpr int kHqStride = kMlpHidden + 8;          /* padded activation row (halves) */

/* Orders a wave's LDS writes before later LDS reads of other lanes of the
 * same wave: the inference kernel's waves each own their LDS tiles, so no
 * block barrier is needed between layers. */
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* The network inputs, throughput and slot of query i of the list. */
struct InferIn {
    float in[kMlpIn];
    float thr[3];
    uint32_t slot;
};

__device__ inline void infer_load(const MlpQueries& q, const uint32_t* index, uint32_t qi, uint32_t n, InferIn* d)
{
    if (qi < n) {
        const uint32_t k = index ? index[qi] : qi;
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) d->in[i] = q.in[(size_t)i * q.stride + k];
        if (q.thr) {
#pragma unroll
            for (int i = 0; i < 3; ++i) d->thr[i] = q.thr[(size_t)i * q.stride + k];
            d->slot = q.slot[k];
        }
    } else {
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) d->in[i] = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; ++i) d->thr[i] = 0.0f;
        d->slot = 0;
    }
}

/* Input scalar c / 3 (c = 4 s + g, chunk of k-step s) from the lane's
 * registers: for a fixed step s it is one of two neighbouring scalars. */
template <int S>
__device__ inline float chunk_input(const InferIn& d, int g)
{
    constexpr int lo = (4 * S) / 3;
    constexpr int hi = lo + 1 < kMlpIn ? lo + 1 : kMlpIn - 1;
    return (4 * S + g) / 3 == lo ? d.in[lo] : d.in[hi];
}

/* Writes relu(acc + bias) of neuron tile mt, column tile ct as BF16
 * activations (four consecutive neurons per lane, one 8-byte store). */
__device__ inline void store_act(uint16_t (*h)[kHqStride], const float* bias, int mt, int ct, int g, int col,
                                 f32x4 acc)
{
    int m = 16 * mt + 4 * g;
    uint32_t lo = cvt2(fmaxf(acc[0] + bias[m], 0.0f), fmaxf(acc[1] + bias[m + 1], 0.0f));
    uint32_t hi = cvt2(fmaxf(acc[2] + bias[m + 2], 0.0f), fmaxf(acc[3] + bias[m + 3], 0.0f));
    *reinterpret_cast<uint2*>(&h[16 * ct + col][m]) = make_uint2(lo, hi);
}

/* Forward pass for queries: the packed BF16 weights and the biases are
 * staged in LDS once per block; each wave then runs 32 queries per
 * iteration (two 16-query column tiles sharing every weight fragment),
 * its activations in its own LDS tile, and prefetches the next
 * iteration's inputs while the matrix cores work. */
__global__ __launch_bounds__(kInferWaves * 64) void k_mlp_infer(MlpWeights W, MlpQueries q, const uint32_t* index,
                                                   uint32_t n_arg, const uint32_t* count, float* out,
                                                   float4* radiance)
{
    const uint32_t n = count ? *count : n_arg;
    if (n == 0) return;
    __shared__ __attribute__((aligned(16))) uint16_t sw[kLdsTotal];
    __shared__ float sb[kBiasTotal];
    __shared__ __attribute__((aligned(16))) uint16_t hq[kInferWaves][kInferWaveQ][kHqStride];

    stage_forward(sw, W.w);
    for (int i = threadIdx.x; i < kBiasTotal; i += blockDim.x) sb[i] = W.bias[i];
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    uint16_t (*h)[kHqStride] = hq[wave];
    const uint32_t step = gridDim.x * kInferBlockQ;
    uint32_t base = blockIdx.x * kInferBlockQ + wave * kInferWaveQ;
    InferIn cur[kInferCols];
#pragma unroll
    for (int ct = 0; ct < kInferCols; ++ct) infer_load(q, index, base + 16 * ct + col, n, &cur[ct]);
    for (; base < n; base += step) {
        InferIn nxt[kInferCols];
#pragma unroll
        for (int ct = 0; ct < kInferCols; ++ct) infer_load(q, index, base + step + 16 * ct + col, n, &nxt[ct]);
        /* layer 0: 4 neuron tiles x 8 feature steps, features made in registers */
        f32x4 acc[4][kInferCols];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int ct = 0; ct < kInferCols; ++ct) acc[mt][ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        auto layer0_step = [&](auto S) {
            constexpr int s = decltype(S)::value;
            const int c = 4 * s + g;
            bf16x8 b[kInferCols];
#pragma unroll
            for (int ct = 0; ct < kInferCols; ++ct) b[ct] = feature_chunk(chunk_input<s>(cur[ct], g), c);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                bf16x8 a = ld8(&sw[(16 * mt + col) * kLdsStride[0] + 32 * s + 8 * g]);
#pragma unroll
                for (int ct = 0; ct < kInferCols; ++ct) acc[mt][ct] = mfma(a, b[ct], acc[mt][ct]);
            }
        };
        layer0_step(std::integral_constant<int, 0>{});
        layer0_step(std::integral_constant<int, 1>{});
        layer0_step(std::integral_constant<int, 2>{});
        layer0_step(std::integral_constant<int, 3>{});
        layer0_step(std::integral_constant<int, 4>{});
        layer0_step(std::integral_constant<int, 5>{});
        layer0_step(std::integral_constant<int, 6>{});
        layer0_step(std::integral_constant<int, 7>{});
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int ct = 0; ct < kInferCols; ++ct) store_act(h, sb, mt, ct, g, col, acc[mt][ct]);
        wave_lds_sync();
        /* hidden layers 1..3 */
#pragma unroll
        for (int L = 1; L <= 3; ++L) {
            const uint16_t* WL = sw + kLdsOff[L];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int ct = 0; ct < kInferCols; ++ct) acc[mt][ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                bf16x8 b[kInferCols];
#pragma unroll
                for (int ct = 0; ct < kInferCols; ++ct) b[ct] = ld8(&h[16 * ct + col][32 * s + 8 * g]);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    bf16x8 a = ld8(&WL[(16 * mt + col) * kLdsStride[L] + 32 * s + 8 * g]);
#pragma unroll
                    for (int ct = 0; ct < kInferCols; ++ct) acc[mt][ct] = mfma(a, b[ct], acc[mt][ct]);
                }
            }
            wave_lds_sync();
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int ct = 0; ct < kInferCols; ++ct) store_act(h, sb + 64 * L, mt, ct, g, col, acc[mt][ct]);
            wave_lds_sync();
        }
        /* output layer: one 16-row tile per column tile, rows 0..2 used */
        f32x4 o[kInferCols];
#pragma unroll
        for (int ct = 0; ct < kInferCols; ++ct) o[ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            bf16x8 a = ld8(&sw[kLdsOff[4] + col * kLdsStride[4] + 32 * s + 8 * g]);
#pragma unroll
            for (int ct = 0; ct < kInferCols; ++ct) o[ct] = mfma(a, ld8(&h[16 * ct + col][32 * s + 8 * g]), o[ct]);
        }
        wave_lds_sync();
        if (g == 0) {
#pragma unroll
            for (int ct = 0; ct < kInferCols; ++ct) {
                uint32_t qi = base + 16 * ct + col;
                if (qi >= n) continue;
                float y0 = o[ct][0] + sb[256], y1 = o[ct][1] + sb[257], y2 = o[ct][2] + sb[258];
                if (out) {
                    out[qi * 3 + 0] = y0;
                    out[qi * 3 + 1] = y1;
                    out[qi * 3 + 2] = y2;
                }
                if (radiance) {
                    float4 r = radiance[cur[ct].slot];
                    r.x = r.x + cur[ct].thr[0] * fmaxf(y0, 0.0f);
                    r.y = r.y + cur[ct].thr[1] * fmaxf(y1, 0.0f);
                    r.z = r.z + cur[ct].thr[2] * fmaxf(y2, 0.0f);
                    radiance[cur[ct].slot] = r;
                }
            }
        }
#pragma unroll
        for (int ct = 0; ct < kInferCols; ++ct) cur[ct] = nxt[ct];
    }
}

/* ---------------- training: forward, loss gradient, backward data ---------------- */

__device__ inline bool bf_positive(uint16_t h) { return h != 0 && (h & 0x8000u) == 0; }

constexpr int kTrainWaves = 8;                      /* waves per training block */
constexpr int kTrainBlockQ = kTrainWaves * kTile;   /* samples per block iteration */

/* relu(acc + bias) of neuron tile mt as BF16: written to the wave's LDS
 * tile (row = sample) and to ht / dzt-style global storage
 * [neuron][sample]; returns the four positive-bits (bit i: neuron
 * 16 mt + 4 g + i). */
__device__ inline uint32_t store_h(uint16_t (*h)[kHqStride], uint16_t* glob, uint32_t cap, uint32_t b,
                                   const float* bias, int mt, int g, int col, f32x4 acc)
{
    const int m = 16 * mt + 4 * g;
    const uint32_t lo = cvt2(fmaxf(acc[0] + bias[m], 0.0f), fmaxf(acc[1] + bias[m + 1], 0.0f));
    const uint32_t hi = cvt2(fmaxf(acc[2] + bias[m + 2], 0.0f), fmaxf(acc[3] + bias[m + 3], 0.0f));
    const uint32_t hv[4] = {lo & 0xffffu, lo >> 16, hi & 0xffffu, hi >> 16};
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        glob[(size_t)(m + i) * cap + b] = (uint16_t)hv[i];
        bits |= (bf_positive((uint16_t)hv[i]) ? 1u : 0u) << i;
    }
    *reinterpret_cast<uint2*>(&h[col][m]) = make_uint2(lo, hi);
    return bits;
}

/* One training step's per-sample work for 16 samples per wave and
 * iteration: forward pass (activations H1..H4 to global for the weight
 * gradient, their positive-bits kept in registers), relative-L2 loss and
 * its scaled gradient, backward data pass (pre-activation gradients dZ to
 * global).  The packed forward and transposed weights are staged in LDS
 * once per block; each wave exchanges layer outputs through its own LDS
 * tile, so only the staging needs a block barrier. */
__global__ __launch_bounds__(kTrainWaves * 64) void k_mlp_train(MlpWeights W, MlpSampleSoA smp, uint32_t n,
                                                                float loss_scale, float inv_valid, MlpTrainBufs tb)
{
    __shared__ float red[kTrainWaves];
    if (blockIdx.x * kTrainBlockQ >= n) {
        if (threadIdx.x == 0) tb.loss_part[blockIdx.x] = 0.0f;
        return;
    }
    __shared__ __attribute__((aligned(16))) uint16_t sw[kLdsTotal];
    __shared__ __attribute__((aligned(16))) uint16_t swt[kLdsTTotal];
    __shared__ float sb[kBiasTotal];
    __shared__ __attribute__((aligned(16))) uint16_t hq[kTrainWaves][kTile][kHqStride];
    stage_forward(sw, W.w);
    stage_transposed(swt, W.wt);
    for (int i = threadIdx.x; i < kBiasTotal; i += blockDim.x) sb[i] = W.bias[i];
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    uint16_t (*h)[kHqStride] = hq[wave];
    const uint32_t cap = tb.capacity;
    float loss_acc = 0.0f;
    for (uint32_t base = blockIdx.x * kTrainBlockQ; base < n; base += gridDim.x * kTrainBlockQ) {
        const uint32_t b = base + wave * kTile + col;   /* n is a multiple of 512: always < n */
        InferIn x;
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) x.in[i] = mlp_clean_input(smp.in[(size_t)i * smp.stride + b]);
        /* forward, layer 0 */
        f32x4 acc[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        auto layer0_step = [&](auto S) {
            constexpr int s = decltype(S)::value;
            bf16x8 bv = feature_chunk(chunk_input<s>(x, g), 4 * s + g);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                acc[mt] = mfma(ld8(&sw[(16 * mt + col) * kLdsStride[0] + 32 * s + 8 * g]), bv, acc[mt]);
        };
        layer0_step(std::integral_constant<int, 0>{});
        layer0_step(std::integral_constant<int, 1>{});
        layer0_step(std::integral_constant<int, 2>{});
        layer0_step(std::integral_constant<int, 3>{});
        layer0_step(std::integral_constant<int, 4>{});
        layer0_step(std::integral_constant<int, 5>{});
        layer0_step(std::integral_constant<int, 6>{});
        layer0_step(std::integral_constant<int, 7>{});
        uint32_t pos[4];   /* positive-bits of H1..H4: bit 4 mt + i */
#pragma unroll
        for (int L = 0; L < 4; ++L) {
            if (L > 0) {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    bf16x8 bv = ld8(&h[col][32 * s + 8 * g]);
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
                        acc[mt] = mfma(ld8(&sw[kLdsOff[L] + (16 * mt + col) * kLdsStride[L] + 32 * s + 8 * g]), bv,
                                       acc[mt]);
                }
                wave_lds_sync();
            }
            /* H_{L+1} = ReLU(Z_L + b_L) */
            uint16_t* glob = tb.ht + (size_t)L * 64 * cap;
            pos[L] = 0;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) pos[L] |= store_h(h, glob, cap, b, sb + 64 * L, mt, g, col, acc[mt]) << (4 * mt);
            wave_lds_sync();
        }
        f32x4 o = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; ++s)
            o = mfma(ld8(&sw[kLdsOff[4] + col * kLdsStride[4] + 32 * s + 8 * g]), ld8(&h[col][32 * s + 8 * g]), o);
        wave_lds_sync();
        /* relative L2 loss gradient, scaled; rows 0..2 of lanes g == 0 */
        float d[3] = {0.0f, 0.0f, 0.0f};
        if (g == 0) {
            float vw = smp.valid[b];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float y = o[c] + sb[256 + c];
                float t = smp.target[(size_t)c * smp.stride + b];
                /* weight-0 samples contribute nothing, whatever their target holds */
                float w = vw != 0.0f ? vw * inv_valid / (y * y + 0.01f) : 0.0f;
                float e = vw != 0.0f ? y - t : 0.0f;
                loss_acc += w * e * e;
                d[c] = 2.0f * w * e * loss_scale;
            }
        }
        uint32_t dz4[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dz4[c] = f2bf(d[c]);
            if (g == 0) tb.dzt[((size_t)4 * 64 + c) * cap + b] = (uint16_t)dz4[c];
        }
        /* dH4 = W4^T dZ4 (K padded to 32: only rows 0..2 of lanes g == 0) */
        bf16x8 bz;
        {
            uint32_t w4[4] = {0u, 0u, 0u, 0u};
            if (g == 0) {
                w4[0] = dz4[0] | (dz4[1] << 16);
                w4[1] = dz4[2];
            }
            __builtin_memcpy(&bz, w4, 16);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
            acc[mt] = mfma(ld8(&swt[kLdsTOff[4] + (16 * mt + col) * kLdsTStride[4] + 8 * g]), bz,
                           f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        /* backward through layers 3..0 */
#pragma unroll
        for (int L = 3; L >= 0; --L) {
            /* dZ_L = dH_{L+1} * [H_{L+1} > 0] */
            uint16_t* glob = tb.dzt + (size_t)L * 64 * cap;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                const int m = 16 * mt + 4 * g;
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = (pos[L] >> (4 * mt + i)) & 1u ? acc[mt][i] : 0.0f;
                const uint32_t lo = cvt2(v[0], v[1]), hi = cvt2(v[2], v[3]);
                glob[(size_t)m * cap + b] = (uint16_t)(lo & 0xffffu);
                glob[(size_t)(m + 1) * cap + b] = (uint16_t)(lo >> 16);
                glob[(size_t)(m + 2) * cap + b] = (uint16_t)(hi & 0xffffu);
                glob[(size_t)(m + 3) * cap + b] = (uint16_t)(hi >> 16);
                if (L > 0) *reinterpret_cast<uint2*>(&h[col][m]) = make_uint2(lo, hi);
            }
            if (L == 0) break;
            wave_lds_sync();
            /* dH_L = W_L^T dZ_L */
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                bf16x8 bv = ld8(&h[col][32 * s + 8 * g]);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
                    acc[mt] = mfma(ld8(&swt[kLdsTOff[L] + (16 * mt + col) * kLdsTStride[L] + 32 * s + 8 * g]), bv,
                                   acc[mt]);
            }
            wave_lds_sync();
        }
    }
    /* block sum of the loss, in a fixed order */
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off);
    if (lane == 0) red[wave] = loss_acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < kTrainWaves; ++w) t += red[w];
        tb.loss_part[blockIdx.x] = t;
    }
}

/* ---------------- weight gradients ---------------- */

/* Sine and cosine of bands b0 and b0 + 1 (b0 even) of input scalar x by
 * the forward pass's own operations (mlp_band_quad: the shared sin/cos of
 * band b0 & ~3, then angle doublings), so bit-identical to its features. */
__device__ inline void band_pair(float x, int b0, float* s0, float* c0, float* s1, float* c1)
{
    float s, c;
    tr_sincos_turns(x * __builtin_ldexpf(1.0f, (b0 & ~3) - 1), &s, &c);
    if (b0 & 2) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const float t = 2.0f * s;
            const float s2 = t * c;
            c = 1.0f - t * s;
            s = s2;
        }
    }
    *s0 = s;
    *c0 = c;
    const float t = 2.0f * s;
    *s1 = t * c;
    *c1 = 1.0f - t * s;
}

/* Sum of a fragment's eight BF16 values in FP32, in order. */
__device__ inline float sum8(bf16x8 v)
{
    uint16_t h[8];
    __builtin_memcpy(h, &v, 16);
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += bf2f(h[j]);
    return s;
}

/* Row sums of a wave's A fragments: each lane holds 8 samples of row
 * col of each tile; the four lane groups are added in a fixed order. */
__device__ inline float rows_total(float v)
{
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

/* dW_L = dZ_L H_L^T and db_L = dZ_L 1 over one slice of kWgradK samples
 * per block; every wave owns one 64-row output block and loads its operand
 * fragments (16 bytes per lane) for the whole slice before the matrix-core
 * work:
 *   waves 0-3  layer 0.  Lane col of wave w owns feature group
 *              q = 16 w + col: input q / 6, bands 2 (q % 6) and +1; column
 *              tile nt holds band 2 (q % 6) + nt / 2, sin (nt even) or cos.
 *              Groups 60..63 are padding.  The features are made from the
 *              inputs here, one sin/cos per sample and group.  Wave 0 also
 *              sums b0's gradient;
 *   waves 4-6  layers 1..3 (64 x 64 each) and their biases;
 *   wave 7     layer 4 (one 16-row tile, rows 0..2 used) and b4.
 * Bias gradients are row sums of the dZ fragments (FP32, fixed order).
 * Tiles are stored whole, 16 bytes per lane, in the gradient-slot layout
 * (mlp_grad_param); each block writes its own partial and k_mlp_adam sums
 * the slices in a fixed order, so training is deterministic. */
constexpr int kWgradK = 128;                 /* samples per slice */
constexpr int kWgradSteps = kWgradK / 32;    /* MFMA k-steps per slice */

__device__ inline void store_tile(float* part, int tile, int lane, f32x4 v)
{
    *reinterpret_cast<f32x4*>(part + tile * 256 + 4 * lane) = v;
}

__global__ __launch_bounds__(512) void k_mlp_wgrad(MlpSampleSoA smp, MlpTrainBufs tb)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const size_t cap = tb.capacity;
    const uint32_t k0 = blockIdx.x * kWgradK + 8 * g;
    float* part = tb.partial + (size_t)blockIdx.x * kGradSlots;
    if (wave < 4) {
        const int q = 16 * wave + col;
        const bool real = q < 60;
        const int inp = real ? q / 6 : 0, b0 = real ? 2 * (q % 6) : 0;
        bf16x8 a[kWgradSteps][4];
        float4 x[kWgradSteps][2];
#pragma unroll
        for (int u = 0; u < kWgradSteps; ++u) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) a[u][mt] = ld8(tb.dzt + (16 * mt + col) * cap + k0 + 32 * u);
            const float4* src = reinterpret_cast<const float4*>(smp.in + (size_t)inp * smp.stride + k0 + 32 * u);
            x[u][0] = src[0];
            x[u][1] = src[1];
        }
        f32x4 acc[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int u = 0; u < kWgradSteps; ++u) {
            float v[8] = {x[u][0].x, x[u][0].y, x[u][0].z, x[u][0].w, x[u][1].x, x[u][1].y, x[u][1].z, x[u][1].w};
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = mlp_clean_input(v[j]);
            float f[4][8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                band_pair(v[j], b0, &f[0][j], &f[1][j], &f[2][j], &f[3][j]);
                if (!real) f[0][j] = f[1][j] = f[2][j] = f[3][j] = 0.0f;
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const bf16x8 b = pack8(f[nt]);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) acc[mt][nt] = mfma(a[u][mt], b, acc[mt][nt]);
            }
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) store_tile(part, 16 * wave + 4 * nt + mt, lane, acc[mt][nt]);
        if (wave == 0) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                float r = 0.0f;
#pragma unroll
                for (int u = 0; u < kWgradSteps; ++u) r += sum8(a[u][mt]);
                r = rows_total(r);
                if (g == 0) part[kGradBiasOff + 16 * mt + col] = r;
            }
        }
    } else if (wave < 7) {
        const int L = wave - 3;
        const uint16_t* dz = tb.dzt + (size_t)L * 64 * cap;
        const uint16_t* h = tb.ht + (size_t)(L - 1) * 64 * cap;
        bf16x8 a[kWgradSteps][4], b[kWgradSteps][4];
#pragma unroll
        for (int u = 0; u < kWgradSteps; ++u)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                a[u][t] = ld8(dz + (16 * t + col) * cap + k0 + 32 * u);
                b[u][t] = ld8(h + (16 * t + col) * cap + k0 + 32 * u);
            }
        f32x4 acc[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int u = 0; u < kWgradSteps; ++u)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = mfma(a[u][mt], b[u][nt], acc[mt][nt]);
        const int base = 64 + 16 * (L - 1);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) store_tile(part, base + 4 * nt + mt, lane, acc[mt][nt]);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            float r = 0.0f;
#pragma unroll
            for (int u = 0; u < kWgradSteps; ++u) r += sum8(a[u][mt]);
            r = rows_total(r);
            if (g == 0) part[kGradBiasOff + 64 * L + 16 * mt + col] = r;
        }
    } else {
        const uint16_t* dz4 = tb.dzt + (size_t)4 * 64 * cap;
        const uint16_t* h4 = tb.ht + (size_t)3 * 64 * cap;
        bf16x8 a[kWgradSteps], b[kWgradSteps][4];
#pragma unroll
        for (int u = 0; u < kWgradSteps; ++u) {
            a[u] = ld8(dz4 + col * cap + k0 + 32 * u);
#pragma unroll
            for (int t = 0; t < 4; ++t) b[u][t] = ld8(h4 + (16 * t + col) * cap + k0 + 32 * u);
        }
        f32x4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int u = 0; u < kWgradSteps; ++u)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = mfma(a[u], b[u][t], acc[t]);
#pragma unroll
        for (int t = 0; t < 4; ++t) store_tile(part, 112 + t, lane, acc[t]);
        float r = 0.0f;
#pragma unroll
        for (int u = 0; u < kWgradSteps; ++u) r += sum8(a[u]);
        r = rows_total(r);
        if (g == 0) part[kGradBiasOff + 256 + col] = r;
    }
}

/* ---------------- Adam and repacking ---------------- */

/* 32 gradient slots per 256-thread block; the eight 32-thread groups sum
 * slices g, g + 8, g + 16, ... (loads coalesced over consecutive slots,
 * eight in flight per thread), then the eight group sums are added in a
 * fixed order: the result does not depend on timing.  The slot's
 * parameter (mlp_grad_param) then takes the Adam step and its BF16 copies
 * are repacked.  With update 0 every parameter is only repacked. */
constexpr int kAdamSlots = 32;
constexpr int kAdamGroups = 8;

__device__ inline void repack_param(int p, float x, uint16_t* w, uint16_t* wt, float* bias)
{
    int L = 4;
    while (L > 0 && p < kMlpOffW[L]) --L;
    const int local = p - kMlpOffW[L];
    const int fan_in = kMlpFanIn[L], fan_out = kMlpFanOut[L];
    const uint16_t h = (uint16_t)f2bf(x);
    if (local < fan_out * fan_in) {
        const int o = local / fan_in, i = local % fan_in;
        w[kPackOff[L] + o * kPackInPad[L] + i] = h;
        if (L > 0) wt[kPackTOff[L] + i * kPackTOutPad[L] + o] = h;
    } else {
        bias[64 * L + (local - fan_out * fan_in)] = x;
    }
}

__global__ __launch_bounds__(256) void k_mlp_adam(float* master, float* mom, float* vel, const float* partial,
                                                  uint32_t slices, float grad_scale, float lr, float bc1, float bc2,
                                                  uint16_t* w, uint16_t* wt, float* bias, int update)
{
    if (!update) {
        const int p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p < kMlpParams) repack_param(p, master[p], w, wt, bias);
        return;
    }
    __shared__ float red[kAdamGroups][kAdamSlots];
    const int q = blockIdx.x * kAdamSlots + (threadIdx.x & (kAdamSlots - 1));
    const int grp = threadIdx.x / kAdamSlots;
    float gsum = 0.0f;
    if (q < kGradSlots) {
        uint32_t s = grp;
        for (; s + 7 * kAdamGroups < slices; s += 8 * kAdamGroups) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partial[(size_t)(s + u * kAdamGroups) * kGradSlots + q];
#pragma unroll
            for (int u = 0; u < 8; ++u) gsum += v[u];
        }
        for (; s < slices; s += kAdamGroups) gsum += partial[(size_t)s * kGradSlots + q];
    }
    red[grp][threadIdx.x & (kAdamSlots - 1)] = gsum;
    __syncthreads();
    if (grp != 0 || q >= kGradSlots) return;
#pragma unroll
    for (int k = 1; k < kAdamGroups; ++k) gsum += red[k][threadIdx.x];
    const int p = mlp_grad_param(q);
    if (p < 0) return;
    float x = master[p];
    const float gr = gsum * grad_scale;
    const float m = 0.9f * mom[p] + 0.1f * gr;
    const float v = 0.99f * vel[p] + 0.01f * gr * gr;
    mom[p] = m;
    vel[p] = v;
    const float mh = m / bc1, vh = v / bc2;
    x = x - lr * mh / (sqrtf(vh) + 1e-8f);
    master[p] = x;
    repack_param(p, x, w, wt, bias);
}

}  // namespace

void launch_mlp_infer(hipStream_t s, const MlpWeights& w, const MlpQueries& q, uint32_t n, float* out, uint32_t grid)
{
    if (n == 0) return;
    uint32_t need = (n + kInferBlockQ - 1) / kInferBlockQ;
    hipLaunchKernelGGL(k_mlp_infer, dim3(need < grid ? need : grid), dim3(kInferWaves * 64), 0, s, w, q,
                       (const uint32_t*)nullptr,
                       n, (const uint32_t*)nullptr, out, (float4*)nullptr);
}

void launch_mlp_infer_counted(hipStream_t s, const MlpWeights& w, const MlpQueries& q, const uint32_t* index,
                              const uint32_t* count, float4* radiance, uint32_t grid)
{
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(kInferWaves * 64), 0, s, w, q, index, 0u, count, (float*)nullptr,
                       radiance);
}

void launch_mlp_train(hipStream_t s, const MlpWeights& w, const MlpSampleSoA& smp, uint32_t n, float loss_scale,
                      float inv_valid, const MlpTrainBufs& tb, uint32_t grid)
{
    hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(kTrainWaves * 64), 0, s, w, smp, n, loss_scale, inv_valid, tb);
}

uint32_t mlp_wgrad_slices(uint32_t n) { return n / kWgradK; }

void launch_mlp_wgrad(hipStream_t s, const MlpSampleSoA& smp, uint32_t n, const MlpTrainBufs& tb)
{
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(n / kWgradK), dim3(512), 0, s, smp, tb);
}

void launch_mlp_adam(hipStream_t s, float* master, float* m, float* v, const float* partial, uint32_t slices,
                     float grad_scale, float lr, float bc1, float bc2, uint16_t* w, uint16_t* wt, float* bias, int update)
{
    const uint32_t grid = update ? (kGradSlots + kAdamSlots - 1) / kAdamSlots : (kMlpParams + 255) / 256;
    hipLaunchKernelGGL(k_mlp_adam, dim3(grid), dim3(256), 0, s, master, m, v, partial, slices, grad_scale, lr, bc1,
                       bc2, w, wt, bias, update);
}

}  // namespace trk
