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
    for (int j = 0; j < 4; ++j) w[j] = f2bf(f[2 * j]) | (f2bf(f[2 * j + 1]) << 16);
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
        int k0 = (c % 3) * 4;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            float t = x * __builtin_ldexpf(1.0f, k0 + b - 1);
            tr_sincos_turns(t, &f[2 * b], &f[2 * b + 1]);
        }
    }
    return pack8(f);
}

constexpr int kWaves = 4;          /* waves per block */
constexpr int kTile = 16;          /* samples per wave tile */
constexpr int kBlockTile = kWaves * kTile;

/* ---------------- inference ---------------- */

__global__ __launch_bounds__(256) void k_mlp_infer(MlpWeights W, const MlpQuery* q, const uint32_t* index,
                                                   uint32_t n_arg, const uint32_t* count, float* out,
                                                   float4* radiance)
{
    const uint32_t n = count ? *count : n_arg;
    if (n == 0) return;
    __shared__ __attribute__((aligned(16))) uint16_t sw[kPackTotal];
    __shared__ float sb[kBiasTotal];
    __shared__ float qin[kWaves][kTile][kMlpIn];
    __shared__ __attribute__((aligned(16))) uint16_t hq[kWaves][kTile][kMlpHidden];

    for (int i = threadIdx.x; i < kPackTotal / 8; i += blockDim.x)
        reinterpret_cast<uint4*>(sw)[i] = reinterpret_cast<const uint4*>(W.w)[i];
    for (int i = threadIdx.x; i < kBiasTotal; i += blockDim.x) sb[i] = W.bias[i];
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    for (uint32_t base = blockIdx.x * kBlockTile; base < n; base += gridDim.x * kBlockTile) {
        uint32_t qi = base + wave * kTile + col;
        bool valid = qi < n;
        if (g == 0) {
#pragma unroll
            for (int i = 0; i < kMlpIn; ++i) qin[wave][col][i] = valid ? q[index ? index[qi] : qi].in[i] : 0.0f;
        }
        __syncthreads();
        /* layer 0: 4 neuron tiles x 8 feature steps */
        f32x4 acc[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            int c = 4 * s + g;
            bf16x8 b = feature_chunk(qin[wave][col][c < 30 ? c / 3 : 0], c);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                acc[mt] = mfma(ld8(&sw[(16 * mt + col) * 256 + 32 * s + 8 * g]), b, acc[mt]);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int m = 16 * mt + 4 * g + i;
                hq[wave][col][m] = (uint16_t)f2bf(fmaxf(acc[mt][i] + sb[m], 0.0f));
            }
        __syncthreads();
        /* hidden layers 1..3 */
#pragma unroll
        for (int L = 1; L <= 3; ++L) {
            const uint16_t* WL = sw + kPackOff[L];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                bf16x8 b = ld8(&hq[wave][col][32 * s + 8 * g]);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma(ld8(&WL[(16 * mt + col) * 64 + 32 * s + 8 * g]), b, acc[mt]);
            }
            __syncthreads();
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    int m = 16 * mt + 4 * g + i;
                    hq[wave][col][m] = (uint16_t)f2bf(fmaxf(acc[mt][i] + sb[64 * L + m], 0.0f));
                }
            __syncthreads();
        }
        /* output layer: one 16-row tile, rows 0..2 used */
        f32x4 o = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; ++s)
            o = mfma(ld8(&sw[kPackOff[4] + col * 64 + 32 * s + 8 * g]), ld8(&hq[wave][col][32 * s + 8 * g]), o);
        if (g == 0 && valid) {
            float y0 = o[0] + sb[256], y1 = o[1] + sb[257], y2 = o[2] + sb[258];
            if (out) {
                out[qi * 3 + 0] = y0;
                out[qi * 3 + 1] = y1;
                out[qi * 3 + 2] = y2;
            }
            if (radiance) {
                const MlpQuery& r = q[index ? index[qi] : qi];
                float4 cur = radiance[r.slot];
                cur.x = cur.x + r.thr[0] * fmaxf(y0, 0.0f);
                cur.y = cur.y + r.thr[1] * fmaxf(y1, 0.0f);
                cur.z = cur.z + r.thr[2] * fmaxf(y2, 0.0f);
                radiance[r.slot] = cur;
            }
        }
        __syncthreads();
    }
}

/* ---------------- training: forward, loss gradient, backward data ---------------- */

__device__ inline bool bf_positive(uint16_t h) { return h != 0 && (h & 0x8000u) == 0; }

__global__ __launch_bounds__(256) void k_mlp_train(MlpWeights W, MlpSampleSoA smp, uint32_t n, float loss_scale,
                                                   float inv_valid, MlpTrainBufs tb)
{
    __shared__ float qin[kWaves][kTile][kMlpIn];
    __shared__ __attribute__((aligned(16))) uint16_t hq[kWaves][4][kTile][kMlpHidden];  /* H1..H4 */
    __shared__ __attribute__((aligned(16))) uint16_t dq[kWaves][kTile][kMlpHidden];
    __shared__ float red[256];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const uint32_t cap = tb.capacity;
    float loss_acc = 0.0f;
    for (uint32_t base = blockIdx.x * kBlockTile; base < n; base += gridDim.x * kBlockTile) {
        uint32_t b = base + wave * kTile + col;   /* n is a multiple of 64: always < n */
        if (g == 0) {
#pragma unroll
            for (int i = 0; i < kMlpIn; ++i) qin[wave][col][i] = smp.in[(size_t)i * smp.stride + b];
        }
        __syncthreads();
        /* forward */
        f32x4 acc[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            int c = 4 * s + g;
            bf16x8 bv = feature_chunk(qin[wave][col][c < 30 ? c / 3 : 0], c);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                acc[mt] = mfma(ld8(W.w + (16 * mt + col) * 256 + 32 * s + 8 * g), bv, acc[mt]);
        }
#pragma unroll
        for (int L = 0; L < 4; ++L) {
            if (L > 0) {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    bf16x8 bv = ld8(&hq[wave][L - 1][col][32 * s + 8 * g]);
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
                        acc[mt] = mfma(ld8(W.w + kPackOff[L] + (16 * mt + col) * 64 + 32 * s + 8 * g), bv, acc[mt]);
                }
            }
            /* H_{L+1} = ReLU(Z_L + b_L): to LDS and to ht[L] */
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    int m = 16 * mt + 4 * g + i;
                    uint16_t h = (uint16_t)f2bf(fmaxf(acc[mt][i] + W.bias[64 * L + m], 0.0f));
                    hq[wave][L][col][m] = h;
                    tb.ht[((size_t)L * 64 + m) * cap + b] = h;
                }
            __syncthreads();
        }
        f32x4 o = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; ++s)
            o = mfma(ld8(W.w + kPackOff[4] + col * 64 + 32 * s + 8 * g), ld8(&hq[wave][3][col][32 * s + 8 * g]), o);
        /* relative L2 loss gradient, scaled; rows 0..2 of lanes g == 0 */
        float d[3] = {0.0f, 0.0f, 0.0f};
        if (g == 0) {
            float vw = smp.valid[b];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float y = o[c] + W.bias[256 + c];
                float t = smp.target[(size_t)c * smp.stride + b];
                float w = vw * inv_valid / (y * y + 0.01f);
                float e = y - t;
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
            acc[mt] = mfma(ld8(W.wt + kPackTOff[4] + (16 * mt + col) * 32 + 8 * g), bz, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        /* backward through layers 3..0 */
#pragma unroll
        for (int L = 3; L >= 0; --L) {
            /* dZ_L = dH_{L+1} * [H_{L+1} > 0] */
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    int m = 16 * mt + 4 * g + i;
                    float v = bf_positive(hq[wave][L][col][m]) ? acc[mt][i] : 0.0f;
                    uint16_t h = (uint16_t)f2bf(v);
                    dq[wave][col][m] = h;
                    tb.dzt[((size_t)L * 64 + m) * cap + b] = h;
                }
            __syncthreads();
            if (L == 0) break;
            /* dH_L = W_L^T dZ_L */
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                bf16x8 bv = ld8(&dq[wave][col][32 * s + 8 * g]);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
                    acc[mt] = mfma(ld8(W.wt + kPackTOff[L] + (16 * mt + col) * 64 + 32 * s + 8 * g), bv, acc[mt]);
            }
            __syncthreads();
        }
    }
    red[threadIdx.x] = loss_acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) tb.loss_part[blockIdx.x] = red[0];
}

/* ---------------- weight gradients ---------------- */

/* Work items of one sample slice: layer 0 feature tiles 0..14 and its bias
 * (item 15), then for layers 1..4 four input tiles and the bias. */
constexpr int kWgradItems = 16 + 4 * 5;

__global__ __launch_bounds__(64) void k_mlp_wgrad(MlpSampleSoA smp, uint32_t n, MlpTrainBufs tb, uint32_t slices)
{
    const int item = blockIdx.x, slice = blockIdx.y;
    const int lane = threadIdx.x, g = lane >> 4, col = lane & 15;
    int L, nt;
    bool bias;
    if (item < 16) {
        L = 0;
        nt = item;
        bias = item == 15;
    } else {
        L = 1 + (item - 16) / 5;
        nt = (item - 16) % 5;
        bias = nt == 4;
    }
    const int mtiles = L == 4 ? 1 : 4;
    const uint32_t cap = tb.capacity;
    const uint32_t per = n / slices;              /* multiple of 32 */
    const uint32_t k_begin = slice * per, k_end = k_begin + per;
    f32x4 acc[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* dz = tb.dzt + (size_t)L * 64 * cap;
    for (uint32_t k0 = k_begin; k0 < k_end; k0 += 32) {
        bf16x8 bv;
        if (bias) {
            uint32_t one = col == 0 ? 0x3f80u : 0u;
            uint32_t w4[4] = {one | (one << 16), one | (one << 16), one | (one << 16), one | (one << 16)};
            __builtin_memcpy(&bv, w4, 16);
        } else if (L == 0) {
            int f = 16 * nt + col;                 /* < 240 */
            int in = f / 24, band = (f % 24) / 2, c = f & 1;
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = mlp_feature(smp.in[(size_t)in * smp.stride + k0 + 8 * g + j], band, c);
            bv = pack8(v);
        } else {
            bv = ld8(tb.ht + ((size_t)(L - 1) * 64 + 16 * nt + col) * cap + k0 + 8 * g);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
            if (mt < mtiles) acc[mt] = mfma(ld8(dz + (size_t)(16 * mt + col) * cap + k0 + 8 * g), bv, acc[mt]);
    }
    float* part = tb.partial + (size_t)slice * kMlpParams;
    const int fan_out = kMlpFanOut[L], fan_in = kMlpFanIn[L];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        if (mt >= mtiles) break;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int m = 16 * mt + 4 * g + i;
            if (m >= fan_out) continue;
            if (bias) {
                if (col == 0) part[kMlpOffB[L] + m] = acc[mt][i];
            } else {
                int nn = 16 * nt + col;
                if (nn < fan_in) part[kMlpOffW[L] + m * fan_in + nn] = acc[mt][i];
            }
        }
    }
}

/* ---------------- Adam and repacking ---------------- */

__global__ __launch_bounds__(256) void k_mlp_adam(float* master, float* mom, float* vel, const float* partial,
                                                  uint32_t slices, float grad_scale, float lr, float bc1, float bc2,
                                                  uint16_t* w, uint16_t* wt, float* bias, int update)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float x = master[p];
    if (update) {
        float gsum = 0.0f;
        for (uint32_t s = 0; s < slices; ++s) gsum += partial[(size_t)s * kMlpParams + p];
        float gr = gsum * grad_scale;
        float m = 0.9f * mom[p] + 0.1f * gr;
        float v = 0.99f * vel[p] + 0.01f * gr * gr;
        mom[p] = m;
        vel[p] = v;
        float mh = m / bc1, vh = v / bc2;
        x = x - lr * mh / (sqrtf(vh) + 1e-8f);
        master[p] = x;
    }
    int L = 4;
    while (L > 0 && p < kMlpOffW[L]) --L;
    int local = p - kMlpOffW[L];
    int fan_in = kMlpFanIn[L], fan_out = kMlpFanOut[L];
    uint16_t h = (uint16_t)f2bf(x);
    if (local < fan_out * fan_in) {
        int o = local / fan_in, i = local % fan_in;
        w[kPackOff[L] + o * kPackInPad[L] + i] = h;
        if (L > 0) wt[kPackTOff[L] + i * kPackTOutPad[L] + o] = h;
    } else {
        int o = local - fan_out * fan_in;
        bias[64 * L + o] = x;
    }
}

}  // namespace

void launch_mlp_infer(hipStream_t s, const MlpWeights& w, const MlpQuery* q, uint32_t n, float* out, float4* radiance,
                      uint32_t grid)
{
    if (n == 0) return;
    uint32_t need = (n + kBlockTile - 1) / kBlockTile;
    hipLaunchKernelGGL(k_mlp_infer, dim3(need < grid ? need : grid), dim3(256), 0, s, w, q, (const uint32_t*)nullptr,
                       n, (const uint32_t*)nullptr, out, radiance);
}

void launch_mlp_infer_counted(hipStream_t s, const MlpWeights& w, const MlpQuery* q, const uint32_t* index,
                              const uint32_t* count, float4* radiance, uint32_t grid)
{
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(256), 0, s, w, q, index, 0u, count, (float*)nullptr, radiance);
}

void launch_mlp_train(hipStream_t s, const MlpWeights& w, const MlpSampleSoA& smp, uint32_t n, float loss_scale,
                      float inv_valid, const MlpTrainBufs& tb, uint32_t grid)
{
    hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(256), 0, s, w, smp, n, loss_scale, inv_valid, tb);
}

void launch_mlp_wgrad(hipStream_t s, const MlpSampleSoA& smp, uint32_t n, const MlpTrainBufs& tb, uint32_t slices)
{
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(kWgradItems, slices), dim3(64), 0, s, smp, n, tb, slices);
}

void launch_mlp_adam(hipStream_t s, float* master, float* m, float* v, const float* partial, uint32_t slices,
                     float grad_scale, float lr, float bc1, float bc2, uint16_t* w, uint16_t* wt, float* bias, int update)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, master, m, v, partial, slices,
                       grad_scale, lr, bc1, bc2, w, wt, bias, update);
}

}  // namespace trk
