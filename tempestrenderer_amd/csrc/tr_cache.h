/*
 * tr_cache.h - device state of the radiance-cache network (FP32 master
 * weights, Adam moments, BF16 operand copies, training buffers).
 */
#ifndef TR_CACHE_H
#define TR_CACHE_H

#include "tr_host.h"
#include "tr_mlp.h"

namespace trk {

struct MlpState {
    static constexpr uint32_t kSlice = 512;      /* samples per weight-gradient slice */
    static constexpr uint32_t kTrainGrid = 256;  /* blocks of the training kernel */
    static constexpr uint32_t kMaxLossSlots = 64; /* per-step loss partials kept per frame */

    hipStream_t stream = nullptr;
    int cu_count = 0;
    DevBuf<float> master, mom, vel, bias;
    DevBuf<uint16_t> w, wt;
    DevBuf<uint16_t> ht, dzt;
    DevBuf<float> partial, loss_part;
    DevBuf<float> s_in, s_target, s_valid;
    uint32_t train_cap = 0;
    uint32_t last_slices = 0;
    uint32_t step = 0;
    float loss_scale = 128.0f;
    float lr = 1e-2f;

    tr_status init(uint64_t seed);
    tr_status set_params(const float* p);
    tr_status ensure_train(uint32_t n);
    tr_status upload_samples(size_t n, const float* inputs, const float* targets, const float* valid,
                             uint32_t* n_valid);
    tr_status enqueue_step(uint32_t n, uint32_t n_valid, hipStream_t s, bool update, uint32_t offset = 0,
                           uint32_t loss_slot = 0);
    /* loss of one step (slot), or the mean of slots [0, n) weighted by `w` */
    float read_loss(uint32_t n_slots = 1, const float* w = nullptr);
    MlpWeights weights() const;
    MlpSampleSoA samples() const;
    MlpTrainBufs train_bufs() const;
};

}  // namespace trk

struct tr_mlp {
    trk::MlpState st;
    trk::DevBuf<float> q;   /* query inputs, [kMlpIn][n] */
    trk::DevBuf<float> out;
};

#endif
