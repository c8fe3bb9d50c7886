/*
 * tr_cache.cpp - the radiance-cache network behind the C ABI: parameter
 * state (FP32 master, Adam moments, BF16 operand copies), inference and one
 * training step over device-resident samples, and host-pointer entry
 * points for checking it.  No reference counterpart.
 */
#include "tr_cache.h"

#include <cmath>
#include <cstring>
#include <vector>

#include "tr_math.h"

using namespace trk;

namespace trk {

tr_status MlpState::init(uint64_t seed)
{
    int dev = 0;
    TR_HIP(hipGetDevice(&dev));
    TR_HIP(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, dev));
    TR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    TR_HIP(master.reserve(kMlpParams));
    TR_HIP(mom.reserve(kMlpParams));
    TR_HIP(vel.reserve(kMlpParams));
    TR_HIP(w.reserve(kPackTotal));
    TR_HIP(wt.reserve(kPackTTotal));
    TR_HIP(bias.reserve(kBiasTotal));
    TR_HIP(hipMemsetAsync(w.ptr, 0, kPackTotal * sizeof(uint16_t), stream));
    TR_HIP(hipMemsetAsync(wt.ptr, 0, kPackTTotal * sizeof(uint16_t), stream));
    TR_HIP(hipMemsetAsync(bias.ptr, 0, kBiasTotal * sizeof(float), stream));
    /* He-uniform weights U(-sqrt(6 / fan_in), +) from Philox(index, seed), zero biases */
    std::vector<float> p(kMlpParams, 0.0f);
    for (int L = 0; L < 5; ++L) {
        float lim = sqrtf(6.0f / (float)kMlpFanIn[L]);
        for (int i = 0; i < kMlpFanOut[L] * kMlpFanIn[L]; ++i) {
            uint32_t idx = (uint32_t)(kMlpOffW[L] + i);
            U4 r = philox(idx, 0x4d4c50u, 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
            p[idx] = (2.0f * u01(r.a) - 1.0f) * lim;
        }
    }
    return set_params(p.data());
}

tr_status MlpState::set_params(const float* p)
{
    TR_HIP(hipMemcpyAsync(master.ptr, p, kMlpParams * sizeof(float), hipMemcpyHostToDevice, stream));
    TR_HIP(hipMemsetAsync(mom.ptr, 0, kMlpParams * sizeof(float), stream));
    TR_HIP(hipMemsetAsync(vel.ptr, 0, kMlpParams * sizeof(float), stream));
    step = 0;
    launch_mlp_adam(stream, master.ptr, mom.ptr, vel.ptr, nullptr, 0, 0.0f, 0.0f, 1.0f, 1.0f, w.ptr, wt.ptr, bias.ptr, 0);
    TR_HIP(hipGetLastError());
    TR_HIP(hipStreamSynchronize(stream));
    return TR_OK;
}

tr_status MlpState::ensure_train(uint32_t n)
{
    if (n <= train_cap) return TR_OK;
    uint32_t cap = (n + kSlice - 1) / kSlice * kSlice;
    TR_HIP(ht.reserve((size_t)4 * 64 * cap));
    TR_HIP(dzt.reserve((size_t)5 * 64 * cap));
    TR_HIP(hipMemsetAsync(dzt.ptr, 0, (size_t)5 * 64 * cap * sizeof(uint16_t), stream));
    TR_HIP(partial.reserve((size_t)mlp_wgrad_slices(cap) * kGradSlots));
    TR_HIP(loss_part.reserve((size_t)kTrainGrid * kMaxLossSlots));
    TR_HIP(s_in.reserve((size_t)kMlpIn * cap));
    TR_HIP(s_target.reserve((size_t)3 * cap));
    TR_HIP(s_valid.reserve(cap));
    /* the zeroed rows are read by steps on other streams (the scene's) */
    TR_HIP(hipStreamSynchronize(stream));
    train_cap = cap;
    return TR_OK;
}

MlpWeights MlpState::weights() const
{
    MlpWeights r;
    r.w = w.ptr;
    r.wt = wt.ptr;
    r.bias = bias.ptr;
    return r;
}

MlpSampleSoA MlpState::samples() const
{
    MlpSampleSoA r;
    r.in = s_in.ptr;
    r.target = s_target.ptr;
    r.valid = s_valid.ptr;
    r.stride = train_cap;
    return r;
}

MlpTrainBufs MlpState::train_bufs() const
{
    MlpTrainBufs r;
    r.ht = ht.ptr;
    r.dzt = dzt.ptr;
    r.partial = partial.ptr;
    r.loss_part = loss_part.ptr;
    r.capacity = train_cap;
    return r;
}

/* One pass over the `n` resident samples (padded with weight-0 samples to
 * a multiple of kSlice): forward, loss gradient, backward, weight gradients
 * into per-slice partials; with `update`, Adam. */
tr_status MlpState::enqueue_step(uint32_t n, uint32_t n_valid, hipStream_t s, bool update, uint32_t offset,
                                 uint32_t loss_slot)
{
    if (loss_slot >= kMaxLossSlots) loss_slot = kMaxLossSlots - 1;
    uint32_t padded = (n + kSlice - 1) / kSlice * kSlice;
    if (offset + padded > train_cap) return fail(TR_ERROR_INVALID_ARG, "training batch beyond the sample buffer");
    float inv_valid = 1.0f / (float)std::max(1u, n_valid);
    MlpTrainBufs tb = train_bufs();
    tb.loss_part += (size_t)loss_slot * kTrainGrid;
    MlpSampleSoA smp = samples();
    smp.in += offset;
    smp.target += offset;
    smp.valid += offset;
    launch_mlp_train(s, weights(), smp, padded, loss_scale, inv_valid, tb, kTrainGrid);
    launch_mlp_wgrad(s, smp, padded, tb);
    last_slices = mlp_wgrad_slices(padded);
    if (update) {
        ++step;
        float bc1 = 1.0f - powf(0.9f, (float)step), bc2 = 1.0f - powf(0.99f, (float)step);
        launch_mlp_adam(s, master.ptr, mom.ptr, vel.ptr, partial.ptr, last_slices, 1.0f / loss_scale, lr, bc1, bc2,
                        w.ptr, wt.ptr, bias.ptr, 1);
    }
    TR_HIP(hipGetLastError());
    return TR_OK;
}

tr_status MlpState::upload_samples(size_t n, const float* inputs, const float* targets, const float* valid,
                                   uint32_t* n_valid)
{
    tr_status st = ensure_train((uint32_t)n);
    if (st != TR_OK) return st;
    uint32_t cap = train_cap;
    std::vector<float> in((size_t)kMlpIn * cap, 0.0f), tg((size_t)3 * cap, 0.0f), vd(cap, 0.0f);
    uint32_t nv = 0;
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < kMlpIn; ++k) in[(size_t)k * cap + i] = inputs[i * kMlpIn + k];
        for (int c = 0; c < 3; ++c) tg[(size_t)c * cap + i] = targets[i * 3 + c];
        vd[i] = valid ? (valid[i] != 0.0f ? 1.0f : 0.0f) : 1.0f;
        nv += vd[i] != 0.0f;
    }
    TR_HIP(hipMemcpyAsync(s_in.ptr, in.data(), in.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    TR_HIP(hipMemcpyAsync(s_target.ptr, tg.data(), tg.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    TR_HIP(hipMemcpyAsync(s_valid.ptr, vd.data(), vd.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    *n_valid = nv;
    return TR_OK;
}

float MlpState::read_loss(uint32_t n_slots, const float* w)
{
    n_slots = std::max(1u, std::min(n_slots, kMaxLossSlots));
    std::vector<float> lp((size_t)kTrainGrid * n_slots);
    if (hipMemcpy(lp.data(), loss_part.ptr, lp.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return NAN;
    double total = 0.0, wsum = 0.0;
    for (uint32_t k = 0; k < n_slots; ++k) {
        double s = 0.0;
        for (uint32_t i = 0; i < kTrainGrid; ++i) s += lp[(size_t)k * kTrainGrid + i];
        double wk = w ? w[k] : 1.0;
        total += wk * s;
        wsum += wk;
    }
    return wsum > 0.0 ? (float)(total / wsum) : 0.0f;
}

}  // namespace trk


extern "C" {

tr_status tr_mlp_create(uint64_t seed, tr_mlp** out)
{
    if (!out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(TR_ERROR_DEVICE, "no HIP device available");
    tr_mlp* m = new tr_mlp();
    tr_status s = m->st.init(seed);
    if (s != TR_OK) {
        delete m;
        return s;
    }
    *out = m;
    return TR_OK;
}

void tr_mlp_destroy(tr_mlp* m)
{
    if (!m) return;
    if (m->st.stream) (void)hipStreamSynchronize(m->st.stream);
    hipStream_t s = m->st.stream;
    delete m;
    if (s) (void)hipStreamDestroy(s);
}

tr_status tr_mlp_get_params(tr_mlp* m, float* params, size_t count)
{
    if (!m || !params || count != (size_t)kMlpParams) return fail(TR_ERROR_INVALID_ARG, "params must hold 28099 floats");
    TR_HIP(hipStreamSynchronize(m->st.stream));
    TR_HIP(hipMemcpy(params, m->st.master.ptr, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return TR_OK;
}

tr_status tr_mlp_set_params(tr_mlp* m, const float* params, size_t count)
{
    if (!m || !params || count != (size_t)kMlpParams) return fail(TR_ERROR_INVALID_ARG, "params must hold 28099 floats");
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(params[i])) return fail(TR_ERROR_INVALID_ARG, "parameters must be finite");
    return m->st.set_params(params);
}

tr_status tr_mlp_infer(tr_mlp* m, size_t n, const float* inputs, float* outputs)
{
    if (!m || (n && (!inputs || !outputs))) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (n == 0) return TR_OK;
    if (n > (1u << 26)) return fail(TR_ERROR_INVALID_ARG, "too many queries");
    std::vector<float> q(n * kMlpIn);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < kMlpIn; ++k) q[(size_t)k * n + i] = inputs[i * kMlpIn + k];
    TR_HIP(m->q.reserve(n * kMlpIn));
    TR_HIP(m->out.reserve(n * 3));
    hipStream_t s = m->st.stream;
    TR_HIP(hipMemcpyAsync(m->q.ptr, q.data(), q.size() * sizeof(float), hipMemcpyHostToDevice, s));
    MlpQueries mq;
    mq.in = m->q.ptr;
    mq.thr = nullptr;
    mq.slot = nullptr;
    mq.stride = (uint32_t)n;
    launch_mlp_infer(s, m->st.weights(), mq, (uint32_t)n, m->out.ptr, (uint32_t)m->st.cu_count);
    TR_HIP(hipGetLastError());
    TR_HIP(hipMemcpyAsync(outputs, m->out.ptr, n * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    TR_HIP(hipStreamSynchronize(s));
    return TR_OK;
}

static tr_status mlp_pass(tr_mlp* m, size_t n, const float* inputs, const float* targets, const float* valid,
                          bool update, float* grad, float* loss)
{
    if (!m || !n || !inputs || !targets) return fail(TR_ERROR_INVALID_ARG, "null argument");
    if (n > (1u << 22)) return fail(TR_ERROR_INVALID_ARG, "at most 2^22 samples per step");
    uint32_t nv = 0;
    tr_status st = m->st.upload_samples(n, inputs, targets, valid, &nv);
    if (st != TR_OK) return st;
    st = m->st.enqueue_step((uint32_t)n, nv, m->st.stream, update);
    if (st != TR_OK) return st;
    TR_HIP(hipStreamSynchronize(m->st.stream));
    if (loss) *loss = m->st.read_loss();
    if (grad) {
        uint32_t slices = m->st.last_slices;
        std::vector<float> part((size_t)slices * kGradSlots);
        TR_HIP(hipMemcpy(part.data(), m->st.partial.ptr, part.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int q = 0; q < kGradSlots; ++q) {
            const int p = mlp_grad_param(q);
            if (p < 0) continue;
            float g = 0.0f;
            for (uint32_t s2 = 0; s2 < slices; ++s2) g += part[(size_t)s2 * kGradSlots + q];
            grad[p] = g * (1.0f / m->st.loss_scale);
        }
    }
    return TR_OK;
}

tr_status tr_mlp_gradient(tr_mlp* m, size_t n, const float* inputs, const float* targets, const float* valid,
                          float* grad, float* loss)
{
    if (!grad) return fail(TR_ERROR_INVALID_ARG, "null gradient");
    return mlp_pass(m, n, inputs, targets, valid, false, grad, loss);
}

tr_status tr_mlp_train_step(tr_mlp* m, size_t n, const float* inputs, const float* targets, const float* valid,
                            float* loss)
{
    return mlp_pass(m, n, inputs, targets, valid, true, nullptr, loss);
}

}  // extern "C"
