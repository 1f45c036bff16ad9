// rc_guide_model.h -- host object of the learned emission proposal: the
// parameters (FP32 master + BF16 copy), Adam state, work buffers of one
// batch, the proposal table and the schedule of the kernels in rc_guide.hip.
#pragma once
#include <stdint.h>

#include <vector>

#include "rc_engine.h"
#include "rc_guide.h"
#include "rc_host.h"

namespace rc {

struct GuideModel {
    // options
    int enabled = 0;       // sample from the table and train every iteration
    int train = 1;         // 0: keep the network fixed
    float alpha = 0.2f;    // defensive share of plain emission
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    int batch = 65536;     // training paths per iteration (the first `batch` path indices)
    int cells = 16;        // position cells per side of a light (a multiple of 8)
    int warmup = 2;        // Adam steps before the table is sampled
    int train_every = 1;   // after warm-up, train on every n-th iteration
    uint64_t seed = 1;     // parameter initialisation

    // state
    bool params_ready = false, table_ready = false;
    int step = 0;
    int slots = 0, table_rows = 0, cells_built = 0;
    int slot_of_light[kGuideMaxLights];
    float center[3] = {0, 0, 0}, inv_half[3] = {1, 1, 1};
    std::vector<float> table_inputs;  // table_rows records
    float last[3] = {0, 0, 0};        // visible luminance sum, loss, training paths with w > 0

    DevBuf<float> master, m, v, grad, scal;
    DevBuf<uint16_t> wbf;
    DevBuf<float> train_x, train_w;
    DevBuf<int> train_bin;
    DevBuf<uint16_t> X, H1, H2, H3, dA1, dA2, dA3, dZ;
    DevBuf<uint16_t> w4frag;  // the output layer's weights in the fused softmax kernel's fragment order
    DevBuf<float> Z, row_loss;
    DevBuf<float> table_x, cdf;
    DevBuf<uint8_t> vis;
    int rows_cap = 0;

    // Deterministic He-uniform initialisation (output layer scaled by 0.1), zero biases.
    bool init_params(hipStream_t s);
    bool set_params(const float* host, hipStream_t s);
    bool get_params(float* host, hipStream_t s);
    bool ensure_rows(int n);
    // Hidden layers of n records (n % 64 == 0) and, with `logits`, the logits into Z (the training step
    // leaves them out: backward makes the output layer again inside its fused softmax kernel).
    bool forward(const float* d_x, int n, hipStream_t s, bool logits = true);
    // Loss and gradients of the last forward (H1..H3 of its rows); weights w, bins b (-1: no target).
    bool backward(const int* d_bin, const float* d_w, int n, hipStream_t s);
    bool adam(hipStream_t s);
    // Slots and table inputs of the scene's lights; center / half extent of the box.
    bool setup_scene(const std::vector<DevLight>& lights, int compat_point, const float lo[3], const float hi[3]);
    bool build_table(hipStream_t s);
    // Training records of this iteration's first paths are written when
    // `records`; the table is sampled when `sample` and it is ready.
    GuideView view(bool records, bool sample) const;
    void release();
};

// Every slot -1: the light pass samples plainly and writes no records.
GuideView plain_guide_view();

}  // namespace rc
