// rc_guide.h -- learned photon-emission proposal (no counterpart in the
// reference, whose emission is plain Area_Light::sample_light,
// src/Lights.cpp:41-55).
//
// Network: x (6 floats: emission point in scene-box coordinates [-1,1]^3 and
// the light's emission axis) -> Fourier features (6 x 16 frequencies x
// {sin, cos} = 192, frequency k is pi * 2^(k/2)) -> 3 hidden layers of 64
// (ReLU) -> 1024 logits over a 32 x 32 grid of the light's two direction
// dimensions (the primary-sample square that the cosine-hemisphere map takes).
// Weights are BF16 with an FP32 master copy; every product accumulates in FP32
// on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Proposal: the network is evaluated at the centres of a G x G grid over
// each light's position square; each row becomes a CDF over the 1024 bins.  A
// light path picks, with probability alpha, the plain direction sample and
// otherwise a bin from the CDF of its position cell (a uniform point inside
// the bin); either way its throughput is divided by
//     q = alpha + (1 - alpha) * 1024 * P(bin | cell),
// the density of this mixture relative to plain sampling, so the estimator
// keeps its expectation for any network.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rc {

constexpr int kGuideIn = 6;
constexpr int kGuideFreqs = 16;
constexpr int kGuideFeat = kGuideIn * kGuideFreqs * 2;  // 192
constexpr int kGuideHidden = 64;
constexpr int kGuideSide = 32;
constexpr int kGuideBins = kGuideSide * kGuideSide;     // 1024
constexpr int kGuideCdfStride = 1032;                   // >= bins + 1, a multiple of 8
constexpr int kGuideMaxLights = 8;
constexpr int kGuideRowFloats = 8;                      // training/table input record (6 used)

// Offsets (floats) of the parameter tensors in the flat vector; weights are
// row-major [in][out].
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kGuideFeat * kGuideHidden;
constexpr int kOffW2 = kOffB1 + kGuideHidden;
constexpr int kOffB2 = kOffW2 + kGuideHidden * kGuideHidden;
constexpr int kOffW3 = kOffB2 + kGuideHidden;
constexpr int kOffB3 = kOffW3 + kGuideHidden * kGuideHidden;
constexpr int kOffW4 = kOffB3 + kGuideHidden;
constexpr int kOffB4 = kOffW4 + kGuideHidden * kGuideBins;
constexpr int kGuideParams = kOffB4 + kGuideBins;       // 87,232

// What the light pass reads (cdf == nullptr: plain emission everywhere).
struct GuideView {
    const float* cdf;        // [slot][cell][kGuideCdfStride], cdf[0] = 0, cdf[1024] = 1
    int cells;               // G
    uint32_t slot_mask;      // bit i: light i has a table; its slot = set bits below bit i
    float alpha;
    // training records of paths < train_paths
    uint32_t train_paths;
    float* train_x;          // kGuideRowFloats per path
    int* train_bin;          // bin index, or -1 when the path has no guidable emission
    float center[3], inv_half[3];
};

// Table slot of light i (lights with a table take slots in light order), or -1.
__host__ __device__ __forceinline__ int guide_slot(const GuideView& g, int i) {
    if (i < 0 || i >= kGuideMaxLights || !((g.slot_mask >> i) & 1u)) return -1;
    return __builtin_popcount(g.slot_mask & ((1u << i) - 1u));
}

// The input record of an emission point: scene-box coordinates and axis.
__host__ __device__ __forceinline__ void guide_record(const GuideView& g, float px, float py, float pz, float ax,
                                                      float ay, float az, float* out) {
    out[0] = (px - g.center[0]) * g.inv_half[0];
    out[1] = (py - g.center[1]) * g.inv_half[1];
    out[2] = (pz - g.center[2]) * g.inv_half[2];
    out[3] = ax;
    out[4] = ay;
    out[5] = az;
    out[6] = 0.0f;
    out[7] = 0.0f;
}

// Matrix product of the training/inference kernels:
// C[M x N] = op(A)[M x K] * op(B)[K x N], A/B BF16, FP32 accumulation.
enum { EPI_BIAS_RELU_BF16 = 0, EPI_BIAS_F32 = 1, EPI_MASK_BF16 = 2, EPI_ATOMIC_F32 = 3 };
struct GemmArgs {
    const uint16_t* A;
    int lda;
    const uint16_t* B;
    int ldb;
    int M, N, K;
    int k_split;          // K range of one blockIdx.z (a multiple of 64)
    const float* bias;    // EPI_BIAS_*: N floats
    const uint16_t* mask; // EPI_MASK_BF16: keep C(m, n) where mask(m, n) > 0 (leading dim ldc)
    void* C;
    int ldc;
    float* colsum;        // optional (B row-major only): += column sums of op(B) (bias gradients)
};
// Requires M % 64 == 0, N % 64 == 0, K % 64 == 0, k_split % 64 == 0, lda/ldb % 8 == 0.
bool launch_gemm(const GemmArgs& g, int epi, bool a_trans, bool b_trans, hipStream_t s);

void launch_guide_encode(const float* x, int n, uint16_t* X, hipStream_t s);
// Output layer + softmax cross-entropy of a training step (n % 32 == 0): dZ, the per-row losses (summed
// into *loss) and the back-propagated dA3 = (dZ W4^T) masked by H3 > 0, from the last hidden layer H3,
// W4 (BF16, [64][1024]) and b4, without writing the logits.  W4f: 64 x 1024 BF16 of work space (W4 in the
// kernel's fragment order).
bool launch_out_softmax(const uint16_t* H3, const uint16_t* W4, uint16_t* W4f, const float* b4, const float* w,
                        const int* bins, int n, const float* wsum, uint16_t* dZ, float* row_loss, float* loss,
                        uint16_t* dA3, hipStream_t s);
// out[0] += sum of v, out[1] += number of entries > 0.
void launch_sum_count(const float* v, int n, float* out, hipStream_t s);
struct AdamArgs {
    float* master;
    float* m;
    float* v;
    const float* grad;
    uint16_t* wbf;
    int n;
    int step;  // 1-based
    float lr, beta1, beta2, eps;
    const float* wsum;  // the step is skipped when *wsum == 0
};
void launch_adam(const AdamArgs& a, hipStream_t s);
void launch_to_bf16(const float* in, uint16_t* out, int n, hipStream_t s);
void launch_guide_table(const float* Z, int rows, float* cdf, hipStream_t s);
// Camera-visible merge cells: one byte per photon-map bucket.
void launch_mark_visible(const float4* qpos, uint32_t nq, float ox, float oy, float oz, float inv_cell, uint32_t mask,
                         uint8_t* vis, hipStream_t s);
// w[path] += luminance of each stored photon of a training path whose cell is visible.
void launch_guide_targets(const float4* rec, uint32_t count, float ox, float oy, float oz, float inv_cell,
                          uint32_t mask, const uint8_t* vis, uint32_t train_paths, float* w, hipStream_t s);

}  // namespace rc
