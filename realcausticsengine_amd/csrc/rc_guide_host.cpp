// rc_guide_host.cpp -- schedule of the learned emission proposal (see
// rc_guide.h for the network and the mixture): initialisation, forward,
// backward, Adam and the per-light proposal table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "rc_guide_model.h"
#include "rc_vec.h"

namespace rc {

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

namespace {

uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

bool zero_f32(float* p, size_t n, hipStream_t s) {
    launch_fill_u32(reinterpret_cast<uint32_t*>(p), 0u, n, s);
    return true;
}

// Split of K = rows for the weight gradients: about 256 workgroups in all.
int wgrad_split(int M, int N, int K) {
    int tiles = (M / 64) * (N / 64);
    int splits = std::max(1, std::min(K / 64, 256 / std::max(1, tiles)));
    int ks = (K + splits - 1) / splits;
    return (ks + 63) / 64 * 64;
}

bool gemm(const uint16_t* A, int lda, bool at, const uint16_t* B, int ldb, bool bt, int M, int N, int K, int epi,
          const float* bias, const uint16_t* mask, void* C, int ldc, hipStream_t s, int k_split = 0,
          float* colsum = nullptr) {
    GemmArgs g{};
    g.A = A;
    g.lda = lda;
    g.B = B;
    g.ldb = ldb;
    g.M = M;
    g.N = N;
    g.K = K;
    g.k_split = k_split > 0 ? k_split : K;
    g.bias = bias;
    g.mask = mask;
    g.C = C;
    g.ldc = ldc;
    g.colsum = colsum;
    if (!launch_gemm(g, epi, at, bt, s)) {
        set_error("guide: matrix product shape not supported");
        return false;
    }
    return true;
}

}  // namespace

GuideView plain_guide_view() {
    GuideView v{};
    v.slot_mask = 0u;
    v.cells = 1;
    return v;
}

bool GuideModel::init_params(hipStream_t s) {
    std::vector<float> p(kGuideParams, 0.0f);
    uint64_t st = seed;
    auto fill = [&](int off, int fan_in, int count, float scale) {
        float a = scale * std::sqrt(6.0f / (float)fan_in);
        for (int i = 0; i < count; ++i) {
            double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
            p[off + i] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(kOffW1, kGuideFeat, kGuideFeat * kGuideHidden, 1.0f);
    fill(kOffW2, kGuideHidden, kGuideHidden * kGuideHidden, 1.0f);
    fill(kOffW3, kGuideHidden, kGuideHidden * kGuideHidden, 1.0f);
    fill(kOffW4, kGuideHidden, kGuideHidden * kGuideBins, 0.1f);
    return set_params(p.data(), s);
}

bool GuideModel::set_params(const float* host, hipStream_t s) {
    if (!master.alloc(kGuideParams) || !m.alloc(kGuideParams) || !v.alloc(kGuideParams) ||
        !grad.alloc(kGuideParams) || !wbf.alloc(kGuideParams) || !scal.alloc(4))
        return false;
    HIP_OK(hipMemcpyAsync(master.p, host, sizeof(float) * kGuideParams, hipMemcpyHostToDevice, s));
    zero_f32(m.p, kGuideParams, s);
    zero_f32(v.p, kGuideParams, s);
    launch_to_bf16(master.p, wbf.p, kGuideParams, s);
    HIP_OK(hipStreamSynchronize(s));
    step = 0;
    params_ready = true;
    table_ready = false;
    return true;
}

bool GuideModel::get_params(float* host, hipStream_t s) {
    if (!params_ready && !init_params(s)) return false;
    HIP_OK(hipMemcpyAsync(host, master.p, sizeof(float) * kGuideParams, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return true;
}

bool GuideModel::ensure_rows(int n) {
    if (n <= rows_cap && X.p) return true;
    size_t r = (size_t)n;
    if (!X.alloc(r * kGuideFeat) || !H1.alloc(r * kGuideHidden) || !H2.alloc(r * kGuideHidden) ||
        !H3.alloc(r * kGuideHidden) || !dA1.alloc(r * kGuideHidden) || !dA2.alloc(r * kGuideHidden) ||
        !dA3.alloc(r * kGuideHidden) || !dZ.alloc(r * kGuideBins) || !row_loss.alloc(r))
        return false;
    rows_cap = n;
    return true;
}

bool GuideModel::forward(const float* d_x, int n, hipStream_t s, bool logits) {
    if (n <= 0 || n % 64) {
        set_error("guide: row count must be a positive multiple of 64");
        return false;
    }
    if (!params_ready && !init_params(s)) return false;
    if (!ensure_rows(n)) return false;
    const uint16_t* w = wbf.p;
    const float* b = master.p;
    if (logits && !Z.alloc((size_t)n * kGuideBins)) return false;
    launch_guide_encode(d_x, n, X.p, s);
    return gemm(X.p, kGuideFeat, false, w + kOffW1, kGuideHidden, false, n, kGuideHidden, kGuideFeat,
                EPI_BIAS_RELU_BF16, b + kOffB1, nullptr, H1.p, kGuideHidden, s) &&
           gemm(H1.p, kGuideHidden, false, w + kOffW2, kGuideHidden, false, n, kGuideHidden, kGuideHidden,
                EPI_BIAS_RELU_BF16, b + kOffB2, nullptr, H2.p, kGuideHidden, s) &&
           gemm(H2.p, kGuideHidden, false, w + kOffW3, kGuideHidden, false, n, kGuideHidden, kGuideHidden,
                EPI_BIAS_RELU_BF16, b + kOffB3, nullptr, H3.p, kGuideHidden, s) &&
           (!logits || gemm(H3.p, kGuideHidden, false, w + kOffW4, kGuideBins, false, n, kGuideBins, kGuideHidden,
                            EPI_BIAS_F32, b + kOffB4, nullptr, Z.p, kGuideBins, s));
}

bool GuideModel::backward(const int* d_bin, const float* d_w, int n, hipStream_t s) {
    const uint16_t* w = wbf.p;
    float* g = grad.p;
    zero_f32(scal.p, 4, s);
    zero_f32(g, kGuideParams, s);
    launch_sum_count(d_w, n, scal.p, s);  // [0] sum of w, [1] rows with w > 0
    // output layer, softmax cross-entropy and the output layer's back-propagated product in one kernel:
    // dZ, the loss and dA3, no logits in memory
    if (!w4frag.alloc((size_t)kGuideHidden * kGuideBins)) return false;
    if (!launch_out_softmax(H3.p, w + kOffW4, w4frag.p, master.p + kOffB4, d_w, d_bin, n, scal.p, dZ.p, row_loss.p,
                            scal.p + 2, dA3.p, s)) {
        set_error("guide: training rows must be a multiple of 32");
        return false;
    }
    // output layer
    if (!gemm(H3.p, kGuideHidden, true, dZ.p, kGuideBins, false, kGuideHidden, kGuideBins, n, EPI_ATOMIC_F32, nullptr,
              nullptr, g + kOffW4, kGuideBins, s, wgrad_split(kGuideHidden, kGuideBins, n), g + kOffB4))
        return false;
    // hidden layer 3
    if (!gemm(H2.p, kGuideHidden, true, dA3.p, kGuideHidden, false, kGuideHidden, kGuideHidden, n, EPI_ATOMIC_F32,
              nullptr, nullptr, g + kOffW3, kGuideHidden, s, wgrad_split(kGuideHidden, kGuideHidden, n), g + kOffB3))
        return false;
    if (!gemm(dA3.p, kGuideHidden, false, w + kOffW3, kGuideHidden, true, n, kGuideHidden, kGuideHidden,
              EPI_MASK_BF16, nullptr, H2.p, dA2.p, kGuideHidden, s))
        return false;
    // hidden layer 2
    if (!gemm(H1.p, kGuideHidden, true, dA2.p, kGuideHidden, false, kGuideHidden, kGuideHidden, n, EPI_ATOMIC_F32,
              nullptr, nullptr, g + kOffW2, kGuideHidden, s, wgrad_split(kGuideHidden, kGuideHidden, n), g + kOffB2))
        return false;
    if (!gemm(dA2.p, kGuideHidden, false, w + kOffW2, kGuideHidden, true, n, kGuideHidden, kGuideHidden,
              EPI_MASK_BF16, nullptr, H1.p, dA1.p, kGuideHidden, s))
        return false;
    // hidden layer 1
    if (!gemm(X.p, kGuideFeat, true, dA1.p, kGuideHidden, false, kGuideFeat, kGuideHidden, n, EPI_ATOMIC_F32, nullptr,
              nullptr, g + kOffW1, kGuideHidden, s, wgrad_split(kGuideFeat, kGuideHidden, n), g + kOffB1))
        return false;
    return true;
}

bool GuideModel::adam(hipStream_t s) {
    ++step;
    AdamArgs a{};
    a.master = master.p;
    a.m = m.p;
    a.v = v.p;
    a.grad = grad.p;
    a.wbf = wbf.p;
    a.n = kGuideParams;
    a.step = step;
    a.lr = lr;
    a.beta1 = beta1;
    a.beta2 = beta2;
    a.eps = eps;
    a.wsum = scal.p;
    launch_adam(a, s);
    return true;
}

bool GuideModel::setup_scene(const std::vector<DevLight>& lights, int compat_point, const float lo[3],
                             const float hi[3]) {
    if (cells < 8 || cells % 8 || cells > 64) {
        set_error("guide_cells must be a multiple of 8 in [8, 64]");
        return false;
    }
    for (int k = 0; k < 3; ++k) {
        center[k] = 0.5f * (lo[k] + hi[k]);
        float h = 0.5f * (hi[k] - lo[k]);
        inv_half[k] = h > 0.0f ? 1.0f / h : 1.0f;
    }
    slots = 0;
    for (int i = 0; i < kGuideMaxLights; ++i) slot_of_light[i] = -1;
    GuideView gv = plain_guide_view();
    for (int k = 0; k < 3; ++k) {
        gv.center[k] = center[k];
        gv.inv_half[k] = inv_half[k];
    }
    table_inputs.clear();
    const int G = cells;
    for (size_t i = 0; i < lights.size() && i < (size_t)kGuideMaxLights; ++i) {
        const DevLight& L = lights[i];
        bool area = L.kind == L_AREA;
        bool sphere = L.kind == L_POINT && !compat_point && L.radius > 0.0f;
        if (!area && !sphere) continue;
        slot_of_light[i] = slots++;
        for (int cy = 0; cy < G; ++cy)
            for (int cx = 0; cx < G; ++cx) {
                float ux = ((float)cx + 0.5f) / (float)G, uy = ((float)cy + 0.5f) / (float)G;
                F3 o, axis;
                if (area) {
                    o = (f3(L.bl[0], L.bl[1], L.bl[2]) + f3(L.h[0], L.h[1], L.h[2]) * ux) +
                        f3(L.v[0], L.v[1], L.v[2]) * uy;
                    axis = f3(L.fz[0], L.fz[1], L.fz[2]);
                } else {
                    float z = 1.0f - 2.0f * ux;
                    float rr = std::sqrt(std::max(0.0f, 1.0f - z * z));
                    float phi = 6.28318530718f * uy;
                    axis = f3(rr * own_cos(phi), rr * own_sin(phi), z);
                    o = f3(L.pos[0], L.pos[1], L.pos[2]) + axis * L.radius;
                }
                float rec[kGuideRowFloats];
                guide_record(gv, o.x, o.y, o.z, axis.x, axis.y, axis.z, rec);
                table_inputs.insert(table_inputs.end(), rec, rec + kGuideRowFloats);
            }
    }
    table_rows = slots * G * G;
    cells_built = G;
    table_ready = false;
    if (table_rows == 0) return true;
    if (!table_x.alloc(table_inputs.size()) || !cdf.alloc((size_t)table_rows * kGuideCdfStride)) return false;
    HIP_OK(hipMemcpy(table_x.p, table_inputs.data(), sizeof(float) * table_inputs.size(), hipMemcpyHostToDevice));
    return true;
}

bool GuideModel::build_table(hipStream_t s) {
    if (table_rows == 0) return true;
    if (!forward(table_x.p, table_rows, s)) return false;
    launch_guide_table(Z.p, table_rows, cdf.p, s);
    table_ready = true;
    return true;
}

GuideView GuideModel::view(bool records, bool sample) const {
    GuideView gv = plain_guide_view();
    for (int i = 0; i < kGuideMaxLights; ++i)
        if (slot_of_light[i] >= 0) gv.slot_mask |= 1u << i;
    gv.cells = cells_built > 0 ? cells_built : 1;
    // alpha on the 2^-24 grid of the uniform draws, so P(u < alpha) = alpha exactly
    gv.alpha = (float)(std::ceil((double)alpha * 16777216.0) / 16777216.0);
    gv.cdf = (sample && table_ready) ? cdf.p : nullptr;
    gv.train_paths = records ? (uint32_t)batch : 0u;
    gv.train_x = train_x.p;
    gv.train_bin = train_bin.p;
    for (int k = 0; k < 3; ++k) {
        gv.center[k] = center[k];
        gv.inv_half[k] = inv_half[k];
    }
    return gv;
}

void GuideModel::release() {
    for (DevBuf<float>* b : {&master, &m, &v, &grad, &scal, &train_x, &train_w, &Z, &row_loss, &table_x, &cdf})
        b->release();
    for (DevBuf<uint16_t>* b : {&wbf, &X, &H1, &H2, &H3, &dA1, &dA2, &dA3, &dZ, &w4frag}) b->release();
    train_bin.release();
    vis.release();
    params_ready = table_ready = false;
    rows_cap = 0;
    step = 0;
}

}  // namespace rc
