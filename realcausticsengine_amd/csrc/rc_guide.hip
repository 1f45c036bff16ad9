// rc_guide.hip -- gfx950 kernels of the learned emission proposal: Fourier
// encoding, an LDS-tiled BF16 matrix-core product with fused epilogues (the
// forward layers, the backward data products and the split-K weight
// gradients), the training step's output layer fused with softmax
// cross-entropy, Adam, the proposal table (softmax -> CDF per position cell)
// and the training targets.
#include "rc_engine.h"
#include "rc_guide.h"
#include "rc_vec.h"

namespace rc {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__constant__ float c_freq[kGuideFreqs] = {
    3.1415927410125732f, 4.442883014678955f,  6.2831854820251465f, 8.88576602935791f,
    12.566370964050293f, 17.77153205871582f,  25.132741928100586f, 35.54306411743164f,
    50.26548385620117f,  71.08612823486328f,  100.53096771240234f, 142.17225646972656f,
    201.0619354248047f,  284.3445129394531f,  402.1238708496094f,  568.6890258789062f};

// ---------------------------------------------------------------------
// Encoding: feature i*32 + 2k + {0: sin, 1: cos} of (pi 2^(k/2)) * x_i.
// ---------------------------------------------------------------------
__global__ void k_guide_encode(const float* __restrict__ x, int n, uint16_t* __restrict__ X) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* xr = x + (size_t)r * kGuideRowFloats;
    uint16_t* out = X + (size_t)r * kGuideFeat;
    for (int i = 0; i < kGuideIn; ++i) {
        float xi = xr[i];
        uint16_t f[32];
#pragma unroll
        for (int k = 0; k < kGuideFreqs; ++k) {
            float s, c;
            own_sincos2(c_freq[k] * xi, &s, &c);
            f[2 * k] = f2bf(s);
            f[2 * k + 1] = f2bf(c);
        }
        uint4* o4 = reinterpret_cast<uint4*>(out + i * 32);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint4 v;
            v.x = f[8 * q] | ((uint32_t)f[8 * q + 1] << 16);
            v.y = f[8 * q + 2] | ((uint32_t)f[8 * q + 3] << 16);
            v.z = f[8 * q + 4] | ((uint32_t)f[8 * q + 5] << 16);
            v.w = f[8 * q + 6] | ((uint32_t)f[8 * q + 7] << 16);
            o4[q] = v;
        }
    }
}

// ---------------------------------------------------------------------
// C = op(A) op(B): a 64 x 64 tile per workgroup of four waves (2 x 2), each
// wave one 32 x 32 v_mfma_f32_32x32x16_bf16 accumulator; K advances 64 at a
// time through LDS (four MFMAs per wave between barriers), and the global
// loads of the next K step are issued before the MFMAs of the current one.
// Both operands are kept k-contiguous in LDS ([m][k] and [n][k]) so every
// lane reads its 8-element fragment with one 16-byte load.
// Fragment layout: lane l holds row (l % 32) and k = 8 (l / 32) .. +7 of A,
// column (l % 32) and the same k of B; accumulator element i is row
// 8 (i / 4) + 4 (l / 32) + (i % 4), column l % 32.
// ---------------------------------------------------------------------
constexpr int kGT = 64, kGK = 64, kGPad = 8;

struct TileRegs {
    uint4 a[2], b[2];
};

// Thread t loads 16 consecutive elements along the contiguous axis of each
// operand: row (or k) t / 4, elements 16 (t % 4) .. +15.
template <bool AT, bool BT>
__device__ __forceinline__ void gemm_load(const GemmArgs& g, int m0, int n0, int k0, int tid, TileRegs& r) {
    const int row = tid >> 2, c = (tid & 3) * 16;
    const uint16_t* pa = AT ? g.A + (size_t)(k0 + row) * g.lda + m0 + c : g.A + (size_t)(m0 + row) * g.lda + k0 + c;
    const uint16_t* pb = BT ? g.B + (size_t)(n0 + row) * g.ldb + k0 + c : g.B + (size_t)(k0 + row) * g.ldb + n0 + c;
    r.a[0] = reinterpret_cast<const uint4*>(pa)[0];
    r.a[1] = reinterpret_cast<const uint4*>(pa)[1];
    r.b[0] = reinterpret_cast<const uint4*>(pb)[0];
    r.b[1] = reinterpret_cast<const uint4*>(pb)[1];
}

template <bool AT, bool BT>
__device__ __forceinline__ void gemm_stage(uint16_t (*As)[kGK + kGPad], uint16_t (*Bs)[kGK + kGPad], int tid,
                                           const TileRegs& r) {
    const int row = tid >> 2, c = (tid & 3) * 16;
    if (!AT) {  // row = m, c = k
        *reinterpret_cast<uint4*>(&As[row][c]) = r.a[0];
        *reinterpret_cast<uint4*>(&As[row][c + 8]) = r.a[1];
    } else {  // row = k, c = m
        const uint16_t* e = reinterpret_cast<const uint16_t*>(r.a);
#pragma unroll
        for (int j = 0; j < 16; ++j) As[c + j][row] = e[j];
    }
    if (BT) {  // row = n, c = k
        *reinterpret_cast<uint4*>(&Bs[row][c]) = r.b[0];
        *reinterpret_cast<uint4*>(&Bs[row][c + 8]) = r.b[1];
    } else {  // row = k, c = n
        const uint16_t* e = reinterpret_cast<const uint16_t*>(r.b);
#pragma unroll
        for (int j = 0; j < 16; ++j) Bs[c + j][row] = e[j];
    }
}

template <int EPI, bool AT, bool BT>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kGT][kGK + kGPad];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kGT][kGK + kGPad];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
    const int kb = blockIdx.z * g.k_split;
    const int ke = min(g.K, kb + g.k_split);
    f32x16_t acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    TileRegs r;
    // column sums of op(B) over this block's K range: the first row of
    // workgroups, waves with wm == 0; lane l adds k = 32 (l / 32) .. +31 of column l % 32
    const bool sums = !BT && g.colsum && blockIdx.y == 0 && wm == 0;
    float csum = 0.0f;
    if (kb < ke) gemm_load<AT, BT>(g, m0, n0, kb, tid, r);
    for (int k0 = kb; k0 < ke; k0 += kGK) {
        gemm_stage<AT, BT>(As, Bs, tid, r);
        __syncthreads();
        if (k0 + kGK < ke) gemm_load<AT, BT>(g, m0, n0, k0 + kGK, tid, r);
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 16) {
            bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&As[wm * 32 + (lane & 31)][kk + (lane >> 5) * 8]);
            bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&Bs[wn * 32 + (lane & 31)][kk + (lane >> 5) * 8]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        if (sums) {
            const uint16_t* row = &Bs[wn * 32 + (lane & 31)][(lane >> 5) * 32];
#pragma unroll
            for (int j = 0; j < 32; ++j) csum += bf2f(row[j]);
        }
        __syncthreads();
    }
    if (sums) {
        csum += __shfl_xor(csum, 32);
        if (lane < 32) atomicAdd(g.colsum + n0 + wn * 32 + lane, csum);
    }
    const int n = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int m = m0 + wm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const size_t at = (size_t)m * g.ldc + n;
        float v = acc[i];
        if (EPI == EPI_BIAS_RELU_BF16) {
            v = v + g.bias[n];
            reinterpret_cast<uint16_t*>(g.C)[at] = f2bf(v > 0.0f ? v : 0.0f);
        } else if (EPI == EPI_BIAS_F32) {
            reinterpret_cast<float*>(g.C)[at] = v + g.bias[n];
        } else if (EPI == EPI_MASK_BF16) {
            reinterpret_cast<uint16_t*>(g.C)[at] = f2bf(bf2f(g.mask[at]) > 0.0f ? v : 0.0f);
        } else {
            atomicAdd(reinterpret_cast<float*>(g.C) + at, v);
        }
    }
}

template <int EPI>
static void gemm_epi(const GemmArgs& g, bool at, bool bt, dim3 grid, hipStream_t s) {
    if (!at && !bt) hipLaunchKernelGGL((k_gemm<EPI, false, false>), grid, dim3(256), 0, s, g);
    else if (!at && bt) hipLaunchKernelGGL((k_gemm<EPI, false, true>), grid, dim3(256), 0, s, g);
    else if (at && !bt) hipLaunchKernelGGL((k_gemm<EPI, true, false>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((k_gemm<EPI, true, true>), grid, dim3(256), 0, s, g);
}

bool launch_gemm(const GemmArgs& g, int epi, bool a_trans, bool b_trans, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.M % kGT || g.N % kGT || g.K % kGK || g.k_split <= 0 ||
        g.k_split % kGK || g.lda % 8 || g.ldb % 8)
        return false;
    if (epi == EPI_MASK_BF16 && !g.mask) return false;
    if (g.colsum && b_trans) return false;
    if ((epi == EPI_BIAS_RELU_BF16 || epi == EPI_BIAS_F32) && !g.bias) return false;
    if (epi != EPI_ATOMIC_F32 && g.k_split < g.K) return false;  // only the atomic epilogue sums partial K
    dim3 grid(g.N / kGT, g.M / kGT, (g.K + g.k_split - 1) / g.k_split);
    switch (epi) {
        case EPI_BIAS_RELU_BF16: gemm_epi<EPI_BIAS_RELU_BF16>(g, a_trans, b_trans, grid, s); break;
        case EPI_BIAS_F32: gemm_epi<EPI_BIAS_F32>(g, a_trans, b_trans, grid, s); break;
        case EPI_MASK_BF16: gemm_epi<EPI_MASK_BF16>(g, a_trans, b_trans, grid, s); break;
        case EPI_ATOMIC_F32: gemm_epi<EPI_ATOMIC_F32>(g, a_trans, b_trans, grid, s); break;
        default: return false;
    }
    return true;
}

// ---------------------------------------------------------------------
// Output layer and softmax cross-entropy of a training step in one kernel:
// the 1024 logits of a row are made, used and dropped in registers, so they
// never reach memory (only dZ does).  A workgroup of 8 waves takes 256 rows,
// 32 per wave.  W4 (64 x 1024) is first copied into LDS in the order of the
// wave's A fragments (k_w4_frag): entry (t * 4 + kk) * 64 + l holds bin
// 32 t + (l % 32) and k = 16 kk + 8 (l / 32) .. +7, one 16-byte read per lane
// and no bank conflicts.  A wave then forms Z^T tiles (32 bins x its 32 rows) with
// v_mfma_f32_32x32x16_bf16, A = W4^T fragments, B = the rows' H3: lane l
// holds bins 32 t + 8 (i / 4) + 4 (l / 32) + (i % 4), i = 0..15, of row l % 32,
// so a row's statistics combine two lanes (l and l ^ 32).  Pass 1 keeps a
// running maximum and sum of exp(z - max) per lane (rescaled when the
// maximum grows); pass 2 makes the tiles again and writes
//   dZ_r = bf16(w_r / W (softmax(z) - onehot(b)))   (4 bins per 8-byte store)
// and the row's loss w_r / W (log sum exp(z) - z_b), W = sum of w, and feeds
// each tile's dZ straight into the back-propagated product
//   dA3_r = bf16((dZ_r W4^T) masked by H3_r > 0)
// (the bins in the same ascending order as a separate product would take).
// ---------------------------------------------------------------------
constexpr int kOutRowsPerWave = 32, kOutWaves = 8;
constexpr int kOutTiles = kGuideBins / 32;  // 32 tiles of 32 bins

// W4 (BF16, [64][1024]) in the A-fragment order of k_out_softmax: chunk
// (t * 4 + kk) * 64 + l holds bin 32 t + (l % 32), k = 16 kk + 8 (l / 32) .. +7.
__global__ __launch_bounds__(256) void k_w4_frag(const uint16_t* __restrict__ W4, uint16_t* __restrict__ out) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= kOutTiles * 4 * 64) return;
    const int t = c >> 8, kk = (c >> 6) & 3, l = c & 63;
    const int bin = 32 * t + (l & 31), k0 = 16 * kk + 8 * (l >> 5);
    uint16_t e[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) e[j] = W4[(size_t)(k0 + j) * kGuideBins + bin];
    reinterpret_cast<uint4*>(out)[c] = make_uint4(e[0] | ((uint32_t)e[1] << 16), e[2] | ((uint32_t)e[3] << 16),
                                                  e[4] | ((uint32_t)e[5] << 16), e[6] | ((uint32_t)e[7] << 16));
}

__global__ __launch_bounds__(kOutWaves * 64) void k_out_softmax(const uint16_t* __restrict__ H3,
                                                                 const uint16_t* __restrict__ W4,
                                                                 const uint16_t* __restrict__ W4f,
                                                                 const float* __restrict__ b4,
                                                                 const float* __restrict__ w,
                                                                 const int* __restrict__ bins, int n,
                                                                 const float* __restrict__ wsum,
                                                                 uint16_t* __restrict__ dZ,
                                                                 float* __restrict__ row_loss,
                                                                 uint16_t* __restrict__ dA3) {
    __shared__ uint16_t frag[kOutTiles * 4 * 64 * 8];  // 128 KB
    const int tid = threadIdx.x, lane = tid & 63;
    // the fragment-ordered copy of W4 (k_w4_frag) into LDS, 16 bytes per thread and step
    for (int c = tid; c < kOutTiles * 4 * 64; c += kOutWaves * 64)
        reinterpret_cast<uint4*>(frag)[c] = reinterpret_cast<const uint4*>(W4f)[c];
    __syncthreads();
    const int r0 = (blockIdx.x * kOutWaves + (tid >> 6)) * kOutRowsPerWave;
    if (r0 >= n) return;
    const int row = r0 + (lane & 31);
    bf16x8_t hb[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
        hb[kk] = *reinterpret_cast<const bf16x8_t*>(H3 + (size_t)row * kGuideHidden + kk * 16 + (lane >> 5) * 8);
    const bf16x8_t* fa = reinterpret_cast<const bf16x8_t*>(frag) + lane;
    auto tile = [&](int t, float (&z)[16]) {
        f32x16_t acc;
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[(t * 4 + kk) * 64], hb[kk], acc, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 b = *reinterpret_cast<const float4*>(b4 + 32 * t + 8 * q + 4 * (lane >> 5));
            z[4 * q] = acc[4 * q] + b.x;
            z[4 * q + 1] = acc[4 * q + 1] + b.y;
            z[4 * q + 2] = acc[4 * q + 2] + b.z;
            z[4 * q + 3] = acc[4 * q + 3] + b.w;
        }
    };
    // pass 1: maximum and sum of exp(z - maximum) over the lane's 512 bins of the row
    float m = -INFINITY, s = 0.0f;
    for (int t = 0; t < kOutTiles; ++t) {
        float z[16];
        tile(t, z);
        float tm = z[0];
#pragma unroll
        for (int i = 1; i < 16; ++i) tm = fmaxf(tm, z[i]);
        const float mn = fmaxf(m, tm);
        s = s * __expf(m - mn);
#pragma unroll
        for (int i = 0; i < 16; ++i) s += __expf(z[i] - mn);
        m = mn;
    }
    {  // the row's other half is in lane l ^ 32
        const float mo = __shfl_xor(m, 32), so = __shfl_xor(s, 32);
        const float mt = fmaxf(m, mo);
        s = s * __expf(m - mt) + so * __expf(mo - mt);
        m = mt;
    }
    // pass 2: dZ, the loss, and the back-propagated dA3 = (dZ W4^T) masked by H3 > 0.  dA3^T (64 x the wave's
    // 32 rows) is two accumulators, A = W4 rows (k; 8 consecutive bins per lane, 16-byte loads), B = dZ^T:
    // the 16-bin K-step s of tile t needs bins 32 t + 16 s + 8 (l / 32) .. +7 of the lane's row, four of
    // which the partner lane l ^ 32 holds
    const float W = *wsum;
    const int b = bins[row];
    const float wr = (W > 0.0f && b >= 0) ? w[row] / W : 0.0f;
    const float inv_s = 1.0f / s;
    const int h = lane >> 5;
    float zb = 0.0f;
    uint16_t* drow = dZ + (size_t)row * kGuideBins;
    f32x16_t da[2];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        da[0][i] = 0.0f;
        da[1][i] = 0.0f;
    }
    for (int t = 0; t < kOutTiles; ++t) {
        float z[16];
        tile(t, z);
        uint2 dq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int bin0 = 32 * t + 8 * q + 4 * h;
            uint16_t d[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float p = __expf(z[4 * q + j] - m) * inv_s - (bin0 + j == b ? 1.0f : 0.0f);
                if (bin0 + j == b) zb = z[4 * q + j];
                d[j] = f2bf(wr * p);
            }
            dq[q] = make_uint2(d[0] | ((uint32_t)d[1] << 16), d[2] | ((uint32_t)d[3] << 16));
            *reinterpret_cast<uint2*>(drow + bin0) = dq[q];
        }
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            const uint2 send = h ? dq[2 * st] : dq[2 * st + 1];
            const uint2 recv = make_uint2((uint32_t)__shfl_xor((int)send.x, 32), (uint32_t)__shfl_xor((int)send.y, 32));
            const uint4 bv = h ? make_uint4(recv.x, recv.y, dq[2 * st + 1].x, dq[2 * st + 1].y)
                               : make_uint4(dq[2 * st].x, dq[2 * st].y, recv.x, recv.y);
            const bf16x8_t bf = __builtin_bit_cast(bf16x8_t, bv);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const bf16x8_t wf = *reinterpret_cast<const bf16x8_t*>(
                    W4 + (size_t)(32 * a + (lane & 31)) * kGuideBins + 32 * t + 16 * st + 8 * h);
                da[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf, bf, da[a], 0, 0, 0);
            }
        }
    }
    zb += __shfl_xor(zb, 32);
    if (lane < 32) row_loss[row] = wr != 0.0f ? wr * (__logf(s) + m - zb) : 0.0f;
    // element i of da[a]: k = 32 a + 8 (i / 4) + 4 h + (i % 4) of the lane's row
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k0 = 32 * a + 8 * q + 4 * h;
            const uint2 hm = *reinterpret_cast<const uint2*>(H3 + (size_t)row * kGuideHidden + k0);
            const uint16_t hv[4] = {(uint16_t)hm.x, (uint16_t)(hm.x >> 16), (uint16_t)hm.y, (uint16_t)(hm.y >> 16)};
            uint16_t o[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) o[j] = f2bf(bf2f(hv[j]) > 0.0f ? da[a][4 * q + j] : 0.0f);
            *reinterpret_cast<uint2*>(dA3 + (size_t)row * kGuideHidden + k0) =
                make_uint2(o[0] | ((uint32_t)o[1] << 16), o[2] | ((uint32_t)o[3] << 16));
        }
    }
}

// out[0] += sum of v; with `count`, out[1] += entries > 0.  One atomic per
// block and value (at most kReduceBlocks blocks).
constexpr int kReduceBlocks = 64;
__global__ __launch_bounds__(256) void k_sum_count(const float* __restrict__ v, int n, float* __restrict__ out,
                                                   int count) {
    __shared__ float part[2][4];
    float s = 0.0f, c = 0.0f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float x = v[i];
        s += x;
        c += x > 0.0f ? 1.0f : 0.0f;
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        c += __shfl_xor(c, off);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][wave] = s;
        part[1][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float ts = (part[0][0] + part[0][1]) + (part[0][2] + part[0][3]);
        float tc = (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]);
        if (ts != 0.0f) atomicAdd(out, ts);
        if (count && tc != 0.0f) atomicAdd(out + 1, tc);
    }
}

// Adam (Kingma & Ba 2015, Algorithm 1) on the FP32 master copy; refreshes the BF16 copy.
__global__ __launch_bounds__(256) void k_adam(AdamArgs a) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    if (a.wsum && !(*a.wsum > 0.0f)) return;
    float g = a.grad[i];
    float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * g;
    float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * (g * g);
    a.m[i] = m;
    a.v[i] = v;
    float bc1 = 1.0f - powf(a.beta1, (float)a.step);
    float bc2 = 1.0f - powf(a.beta2, (float)a.step);
    float mh = m / bc1, vh = v / bc2;
    float w = a.master[i] - a.lr * mh / (sqrtf(vh) + a.eps);
    a.master[i] = w;
    a.wbf[i] = f2bf(w);
}

__global__ void k_to_bf16(const float* __restrict__ in, uint16_t* __restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// Proposal table: one wave per row, p = exp(z - max), CDF of the prefix sums
// (16 consecutive bins per lane, then a wave scan of the lane totals), each
// divided by the total; cdf[0] = 0 and cdf[1024] = 1 exactly.
__global__ __launch_bounds__(256) void k_guide_table(const float* __restrict__ Z, int rows, float* __restrict__ cdf) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    const float* zr = Z + (size_t)r * kGuideBins + lane * 16;
    float z[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) z[j] = zr[j];
    float m = z[0];
#pragma unroll
    for (int j = 1; j < 16; ++j) m = fmaxf(m, z[j]);
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    float pre[16], s = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        s += __expf(z[j] - m);
        pre[j] = s;
    }
    // inclusive scan of the lane totals
    float incl = s;
    for (int off = 1; off < 64; off <<= 1) {
        float t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
    }
    float total = __shfl(incl, 63);
    float base = incl - s;
    float* out = cdf + (size_t)r * kGuideCdfStride;
#pragma unroll
    for (int j = 0; j < 16; ++j) out[lane * 16 + j + 1] = (base + pre[j]) / total;
    if (lane == 0) out[0] = 0.0f;
    if (lane == 63) out[kGuideBins] = 1.0f;
}

__device__ __forceinline__ uint32_t bucket_of(float4 p, float ox, float oy, float oz, float inv_cell, uint32_t mask) {
    int cx = (int)floorf((p.x - ox) * inv_cell);
    int cy = (int)floorf((p.y - oy) * inv_cell);
    int cz = (int)floorf((p.z - oz) * inv_cell);
    return cell_bucket(cx, cy, cz, mask);
}

__global__ void k_mark_visible(const float4* __restrict__ qpos, uint32_t n, float ox, float oy, float oz,
                               float inv_cell, uint32_t mask, uint8_t* __restrict__ vis) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        vis[bucket_of(qpos[i], ox, oy, oz, inv_cell, mask)] = 1;
}

__global__ void k_guide_targets(const float4* __restrict__ rec, uint32_t n,
                                float ox, float oy, float oz, float inv_cell, uint32_t mask,
                                const uint8_t* __restrict__ vis, uint32_t train_paths, float* __restrict__ w) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4* r = rec + kPhotonRecord * (size_t)i;
        const float4 r2 = r[2];
        uint32_t path = __float_as_uint(r2.z);
        if (path >= train_paths) continue;
        if (!vis[bucket_of(r[0], ox, oy, oz, inv_cell, mask)]) continue;
        float lum = (r[1].w + r2.x + r2.y) * (1.0f / 3.0f);
        atomicAdd(w + path, lum);
    }
}

__global__ void k_mark_visible_batch(const float4* __restrict__ qpos, const float4* __restrict__ qgn, uint32_t n,
                                     BatchView b, uint8_t* __restrict__ vis) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int sc = __float_as_int(qgn[i].w);
        const PhotonMapView& pm = b.pm[sc];
        vis[b.bs[sc].bucket_base +
            bucket_of(qpos[i], pm.origin[0], pm.origin[1], pm.origin[2], pm.inv_cell, pm.mask)] = 1;
    }
}

__global__ void k_guide_targets_batch(const float4* __restrict__ rec, uint32_t n,
                                      BatchView b, const uint8_t* __restrict__ vis,
                                      const GuideTargets* __restrict__ t) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4* r = rec + kPhotonRecord * (size_t)i;
        const float4 r2 = r[2];
        uint32_t path = __float_as_uint(r2.z);
        int sc = batch_scene_of(b.light_starts, b.n, path);
        const BatchScene& bs = b.bs[sc];
        uint32_t local = path - bs.light_base;
        GuideTargets tg = t[sc];
        if (local >= tg.rows) continue;
        const PhotonMapView& pm = b.pm[sc];
        if (!vis[bs.bucket_base + bucket_of(r[0], pm.origin[0], pm.origin[1], pm.origin[2], pm.inv_cell, pm.mask)])
            continue;
        float lum = (r[1].w + r2.x + r2.y) * (1.0f / 3.0f);
        atomicAdd(tg.w + local, lum);
    }
}

// ---- launchers ----
void launch_guide_encode(const float* x, int n, uint16_t* X, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_guide_encode, dim3((n + 127) / 128), dim3(128), 0, s, x, n, X);
}
bool launch_out_softmax(const uint16_t* H3, const uint16_t* W4, uint16_t* W4f, const float* b4, const float* w,
                        const int* bins, int n, const float* wsum, uint16_t* dZ, float* row_loss, float* loss,
                        uint16_t* dA3, hipStream_t s) {
    if (n <= 0 || n % kOutRowsPerWave) return false;
    const int rows = kOutWaves * kOutRowsPerWave;
    hipLaunchKernelGGL(k_w4_frag, dim3(kOutTiles * 4 * 64 / 256), dim3(256), 0, s, W4, W4f);
    hipLaunchKernelGGL(k_out_softmax, dim3((n + rows - 1) / rows), dim3(kOutWaves * 64), 0, s, H3, W4, W4f, b4, w,
                       bins, n, wsum, dZ, row_loss, dA3);
    hipLaunchKernelGGL(k_sum_count, dim3(std::min(kReduceBlocks, (n + 255) / 256)), dim3(256), 0, s, row_loss, n,
                       loss, 0);
    return true;
}
void launch_sum_count(const float* v, int n, float* out, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_sum_count, dim3(std::min(kReduceBlocks, (n + 255) / 256)), dim3(256), 0, s, v, n, out, 1);
}
void launch_adam(const AdamArgs& a, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_adam, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}
void launch_to_bf16(const float* in, uint16_t* out, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n);
}
void launch_guide_table(const float* Z, int rows, float* cdf, hipStream_t s) {
    if (rows > 0) hipLaunchKernelGGL(k_guide_table, dim3((rows + 3) / 4), dim3(256), 0, s, Z, rows, cdf);
}
void launch_mark_visible(const float4* qpos, uint32_t nq, float ox, float oy, float oz, float inv_cell, uint32_t mask,
                         uint8_t* vis, hipStream_t s) {
    if (nq > 0)
        hipLaunchKernelGGL(k_mark_visible, dim3(std::min<uint32_t>(4096, (nq + 255) / 256)), dim3(256), 0, s, qpos, nq,
                           ox, oy, oz, inv_cell, mask, vis);
}
void launch_guide_targets(const float4* rec, uint32_t count, float ox, float oy, float oz, float inv_cell,
                          uint32_t mask, const uint8_t* vis, uint32_t train_paths, float* w, hipStream_t s) {
    if (count > 0 && train_paths > 0)
        hipLaunchKernelGGL(k_guide_targets, dim3(std::min<uint32_t>(8192, (count + 255) / 256)), dim3(256), 0, s, rec,
                           count, ox, oy, oz, inv_cell, mask, vis, train_paths, w);
}

void launch_mark_visible_batch(const float4* qpos, const float4* qgn, uint32_t nq, const BatchView& b, uint8_t* vis,
                               hipStream_t s) {
    if (nq > 0)
        hipLaunchKernelGGL(k_mark_visible_batch, dim3(std::min<uint32_t>(4096, (nq + 255) / 256)), dim3(256), 0, s,
                           qpos, qgn, nq, b, vis);
}
void launch_guide_targets_batch(const float4* rec, uint32_t count, const BatchView& b, const uint8_t* vis,
                                const GuideTargets* t, hipStream_t s) {
    if (count > 0)
        hipLaunchKernelGGL(k_guide_targets_batch, dim3(std::min<uint32_t>(8192, (count + 255) / 256)), dim3(256), 0,
                           s, rec, count, b, vis, t);
}

}  // namespace rc
