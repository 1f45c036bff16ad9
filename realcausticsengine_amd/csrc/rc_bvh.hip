// rc_bvh.hip -- GPU build of a linear BVH (Karras, "Maximizing Parallelism
// in the Construction of BVHs, Octrees, and k-d Trees", HPG 2012) in the
// engine's binary node layout: Morton codes of the triangle box centres
// (10 bits per axis) above the triangle index (so that every key is
// distinct), one radix sort, one thread per internal node for the hierarchy, and a bottom-up pass
// in which the second child to arrive at a node forms its box; then,
// optionally, the collapse to 4-wide nodes.  One triangle per leaf.  The
// reference builds its BVH on the CPU (BVH_mesh.cpp:28-155); the GPU build
// serves scenes of millions of triangles.
#include "rc_engine.h"
#include "rc_vec.h"

namespace rc {

__device__ __forceinline__ void tri_box(const float4* tri, uint32_t i, float lo[3], float hi[3]) {
    float4 a = tri[3 * i], b = tri[3 * i + 1], c = tri[3 * i + 2];
    float p0[3] = {a.x, a.y, a.z};
    float p1[3] = {a.x + b.x, a.y + b.y, a.z + b.z};
    float p2[3] = {a.x + c.x, a.y + c.y, a.z + c.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(fminf(p0[k], p1[k]), p2[k]);
        hi[k] = fmaxf(fmaxf(p0[k], p1[k]), p2[k]);
    }
}

__global__ void k_lbvh_keys(const float4* tri, uint32_t n, int index_bits, float lx, float ly, float lz, float sx,
                            float sy, float sz, unsigned long long* keys, uint32_t* vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lo[3], hi[3];
    tri_box(tri, i, lo, hi);
    float c[3] = {(lo[0] + hi[0]) * 0.5f, (lo[1] + hi[1]) * 0.5f, (lo[2] + hi[2]) * 0.5f};
    float o[3] = {lx, ly, lz}, sc[3] = {sx, sy, sz};
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float t = (c[k] - o[k]) * sc[k];
        t = fminf(fmaxf(t, 0.0f), 1023.0f);
        q[k] = (uint32_t)t;
    }
    uint32_t code = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    keys[i] = ((unsigned long long)code << index_bits) | i;
    vals[i] = i;
}

// Triangles into leaf (sorted) order.
__global__ void k_lbvh_permute(const float4* tri_in, const float4* shade_in, const int* id_in, const uint32_t* order,
                               uint32_t n, float4* tri_out, float4* shade_out, int* id_out) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t j = order[k];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        tri_out[3 * k + r] = tri_in[3 * j + r];
        shade_out[3 * k + r] = shade_in[3 * j + r];
    }
    id_out[k] = id_in[j];
}

__device__ __forceinline__ int common_prefix(const unsigned long long* keys, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    return __clzll(keys[i] ^ keys[j]);
}

__device__ __forceinline__ int leaf_ref(int k) { return ~(k << 3); }

// Internal node i of n - 1 (Karras 2012, Figure 4): its key range, split and children.
__global__ void k_lbvh_internal(const unsigned long long* keys, int n, float4* nodes, int* parent) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = common_prefix(keys, n, i, i + 1) - common_prefix(keys, n, i, i - 1) >= 0 ? 1 : -1;
    int dmin = common_prefix(keys, n, i, i - d);
    int lmax = 2;
    while (common_prefix(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (common_prefix(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = common_prefix(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div <<= 1) {
        int t = (l + div - 1) / div;
        if (common_prefix(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    int gamma = i + s * d + min(d, 0);
    int left = min(i, j) == gamma ? leaf_ref(gamma) : gamma;
    int right = max(i, j) == gamma + 1 ? leaf_ref(gamma + 1) : gamma + 1;
    nodes[4 * i + 3] = make_float4(__int_as_float(left), __int_as_float(right), 0.0f, 0.0f);
    if (left >= 0) parent[left] = i;
    else parent[n + gamma] = i;
    if (right >= 0) parent[right] = i;
    else parent[n + gamma + 1] = i;
    if (i == 0) parent[0] = -1;
}

// Child box slot of node p: slot 0 is the left child.  The node keeps the
// box padded outwards by 2^-20 of its magnitude (the host builder's padding),
// boxes[] keeps it unpadded for the parent's union.
__device__ __forceinline__ void put_child_box(float4* nodes, float* boxes, int p, int slot, const float lo[3],
                                              const float hi[3]) {
    float pl[3], ph[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float m = fmaxf(fabsf(lo[k]), fabsf(hi[k]));
        float pad = m * 9.5367431640625e-07f + 1e-30f;
        pl[k] = lo[k] - pad;
        ph[k] = hi[k] + pad;
    }
    float* nf = reinterpret_cast<float*>(nodes + 4 * p);
    if (slot == 0) {
        nf[0] = pl[0];
        nf[1] = ph[0];
        nf[2] = pl[1];
        nf[3] = ph[1];
        nf[4] = pl[2];
        nf[5] = ph[2];
    } else {
        nf[6] = pl[0];
        nf[7] = ph[0];
        nf[8] = pl[1];
        nf[9] = ph[1];
        nf[10] = pl[2];
        nf[11] = ph[2];
    }
    float* b = boxes + 12 * (size_t)p + 6 * slot;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        __hip_atomic_store(b + k, lo[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(b + 3 + k, hi[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_lbvh_boxes(const float4* tri, int n, float4* nodes, const int* parent, uint32_t* flags,
                             float* boxes) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float lo[3], hi[3];
    tri_box(tri, (uint32_t)k, lo, hi);
    int child = leaf_ref(k);
    int p = parent[n + k];
    for (int steps = 0; p >= 0 && steps < n; ++steps) {  // a path to the root has at most n - 1 nodes
        int slot = __float_as_int(nodes[4 * p + 3].x) == child ? 0 : 1;
        put_child_box(nodes, boxes, p, slot, lo, hi);
        __threadfence();
        if (atomicAdd(&flags[p], 1u) == 0u) return;  // the sibling's thread finishes this node
        __threadfence();
        const float* b = boxes + 12 * (size_t)p + 6 * (1 - slot);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], __hip_atomic_load(b + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            hi[a] = fmaxf(hi[a], __hip_atomic_load(b + 3 + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
        child = p;
        p = parent[p];
    }
}

// Leaf depth (number of nodes from the root): the traversal stack must hold it.
__global__ void k_lbvh_depth(int n, const int* parent, uint32_t* depth) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t d = 0;
    for (int p = parent[n + k]; p >= 0 && d <= (uint32_t)(2 * kMaxDepth); p = parent[p]) ++d;
    atomicMax(depth, d);
}

// 4-wide nodes of the linear BVH: the internal nodes at even depth (edges
// from the root).  A 4-wide node takes its binary node's children, and in
// place of an internal child (odd depth) that child's two children, so every
// internal node of odd depth is opened by its parent and every grandchild
// that is internal is again a 4-wide node.  Layout as the host-collapsed tree.
__global__ void k_lbvh_wide_flags(int n, const int* parent, uint32_t* wide) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    uint32_t d = 0;
    for (int p = parent[i]; p >= 0 && d <= (uint32_t)(2 * kMaxDepth); p = parent[p]) ++d;
    wide[i] = (d & 1u) ? 0u : 1u;
}

// Child `side` (0 left, 1 right) of binary node `node`: box and ref.
__device__ __forceinline__ int binary_child(const float4* nodes, int node, int side, float lo[3], float hi[3]) {
    const float4* b = nodes + 4 * node;
    if (side == 0) {
        lo[0] = b[0].x; hi[0] = b[0].y; lo[1] = b[0].z; hi[1] = b[0].w; lo[2] = b[1].x; hi[2] = b[1].y;
        return __float_as_int(b[3].x);
    }
    lo[0] = b[1].z; hi[0] = b[1].w; lo[1] = b[2].x; hi[1] = b[2].y; lo[2] = b[2].z; hi[2] = b[2].w;
    return __float_as_int(b[3].y);
}

__global__ void k_lbvh_collapse4(int n, const float4* nodes, const uint32_t* wide, const uint32_t* wide_idx,
                                 float4* nodes4) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1 || !wide[i]) return;
    float lo[3][4] = {}, hi[3][4] = {};
    int refs[4] = {kEmptyChild, kEmptyChild, kEmptyChild, kEmptyChild};
    int c = 0;
    for (int side = 0; side < 2; ++side) {
        float bl[3], bh[3];
        const int ref = binary_child(nodes, i, side, bl, bh);
        if (ref < 0) {
            for (int k = 0; k < 3; ++k) {
                lo[k][c] = bl[k];
                hi[k][c] = bh[k];
            }
            refs[c++] = ref;
            continue;
        }
        for (int s2 = 0; s2 < 2; ++s2) {
            const int g = binary_child(nodes, ref, s2, bl, bh);
            for (int k = 0; k < 3; ++k) {
                lo[k][c] = bl[k];
                hi[k][c] = bh[k];
            }
            refs[c++] = g < 0 ? g : (int)wide_idx[g];
        }
    }
    float4* o = nodes4 + 8 * (size_t)wide_idx[i];
    for (int k = 0; k < 3; ++k) {
        o[2 * k] = make_float4(lo[k][0], lo[k][1], lo[k][2], lo[k][3]);
        o[2 * k + 1] = make_float4(hi[k][0], hi[k][1], hi[k][2], hi[k][3]);
    }
    o[6] = make_float4(__int_as_float(refs[0]), __int_as_float(refs[1]), __int_as_float(refs[2]),
                       __int_as_float(refs[3]));
    o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

int lbvh_wide_count(const LbvhBuffers& b, uint32_t n, hipStream_t s) {
    if (n < 2) return -1;
    unsigned g = (n + 255) / 256;
    hipLaunchKernelGGL(k_lbvh_wide_flags, dim3(g), dim3(256), 0, s, (int)n, b.parent, b.flags);
    launch_exclusive_scan_u32(b.flags, b.vals_a, n - 1, b.hist, s);
    uint32_t last[2] = {0, 0};
    if (hipMemcpyAsync(&last[0], b.vals_a + (n - 2), 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&last[1], b.flags + (n - 2), 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
        return -1;
    return (int)(last[0] + last[1]);
}

bool lbvh_collapse4(const LbvhBuffers& b, uint32_t n, float4* nodes4, hipStream_t s) {
    unsigned g = (n + 255) / 256;
    hipLaunchKernelGGL(k_lbvh_collapse4, dim3(g), dim3(256), 0, s, (int)n, b.nodes, b.flags, b.vals_a, nodes4);
    return hipStreamSynchronize(s) == hipSuccess && hipGetLastError() == hipSuccess;
}

int build_lbvh(const LbvhBuffers& b, uint32_t n, const float lo[3], const float hi[3], hipStream_t s) {
    if (n < 2) return -1;
    unsigned g = (n + 255) / 256;
    int index_bits = 1;
    while ((1u << index_bits) < n) ++index_bits;
    float sc[3];
    for (int k = 0; k < 3; ++k) {
        float ext = hi[k] - lo[k];
        sc[k] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    hipLaunchKernelGGL(k_lbvh_keys, dim3(g), dim3(256), 0, s, b.tri, n, index_bits, lo[0], lo[1], lo[2], sc[0], sc[1],
                       sc[2], b.keys_a, b.vals_a);
    unsigned long long* keys;
    uint32_t* order;
    radix_sort_pairs(b.keys_a, b.keys_b, b.vals_a, b.vals_b, b.hist, b.hist_sums, nullptr, n, 30 + index_bits, &keys,
                     &order, s);
    // reorder the triangle arrays into leaf order (through the scratch copies)
    float4* shade_tmp = b.tmp4 + 3 * (size_t)n;
    if (hipMemcpyAsync(b.tmp4, b.tri, sizeof(float4) * 3 * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b.tmp_id, b.tri_id, sizeof(int) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(shade_tmp, b.shade, sizeof(float4) * 3 * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(k_lbvh_permute, dim3(g), dim3(256), 0, s, b.tmp4, shade_tmp, b.tmp_id, order, n, b.tri, b.shade,
                       b.tri_id);
    launch_fill_u32(b.flags, 0u, n, s);
    launch_fill_u32(b.depth, 0u, 1, s);
    hipLaunchKernelGGL(k_lbvh_internal, dim3(g), dim3(256), 0, s, keys, (int)n, b.nodes, b.parent);
    hipLaunchKernelGGL(k_lbvh_boxes, dim3(g), dim3(256), 0, s, b.tri, (int)n, b.nodes, b.parent, b.flags, b.boxes);
    hipLaunchKernelGGL(k_lbvh_depth, dim3(g), dim3(256), 0, s, (int)n, b.parent, b.depth);
    uint32_t depth = 0;
    if (hipMemcpyAsync(&depth, b.depth, 4, hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) return -1;
    return (int)depth;
}

}  // namespace rc
