// rc_ingest.hip -- scene ingest on the GPU.  The host copies each mesh's
// Blender arrays to the device as raw bytes (MVert, MPoly, MLoop); two
// HBM-bound kernels make the world-space vertices and normals (one thread per
// vertex) and then the triangles the BVH build and traversal read (one thread
// per triangle: the first three loops of each polygon, its material, the edge
// vectors, the corner normals) with the scene bounds, reduced per wave and
// folded in with integer atomics on an order-preserving encoding (exact, so
// independent of the order).  The arithmetic is the host ingest's, term for
// term (Transform::ApplyToPoint / ApplyToNormal, Transform.cpp:3-31;
// normal_short_to_float, Triangle.cpp:11-16; Mesh::Mesh, mesh.cpp:12-26), so
// both give the same bits.
#include "rc_engine.h"
#include "rc_vec.h"

namespace rc {

// The mesh holding global element i: the last mesh whose first element is <= i
// (meshes with no elements share their first element with the next mesh).
__device__ __forceinline__ int ingest_mesh_of(const IngestMesh* m, int n, uint32_t i, bool tris) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const uint32_t first = tris ? m[mid].tbase : m[mid].vbase;
        if (first <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_ingest_verts(const IngestMesh* __restrict__ meshes, int n_meshes, uint32_t nv,
                               float4* __restrict__ wv, float4* __restrict__ wn, uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const IngestMesh& m = meshes[ingest_mesh_of(meshes, n_meshes, i, false)];
    const uint32_t* r = reinterpret_cast<const uint32_t*>(m.mvert + 20 * (size_t)(i - m.vbase));
    const float x = __uint_as_float(r[0]), y = __uint_as_float(r[1]), z = __uint_as_float(r[2]);
    const uint32_t n01 = r[3], n2 = r[4];
    if (!isfinite(x) || !isfinite(y) || !isfinite(z)) atomicOr(flags, kIngestNonFinite);
    const float* M = m.M;
    wv[i] = make_float4(M[0] * x + M[1] * y + M[2] * z + M[3], M[4] * x + M[5] * y + M[6] * z + M[7],
                        M[8] * x + M[9] * y + M[10] * z + M[11], 0.0f);
    const float s = 1.0f / 32767.0f;
    const float nx = (float)(int16_t)(n01 & 0xffffu) * s, ny = (float)(int16_t)(n01 >> 16) * s,
                nz = (float)(int16_t)(n2 & 0xffffu) * s;
    const float* I = m.inv;  // I[3 r + c] = inv[r][c]
    F3 n = f3(I[0] * nx + I[3] * ny + I[6] * nz, I[1] * nx + I[4] * ny + I[7] * nz, I[2] * nx + I[5] * ny + I[8] * nz);
    n = normalize(n);
    wn[i] = make_float4(n.x, n.y, n.z, 0.0f);
}

__global__ void k_ingest_tris(const IngestMesh* __restrict__ meshes, int n_meshes, uint32_t nt,
                              const float4* __restrict__ wv, const float4* __restrict__ wn, float4* __restrict__ tri,
                              float4* __restrict__ shade, int* __restrict__ tri_id, int* bounds, uint32_t* flags) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (t < nt) {
        const IngestMesh& m = meshes[ingest_mesh_of(meshes, n_meshes, t, true)];
        const int* poly = reinterpret_cast<const int*>(m.mpoly + 12 * (size_t)(t - m.tbase));
        const int loopstart = poly[0];  // >= 0 and its three loops in the copy: checked by the host
        const int mat_nr = (int)(int16_t)((uint32_t)poly[2] & 0xffffu);
        const uint32_t* loop = reinterpret_cast<const uint32_t*>(m.mloop) + 2 * (size_t)loopstart;
        uint32_t vid[3];
        bool bad = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            vid[k] = loop[2 * k];
            bad |= vid[k] >= m.nv;
        }
        tri_id[t] = (int)t;
        if (bad) {  // reported through the flags; init fails, the triangle is zero
            atomicOr(flags, kIngestLoopVertex);
            for (int k = 0; k < 3; ++k) {
                tri[3 * (size_t)t + k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                shade[3 * (size_t)t + k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        } else {
            const int mat = m.mat_map[mat_nr < 0 ? 0 : mat_nr];
            const float4 a = wv[m.vbase + vid[0]], b = wv[m.vbase + vid[1]], c = wv[m.vbase + vid[2]];
            const F3 p0 = f3(a.x, a.y, a.z), p1 = f3(b.x, b.y, b.z), p2 = f3(c.x, c.y, c.z);
            const F3 e1 = p1 - p0, e2 = p2 - p0;
            tri[3 * (size_t)t] = make_float4(p0.x, p0.y, p0.z, __int_as_float(mat));
            tri[3 * (size_t)t + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
            tri[3 * (size_t)t + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
#pragma unroll
            for (int k = 0; k < 3; ++k) shade[3 * (size_t)t + k] = wn[m.vbase + vid[k]];
            const float pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z}, pc[3] = {c.x, c.y, c.z};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lo[k] = smin(smin(pa[k], pb[k]), pc[k]);
                hi[k] = smax(smax(pa[k], pb[k]), pc[k]);
            }
        }
    }
    // the wave's bounds, then one atomic per wave and coordinate
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float l = lo[k], h = hi[k];
        for (int off = 32; off > 0; off >>= 1) {
            l = fminf(l, __shfl_xor(l, off));
            h = fmaxf(h, __shfl_xor(h, off));
        }
        if ((threadIdx.x & 63u) == 0) {
            atomicMin(bounds + k, ordered_float_bits(__float_as_int(l)));
            atomicMax(bounds + 3 + k, ordered_float_bits(__float_as_int(h)));
        }
    }
}

void launch_ingest(const IngestMesh* meshes, int n_meshes, uint32_t nv, uint32_t nt, float4* wv, float4* wn,
                   float4* tri, float4* shade, int* tri_id, int* bounds, uint32_t* flags, hipStream_t s) {
    if (nv > 0)
        hipLaunchKernelGGL(k_ingest_verts, dim3((nv + 255) / 256), dim3(256), 0, s, meshes, n_meshes, nv, wv, wn, flags);
    if (nt > 0)
        hipLaunchKernelGGL(k_ingest_tris, dim3((nt + 255) / 256), dim3(256), 0, s, meshes, n_meshes, nt, wv, wn, tri,
                           shade, tri_id, bounds, flags);
}

}  // namespace rc
