// rc_kernels.hip -- gfx950 kernels of the vertex-merging path.
//
//  k_light_pass   persistent light tracer (Vertex_Merging.cpp:140-194, 231-276):
//                 each wave keeps 64 paths in flight and refills finished
//                 lanes from a global path counter (ballot + mbcnt), traces
//                 one bounce per step and appends photons with one atomic per
//                 wave (ballot + prefix).
//  photon map     hash-grid keys -> LSD radix sort of (bucket, uid) ->
//                 permute -> per-bucket [start, end) (replaces the KD tree of
//                 KD_Tree.cpp:110-190; the radius sets are identical).
//  k_camera_pass  persistent camera tracer with the merge at every non-delta
//                 vertex (Vertex_Merging.cpp:309-363, Path_Vertices.h:46-68).
#include "rc_device.h"

namespace rc {

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// One atomic per wave: lanes with `want` receive consecutive slots.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t* counter, bool want, uint64_t* mask_out) {
    uint64_t m = __ballot(want);
    *mask_out = m;
    if (m == 0) return 0;
    uint32_t leader = (uint32_t)__builtin_ctzll(m);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__builtin_popcountll(m));
    base = (uint32_t)__shfl((int)base, (int)leader);
    return base + lanes_below(m);
}

// Adds per-lane event counts into a global Counters with one atomic per wave and field.
__device__ __forceinline__ void flush_counters(Counters* dst, unsigned long long (&c)[kCounterFields]) {
#pragma unroll
    for (int f = 0; f < kCounterFields; ++f) {
        unsigned long long v = c[f];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (lane_id() == 0 && v) atomicAdd(reinterpret_cast<unsigned long long*>(dst) + f, v);
    }
}
enum { C_PATHS, C_NODES, C_TRIS, C_HITS, C_STORED, C_GATHERS, C_CELLS, C_TESTED, C_ACCEPTED };

// ---------------------------------------------------------------------
// Light pass
// ---------------------------------------------------------------------
struct LightState {
    F3 o, d, tp;
    float dvm;
    int pl, diffuse, last;
    uint32_t path;
};

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_light_pass(LightPassArgs a) {
    __shared__ int stack_lds[kLdsStack * kBlock];
    int* my_stack = stack_lds + threadIdx.x;
    const SceneView& sc = a.sc;
    unsigned long long cnt[kCounterFields] = {0};
    uint32_t n_nodes = 0, n_tris = 0;
    LightState st;
    bool active = false, exhausted = false;
    PathRng rng{a.k0, a.k1, 0u, a.iteration, a.stream};
    for (;;) {
        // refill finished lanes
        bool need = !active && !exhausted;
        uint64_t m;
        uint32_t slot = wave_alloc(a.path_counter, need, &m);
        if (need) {
            if (slot < a.n_paths) {
                if (kCount) cnt[C_PATHS]++;
                st.path = slot;
                rng.path = slot;
                // Generate_light_sample (Vertex_Merging.cpp:18-25)
                active = sample_light(sc, rng, st.o, st.d, st.tp);
                st.pl = 0;
                st.diffuse = 0;
                st.last = EV_DIFFUSE;
                st.dvm = 0.0f;
            } else {
                exhausted = true;
            }
        }
        if (__ballot(active || !exhausted) == 0) break;
        if (__ballot(active) == 0) continue;
        n_nodes = n_tris = 0;
        // one bounce
        bool alive = active;
        Hit h;
        Surface sf;
        Bsdf b;
        if (alive) {
            F3 ro = st.o + st.d * 0.001f;
            bool hit = trace_closest<kCount>(sc, ro, st.d, 0.001f, h, my_stack, n_nodes, n_tris);
            if (kCount) {
                cnt[C_NODES] += n_nodes;
                cnt[C_TRIS] += n_tris;
                cnt[C_HITS] += hit;
            }
            if (!hit) {
                alive = false;
            } else {
                sf = surface_at(sc, h, ro, st.d);
                if (!bsdf_setup(b, sc.mats[sf.mat], sf.sn, st.d, false)) alive = false;
                else st.dvm /= fabsf(fabsf(b.wo.z));
            }
        }
        bool store = alive && !bsdf_delta(b) && (st.last & EV_GLOSSY);
        uint32_t at = wave_alloc(a.photon_counter, store, &m);
        if (kCount) cnt[C_STORED] += store;
        if (store && at < a.capacity) {
            a.out_pos[at] = make_float4(sf.pos.x, sf.pos.y, sf.pos.z, st.dvm);
            a.out_dir[at] = make_float4(-st.d.x, -st.d.y, -st.d.z, st.tp.x);
            a.out_tp[at] = make_float2(st.tp.y, st.tp.z);
            a.out_uid[at] = st.path * 8u + (uint32_t)st.pl;
        }
        if (alive) {
            if (st.pl > sc.max_path_length || st.diffuse > 2) {
                alive = false;
            } else {
                // Sample_Scattering_light (Vertex_Merging.cpp:48-96)
                float u[3];
                rng.path = st.path;
                rng.u3(5 + 3 * st.pl, u);
                BsdfSample bs;
                F3 f = bsdf_sample(b, sf.gn, u[0], u[1], u[2], bs);
                if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) {
                    alive = false;
                } else {
                    if (st.pl > 2) {
                        if (rng.u(DIM_RR + (st.pl & 7)) > bs.cont) alive = false;
                        else f = div_each(f, bs.cont);
                    }
                    if (alive) {
                        st.o = sf.pos;
                        st.d = bs.wi;
                        if (bs.event & EV_DIFFUSE) st.diffuse++;
                        st.tp = st.tp * f;
                        st.last = bs.event;
                        if (!store) st.dvm *= (bs.cos_out / bs.fwd) * bs.rev;
                        else st.dvm = (bs.cos_out / bs.fwd) * (1.0f + bs.rev * st.dvm);
                        st.pl++;
                    }
                }
            }
        }
        active = alive;
    }
    if (kCount) flush_counters(a.counters, cnt);
}

void launch_light_pass(const LightPassArgs& a, int blocks, hipStream_t s) {
    if (a.counters) hipLaunchKernelGGL(k_light_pass<true>, dim3(blocks), dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL(k_light_pass<false>, dim3(blocks), dim3(kBlock), 0, s, a);
}

// ---------------------------------------------------------------------
// Photon-map radius query
// ---------------------------------------------------------------------
template <bool kCount, typename F>
__device__ __forceinline__ void for_each_in_radius(const PhotonMapView& pm, F3 q, F&& fn, uint32_t& n_cells,
                                                   uint32_t& n_tested) {
    int lo[3], hi[3];
    float qq[3] = {q.x, q.y, q.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = (int)floorf(((qq[k] - pm.r_query) - pm.origin[k]) * pm.inv_cell);
        hi[k] = (int)floorf(((qq[k] + pm.r_query) - pm.origin[k]) * pm.inv_cell);
    }
    uint32_t seen[27];
    int nseen = 0;
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                uint32_t bkt = cell_hash(x, y, z) & pm.mask;
                bool dup = false;
                for (int i = 0; i < nseen; ++i) dup |= seen[i] == bkt;
                if (dup) continue;
                if (nseen < 27) seen[nseen++] = bkt;
                uint2 range = pm.cells[bkt];
                if (kCount) {
                    ++n_cells;
                    n_tested += range.y - range.x;
                }
                for (uint32_t i = range.x; i < range.y; ++i) {
                    float4 p = pm.pos[i];
                    F3 dd = q - f3(p.x, p.y, p.z);
                    float d2 = len2(dd);
                    if (d2 < pm.r2) fn(i, p);
                }
            }
}

// ---------------------------------------------------------------------
// Camera pass
// ---------------------------------------------------------------------
template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_camera_pass(CameraPassArgs a) {
    __shared__ int stack_lds[kLdsStack * kBlock];
    int* my_stack = stack_lds + threadIdx.x;
    const SceneView& sc = a.sc;
    PathRng rng{a.k0, a.k1, 0u, a.iteration, a.stream};
    unsigned long long cnt[kCounterFields] = {0};
    for (;;) {
        uint64_t m;
        uint32_t path = wave_alloc(a.path_counter, true, &m);
        if (__ballot(path < a.n_paths) == 0) break;
        if (path < a.n_paths) {
            if (kCount) cnt[C_PATHS]++;
            rng.path = path;
            const int W = sc.W, H = sc.H;
            // Generate_camera_sample (Vertex_Merging.cpp:27-46), Camera::get_ray (Camera.h:29-34)
            float sx = (float)(int)(path % (uint32_t)W) + 0.0f + 1.0f * rng.u(DIM_JITTER);
            float sy = (float)(int)(path / (uint32_t)W) + 0.0f + 1.0f * rng.u(DIM_JITTER + 1);
            F3 dir = (ld3(sc.cam_bl) + ld3(sc.cam_h) * (sx / (float)W)) + ld3(sc.cam_v) * (sy / (float)H);
            dir = dir - ld3(sc.cam_o);
            F3 o = ld3(sc.cam_o), d = normalize(dir), tp = f3s(1.0f), color = f3s(0.0f);
            float dvm = 0.0f;
            for (int pl = 0;; ++pl) {
                F3 ro = o + d * 0.001f;
                Hit h;
                uint32_t n_nodes = 0, n_tris = 0;
                bool hit = trace_closest<kCount>(sc, ro, d, 0.001f, h, my_stack, n_nodes, n_tris);
                if (kCount) {
                    cnt[C_NODES] += n_nodes;
                    cnt[C_TRIS] += n_tris;
                    cnt[C_HITS] += hit;
                }
                if (!hit) break;
                Surface sf = surface_at(sc, h, ro, d);
                Bsdf b;
                if (!bsdf_setup(b, sc.mats[sf.mat], sf.sn, d, true)) break;
                dvm /= fabsf(b.wo.z);
                if (!bsdf_delta(b)) {
                    F3 c = f3s(0.0f);
                    uint32_t n_cells = 0, n_tested = 0, n_acc = 0;
                    for_each_in_radius<kCount>(a.pm, sf.pos, [&](uint32_t i, float4 p) {
                        float4 dt = a.pm.dir_tr[i];
                        float2 gb = a.pm.tp_gb[i];
                        float fwd, rev;
                        F3 e = bsdf_eval_pdf(b, f3(dt.x, dt.y, dt.z), sf.gn, fwd, rev);
                        float w = 1.0f / (1.0f + dvm * rev + p.w * fwd);
                        c = c + (e * f3(dt.w, gb.x, gb.y)) * w;
                        if (kCount) ++n_acc;
                    }, n_cells, n_tested);
                    if (kCount) {
                        cnt[C_GATHERS]++;
                        cnt[C_CELLS] += n_cells;
                        cnt[C_TESTED] += n_tested;
                        cnt[C_ACCEPTED] += n_acc;
                    }
                    color = color + (tp * a.vm_norm) * c;
                }
                if (pl > sc.max_path_length) break;
                float u[3];
                rng.u3(3 * pl, u);
                BsdfSample bs;
                F3 f = bsdf_sample(b, sf.gn, u[0], u[1], u[2], bs);
                if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) break;
                if (pl > 2) {
                    if (rng.u(DIM_RR + (pl & 7)) > bs.cont) break;
                    f = div_each(f, bs.cont);
                }
                o = sf.pos;
                d = bs.wi;
                tp = tp * f;
                if (bs.event & EV_SPECULAR) dvm *= (bs.cos_out / bs.fwd) * bs.rev;
                else dvm = (bs.cos_out / bs.fwd) * (1.0f + bs.rev * dvm);
            }
            // Image_Buffer::add_color (Image_Buffer.h:24-37)
            if (!(sx < 0.0f || sx >= (float)W) && !(sy < 0.0f || sy >= (float)H)) {
                int X = (int)sx;
                int Y = sc.compat_rows ? H - (int)sy : H - 1 - (int)sy;
                long long at = (long long)X + (long long)Y * W;
                if (at >= 0 && at <= (long long)W * H - 1) {
                    unsafeAtomicAdd(&a.image[3 * at], color.x);
                    unsafeAtomicAdd(&a.image[3 * at + 1], color.y);
                    unsafeAtomicAdd(&a.image[3 * at + 2], color.z);
                }
            }
        }
    }
    if (kCount) flush_counters(a.counters, cnt);
}

void launch_camera_pass(const CameraPassArgs& a, int blocks, hipStream_t s) {
    if (a.counters) hipLaunchKernelGGL(k_camera_pass<true>, dim3(blocks), dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL(k_camera_pass<false>, dim3(blocks), dim3(kBlock), 0, s, a);
}

// ---------------------------------------------------------------------
// First hit (parity diagnostic)
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_first_hit(SceneView sc, const float* rays, int n, float tmin, int* tri,
                                                      float* t) {
    __shared__ int stack_lds[kLdsStack * kBlock];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 o = f3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
    F3 d = f3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
    Hit h;
    uint32_t n_nodes = 0, n_tris = 0;
    bool hit = trace_closest<false>(sc, o, d, tmin, h, stack_lds + threadIdx.x, n_nodes, n_tris);
    tri[i] = hit ? sc.tri_id[h.tri] : -1;
    t[i] = hit ? h.t : INFINITY;
}

void launch_first_hit(const SceneView& sc, const float* rays, int n, float tmin, int* tri, float* t, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_first_hit, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, sc, rays, n, tmin, tri, t);
}

// ---------------------------------------------------------------------
// Utilities: fill, exclusive scan
// ---------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kScanThreads / 64];
    uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t k = 0; k < kScanThreads / 64; ++k) {
        if (k < w) before += wsum[k];
        all += wsum[k];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

__global__ void k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* partials) {
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems], sum = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t pre = block_exclusive_scan(sum, &total);
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0 && partials) partials[blockIdx.x] = total;
}

__global__ void k_scan_partials(uint32_t* partials, uint32_t n) {
    uint32_t carry = 0;
    for (uint32_t start = 0; start < n; start += kScanTile) {
        uint32_t base = start + threadIdx.x * kScanItems;
        uint32_t v[kScanItems], sum = 0;
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = (base + k < n) ? partials[base + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t pre = block_exclusive_scan(sum, &total) + carry;
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < n) partials[base + k] = pre;
            pre += v[k];
        }
        carry += total;
        __syncthreads();
    }
}

__global__ void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* partials) {
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t add = partials[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) out[base + k] += add;
}

void launch_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* partials, hipStream_t s) {
    if (!n) return;
    uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, s, in, out, n, partials);
    if (tiles > 1) {
        hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, s, partials, tiles);
        hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kScanThreads), 0, s, out, n, partials);
    }
}

// ---------------------------------------------------------------------
// Photon map build: keys, LSD radix sort (8-bit digits, stable), permute,
// bucket ranges.
// ---------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;

size_t photon_map_hist_entries(uint32_t max_count) {
    size_t blocks = (max_count + kSortTile - 1) / kSortTile;
    if (blocks == 0) blocks = 1;
    return 256 * blocks;
}
size_t photon_map_sum_entries(uint32_t max_count) {
    size_t h = photon_map_hist_entries(max_count);
    return (h + kScanTile - 1) / kScanTile + 1;
}

__global__ void k_map_keys(const float4* pos, const uint32_t* uid, uint32_t n, float ox, float oy, float oz,
                           float inv_cell, uint32_t mask, int uid_bits, unsigned long long* keys, uint32_t* vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i];
    int cx = (int)floorf((p.x - ox) * inv_cell);
    int cy = (int)floorf((p.y - oy) * inv_cell);
    int cz = (int)floorf((p.z - oz) * inv_cell);
    uint32_t bkt = cell_hash(cx, cy, cz) & mask;
    keys[i] = ((unsigned long long)bkt << uid_bits) | (unsigned long long)uid[i];
    vals[i] = i;
}

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const unsigned long long* keys, uint32_t n, int shift,
                                                            uint32_t* hist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const unsigned long long* kin, const uint32_t* vin,
                                                               uint32_t n, int shift, const uint32_t* offsets,
                                                               uint32_t nblocks, unsigned long long* kout,
                                                               uint32_t* vout) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcount[kSortThreads / 64][256];
    const uint32_t wave = threadIdx.x >> 6;
    running[threadIdx.x] = offsets[threadIdx.x * nblocks + blockIdx.x];
    uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        for (int k = 0; k < kSortThreads / 64; ++k) wcount[k][threadIdx.x] = 0;
        __syncthreads();
        uint32_t i = base + r * kSortThreads + threadIdx.x;
        bool valid = i < n;
        unsigned long long key = valid ? kin[i] : 0ull;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t digit = (uint32_t)(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            bool b = (digit >> bit) & 1u;
            uint64_t bb = __ballot(valid && b);
            peers &= b ? bb : ~bb;
        }
        uint32_t rank = lanes_below(peers);
        if (valid && rank == 0) wcount[wave][digit] = (uint32_t)__builtin_popcountll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[digit] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += wcount[w][digit];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int k = 0; k < kSortThreads / 64; ++k) add += wcount[k][threadIdx.x];
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

__global__ void k_map_permute(const float4* pos, const float4* dir, const float2* tp, const uint32_t* vals, uint32_t n,
                              float4* opos, float4* odir, float2* otp, uint32_t* oidx) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t j = vals[i];
    opos[i] = pos[j];
    odir[i] = dir[j];
    otp[i] = tp[j];
    oidx[i] = j;
}

__global__ void k_map_cells(const unsigned long long* keys, uint32_t n, int uid_bits, uint2* cells) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t b = (uint32_t)(keys[i] >> uid_bits);
    if (i == 0 || (uint32_t)(keys[i - 1] >> uid_bits) != b) cells[b].x = i;
    if (i == n - 1 || (uint32_t)(keys[i + 1] >> uid_bits) != b) cells[b].y = i + 1;
}

void build_photon_map(const MapBuildArgs& a, hipStream_t s) {
    launch_fill_u32(reinterpret_cast<uint32_t*>(a.cells), 0u, 2 * ((size_t)a.mask + 1), s);
    uint32_t n = a.count;
    if (n == 0) return;
    unsigned g = (n + 255) / 256;
    hipLaunchKernelGGL(k_map_keys, dim3(g), dim3(256), 0, s, a.in_pos, a.in_uid, n, a.origin[0], a.origin[1],
                       a.origin[2], a.inv_cell, a.mask, a.uid_bits, a.keys_a, a.vals_a);
    uint32_t nblocks = (n + kSortTile - 1) / kSortTile;
    int bits = a.bucket_bits + a.uid_bits;
    unsigned long long* kin = a.keys_a;
    unsigned long long* kout = a.keys_b;
    uint32_t* vin = a.vals_a;
    uint32_t* vout = a.vals_b;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(kSortThreads), 0, s, kin, n, shift, a.hist, nblocks);
        launch_exclusive_scan_u32(a.hist, a.hist, 256 * nblocks, a.hist_sums, s);
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(kSortThreads), 0, s, kin, vin, n, shift, a.hist,
                           nblocks, kout, vout);
        unsigned long long* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    hipLaunchKernelGGL(k_map_permute, dim3(g), dim3(256), 0, s, a.in_pos, a.in_dir, a.in_tp, vin, n, a.out_pos,
                       a.out_dir, a.out_tp, a.out_index);
    hipLaunchKernelGGL(k_map_cells, dim3(g), dim3(256), 0, s, kin, n, a.uid_bits, a.cells);
}

// ---------------------------------------------------------------------
// Radius sets (count, then fill) for the parity tests
// ---------------------------------------------------------------------
__global__ void k_gather_count(PhotonMapView pm, const float* q, int nq, uint32_t* counts) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    uint32_t c = 0, n_cells = 0, n_tested = 0;
    for_each_in_radius<false>(pm, f3(q[3 * i], q[3 * i + 1], q[3 * i + 2]), [&](uint32_t, float4) { ++c; }, n_cells,
                              n_tested);
    counts[i] = c;
}
__global__ void k_gather_fill(PhotonMapView pm, const float* q, int nq, const uint32_t* orig,
                              const unsigned long long* offsets, int* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    unsigned long long at = offsets[i];
    uint32_t n_cells = 0, n_tested = 0;
    for_each_in_radius<false>(pm, f3(q[3 * i], q[3 * i + 1], q[3 * i + 2]),
                              [&](uint32_t j, float4) { out[at++] = (int)orig[j]; }, n_cells, n_tested);
}
void launch_gather_count(const PhotonMapView& pm, const float* q, int nq, const uint32_t* orig, uint32_t* counts,
                         hipStream_t s) {
    (void)orig;
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_gather_count, dim3((nq + 127) / 128), dim3(128), 0, s, pm, q, nq, counts);
}
void launch_gather_fill(const PhotonMapView& pm, const float* q, int nq, const uint32_t* orig,
                        const unsigned long long* offsets, int* out, hipStream_t s) {
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_gather_fill, dim3((nq + 127) / 128), dim3(128), 0, s, pm, q, nq, orig, offsets, out);
}

}  // namespace rc
