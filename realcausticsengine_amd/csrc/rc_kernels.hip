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
//  k_camera_trace persistent camera tracer (Vertex_Merging.cpp:309-363); every
//                 non-delta vertex becomes a merge query.
//  k_gather       the merge of each query (Path_Vertices.h:46-68), one wave
//                 per query with ballot-compacted BSDF evaluation.
//  k_resolve      per-path sum of the vertex contributions into the film.
#include <mutex>

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

// Per-wave pool of path indices taken from the global counter in chunks of
// `chunk` (>= 64, the autotuned queue granularity), so the refill of finished
// lanes costs one atomic per chunk.  Indices are handed out in increasing
// order; a wave only stops when its lanes draw indices past the end, so no
// index below the end is skipped.
struct WavePool {
    uint32_t chunk;
    uint32_t next = 0, end = 0;  // wave-uniform
    __device__ __forceinline__ uint32_t draw(uint32_t* counter, bool want) {
        uint64_t m = __ballot(want);
        if (m == 0) return 0;
        uint32_t n = (uint32_t)__builtin_popcountll(m);
        uint32_t r = lanes_below(m);
        uint32_t rem = end - next;
        uint32_t slot;
        if (rem >= n) {
            slot = next + r;
            next += n;
        } else {
            uint32_t base = 0;
            if (lane_id() == 0) base = atomicAdd(counter, chunk);
            base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
            slot = r < rem ? next + r : base + (r - rem);
            next = base + (n - rem);
            end = base + chunk;
        }
        return slot;
    }
};

// Copies of the sort's digit counts: a count is added to copy (workgroup index
// modulo kDigitCopies), so that no address takes more than a few hundred atomics.
constexpr int kDigitCopies = 8;

// The photon-map key (bucket, light path) of a photon at `p` (as k_map_keys).
__device__ __forceinline__ unsigned long long light_photon_key(const LightPassArgs& a, float4 p, uint32_t path) {
    float ox = a.map.origin[0], oy = a.map.origin[1], oz = a.map.origin[2], inv_cell = a.map.inv_cell;
    uint32_t mask = a.map.mask, base = 0;
    if (a.batch.n > 0) {
        const int sc = batch_scene_of(a.batch.light_starts, a.batch.n, path);
        const PhotonMapView& pm = a.batch.pm[sc];
        ox = pm.origin[0];
        oy = pm.origin[1];
        oz = pm.origin[2];
        inv_cell = pm.inv_cell;
        mask = pm.mask;
        base = a.batch.bs[sc].bucket_base;
    } else if (a.dyn) {
        ox = a.dyn->origin[0];
        oy = a.dyn->origin[1];
        oz = a.dyn->origin[2];
        inv_cell = a.dyn->inv_cell;
    }
    int cx = (int)floorf((p.x - ox) * inv_cell);
    int cy = (int)floorf((p.y - oy) * inv_cell);
    int cz = (int)floorf((p.z - oz) * inv_cell);
    const uint32_t bkt = base + cell_bucket(cx, cy, cz, mask);
    return ((unsigned long long)bkt << a.map.uid_bits) | (unsigned long long)path;
}

// Photons of a wave are gathered in an LDS ring and written out 64 at a time:
// one atomic on the shared photon counter per 64 photons instead of one per
// bounce step (the counter is the light pass's most contended address), and
// coalesced stores.  A ring holds fewer than 64 entries before a push, so 128
// entries never overflow.
constexpr uint32_t kPhotonRing = 128;
struct PhotonRing {
    float4 pos[kPhotonRing];
    float4 dir[kPhotonRing];
    float2 tp[kPhotonRing];
    uint32_t path[kPhotonRing];
    uint8_t len[kPhotonRing];
};
struct PhotonQueue {
    PhotonRing* r;
    uint32_t head = 0, count = 0;  // wave-uniform
    __device__ __forceinline__ void flush(const LightPassArgs& a, uint32_t n) {
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(a.photon_counter, n);
        base = (uint32_t)__shfl((int)base, 0);
        const uint32_t l = lane_id();
        if (l < n) {
            const uint32_t e = (head + l) & (kPhotonRing - 1u);
            const uint32_t at = base + l;
            if (at < a.capacity) {
                float4* o = a.out_rec + kPhotonRecord * (size_t)at;
                const float2 tp = r->tp[e];
                const float4 pos = r->pos[e];
                o[0] = pos;
                o[1] = r->dir[e];
                o[2] = make_float4(tp.x, tp.y, __uint_as_float(r->path[e]), __uint_as_float((uint32_t)r->len[e]));
                if (a.map.keys) {  // the photon map's key and its first digit's count
                    const unsigned long long k = light_photon_key(a, pos, r->path[e]);
                    a.map.keys[at] = k;
                    a.map.vals[at] = at;
                    atomicAdd(&a.map.dig[(blockIdx.x % kDigitCopies) * 256 + ((uint32_t)k & 255u)], 1u);
                }
            }
        }
        head = (head + n) & (kPhotonRing - 1u);
        count -= n;
        __builtin_amdgcn_wave_barrier();
    }
    __device__ __forceinline__ void push(const LightPassArgs& a, bool store, float4 pos, float4 dir, float2 tp,
                                         uint32_t path, int len) {
        const uint64_t m = __ballot(store);
        if (m == 0) return;
        if (store) {
            const uint32_t e = (head + count + lanes_below(m)) & (kPhotonRing - 1u);
            r->pos[e] = pos;
            r->dir[e] = dir;
            r->tp[e] = tp;
            r->path[e] = path;
            r->len[e] = (uint8_t)len;
        }
        count += (uint32_t)__builtin_popcountll(m);
        __builtin_amdgcn_wave_barrier();
        if (count >= 64u) flush(a, 64u);
    }
    __device__ __forceinline__ void drain(const LightPassArgs& a) {
        while (count > 0) flush(a, count < 64u ? count : 64u);
    }
};

// Adds per-lane event counts into a global Counters with one atomic per wave and field.
__device__ __forceinline__ void flush_counters(Counters* dst, unsigned long long (&c)[kCounterFields]) {
#pragma unroll
    for (int f = 0; f < kCounterFields; ++f) {
        unsigned long long v = c[f];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (lane_id() == 0 && v) atomicAdd(reinterpret_cast<unsigned long long*>(dst) + f, v);
    }
}
enum { C_PATHS, C_NODES, C_TRIS, C_HITS, C_STORED, C_GATHERS, C_CELLS, C_TESTED, C_ACCEPTED, C_STREAMED };

// ---------------------------------------------------------------------
// Light pass
// ---------------------------------------------------------------------
struct LightState {
    F3 o, d, tp;
    float dvm;
    int pl, diffuse, last;
    uint32_t path;   // global light-path index (photon uid = path * 8 + length)
    uint32_t local;  // index within its scene (the Philox counter)
    int scene;
};

// Generate_light_sample (Vertex_Merging.cpp:18-25) for the light-path slot
// `slot`: the path, its scene and Philox counter, the emitted ray and the
// proposal's training record.  Returns whether the light emits.
template <bool kBatch>
__device__ __forceinline__ bool emit_light_path(const LightPassArgs& a, uint32_t slot, LightState& st, PathRng& rng) {
    const uint32_t path = a.path_order ? a.path_order[slot] : slot;
    st.path = path;
    st.local = path;
    st.scene = 0;
    if (kBatch) {
        st.scene = batch_scene_of(a.batch.light_starts, a.batch.n, path);
        const BatchScene& bs = a.batch.bs[st.scene];
        st.local = path - bs.light_base;
        rng.k0 = bs.k0;
        rng.k1 = bs.k1;
        rng.iter = bs.iteration;
    }
    rng.path = st.local;
    st.pl = 0;
    st.diffuse = 0;
    st.last = EV_DIFFUSE;
    st.dvm = 0.0f;
    if (a.emitted) return false;  // the emission kernel made this slot's ray: the caller loads it
    const SceneView& sc = kBatch ? a.batch.sc[st.scene] : a.sc;
    const GuideView& gv = kBatch ? a.batch.guides[st.scene] : a.guide;
    int bin;
    F3 axis;
    const bool emits = sample_light(sc, rng, gv, st.o, st.d, st.tp, bin, axis);
    if (st.local < gv.train_paths) {
        float rec[kGuideRowFloats];
        guide_record(gv, st.o.x, st.o.y, st.o.z, axis.x, axis.y, axis.z, rec);
        float4* dst = reinterpret_cast<float4*>(gv.train_x + (size_t)st.local * kGuideRowFloats);
        dst[0] = make_float4(rec[0], rec[1], rec[2], rec[3]);
        dst[1] = make_float4(rec[4], rec[5], rec[6], rec[7]);
        gv.train_bin[st.local] = emits ? bin : -1;
    }
    return emits;
}

// The emission of every light-path slot in one coherent pass (option
// light_emit): the persistent pass then loads its refills instead of
// sampling the lights on the few lanes a refill serves.
template <bool kBatch>
__global__ __launch_bounds__(256) void k_light_emit(LightPassArgs a) {
    LightPassArgs b = a;
    b.emitted = nullptr;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < a.n_paths; slot += stride) {
        LightState st;
        PathRng rng{a.k0, a.k1, 0u, a.dyn ? a.dyn->iteration : a.iteration, a.stream};
        const bool emits = emit_light_path<kBatch>(b, slot, st, rng);
        float4* e = a.emitted + 3 * (size_t)slot;
        e[0] = make_float4(st.o.x, st.o.y, st.o.z, st.tp.x);
        e[1] = make_float4(st.d.x, st.d.y, st.d.z, st.tp.y);
        e[2] = make_float4(st.tp.z, emits ? 1.0f : 0.0f, 0.0f, 0.0f);
    }
}

template <bool kCount, bool kBatch>
__global__ __launch_bounds__(kBlock) void k_light_pass(LightPassArgs a) {
    __shared__ int2 stack_lds[kLdsStack * kBlock];
    __shared__ PhotonRing rings[kBlock / 64];
    int2* my_stack = stack_lds + threadIdx.x;
    PhotonQueue photons{&rings[threadIdx.x >> 6]};
    unsigned long long cnt[kCounterFields] = {0};
    uint32_t n_nodes = 0, n_tris = 0;
    LightState st;
    bool active = false, exhausted = false;
    PathRng rng{a.k0, a.k1, 0u, a.dyn ? a.dyn->iteration : a.iteration, a.stream};
    WavePool pool{a.path_chunk};
    for (;;) {
        // refill finished lanes
        bool need = !active && !exhausted;
        // refill once at least refill_min lanes are idle (or none is live):
        // the refill code then runs for many lanes at once
        if ((uint32_t)__builtin_popcountll(__ballot(need)) < a.refill_min && __ballot(active) != 0) need = false;
        uint32_t slot = pool.draw(a.path_counter, need);
        if (need) {
            if (slot < a.n_paths) {
                if (kCount) cnt[C_PATHS]++;
                active = emit_light_path<kBatch>(a, slot, st, rng);
                if (a.emitted) {
                    const float4* e = a.emitted + 3 * (size_t)slot;
                    const float4 e0 = e[0], e1 = e[1], e2 = e[2];
                    st.o = f3(e0.x, e0.y, e0.z);
                    st.d = f3(e1.x, e1.y, e1.z);
                    st.tp = f3(e0.w, e1.w, e2.x);
                    active = e2.y != 0.0f;
                }
            } else {
                exhausted = true;
            }
        }
        if (__ballot(active || !exhausted) == 0) break;
        if (__ballot(active) == 0) continue;
        n_nodes = n_tris = 0;
        const SceneView& sc = kBatch ? a.batch.sc[st.scene] : a.sc;
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
        if (kCount) cnt[C_STORED] += store;
        photons.push(a, store, make_float4(sf.pos.x, sf.pos.y, sf.pos.z, st.dvm),
                     make_float4(-st.d.x, -st.d.y, -st.d.z, st.tp.x), make_float2(st.tp.y, st.tp.z),
                     st.path, st.pl);
        if (alive) {
            if (st.pl > sc.max_path_length || st.diffuse > 2) {
                alive = false;
            } else {
                // Sample_Scattering_light (Vertex_Merging.cpp:48-96)
                rng.path = st.local;
                U4 w = rng.block((uint32_t)bounce_block(STREAM_LIGHT, st.pl));
                BsdfSample bs;
                F3 f = bsdf_sample(b, sf.gn, u24(w.a), u24(w.b), u24(w.c), bs);
                if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) {
                    alive = false;
                } else {
                    if (st.pl > 2) {
                        if (u24(w.d) > bs.cont) alive = false;
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
    photons.drain(a);
    if (kCount) flush_counters(a.counters, cnt);
}

// ---------------------------------------------------------------------
// Adaptive MLT light pass (Vertex_Merging.cpp:231-282, MLT_Sampler.cpp:7-67)
// with one Markov chain per lane instead of the reference's single serial
// chain.  Chain c takes the path indices c, c + chains, c + 2 chains, ...;
// for each index it traces the uniform sample of the index (the plain
// estimator's path); a path that stores no photon is "invisible" and is
// followed by a mutation of the chain's current state, accepted when the
// mutated path stores a photon.  Visible uniform samples become the chain's
// state; every path that stores photons keeps them.  The mutation size adapts
// towards an acceptance rate of 0.234 as in the reference, per chain.  The
// host scales the merge by (visible uniform paths) / N, the reference's
// visibility normalisation.
//
// Primary samples are Philox words; the chain's state is either the Philox
// sample of (base_path, base_stream) -- its last accepted uniform sample or
// its initial state -- or, after an accepted mutation, explicit words in
// memory.  A mutation moves each 24-bit value by +-e (1/e)^(1/size) modulo 1
// (MLT_Sampler::Mutate), e and the sign drawn per dimension; a size that has
// adapted to zero or below moves by up to +-e (the reference's huge steps
// modulo 1 are likewise jumps across the whole interval).
// ---------------------------------------------------------------------
struct MltRng {
    uint32_t k0, k1, iter, path;
    bool mutate;
    const uint4* state;  // explicit state (block b at state[b * stride]) or null
    uint32_t stride, base_path, base_stream;
    float step;  // powf(1/e, 1/mutation_size)
    __device__ __forceinline__ U4 current(uint32_t b) const {
        if (state) {
            uint4 v = state[(size_t)b * stride];
            return U4{v.x, v.y, v.z, v.w};
        }
        return philox(b, base_path, iter, base_stream, k0, k1);
    }
    __device__ __forceinline__ uint32_t moved(uint32_t w, uint32_t e_word, uint32_t sign_word) const {
        float dlt = u24(e_word) * step;
        dlt -= floorf(dlt);
        uint32_t d24 = (uint32_t)(dlt * 16777216.0f) & 0xFFFFFFu;
        uint32_t v = w >> 8;
        v = u24(sign_word) < 0.5f ? v + d24 : v - d24;
        return (v & 0xFFFFFFu) << 8;
    }
    __device__ __forceinline__ U4 block(uint32_t b) const {
        if (!mutate) return philox(b, path, iter, STREAM_LIGHT, k0, k1);
        U4 s = current(b);
        U4 r0 = philox(2 * b, path, iter, STREAM_MLT_DELTA, k0, k1);
        U4 r1 = philox(2 * b + 1, path, iter, STREAM_MLT_DELTA, k0, k1);
        return U4{moved(s.a, r0.a, r0.b), moved(s.b, r0.c, r0.d), moved(s.c, r1.a, r1.b), moved(s.d, r1.c, r1.d)};
    }
};

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_light_pass_mlt(LightPassArgs a) {
    __shared__ int2 stack_lds[kLdsStack * kBlock];
    __shared__ PhotonRing rings[kBlock / 64];
    int2* my_stack = stack_lds + threadIdx.x;
    PhotonQueue photons{&rings[threadIdx.x >> 6]};
    unsigned long long cnt[kCounterFields] = {0};
    uint32_t n_nodes = 0, n_tris = 0;
    LightState st;
    const uint32_t lanes = gridDim.x * blockDim.x;
    const uint32_t C = a.mlt_chains;
    GuideView plain{};  // no proposal: every light samples plain emission
    plain.cells = 1;
    uint32_t chain = blockIdx.x * blockDim.x + threadIdx.x;  // this lane's first chain
    uint32_t step_k = 0;                                      // position of the current index in the chain
    float msize = 1.0f;                                       // MLT_Sampler(max_path_length, 1.0f)
    int acc = 1, mut = 0;                                     // Vertex_Merging.cpp:227-228
    bool expl = false;
    uint32_t base_path = chain, base_stream = STREAM_MLT_INIT;
    bool mutating = false, started = false, visible = false;
    bool active = false, exhausted = chain >= C || chain >= a.n_paths;
    uint32_t n_visible = 0, n_mut = 0, n_acc = 0;
    MltRng rng{a.k0, a.k1, a.iteration, 0u, false, nullptr, C, 0u, 0u, 1.0f};
    for (;;) {
        if (!active && !exhausted) {
            bool next = started;
            if (started) {
                if (!mutating) {
                    if (visible) {  // the uniform sample becomes the chain's state
                        ++n_visible;
                        expl = false;
                        base_path = st.path;
                        base_stream = STREAM_LIGHT;
                    } else {  // Reject + Mutate: a mutation of the state for the same index
                        mutating = true;
                        next = false;
                    }
                } else {
                    ++mut;
                    ++n_mut;
                    if (visible) {  // accepted: the proposal becomes explicit state
                        ++acc;
                        ++n_acc;
#pragma unroll 1
                        for (uint32_t b = 0; b < (uint32_t)kMltBlocks; ++b) {
                            U4 w = rng.block(b);
                            a.mlt_state[(size_t)b * C + chain] = make_uint4(w.a, w.b, w.c, w.d);
                        }
                        expl = true;
                    }
                    msize += ((float)acc / (float)mut - 0.234f) / (float)mut;
                    mutating = false;
                }
            }
            if (next) {
                ++step_k;
                if ((uint64_t)chain + (uint64_t)step_k * C >= a.n_paths) {  // the lane's next chain
                    chain += lanes;
                    step_k = 0;
                    msize = 1.0f;
                    acc = 1;
                    mut = 0;
                    expl = false;
                    base_path = chain;
                    base_stream = STREAM_MLT_INIT;
                    if (chain >= C || chain >= a.n_paths) exhausted = true;
                }
            }
            if (!exhausted) {
                if (kCount) cnt[C_PATHS]++;
                st.path = chain + step_k * C;
                st.local = st.path;
                st.scene = 0;
                rng.path = st.path;
                rng.mutate = mutating;
                rng.state = expl ? a.mlt_state + chain : nullptr;
                rng.base_path = base_path;
                rng.base_stream = base_stream;
                rng.step = mlt_step(msize);
                int bin;
                F3 axis;
                active = sample_light(a.sc, rng, plain, st.o, st.d, st.tp, bin, axis);
                st.pl = 0;
                st.diffuse = 0;
                st.last = EV_DIFFUSE;
                st.dvm = 0.0f;
                started = true;
                visible = false;
            }
        }
        if (__ballot(active || !exhausted) == 0) break;
        if (__ballot(active) == 0) continue;
        n_nodes = n_tris = 0;
        const SceneView& sc = a.sc;
        // one bounce, as k_light_pass
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
        if (kCount) cnt[C_STORED] += store;
        if (store) visible = true;
        photons.push(a, store, make_float4(sf.pos.x, sf.pos.y, sf.pos.z, st.dvm),
                     make_float4(-st.d.x, -st.d.y, -st.d.z, st.tp.x), make_float2(st.tp.y, st.tp.z),
                     st.path, st.pl);
        if (alive) {
            if (st.pl > sc.max_path_length || st.diffuse > 2) {
                alive = false;
            } else {
                U4 w = rng.block((uint32_t)bounce_block(STREAM_LIGHT, st.pl));
                BsdfSample bs;
                F3 f = bsdf_sample(b, sf.gn, u24(w.a), u24(w.b), u24(w.c), bs);
                if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) {
                    alive = false;
                } else {
                    if (st.pl > 2) {
                        if (u24(w.d) > bs.cont) alive = false;
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
    photons.drain(a);
    // chain statistics: one atomic per wave and counter
    unsigned long long v3[3] = {n_visible, n_mut, n_acc};
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        unsigned long long v = v3[f];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (lane_id() == 0 && v) atomicAdd(a.mlt_counts + f, (uint32_t)v);
    }
    if (kCount) flush_counters(a.counters, cnt);
}

static unsigned emit_blocks(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 8192u)); }

void launch_light_pass(const LightPassArgs& a, int blocks, hipStream_t s) {
    if (a.mlt_state) {
        if (a.counters) hipLaunchKernelGGL((k_light_pass_mlt<true>), dim3(blocks), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_light_pass_mlt<false>), dim3(blocks), dim3(kBlock), 0, s, a);
        return;
    }
    if (a.batch.n > 0) {
        if (a.emitted) hipLaunchKernelGGL(k_light_emit<true>, dim3(emit_blocks(a.n_paths)), dim3(256), 0, s, a);
        if (a.counters) hipLaunchKernelGGL((k_light_pass<true, true>), dim3(blocks), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_light_pass<false, true>), dim3(blocks), dim3(kBlock), 0, s, a);
    } else {
        if (a.emitted) hipLaunchKernelGGL(k_light_emit<false>, dim3(emit_blocks(a.n_paths)), dim3(256), 0, s, a);
        if (a.counters) hipLaunchKernelGGL((k_light_pass<true, false>), dim3(blocks), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_light_pass<false, false>), dim3(blocks), dim3(kBlock), 0, s, a);
    }
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
    uint32_t seen[64];
    int nseen = 0;
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                uint32_t bkt = cell_bucket(x, y, z, pm.mask);
                bool dup = false;
                for (int i = 0; i < nseen; ++i) dup |= seen[i] == bkt;
                if (dup) continue;
                if (nseen < 64) seen[nseen++] = bkt;
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
// Camera pass, part 1: trace camera paths (Vertex_Merging.cpp:309-363) with
// per-bounce refill like the light pass; every non-delta vertex becomes a
// merge query instead of gathering inline.
// ---------------------------------------------------------------------
struct CameraState {
    F3 o, d, tp;
    float dvm;
    int pl;
    uint32_t path, mask;  // global camera-path index, vertices with a query
    uint32_t local;       // pixel-sample index within its scene (the Philox counter)
    int scene;
    float vm_norm;
    int pixel;
    uint32_t qmc_u0;  // camera_qmc: the first bounce's stratified first word
};

template <bool kCount, bool kBatch>
__global__ __launch_bounds__(kBlock) void k_camera_trace(CameraPassArgs a) {
    __shared__ int2 stack_lds[kLdsStack * kBlock];
    int2* my_stack = stack_lds + threadIdx.x;
    if (a.dyn) {
        a.iteration = a.dyn->iteration;
        a.vm_norm = a.dyn->vm_norm;
    }
    PathRng rng{a.k0, a.k1, 0u, a.iteration, a.stream};
    unsigned long long cnt[kCounterFields] = {0};
    CameraState st;
    bool active = false, exhausted = false;
    WavePool pool{a.path_chunk};
    for (;;) {
        bool need = !active && !exhausted;
        // refill once at least refill_min lanes are idle (or none is live):
        // the refill code then runs for many lanes at once
        if ((uint32_t)__builtin_popcountll(__ballot(need)) < a.refill_min && __ballot(active) != 0) need = false;
        uint64_t m;
        uint32_t slot = pool.draw(a.path_counter, need);
        if (need) {
            if (slot < a.n_paths) {
                if (kCount) cnt[C_PATHS]++;
                st.path = slot;
                st.local = slot;
                st.scene = 0;
                st.vm_norm = a.vm_norm;
                if (kBatch) {
                    st.scene = batch_scene_of(a.batch.camera_starts, a.batch.n, slot);
                    const BatchScene& bs = a.batch.bs[st.scene];
                    st.local = slot - bs.camera_base;
                    st.vm_norm = bs.vm_norm;
                    rng.k0 = bs.k0;
                    rng.k1 = bs.k1;
                    rng.iter = bs.iteration;
                }
                rng.path = st.local;
                const SceneView& sc = kBatch ? a.batch.sc[st.scene] : a.sc;
                const int W = sc.W, H = sc.H;
                const uint32_t px = st.local;
                // Generate_camera_sample (Vertex_Merging.cpp:27-46), Camera::get_ray (Camera.h:29-34)
                U4 wj = rng.block(0);
                if (a.camera_qmc) {  // stratified across iterations (sample index = iteration)
                    U4 sh = philox(0u, px, 0xFFFFFFFFu, STREAM_QMC_SHIFT, rng.k0, rng.k1);
                    wj.a = qmc_word(qmc_vdc(rng.iter) >> 8, sh.a);
                    wj.b = qmc_word(qmc_sobol2(rng.iter) >> 8, sh.b);
                    st.qmc_u0 = qmc_word(qmc_ri3_24(rng.iter), sh.c);
                }
                float sx = (float)(int)(px % (uint32_t)W) + 0.0f + 1.0f * u24(wj.a);
                float sy = (float)(int)(px / (uint32_t)W) + 0.0f + 1.0f * u24(wj.b);
                F3 dir = (ld3(sc.cam_bl) + ld3(sc.cam_h) * (sx / (float)W)) + ld3(sc.cam_v) * (sy / (float)H);
                dir = dir - ld3(sc.cam_o);
                st.o = ld3(sc.cam_o);
                st.d = normalize(dir);
                st.tp = f3s(1.0f);
                st.dvm = 0.0f;
                st.pl = 0;
                st.mask = 0;
                // Image_Buffer::add_color (Image_Buffer.h:24-37): target buffer element
                st.pixel = -1;
                if (!(sx < 0.0f || sx >= (float)W) && !(sy < 0.0f || sy >= (float)H)) {
                    int X = (int)sx;
                    int Y = sc.compat_rows ? H - (int)sy : H - 1 - (int)sy;
                    long long at = (long long)X + (long long)Y * W;
                    if (at >= 0 && at <= (long long)W * H - 1) st.pixel = (int)at;
                }
                active = true;
            } else {
                exhausted = true;
            }
        }
        if (__ballot(active || !exhausted) == 0) break;
        if (__ballot(active) == 0) continue;
        const SceneView& sc = kBatch ? a.batch.sc[st.scene] : a.sc;
        bool alive = active;
        Hit h;
        Surface sf;
        Bsdf b;
        if (alive) {
            F3 ro = st.o + st.d * 0.001f;
            uint32_t n_nodes = 0, n_tris = 0;
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
                if (!bsdf_setup(b, sc.mats[sf.mat], sf.sn, st.d, true)) alive = false;
                else st.dvm /= fabsf(b.wo.z);
            }
        }
        bool query = alive && !bsdf_delta(b);
        uint32_t at = wave_alloc(a.path_counter + 1, query, &m);
        if (query && at < a.q.capacity) {
            F3 tpn = st.tp * st.vm_norm;
            a.q.pos[at] = make_float4(sf.pos.x, sf.pos.y, sf.pos.z, st.dvm);
            a.q.dir[at] = make_float4(st.d.x, st.d.y, st.d.z, __int_as_float(sf.mat));
            a.q.sn[at] = make_float4(sf.sn.x, sf.sn.y, sf.sn.z, __uint_as_float(st.path * 8u + (uint32_t)st.pl));
            a.q.gn[at] = make_float4(sf.gn.x, sf.gn.y, sf.gn.z, __int_as_float(st.scene));
            a.q.tpn[at] = make_float4(tpn.x, tpn.y, tpn.z, 0.0f);
        }
        if (query) st.mask |= 1u << st.pl;
        if (alive) {
            if (st.pl > sc.max_path_length) {
                alive = false;
            } else {
                // Sample_Scattering_camera (Vertex_Merging.cpp:98-135)
                rng.path = st.local;
                U4 w = rng.block((uint32_t)bounce_block(STREAM_CAMERA, st.pl));
                if (a.camera_qmc && st.pl == 0) w.a = st.qmc_u0;
                BsdfSample bs;
                F3 f = bsdf_sample(b, sf.gn, u24(w.a), u24(w.b), u24(w.c), bs);
                if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) {
                    alive = false;
                } else {
                    if (st.pl > 2) {
                        if (u24(w.d) > bs.cont) alive = false;
                        else f = div_each(f, bs.cont);
                    }
                    if (alive) {
                        st.o = sf.pos;
                        st.d = bs.wi;
                        st.tp = st.tp * f;
                        if (bs.event & EV_SPECULAR) st.dvm *= (bs.cos_out / bs.fwd) * bs.rev;
                        else st.dvm = (bs.cos_out / bs.fwd) * (1.0f + bs.rev * st.dvm);
                        st.pl++;
                    }
                }
            }
        }
        if (active && !alive) a.path_info[st.path] = make_int2(st.pixel, (int)st.mask);
        active = alive;
    }
    if (kCount) flush_counters(a.counters, cnt);
}

// ---------------------------------------------------------------------
// Camera pass, part 2: the merge (Path_Vertices.h:46-68).  One wave per
// query: lanes stride over the query's photon cells with coalesced loads,
// accepted photons are compacted (ballot + prefix) into an LDS queue and
// evaluated 64 at a time with every lane busy; the lane sums are reduced
// with a fixed butterfly, so a query's result does not depend on timing.
// ---------------------------------------------------------------------
constexpr int kGatherBlock = 256;
constexpr int kGatherWaves = kGatherBlock / 64;

// Wave-wide inclusive prefix sum.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)v, off);
        if (lane >= (uint32_t)off) v += y;
    }
    return v;
}

// The gather grid is a multiple of kXcds blocks; the dispatcher places block b
// on XCD b % kXcds, so XCD x takes the x-th contiguous segment of the query
// list (queries come out of the trace in pixel order, so a segment is a band
// of the image) and its L2 holds that band's photons.
constexpr int kXcds = 8;

// Fingerprint of a merge's radius set (counting launches): photon count and
// the sum of mix64 of the input indices, per contribution slot.
__device__ __forceinline__ void record_merge(const CameraPassArgs& a, uint32_t slot, uint32_t sorted_i) {
    atomicAdd(a.merge_sets + 2 * (size_t)slot, 1ull);
    atomicAdd(a.merge_sets + 2 * (size_t)slot + 1, mix64((unsigned long long)a.map_index[sorted_i]));
}

template <bool kCount, bool kFast, bool kBatch>
__global__ __launch_bounds__(kGatherBlock) void k_gather(CameraPassArgs a) {
    __shared__ uint32_t queue_lds[kGatherWaves][128];
    __shared__ uint32_t rfirst_lds[kGatherWaves][64];
    __shared__ uint32_t rbase_lds[kGatherWaves][64];
    const uint32_t wv = threadIdx.x >> 6;
    uint32_t* queue = queue_lds[wv];
    const uint32_t lane = lane_id();
    unsigned long long cnt[kCounterFields] = {0};
    PhotonMapView pm0 = a.pm;
    if (a.dyn) {
        pm0.r2 = a.dyn->r2;
        pm0.r_query = a.dyn->r_query;
        pm0.inv_cell = a.dyn->inv_cell;
        pm0.origin[0] = a.dyn->origin[0];
        pm0.origin[1] = a.dyn->origin[1];
        pm0.origin[2] = a.dyn->origin[2];
    }
    const uint32_t nq = min(*(volatile uint32_t*)(a.path_counter + 1), a.q.capacity);
    const uint32_t xcd = blockIdx.x % kXcds;
    const uint32_t seg = (nq + kXcds - 1) / kXcds;
    const uint32_t q_begin = min(nq, xcd * seg), q_end = min(nq, q_begin + seg);
    const uint32_t local_waves = (gridDim.x / kXcds) * kGatherWaves;
    for (uint32_t qj = q_begin + (blockIdx.x / kXcds) * kGatherWaves + wv; qj < q_end; qj += local_waves) {
        const uint32_t qi = a.q_order ? a.q_order[qj] : qj;
        float4 qp = a.q.pos[qi], qd = a.q.dir[qi], qs = a.q.sn[qi], qg = a.q.gn[qi];
        // the query's scene is the same in every lane
        const int scene = kBatch ? __builtin_amdgcn_readfirstlane(__float_as_int(qg.w)) : 0;
        const PhotonMapView& pm = kBatch ? a.batch.pm[scene] : pm0;
        const DevMaterial* mats = kBatch ? a.batch.sc[scene].mats : a.sc.mats;
        F3 q = f3(qp.x, qp.y, qp.z);
        float cdvm = qp.w;
        Bsdf b;
        bsdf_setup(b, mats[__float_as_int(qd.w)], f3(qs.x, qs.y, qs.z), f3(qd.x, qd.y, qd.z), true);
        const MergeConst mc = merge_const(b, f3(qg.x, qg.y, qg.z));
        F3 c = f3s(0.0f);
        auto eval_one = [&](uint32_t i) {
            float4 p = pm.pos[i];
            float4 dt = pm.dir_tr[i];
            float2 gb = pm.tp_gb[i];
            float fwd, rev;
            F3 e = bsdf_eval_pdf<kFast>(b, mc, f3(dt.x, dt.y, dt.z), fwd, rev);
            float w = mdiv<kFast>(1.0f, 1.0f + cdvm * rev + p.w * fwd);
            c = c + (e * f3(dt.w, gb.x, gb.y)) * w;
        };
        // cells overlapping the ball: at most 4 per axis (cell >= 2/3 of the
        // query radius), so lane l takes cell (l & 3, l >> 2 & 3, l >> 4); those
        // cells never share a bucket (cell_bucket)
        int lo[3], hi[3];
        float qq[3] = {q.x, q.y, q.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = (int)floorf(((qq[k] - pm.r_query) - pm.origin[k]) * pm.inv_cell);
            hi[k] = (int)floorf(((qq[k] + pm.r_query) - pm.origin[k]) * pm.inv_cell);
        }
        int cx = lo[0] + (int)(lane & 3u), cy = lo[1] + (int)((lane >> 2) & 3u), cz = lo[2] + (int)(lane >> 4);
        bool valid = cx <= hi[0] && cy <= hi[1] && cz <= hi[2];
        uint2 range = valid ? pm.cells[cell_bucket(cx, cy, cz, pm.mask)] : make_uint2(0u, 0u);
        uint32_t len = range.y - range.x;
        uint64_t ne = __ballot(len > 0u);
        const uint32_t nr = (uint32_t)__builtin_popcountll(ne);
        uint32_t incl = wave_incl_scan(len);
        const uint32_t total = (uint32_t)__shfl((int)incl, 63);
        // compact the non-empty ranges so that lane r holds range r
        if (len > 0u) {
            uint32_t r = lanes_below(ne);
            rfirst_lds[wv][r] = range.x;
            rbase_lds[wv][r] = incl - len;
        }
        __builtin_amdgcn_wave_barrier();
        const uint32_t r_first = rfirst_lds[wv][lane];
        const uint32_t r_base = rbase_lds[wv][lane];
        if (kCount) {
            uint64_t lookups = __ballot(valid);
            if (lane == 0) {
                cnt[C_CELLS] += (uint32_t)__builtin_popcountll(lookups);
                cnt[C_TESTED] += total;
                cnt[C_STREAMED] += total;
            }
        }
        uint32_t qhead = 0, qcount = 0;  // wave-uniform queue state
        uint32_t next_r = 0;             // wave-uniform: next range start to apply
        uint32_t cur_first = 0, cur_base = 0;
        // stream position s of this lane -> photon index.  Range starts below
        // the chunk's end are applied in increasing order with uniform
        // read-lanes; a start is retired (next_r) once the whole chunk is past it.
        auto locate = [&](uint32_t chunk_base, uint32_t s) -> uint32_t {
            for (uint32_t r = next_r; r < nr; ++r) {
                uint32_t bnd = (uint32_t)__builtin_amdgcn_readlane((int)r_base, (int)r);
                if (bnd >= chunk_base + 64u) break;
                uint32_t fst = (uint32_t)__builtin_amdgcn_readlane((int)r_first, (int)r);
                if (s >= bnd) {
                    cur_first = fst;
                    cur_base = bnd;
                }
                if (bnd <= chunk_base) next_r = r + 1;
            }
            return cur_first + (s - cur_base);
        };
        auto push = [&](bool acc, uint32_t i) {
            uint64_t mk = __ballot(acc);
            uint32_t n = (uint32_t)__builtin_popcountll(mk);
            if (acc) queue[(qhead + qcount + lanes_below(mk)) & 127u] = i;
            qcount += n;
            if (kCount && lane == 0) cnt[C_ACCEPTED] += n;
            if (kCount && a.merge_sets && acc) record_merge(a, __float_as_uint(qs.w), i);
            __builtin_amdgcn_wave_barrier();
            if (qcount >= 64) {
                eval_one(queue[(qhead + lane) & 127u]);
                qhead = (qhead + 64) & 127u;
                qcount -= 64;
                __builtin_amdgcn_wave_barrier();
            }
        };
        for (uint32_t sb = 0; sb < total; sb += 128) {
            // two chunks of the stream in flight per lane
            uint32_t s0 = sb + lane, s1 = sb + 64 + lane;
            uint32_t i0 = locate(sb, s0);
            uint32_t i1 = locate(sb + 64, s1);
            bool v0 = s0 < total, v1 = s1 < total;
            float4 p0 = v0 ? pm.pos[i0] : make_float4(0, 0, 0, 0);
            float4 p1 = v1 ? pm.pos[i1] : make_float4(0, 0, 0, 0);
            bool acc0 = v0 && len2(q - f3(p0.x, p0.y, p0.z)) < pm.r2;
            bool acc1 = v1 && len2(q - f3(p1.x, p1.y, p1.z)) < pm.r2;
            push(acc0, i0);
            push(acc1, i1);
        }
        if (lane < qcount) eval_one(queue[(qhead + lane) & 127u]);
        __builtin_amdgcn_wave_barrier();
        // fixed butterfly reduction over the 64 lanes
        for (int off = 32; off > 0; off >>= 1) {
            c.x += __shfl_xor(c.x, off);
            c.y += __shfl_xor(c.y, off);
            c.z += __shfl_xor(c.z, off);
        }
        if (lane == 0) {
            float4 t = a.q.tpn[qi];
            F3 r = f3(t.x, t.y, t.z) * c;
            uint32_t slot = __float_as_uint(qs.w);
            a.contrib[3 * (size_t)slot] = r.x;
            a.contrib[3 * (size_t)slot + 1] = r.y;
            a.contrib[3 * (size_t)slot + 2] = r.z;
            if (kCount) cnt[C_GATHERS]++;
        }
    }
    if (kCount) flush_counters(a.counters, cnt);
}

// ---------------------------------------------------------------------
// Grouped merge: a wave takes kQ (4, 8 or 16) consecutive queries of the spatial
// order.  Queries whose candidate cells are the same window (sub-runs of the
// group) share one pass over the window's photons: each lane loads one
// photon (position, direction, throughput) per step and tests it against
// every query of the sub-run, evaluating the accepted ones.  A photon is read
// once per sub-run instead of once per query.  Per-query state (BSDF frame,
// merge constants) sits in LDS and is read as broadcasts; the per-lane
// partial sums of the sub-run's queries are registers (the query loop is
// unrolled), summed over the lanes in a fixed order when the sub-run ends.  A
// query's sum depends only on its window's photon stream (photon s goes to
// lane s mod 64), not on which queries share its wave; the terms are the
// per-query kernel's, added in another order.
// ---------------------------------------------------------------------
struct GroupQuery {
    Bsdf b;
    MergeConst mc;
    F3 q;
    float cdvm;
};

// What the merge needs of a query whose surface is a diffuse-only plastic
// (Fresnel-microfacet lobe with reflectance 0, fast arithmetic): the lobe is
// the constant R / pi with pdf |cos| / pi, so a photon needs only its cosine
// against the shading normal and the same-hemisphere tests, and R / pi
// multiplies the query's sum once, after the reduction.  k1 = 1 + camera
// dVM x reverse pdf is the left part of the MIS denominator, rounded as
// there.  Every query carries its position here; a per-group bit mask marks
// the ones that take this form.  Three float4 (one LDS read each):
// {q, k1}, {shading normal, side_wo}, {geometric normal, wo.z}.
struct DiffuseQuery {
    float4 qk, ns, gw;
};

// Sum of a float over the 64 lanes of a wave, the same bits in every lane, in
// a fixed order: pairs and quads (quad permutes), eights (half-row mirror),
// rows of 16 (row mirror) -- each step adds a lane's value and its partner's,
// so partners agree bit for bit -- then the four row sums read from lanes 15,
// 31, 47 and 63 and added as (r0 + r1) + (r2 + r3).  Every lane must be active.
template <int kCtrl>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), kCtrl, 0xf, 0xf, false));
}
__device__ __forceinline__ float wave_total(float v) {
    v += dpp_f<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_f<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 47));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    return (r0 + r1) + (r2 + r3);
}

__device__ __forceinline__ float uniform_f(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

__device__ __forceinline__ float4 uniform_f4(float4 v) {
    return make_float4(uniform_f(v.x), uniform_f(v.y), uniform_f(v.z), uniform_f(v.w));
}
// Wave-uniform copy of a DiffuseQuery read from LDS: scalar registers.
// The diffuse-only merge term of one photon at a query (kFast): the two
// same-hemisphere tests of bsdf_eval_pdf's diffuse branch in the unfused
// arithmetic of the reference (they decide whether the term counts), then the
// MIS weight and the sum with fused multiply-adds (the query's sum is
// compared with the reference's to a relative tolerance).  pw_inv_pi is the
// photon's dVM over pi.
__device__ __forceinline__ void add_diffuse_term(F3& c, const DiffuseQuery& d, bool acc, float4 dt, float2 gb,
                                                 float pw_inv_pi) {
    const F3 wi = f3(dt.x, dt.y, dt.z);
    const float wz = dot(wi, f3(d.ns.x, d.ns.y, d.ns.z));
    const bool keep = acc & !(dot(wi, f3(d.gw.x, d.gw.y, d.gw.z)) * d.ns.w <= 0.0f) & !(wz * d.gw.w <= 0.0f);
    // without a branch: a term that does not count adds 0 (the photon's throughput is finite)
    const float w = keep ? __builtin_amdgcn_rcpf(fmaf(pw_inv_pi, fabsf(wz), d.qk.w)) : 0.0f;
    c.x = fmaf(dt.w, w, c.x);
    c.y = fmaf(gb.x, w, c.y);
    c.z = fmaf(gb.y, w, c.z);
}

__device__ __forceinline__ DiffuseQuery uniform_dq(const DiffuseQuery& d) {
    return DiffuseQuery{uniform_f4(d.qk), uniform_f4(d.ns), uniform_f4(d.gw)};
}

// The cell window of a query: cells [lo, hi] per axis overlapping its ball.
__device__ __forceinline__ void query_window(const PhotonMapView& pm, F3 q, int lo[3], int hi[3]) {
    float qq[3] = {q.x, q.y, q.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = (int)floorf(((qq[k] - pm.r_query) - pm.origin[k]) * pm.inv_cell);
        hi[k] = (int)floorf(((qq[k] + pm.r_query) - pm.origin[k]) * pm.inv_cell);
    }
}

// kB = 0: the query loop is unrolled and each query's per-lane sums are
// registers.  kB > 0: the sub-run's photons are taken kB steps (64 kB
// photons) at a time into registers and the query loop is a scalar loop over
// them; each query's per-lane sums live in LDS between blocks.  A lane adds a
// query's terms in the same order either way (photon sb + lane, sb rising), so
// the two give the same bits.
template <bool kCount, bool kFast, bool kBatch, int kQ, int kB>
__global__ __launch_bounds__(kGatherBlock) void k_gather_grouped(CameraPassArgs a) {
    __shared__ uint32_t rfirst_lds[kGatherWaves][64];
    __shared__ uint32_t rbase_lds[kGatherWaves][64];
    __shared__ GroupQuery gq_lds[kGatherWaves][kQ];
    __shared__ DiffuseQuery dq_lds[kGatherWaves][kQ];
    __shared__ float acc_lds[kGatherWaves][kB > 0 ? 3 * kQ : 1][64];
    // up to 4 queries' DiffuseQuery constants stay in scalar registers over a sub-run
    constexpr bool kScalarQ = kQ <= 4;
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    unsigned long long cnt[kCounterFields] = {0};
    PhotonMapView pm0 = a.pm;
    if (a.dyn) {
        pm0.r2 = a.dyn->r2;
        pm0.r_query = a.dyn->r_query;
        pm0.inv_cell = a.dyn->inv_cell;
        pm0.origin[0] = a.dyn->origin[0];
        pm0.origin[1] = a.dyn->origin[1];
        pm0.origin[2] = a.dyn->origin[2];
    }
    const uint32_t nq = min(*(volatile uint32_t*)(a.path_counter + 1), a.q.capacity);
    const uint32_t nblk = (nq + kQ - 1) / kQ;
    const uint32_t xcd = blockIdx.x % kXcds;
    // XCD x takes groups [b_begin, b_end) of the order, dealt to its waves in
    // turn (gather_tile = 0), or the tiles t = x, x + 8, ... of `tile` groups
    // each, taken one group at a time from the XCD's counter by whichever of its
    // waves is free (the counters start at zero with the pass's path counters)
    const bool dyn = a.gather_tile != 0;
    const uint32_t tile = dyn ? max(1u, a.gather_tile / kQ) : (nblk + kXcds - 1) / kXcds;
    const uint32_t b_begin = dyn ? 0u : min(nblk, xcd * tile);
    const uint32_t b_end = dyn ? 0xffffffffu : min(nblk, b_begin + tile);
    const uint32_t local_waves = (gridDim.x / kXcds) * kGatherWaves;
    // a wave takes `grab` consecutive groups per atomic on its XCD's counter
    // (about eight takes per wave), so that the counter is not the critical
    // path when groups are short
    uint32_t* take = a.path_counter + 4 + xcd;
    const uint32_t grab = max(1u, nblk / (kXcds * local_waves * 8u));
    auto take_groups = [&]() -> uint32_t {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(take, grab);
        return (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
    };
    uint32_t k = dyn ? take_groups() : (blockIdx.x / kXcds) * kGatherWaves + wv;
    uint32_t left = grab;
    for (;;) {
        const uint32_t bj = dyn ? ((k / tile) * kXcds + xcd) * tile + k % tile : b_begin + k;
        if (bj >= b_end || bj >= nblk) break;
        const uint32_t j0 = bj * kQ;
        const uint32_t nj = min((uint32_t)kQ, nq - j0);
        // lane j < nj sets up query j of the group
        const bool mine = lane < nj;
        uint32_t qi = 0, slot = 0;
        F3 tpn = f3s(0.0f);
        int scene = -1;
        bool fast = false;
        int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
        if (mine) {
            qi = a.q_order ? a.q_order[j0 + lane] : j0 + lane;
            float4 qp = a.q.pos[qi], qd = a.q.dir[qi], qs = a.q.sn[qi], qg = a.q.gn[qi];
            const float4 t = a.q.tpn[qi];
            tpn = f3(t.x, t.y, t.z);
            slot = __float_as_uint(qs.w);
            scene = kBatch ? __float_as_int(qg.w) : 0;
            const PhotonMapView& pm = kBatch ? a.batch.pm[scene] : pm0;
            const DevMaterial* mats = kBatch ? a.batch.sc[scene].mats : a.sc.mats;
            GroupQuery g;
            g.q = f3(qp.x, qp.y, qp.z);
            g.cdvm = qp.w;
            bsdf_setup(g.b, mats[__float_as_int(qd.w)], f3(qs.x, qs.y, qs.z), f3(qd.x, qd.y, qd.z), true);
            g.mc = merge_const(g.b, f3(qg.x, qg.y, qg.z));
            gq_lds[wv][lane] = g;
            DiffuseQuery d;
            d.qk = make_float4(g.q.x, g.q.y, g.q.z, 1.0f + g.cdvm * g.mc.rd);
            d.ns = make_float4(g.b.fr.z.x, g.b.fr.z.y, g.b.fr.z.z, g.mc.side_wo);
            d.gw = make_float4(g.mc.gn.x, g.mc.gn.y, g.mc.gn.z, g.b.wo.z);
            dq_lds[wv][lane] = d;
            fast = kFast && g.b.kind == BK_FMG && g.mc.f == 0.0f;
            query_window(pm, g.q, lo, hi);
        }
        // bit j: query j of the group takes the diffuse-only form
        const uint64_t fastmask = __ballot(fast);
        // sub-runs: query j starts one when its window (or scene) differs from query j - 1's
        bool same = lane > 0;
        {
            int p;
            p = __shfl_up(scene, 1);
            same = same && p == scene;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                p = __shfl_up(lo[k], 1);
                same = same && p == lo[k];
                p = __shfl_up(hi[k], 1);
                same = same && p == hi[k];
            }
        }
        uint64_t starts = __ballot(mine && !same);
        __builtin_amdgcn_wave_barrier();
        while (starts) {
            const uint32_t s0 = (uint32_t)__builtin_ctzll(starts);
            starts &= starts - 1;
            const uint32_t e0 = starts ? (uint32_t)__builtin_ctzll(starts) : nj;
            const uint32_t nrun = e0 - s0;
            const int sc = __builtin_amdgcn_readlane(scene, (int)s0);
            const PhotonMapView& pm = kBatch ? a.batch.pm[sc] : pm0;
            int wlo[3], whi[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                wlo[k] = __builtin_amdgcn_readlane(lo[k], (int)s0);
                whi[k] = __builtin_amdgcn_readlane(hi[k], (int)s0);
            }
            // lane l takes cell (l & 3, l >> 2 & 3, l >> 4) of the window, as the per-query merge
            int cx = wlo[0] + (int)(lane & 3u), cy = wlo[1] + (int)((lane >> 2) & 3u), cz = wlo[2] + (int)(lane >> 4);
            bool valid = cx <= whi[0] && cy <= whi[1] && cz <= whi[2];
            uint2 range = valid ? pm.cells[cell_bucket(cx, cy, cz, pm.mask)] : make_uint2(0u, 0u);
            uint32_t len = range.y - range.x;
            uint64_t ne = __ballot(len > 0u);
            const uint32_t nr = (uint32_t)__builtin_popcountll(ne);
            uint32_t incl = wave_incl_scan(len);
            const uint32_t total = (uint32_t)__shfl((int)incl, 63);
            if (len > 0u) {
                uint32_t r = lanes_below(ne);
                rfirst_lds[wv][r] = range.x;
                rbase_lds[wv][r] = incl - len;
            }
            __builtin_amdgcn_wave_barrier();
            const uint32_t r_first = rfirst_lds[wv][lane];
            const uint32_t r_base = rbase_lds[wv][lane];
            __builtin_amdgcn_wave_barrier();
            if (kCount) {
                const uint64_t lookups = __ballot(valid);
                if (lane == 0) {
                    cnt[C_CELLS] += (uint32_t)__builtin_popcountll(lookups);
                    cnt[C_TESTED] += (unsigned long long)total * nrun;
                    cnt[C_STREAMED] += total;
                }
            }
            if constexpr (kB > 0) {
                for (uint32_t jj = 0; jj < nrun; ++jj) {
                    acc_lds[wv][3 * jj][lane] = 0.0f;
                    acc_lds[wv][3 * jj + 1][lane] = 0.0f;
                    acc_lds[wv][3 * jj + 2][lane] = 0.0f;
                }
                uint32_t next_r = 0, cur_first = 0, cur_base = 0;
                for (uint32_t sb0 = 0; sb0 < total; sb0 += 64u * kB) {
                    float4 P[kB], DT[kB];
                    float2 GB[kB];
                    float PW[kB];  // the photons' dVM / pi
                    bool V[kB];
                    uint32_t I[kB];
#pragma unroll
                    for (int bb = 0; bb < kB; ++bb) {
                        const uint32_t sb = sb0 + 64u * bb;
                        const uint32_t sp = sb + lane;
                        V[bb] = false;
                        I[bb] = 0u;
                        P[bb] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        DT[bb] = P[bb];
                        GB[bb] = make_float2(0.0f, 0.0f);
                        if (sb < total) {
                            for (uint32_t r = next_r; r < nr; ++r) {
                                uint32_t bnd = (uint32_t)__builtin_amdgcn_readlane((int)r_base, (int)r);
                                if (bnd >= sb + 64u) break;
                                uint32_t fst = (uint32_t)__builtin_amdgcn_readlane((int)r_first, (int)r);
                                if (sp >= bnd) {
                                    cur_first = fst;
                                    cur_base = bnd;
                                }
                                if (bnd <= sb) next_r = r + 1;
                            }
                            V[bb] = sp < total;
                            const uint32_t i = cur_first + (sp - cur_base);
                            I[bb] = i;
                            if (V[bb]) {
                                P[bb] = pm.pos[i];
                                DT[bb] = pm.dir_tr[i];
                                GB[bb] = pm.tp_gb[i];
                            }
                        }
                        PW[bb] = P[bb].w * kInvPI;
                    }
                    for (uint32_t jj = 0; jj < nrun; ++jj) {
                        const uint32_t qj = s0 + jj;
                        const uint32_t qslot = kCount ? (uint32_t)__builtin_amdgcn_readlane((int)slot, (int)qj) : 0u;
                        F3 c = f3(acc_lds[wv][3 * jj][lane], acc_lds[wv][3 * jj + 1][lane], acc_lds[wv][3 * jj + 2][lane]);
                        const DiffuseQuery d = dq_lds[wv][qj];
                        const F3 dq = f3(d.qk.x, d.qk.y, d.qk.z);
                        if (kFast && ((fastmask >> qj) & 1u)) {
#pragma unroll
                            for (int bb = 0; bb < kB; ++bb) {
                                const bool acc = V[bb] && len2(dq - f3(P[bb].x, P[bb].y, P[bb].z)) < pm.r2;
                                if (kCount) {
                                    const uint64_t am = __ballot(acc);
                                    if (lane == 0) cnt[C_ACCEPTED] += (uint32_t)__builtin_popcountll(am);
                                    if (a.merge_sets && acc) record_merge(a, qslot, I[bb]);
                                }
                                // a photon on the wrong side adds nothing (its term is 0 there)
                                add_diffuse_term(c, d, acc, DT[bb], GB[bb], PW[bb]);
                            }
                        } else {
                            const GroupQuery& g = gq_lds[wv][qj];
                            Bsdf b = g.b;
                            b.kind = __builtin_amdgcn_readfirstlane(b.kind);
#pragma unroll
                            for (int bb = 0; bb < kB; ++bb) {
                                const bool acc = V[bb] && len2(dq - f3(P[bb].x, P[bb].y, P[bb].z)) < pm.r2;
                                if (kCount) {
                                    const uint64_t am = __ballot(acc);
                                    if (lane == 0) cnt[C_ACCEPTED] += (uint32_t)__builtin_popcountll(am);
                                    if (a.merge_sets && acc) record_merge(a, qslot, I[bb]);
                                }
                                if (acc) {
                                    float fwd, rev;
                                    F3 e = bsdf_eval_pdf<kFast>(b, g.mc, f3(DT[bb].x, DT[bb].y, DT[bb].z), fwd, rev);
                                    float w = mdiv<kFast>(1.0f, 1.0f + g.cdvm * rev + P[bb].w * fwd);
                                    c = c + (e * f3(DT[bb].w, GB[bb].x, GB[bb].y)) * w;
                                }
                            }
                        }
                        acc_lds[wv][3 * jj][lane] = c.x;
                        acc_lds[wv][3 * jj + 1][lane] = c.y;
                        acc_lds[wv][3 * jj + 2][lane] = c.z;
                    }
                }
                // the sub-run's queries: fixed-order sum over the lanes, then the weighted result into its slot
                for (uint32_t jj = 0; jj < nrun; ++jj) {
                    F3 r = f3(wave_total(acc_lds[wv][3 * jj][lane]), wave_total(acc_lds[wv][3 * jj + 1][lane]),
                              wave_total(acc_lds[wv][3 * jj + 2][lane]));
                    // every lane holds the same total: the query's own lane weights and writes it
                    if (lane == s0 + jj) {
                        if (kFast && ((fastmask >> lane) & 1u)) r = gq_lds[wv][lane].mc.diffuse * r;
                        const F3 o = tpn * r;
                        a.contrib[3 * (size_t)slot] = o.x;
                        a.contrib[3 * (size_t)slot + 1] = o.y;
                        a.contrib[3 * (size_t)slot + 2] = o.z;
                        if (kCount) cnt[C_GATHERS]++;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            } else {
                F3 c[kQ];
                DiffuseQuery sq[kQ];
#pragma unroll
                for (int jj = 0; jj < kQ; ++jj) {
                    c[jj] = f3s(0.0f);
                    if (kScalarQ && (uint32_t)jj < nrun) sq[jj] = uniform_dq(dq_lds[wv][s0 + jj]);
                }
                uint32_t next_r = 0, cur_first = 0, cur_base = 0;
                for (uint32_t sb = 0; sb < total; sb += 64) {
                    const uint32_t sp = sb + lane;
                    for (uint32_t r = next_r; r < nr; ++r) {
                        uint32_t bnd = (uint32_t)__builtin_amdgcn_readlane((int)r_base, (int)r);
                        if (bnd >= sb + 64u) break;
                        uint32_t fst = (uint32_t)__builtin_amdgcn_readlane((int)r_first, (int)r);
                        if (sp >= bnd) {
                            cur_first = fst;
                            cur_base = bnd;
                        }
                        if (bnd <= sb) next_r = r + 1;
                    }
                    const bool v = sp < total;
                    const uint32_t i = cur_first + (sp - cur_base);
                    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f), dt = p;
                    float2 gb = make_float2(0.0f, 0.0f);
                    if (v) {
                        p = pm.pos[i];
                        dt = pm.dir_tr[i];
                        gb = pm.tp_gb[i];
                    }
                    const float pw = p.w * kInvPI;
#pragma unroll
                    for (int jj = 0; jj < kQ; ++jj) {
                        if ((uint32_t)jj < nrun) {
                            const DiffuseQuery& d = kScalarQ ? sq[jj] : dq_lds[wv][s0 + jj];
                            const bool acc = v && len2(f3(d.qk.x, d.qk.y, d.qk.z) - f3(p.x, p.y, p.z)) < pm.r2;
                            if (kCount) {
                                const uint64_t am = __ballot(acc);
                                if (lane == 0) cnt[C_ACCEPTED] += (uint32_t)__builtin_popcountll(am);
                                const uint32_t qslot = (uint32_t)__builtin_amdgcn_readlane((int)slot, (int)(s0 + jj));
                                if (a.merge_sets && acc) record_merge(a, qslot, i);
                            }
                            if (kFast && ((fastmask >> (s0 + jj)) & 1u)) {
                                // a photon on the wrong side adds nothing (its term is 0 there)
                                add_diffuse_term(c[jj], d, acc, dt, gb, pw);
                            } else if (acc) {
                                const GroupQuery& g = gq_lds[wv][s0 + jj];
                                Bsdf b = g.b;
                                b.kind = __builtin_amdgcn_readfirstlane(b.kind);
                                float fwd, rev;
                                F3 e = bsdf_eval_pdf<kFast>(b, g.mc, f3(dt.x, dt.y, dt.z), fwd, rev);
                                float w = mdiv<kFast>(1.0f, 1.0f + g.cdvm * rev + p.w * fwd);
                                c[jj] = c[jj] + (e * f3(dt.w, gb.x, gb.y)) * w;
                            }
                        }
                    }
                }
                // the sub-run's queries: fixed-order sum over the lanes, then the weighted result into its slot
#pragma unroll
                for (int jj = 0; jj < kQ; ++jj) {
                    if ((uint32_t)jj < nrun) {
                        F3 r = f3(wave_total(c[jj].x), wave_total(c[jj].y), wave_total(c[jj].z));
                        // every lane holds the same total: the query's own lane weights and writes it
                        if (lane == s0 + jj) {
                            if (kFast && ((fastmask >> lane) & 1u)) r = gq_lds[wv][lane].mc.diffuse * r;
                            const F3 o = tpn * r;
                            a.contrib[3 * (size_t)slot] = o.x;
                            a.contrib[3 * (size_t)slot + 1] = o.y;
                            a.contrib[3 * (size_t)slot + 2] = o.z;
                            if (kCount) cnt[C_GATHERS]++;
                        }
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (!dyn) {
            k += local_waves;
        } else if (--left == 0) {
            k = take_groups();
            left = grab;
        } else {
            ++k;
        }
    }
    if (kCount) flush_counters(a.counters, cnt);
}

// One lane per query (group_queries = 1), for sparse photon maps: at a few
// tens of photons per query window the wave-per-group merge spends most of
// its instructions on per-query setup, per-sub-run cell lookups and the
// lane reductions of every query's sum, whereas here a lane sets up its own
// query, walks its cell window (z, then y, then x, each cell's photons in
// map order) and adds the accepted photons' terms in that order -- no
// reduction, no LDS.  Lanes take consecutive queries of the spatial order,
// so a wave's windows overlap and its photon loads hit the same lines.
// The radius test is the one of the other merges (same sets); the sum of a
// query is the same terms in another order.
template <bool kCount, bool kFast, bool kBatch>
__global__ __launch_bounds__(kGatherBlock) void k_gather_lane(CameraPassArgs a) {
    unsigned long long cnt[kCounterFields] = {0};
    PhotonMapView pm0 = a.pm;
    if (a.dyn) {
        pm0.r2 = a.dyn->r2;
        pm0.r_query = a.dyn->r_query;
        pm0.inv_cell = a.dyn->inv_cell;
        pm0.origin[0] = a.dyn->origin[0];
        pm0.origin[1] = a.dyn->origin[1];
        pm0.origin[2] = a.dyn->origin[2];
    }
    const uint32_t nq = min(*(volatile uint32_t*)(a.path_counter + 1), a.q.capacity);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nq; j += gridDim.x * blockDim.x) {
        const uint32_t qi = a.q_order ? a.q_order[j] : j;
        const float4 qp = a.q.pos[qi], qd = a.q.dir[qi], qs = a.q.sn[qi], qg = a.q.gn[qi], t = a.q.tpn[qi];
        const uint32_t slot = __float_as_uint(qs.w);
        const int scene = kBatch ? __float_as_int(qg.w) : 0;
        const PhotonMapView& pm = kBatch ? a.batch.pm[scene] : pm0;
        const DevMaterial* mats = kBatch ? a.batch.sc[scene].mats : a.sc.mats;
        GroupQuery g;
        g.q = f3(qp.x, qp.y, qp.z);
        g.cdvm = qp.w;
        bsdf_setup(g.b, mats[__float_as_int(qd.w)], f3(qs.x, qs.y, qs.z), f3(qd.x, qd.y, qd.z), true);
        g.mc = merge_const(g.b, f3(qg.x, qg.y, qg.z));
        const bool fast = kFast && g.b.kind == BK_FMG && g.mc.f == 0.0f;
        DiffuseQuery d;
        d.qk = make_float4(g.q.x, g.q.y, g.q.z, 1.0f + g.cdvm * g.mc.rd);
        d.ns = make_float4(g.b.fr.z.x, g.b.fr.z.y, g.b.fr.z.z, g.mc.side_wo);
        d.gw = make_float4(g.mc.gn.x, g.mc.gn.y, g.mc.gn.z, g.b.wo.z);
        int lo[3], hi[3];
        query_window(pm, g.q, lo, hi);
        F3 c = f3s(0.0f);
        // at most 4 cells per axis (cell >= 2/3 of the query radius): their buckets differ (cell_bucket)
        for (int cz = lo[2]; cz <= hi[2]; ++cz)
            for (int cy = lo[1]; cy <= hi[1]; ++cy)
                for (int cx = lo[0]; cx <= hi[0]; ++cx) {
                    const uint2 range = pm.cells[cell_bucket(cx, cy, cz, pm.mask)];
                    if (kCount) {
                        cnt[C_CELLS]++;
                        cnt[C_TESTED] += range.y - range.x;
                        cnt[C_STREAMED] += range.y - range.x;
                    }
                    auto photon = [&](uint32_t i, const float4& p) {
                        if (!(len2(g.q - f3(p.x, p.y, p.z)) < pm.r2)) return;
                        if (kCount) {
                            cnt[C_ACCEPTED]++;
                            if (a.merge_sets) record_merge(a, slot, i);
                        }
                        const float4 dt = pm.dir_tr[i];
                        const float2 gb = pm.tp_gb[i];
                        if (fast) {
                            add_diffuse_term(c, d, true, dt, gb, p.w * kInvPI);
                        } else {
                            float fwd, rev;
                            const F3 e = bsdf_eval_pdf<kFast>(g.b, g.mc, f3(dt.x, dt.y, dt.z), fwd, rev);
                            const float w = mdiv<kFast>(1.0f, 1.0f + g.cdvm * rev + p.w * fwd);
                            c = c + (e * f3(dt.w, gb.x, gb.y)) * w;
                        }
                    };
                    // four positions in flight at a time; the terms are still added in map order
                    uint32_t i = range.x;
                    for (; i + 4u <= range.y; i += 4u) {
                        const float4 p0 = pm.pos[i], p1 = pm.pos[i + 1], p2 = pm.pos[i + 2], p3 = pm.pos[i + 3];
                        photon(i, p0);
                        photon(i + 1, p1);
                        photon(i + 2, p2);
                        photon(i + 3, p3);
                    }
                    for (; i < range.y; ++i) photon(i, pm.pos[i]);
                }
        const F3 r = fast ? g.mc.diffuse * c : c;
        const F3 o = f3(t.x, t.y, t.z) * r;
        a.contrib[3 * (size_t)slot] = o.x;
        a.contrib[3 * (size_t)slot + 1] = o.y;
        a.contrib[3 * (size_t)slot + 2] = o.z;
        if (kCount) cnt[C_GATHERS]++;
    }
    if (kCount) flush_counters(a.counters, cnt);
}

// Camera pass, part 3: color = sum of the path's vertex contributions in
// vertex order, added to the film (Vertex_Merging.cpp:347, :361).
__global__ void k_resolve(CameraPassArgs a) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_paths) return;
    int2 info = a.path_info[p];
    if (info.x < 0) return;
    float* image = a.batch.n > 0 ? a.batch.bs[batch_scene_of(a.batch.camera_starts, a.batch.n, p)].image : a.image;
    F3 color = f3s(0.0f);
    uint32_t mask = (uint32_t)info.y;
    while (mask) {
        int d = __builtin_ctz(mask);
        mask &= mask - 1;
        size_t s = ((size_t)p * 8 + d) * 3;
        color = color + f3(a.contrib[s], a.contrib[s + 1], a.contrib[s + 2]);
    }
    if (color.x != 0.0f || color.y != 0.0f || color.z != 0.0f || (uint32_t)info.y != 0) {
        unsafeAtomicAdd(&image[3 * (size_t)info.x], color.x);
        unsafeAtomicAdd(&image[3 * (size_t)info.x + 1], color.y);
        unsafeAtomicAdd(&image[3 * (size_t)info.x + 2], color.z);
    }
}

void launch_camera_trace(const CameraPassArgs& a, int blocks, hipStream_t s) {
    if (a.batch.n > 0) {
        if (a.counters) hipLaunchKernelGGL((k_camera_trace<true, true>), dim3(blocks), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_camera_trace<false, true>), dim3(blocks), dim3(kBlock), 0, s, a);
    } else {
        if (a.counters) hipLaunchKernelGGL((k_camera_trace<true, false>), dim3(blocks), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_camera_trace<false, false>), dim3(blocks), dim3(kBlock), 0, s, a);
    }
}
template <bool kBatch, int kQ, int kB>
static void grouped_variant(const CameraPassArgs& a, int blocks, hipStream_t s) {
    dim3 g(blocks), b(kGatherBlock);
    if (a.exact_merge) {
        if (a.counters) hipLaunchKernelGGL((k_gather_grouped<true, false, kBatch, kQ, kB>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_gather_grouped<false, false, kBatch, kQ, kB>), g, b, 0, s, a);
    } else {
        if (a.counters) hipLaunchKernelGGL((k_gather_grouped<true, true, kBatch, kQ, kB>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_gather_grouped<false, true, kBatch, kQ, kB>), g, b, 0, s, a);
    }
}
template <bool kBatch, int kQ>
static void blocked_variant(const CameraPassArgs& a, int blocks, hipStream_t s) {
    if (a.gather_block == 4) grouped_variant<kBatch, kQ, 4>(a, blocks, s);
    else if (a.gather_block == 2) grouped_variant<kBatch, kQ, 2>(a, blocks, s);
    else if (a.gather_block == 1) grouped_variant<kBatch, kQ, 1>(a, blocks, s);
    else grouped_variant<kBatch, kQ, 0>(a, blocks, s);
}
template <bool kBatch>
static void gather_variant(const CameraPassArgs& a, int blocks, hipStream_t s) {
    dim3 g(blocks), b(kGatherBlock);
    if (a.group_queries == 1) {
        if (a.exact_merge) {
            if (a.counters) hipLaunchKernelGGL((k_gather_lane<true, false, kBatch>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_gather_lane<false, false, kBatch>), g, b, 0, s, a);
        } else {
            if (a.counters) hipLaunchKernelGGL((k_gather_lane<true, true, kBatch>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_gather_lane<false, true, kBatch>), g, b, 0, s, a);
        }
        return;
    }
    if (a.group_queries == 4) blocked_variant<kBatch, 4>(a, blocks, s);
    else if (a.group_queries == 16) blocked_variant<kBatch, 16>(a, blocks, s);
    else if (a.group_queries) blocked_variant<kBatch, 8>(a, blocks, s);
    if (a.group_queries) return;
    if (a.exact_merge) {
        if (a.counters) hipLaunchKernelGGL((k_gather<true, false, kBatch>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_gather<false, false, kBatch>), g, b, 0, s, a);
    } else {
        if (a.counters) hipLaunchKernelGGL((k_gather<true, true, kBatch>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_gather<false, true, kBatch>), g, b, 0, s, a);
    }
}
void launch_gather(const CameraPassArgs& a, int blocks, hipStream_t s) {
    if (a.batch.n > 0) gather_variant<true>(a, blocks, s);
    else gather_variant<false>(a, blocks, s);
}
void launch_resolve(const CameraPassArgs& a, hipStream_t s) {
    if (a.n_paths == 0) return;
    hipLaunchKernelGGL(k_resolve, dim3((a.n_paths + 255) / 256), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------
// Debug AOVs (Vertex_Merging::run_debug_normals / _facing / _materials,
// src/Debug_scene.cpp:21-119, 156-205): one unjittered camera ray per
// pixel from the camera position with tmin 1e-4; mode 2 the interpolated
// shading normal, 3 blue (front) or red (back) times |d . n|, 4 the
// dielectric Fresnel reflectance of the view direction for IOR `ior`.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_debug_aov(SceneView sc, int mode, float ior, float* image) {
    __shared__ int2 stack_lds[kLdsStack * kBlock];
    const int W = sc.W, H = sc.H;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= W * H) return;
    float sx = (float)(p % W), sy = (float)(p / W);
    F3 dir = (ld3(sc.cam_bl) + ld3(sc.cam_h) * (sx / (float)W)) + ld3(sc.cam_v) * (sy / (float)H);
    dir = dir - ld3(sc.cam_o);
    F3 o = ld3(sc.cam_o), d = normalize(dir);
    Hit h;
    uint32_t n_nodes = 0, n_tris = 0;
    if (!trace_closest<false>(sc, o, d, 0.0001f, h, stack_lds + threadIdx.x, n_nodes, n_tris)) return;
    Surface sf = surface_at(sc, h, o, d);
    F3 c;
    if (mode == 2) {
        c = sf.sn;
    } else if (mode == 3) {
        bool front = dot(sf.sn, d) < 0.0f;
        float factor = fabsf(dot(d, sf.sn));
        c = (front ? f3(0.0f, 0.0f, 1.0f) : f3(1.0f, 0.0f, 0.0f)) * factor;
    } else {
        Frame fr = frame_from_z(sf.sn);
        F3 wl = to_local(fr, -d);
        float eta = wl.z > 0.0f ? 1.0f / ior : ior;
        c = f3s(fresnel_eta(fabsf(wl.z), eta));
    }
    // Image_Buffer::add_color (Image_Buffer.h:24-37); one camera sample per pixel
    int X = (int)sx;
    int Y = sc.compat_rows ? H - (int)sy : H - 1 - (int)sy;
    long long at = (long long)X + (long long)Y * W;
    if (at < 0 || at > (long long)W * H - 1) return;
    image[3 * at] += c.x;
    image[3 * at + 1] += c.y;
    image[3 * at + 2] += c.z;
}

void launch_debug_aov(const SceneView& sc, int mode, float ior, float* image, hipStream_t s) {
    int n = sc.W * sc.H;
    if (n > 0) hipLaunchKernelGGL(k_debug_aov, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, sc, mode, ior, image);
}

// ---------------------------------------------------------------------
// First hit (parity diagnostic)
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_first_hit(SceneView sc, const float* rays, int n, float tmin, int* tri,
                                                      float* t) {
    __shared__ int2 stack_lds[kLdsStack * kBlock];
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
__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}
void launch_copy_u32(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s) {
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
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
// The sort (radix_sort_pairs): one-sweep LSD radix sort of (key, value)
// pairs on 8-bit digits (Adinets and Merrill, "Onesweep: A Faster Least
// Significant Digit Radix Sort for GPUs", 2022).  The digit counts of the
// first pass come from a counting kernel (or the kernel that makes the keys);
// then each pass is one kernel: a workgroup takes the next tile by ticket,
// ranks its keys by digit (stable: wave-ordered chunks, ballot matches within
// a wave), counts the next pass's digits, publishes its digit counts and finds
// the counts of the tiles before it by decoupled look-back, reorders the tile
// by digit in LDS and writes each digit's run contiguously.  A tile is 16
// keys per thread of a workgroup of 256, 512 or 1024 threads (by key count):
// a 4 M-key pass is 256 tiles of 16,384 keys (one 1024-thread workgroup per
// CU, 150 KB of LDS) in one round, with short look-back chains.
// Scratch (u32 units): two look-back status arrays of tiles x 256 64-bit
// words (a pass reads one and clears the other for the next pass), the digit
// counts of 8 passes (each in kOsCopies copies, added to by workgroup index
// modulo kOsCopies, so that no count takes more than a few hundred atomics),
// and the tickets.
constexpr int kOsItems = 16;
constexpr uint32_t kOsChunk = 64 * kOsItems;  // keys per wave, in index order
constexpr uint32_t kOsMinTile = 256 * kOsItems;
constexpr int kOsMaxPasses = 8;
constexpr int kOsCopies = kDigitCopies;

// Tiles of the smallest tile size (the scratch holds that many status rows).
__host__ __device__ constexpr size_t onesweep_tiles(size_t n) { return n == 0 ? 1 : (n + kOsMinTile - 1) / kOsMinTile; }
// Threads per workgroup of the passes over n keys (the fastest of 256, 512 and
// 1024 measured at 0.5, 1, 2, 4 and 8 M keys: one round of about 256 tiles of
// 16,384 keys is best near 4 M keys, 8,192-key tiles above and below).
static int onesweep_threads(uint32_t n) {
    return n < (3u << 18) ? 256 : n < (3u << 20) ? 512 : n < (6u << 20) ? 1024 : 512;
}
__host__ __device__ constexpr size_t onesweep_scratch_u32(size_t n) {
    return 2 * (2 * 256 * onesweep_tiles(n)) + kOsMaxPasses * kOsCopies * 256 + 16;
}

size_t photon_map_hist_entries(uint32_t max_count) { return onesweep_scratch_u32(max_count); }
size_t photon_map_sum_entries(uint32_t) { return 1; }  // the one-sweep sort needs no scan partials

// Element count of the map kernels: n, or the count in device memory (at most n) when n_dev is given.
__device__ __forceinline__ uint32_t live_count(const uint32_t* n_dev, uint32_t n) { return n_dev ? min(*n_dev, n) : n; }

// Pointers into the sort's scratch for up to n keys.
struct OnesweepScratch {
    unsigned long long* status[2];
    uint32_t* dig;      // kOsMaxPasses x kOsCopies x 256 digit counts
    uint32_t* tickets;  // one per pass, then the failure flag
    size_t tiles;
};
static OnesweepScratch onesweep_scratch(uint32_t* hist, uint32_t n) {
    OnesweepScratch o;
    o.tiles = onesweep_tiles(n);
    o.status[0] = reinterpret_cast<unsigned long long*>(hist);
    o.status[1] = o.status[0] + 256 * o.tiles;
    o.dig = hist + 2 * (2 * 256 * o.tiles);
    o.tickets = o.dig + kOsMaxPasses * kOsCopies * 256;
    return o;
}
// Zeroes the digit counts, the tickets, the flag and the first pass's status rows.
static void onesweep_clear(const OnesweepScratch& o, hipStream_t s) {
    launch_fill_u32(o.dig, 0u, kOsMaxPasses * kOsCopies * 256 + 16, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(o.status[0]), 0u, 2 * 256 * o.tiles, s);
}
void clear_map_sort(uint32_t* hist, uint32_t n, hipStream_t s) { onesweep_clear(onesweep_scratch(hist, n), s); }
uint32_t* map_sort_digits(uint32_t* hist, uint32_t n) { return onesweep_scratch(hist, n).dig; }

// Keys of the photon map, (bucket, light path), and the counts of their low
// digits (the sort's first pass): 16 photons per thread, all record loads
// issued first, one LDS count per photon and one global add per digit and
// workgroup.
constexpr int kKeyThreads = 256, kKeyItems = 16;
__device__ __forceinline__ void flush_digit_counts(const uint32_t* h, uint32_t* dig) {
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&dig[(blockIdx.x % kOsCopies) * 256 + threadIdx.x], h[threadIdx.x]);
}
__global__ __launch_bounds__(kKeyThreads) void k_map_keys(const float4* rec, const uint32_t* n_dev, uint32_t n,
                                                          const IterParams* dyn, float ox, float oy, float oz,
                                                          float inv_cell, uint32_t mask, int uid_bits,
                                                          unsigned long long* keys, uint32_t* vals, uint32_t* dig) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    n = live_count(n_dev, n);
    if (dyn) {
        ox = dyn->origin[0];
        oy = dyn->origin[1];
        oz = dyn->origin[2];
        inv_cell = dyn->inv_cell;
    }
    const uint32_t i0 = blockIdx.x * (kKeyThreads * kKeyItems) + threadIdx.x;
    float4 p[kKeyItems];
    uint32_t path[kKeyItems];
#pragma unroll
    for (int j = 0; j < kKeyItems; ++j) {
        const uint32_t i = i0 + j * kKeyThreads;
        p[j] = i < n ? rec[kPhotonRecord * (size_t)i] : make_float4(0.f, 0.f, 0.f, 0.f);
        path[j] = i < n ? __float_as_uint(rec[kPhotonRecord * (size_t)i + 2].z) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kKeyItems; ++j) {
        const uint32_t i = i0 + j * kKeyThreads;
        if (i < n) {
            int cx = (int)floorf((p[j].x - ox) * inv_cell);
            int cy = (int)floorf((p[j].y - oy) * inv_cell);
            int cz = (int)floorf((p[j].z - oz) * inv_cell);
            uint32_t bkt = cell_bucket(cx, cy, cz, mask);
            const unsigned long long k = ((unsigned long long)bkt << uid_bits) | (unsigned long long)path[j];
            keys[i] = k;
            vals[i] = i;
            atomicAdd(&h[(uint32_t)k & 255u], 1u);
        }
    }
    flush_digit_counts(h, dig);
}

// Counts of the low digits of the keys (the sort's first pass), for keys
// made by a kernel that does not count them.
constexpr int kOsHistItems = 8;  // keys per lane loaded before they are counted
__global__ __launch_bounds__(256) void k_os_hist(const unsigned long long* keys, const uint32_t* n_dev, uint32_t n,
                                                 uint32_t* dig) {
    __shared__ uint32_t h[4][256];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    for (int w = 0; w < 4; ++w) h[w][threadIdx.x] = 0;
    __syncthreads();
    n = live_count(n_dev, n);
    // each wave takes chunks of 64 x kOsHistItems keys, all loads issued before the counting;
    // a digit shared by the whole wave is added once by the first lane
    constexpr uint32_t chunk = 64u * kOsHistItems;
    for (uint32_t c0 = (blockIdx.x * 4 + wave) * chunk; c0 < n; c0 += gridDim.x * 4 * chunk) {
        unsigned long long k[kOsHistItems];
#pragma unroll
        for (int j = 0; j < kOsHistItems; ++j) {
            const uint32_t i = c0 + j * 64 + lane;
            k[j] = i < n ? keys[i] : 0ull;
        }
#pragma unroll
        for (int j = 0; j < kOsHistItems; ++j) {
            const bool valid = c0 + j * 64 + lane < n;
            const uint64_t live = __ballot(valid);
            if (live == 0) break;
            const uint32_t d = (uint32_t)k[j] & 255u;
            const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
            if (__ballot(valid && d == d0) == live) {
                if (lane == 0) h[wave][d0] += (uint32_t)__builtin_popcountll(live);
            } else if (valid) {
                atomicAdd(&h[wave][d], 1u);
            }
        }
    }
    __syncthreads();
    const uint32_t c = h[0][threadIdx.x] + h[1][threadIdx.x] + h[2][threadIdx.x] + h[3][threadIdx.x];
    if (c) atomicAdd(&dig[(blockIdx.x % kOsCopies) * 256 + threadIdx.x], c);
}

// Look-back status word of (tile, digit): flag in the high half (1: the
// tile's own count, 2: the inclusive count of every tile up to it), count in
// the low half; 0 = not yet published.
constexpr unsigned long long kOsAggregate = 1ull << 32, kOsInclusive = 2ull << 32;
// Polls of a predecessor's status before a tile gives up (about 2^24 short
// sleeps): a bound on the wait.  A tile that gives up flags the failure, and
// every write of the pass stays below n.
constexpr uint32_t kOsSpinLimit = 1u << 24;
constexpr int kOsLookback = 4;  // predecessors read per look-back round trip

// One pass: digit (key >> shift) & 255 of keys of type KIn, whose counts over
// all keys are the sums of the kOsCopies copies of dig[0..255]; the keys are written as KOut(key >> drop) (drop:
// low bits no later pass or caller reads; 32-bit keys halve the pass's
// bytes), and the counts of the next pass's digit, (key >> (shift + 8)) & 255,
// are added to dig_next (when given).  The tile is reordered in LDS through
// two 32-bit words per key: first the output index and the value, then the
// output key.
template <int kOsThreads, typename KIn, typename KOut>
__global__ __launch_bounds__(kOsThreads) void k_os_pass(const KIn* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                        const uint32_t* n_dev, uint32_t n, int shift, int drop,
                                                        const uint32_t* __restrict__ dig, uint32_t* dig_next,
                                                        unsigned long long* status, unsigned long long* status_next,
                                                        uint32_t* ticket, uint32_t* fail, KOut* __restrict__ kout,
                                                        uint32_t* __restrict__ vout) {
    constexpr uint32_t kOsTile = kOsThreads * kOsItems;
    constexpr int kOsWaves = kOsThreads / 64;
    __shared__ uint32_t s_a[kOsTile];
    __shared__ uint32_t s_b[kOsTile];
    __shared__ uint32_t s_cnt[kOsWaves][256];  // per wave: digit counts, then the waves-before prefix
    __shared__ uint32_t s_excl[256];           // tile-exclusive digit offsets
    __shared__ uint32_t s_glob[256];           // scan of the global counts, then the output index of position 0
    __shared__ uint32_t s_next[256];           // the tile's counts of the next pass's digits
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const bool digit_thread = tid < 256;  // thread tid < 256 handles digit tid
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    for (uint32_t i = tid; i < kOsWaves * 256; i += kOsThreads) s_cnt[i >> 8][i & 255u] = 0;
    if (digit_thread) s_next[tid] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    n = live_count(n_dev, n);
    const uint32_t base = tile * kOsTile;
    if (base >= n) return;  // no keys: no tile holding keys reads this one's status
    if (digit_thread) status_next[(size_t)tile * 256 + tid] = 0ull;  // read after this kernel ends
    // 1. ranks within the wave's chunk: kOsItems rounds of 64 keys in index order
    KIn key[kOsItems];
    uint32_t val[kOsItems], at[kOsItems];
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {  // every load issued before the ranking's wave barriers
        const uint32_t i = base + wave * kOsChunk + j * 64 + lane;
        key[j] = i < n ? kin[i] : (KIn)0;
        val[j] = i < n ? vin[i] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        const bool valid = base + wave * kOsChunk + j * 64 + lane < n;
        const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool b = (d >> bit) & 1u;
            const uint64_t bb = __ballot(valid && b);
            peers &= b ? bb : ~bb;
        }
        const uint32_t r = lanes_below(peers);
        const uint32_t before = valid ? s_cnt[wave][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && r == 0) s_cnt[wave][d] = before + (uint32_t)__builtin_popcountll(peers);
        __builtin_amdgcn_wave_barrier();
        at[j] = before + r;  // rank among the wave's keys of digit d
        if (dig_next && valid) atomicAdd(&s_next[(uint32_t)(key[j] >> (shift + 8)) & 255u], 1u);
    }
    __syncthreads();
    // 2. digit tid: the tile's count, the waves-before prefixes; scans of the tile's counts and of the global counts
    uint32_t h = 0, g = 0;
    if (digit_thread) {
#pragma unroll
        for (int w = 0; w < kOsWaves; ++w) {
            const uint32_t c = s_cnt[w][tid];
            s_cnt[w][tid] = h;
            h += c;
        }
#pragma unroll
        for (int c = 0; c < kOsCopies; ++c) g += dig[c * 256 + tid];
        s_excl[tid] = h;
        s_glob[tid] = g;
    }
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        uint32_t a = 0, b = 0;
        if (digit_thread && tid >= off) {
            a = s_excl[tid - off];
            b = s_glob[tid - off];
        }
        __syncthreads();
        if (digit_thread) {
            s_excl[tid] += a;
            s_glob[tid] += b;
        }
        __syncthreads();
    }
    uint32_t excl = 0, gexcl = 0;
    if (digit_thread) {
        excl = s_excl[tid] - h;
        gexcl = s_glob[tid] - g;
    }
    __syncthreads();
    if (digit_thread) {
        s_excl[tid] = excl;
        if (dig_next && s_next[tid]) atomicAdd(&dig_next[(tile % kOsCopies) * 256 + tid], s_next[tid]);
        // 3. decoupled look-back for digit tid
        unsigned long long* st = status + (size_t)tile * 256 + tid;
        __hip_atomic_store(st, (tile == 0 ? kOsInclusive : kOsAggregate) | h, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        uint32_t prefix = 0;
        bool gave_up = false;
        if (tile > 0) {
            // a window of kOsLookback predecessors per round trip
            int j = (int)tile - 1;
            uint32_t spins = 0;
            bool done = false;
            while (!done) {
                unsigned long long w[kOsLookback];
#pragma unroll
                for (int k = 0; k < kOsLookback; ++k)
                    w[k] = j - k >= 0 ? __hip_atomic_load(status + (size_t)(j - k) * 256 + tid, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : kOsInclusive;  // below tile 0: nothing more to add
                int k = 0;
                for (; k < kOsLookback; ++k) {
                    if ((w[k] >> 32) == 0ull) break;  // not yet published: read again from here
                    prefix += (uint32_t)w[k];
                    if ((w[k] >> 32) == 2ull) {
                        done = true;
                        break;
                    }
                }
                if (done) break;
                j -= k;
                if (k < kOsLookback) {
                    if (++spins >= kOsSpinLimit) {
                        gave_up = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __hip_atomic_store(st, kOsInclusive | (prefix + h), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (gave_up) atomicOr(fail, 1u);
        s_glob[tid] = gexcl + prefix - excl;
    }
    __syncthreads();
    // 4. the tile in digit order in LDS (position at = digit offset + waves before + rank): first each key's
    // output index and value, then its output key; each digit's run is written contiguously
    const uint32_t cnt = min(kOsTile, n - base);
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        if (base + wave * kOsChunk + j * 64 + lane < n) {
            const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
            at[j] += s_excl[d] + s_cnt[wave][d];
            s_a[at[j]] = s_glob[d] + at[j];
            s_b[at[j]] = val[j];
        }
    }
    __syncthreads();
    uint32_t out[kOsItems];
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        const uint32_t e = j * kOsThreads + tid;
        out[j] = e < cnt ? s_a[e] : n;
        if (out[j] < n) vout[out[j]] = s_b[e];
    }
    if (kout == nullptr) return;  // the caller reads no keys after this pass (uniform over the workgroup)
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        if (base + wave * kOsChunk + j * 64 + lane < n) {
            const KOut k = (KOut)(key[j] >> drop);
            s_a[at[j]] = (uint32_t)k;
            if (sizeof(KOut) == 8) s_b[at[j]] = (uint32_t)((unsigned long long)k >> 32);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        if (out[j] < n) {
            const uint32_t e = j * kOsThreads + tid;
            if (sizeof(KOut) == 8)
                kout[out[j]] = (KOut)((unsigned long long)s_a[e] | ((unsigned long long)s_b[e] << 32));
            else
                kout[out[j]] = (KOut)s_a[e];
        }
    }
}

__global__ void k_map_permute(const float4* rec, const uint32_t* vals, const uint32_t* n_dev, uint32_t n,
                              float4* opos, float4* odir, float2* otp, uint32_t* oidx) {
    n = live_count(n_dev, n);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = vals[i];
    const float4* r = rec + kPhotonRecord * (size_t)j;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2];
    opos[i] = r0;
    odir[i] = r1;
    otp[i] = make_float2(r2.x, r2.y);
    oidx[i] = j;
}

// Bucket ranges from the sorted keys; the bucket is key >> shift.
template <typename K>
__global__ void k_map_cells(const K* keys, const uint32_t* n_dev, uint32_t n, int shift, uint2* cells) {
    n = live_count(n_dev, n);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t b = (uint32_t)(keys[i] >> shift);
    if (i == 0 || (uint32_t)(keys[i - 1] >> shift) != b) cells[b].x = i;
    if (i == n - 1 || (uint32_t)(keys[i + 1] >> shift) != b) cells[b].y = i + 1;
}

__global__ __launch_bounds__(kKeyThreads) void k_map_keys_batch(const float4* rec, const uint32_t* n_dev, uint32_t n,
                                                                BatchView b, int uid_bits, unsigned long long* keys,
                                                                uint32_t* vals, uint32_t* dig) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    n = live_count(n_dev, n);
    const uint32_t i0 = blockIdx.x * (kKeyThreads * kKeyItems) + threadIdx.x;
    float4 p[kKeyItems];
    uint32_t path[kKeyItems];
#pragma unroll
    for (int j = 0; j < kKeyItems; ++j) {
        const uint32_t i = i0 + j * kKeyThreads;
        p[j] = i < n ? rec[kPhotonRecord * (size_t)i] : make_float4(0.f, 0.f, 0.f, 0.f);
        path[j] = i < n ? __float_as_uint(rec[kPhotonRecord * (size_t)i + 2].z) : 0u;
    }
    __syncthreads();
    for (int j = 0; j < kKeyItems; ++j) {
        const uint32_t i = i0 + j * kKeyThreads;
        if (i < n) {
            const uint32_t u = path[j];
            int sc = batch_scene_of(b.light_starts, b.n, u);
            const PhotonMapView& pm = b.pm[sc];
            int cx = (int)floorf((p[j].x - pm.origin[0]) * pm.inv_cell);
            int cy = (int)floorf((p[j].y - pm.origin[1]) * pm.inv_cell);
            int cz = (int)floorf((p[j].z - pm.origin[2]) * pm.inv_cell);
            uint32_t bkt = b.bs[sc].bucket_base + cell_bucket(cx, cy, cz, pm.mask);
            const unsigned long long k = ((unsigned long long)bkt << uid_bits) | (unsigned long long)u;
            keys[i] = k;
            vals[i] = i;
            atomicAdd(&h[(uint32_t)k & 255u], 1u);
        }
    }
    flush_digit_counts(h, dig);
}

static void sort_permute_cells(const MapBuildArgs& a, const OnesweepScratch& o, hipStream_t s);
static SortedPairs onesweep_passes(unsigned long long* keys_a, unsigned long long* keys_b, uint32_t* vals_a,
                                   uint32_t* vals_b, const OnesweepScratch& o, const uint32_t* count_dev, uint32_t n,
                                   int bits, int droppable, hipStream_t s);

static unsigned map_key_blocks(uint32_t n) { return (n + kKeyThreads * kKeyItems - 1) / (kKeyThreads * kKeyItems); }

void build_photon_map(const MapBuildArgs& a, hipStream_t s) {
    launch_fill_u32(reinterpret_cast<uint32_t*>(a.cells), 0u, 2 * ((size_t)a.mask + 1), s);
    uint32_t n = a.count;
    if (n == 0) return;
    const OnesweepScratch o = onesweep_scratch(a.hist, a.scratch_n ? a.scratch_n : n);
    if (!a.keys_ready) {
        onesweep_clear(o, s);
        hipLaunchKernelGGL(k_map_keys, dim3(map_key_blocks(n)), dim3(kKeyThreads), 0, s, a.in_rec, a.count_dev, n,
                           a.dyn, a.origin[0], a.origin[1], a.origin[2], a.inv_cell, a.mask, a.uid_bits, a.keys_a,
                           a.vals_a, o.dig);
    }
    sort_permute_cells(a, o, s);
}

// a.mask + 1 = the total bucket count of all scenes.
void build_photon_map_batch(const MapBuildArgs& a, const BatchView& b, hipStream_t s) {
    launch_fill_u32(reinterpret_cast<uint32_t*>(a.cells), 0u, 2 * ((size_t)a.mask + 1), s);
    uint32_t n = a.count;
    if (n == 0) return;
    const OnesweepScratch o = onesweep_scratch(a.hist, a.scratch_n ? a.scratch_n : n);
    if (!a.keys_ready) {
        onesweep_clear(o, s);
        hipLaunchKernelGGL(k_map_keys_batch, dim3(map_key_blocks(n)), dim3(kKeyThreads), 0, s, a.in_rec, a.count_dev,
                           n, b, a.uid_bits, a.keys_a, a.vals_a, o.dig);
    }
    sort_permute_cells(a, o, s);
}

template <int T>
static void launch_os_pass(bool in32, bool out32, dim3 g, dim3 b, hipStream_t s, const void* kin, const uint32_t* vin,
                           const uint32_t* count_dev, uint32_t n, int shift, int drop, const uint32_t* dp,
                           uint32_t* dnext, unsigned long long* st, unsigned long long* sn, uint32_t* tk, uint32_t* fl,
                           void* ko, uint32_t* vout) {
    using u64 = unsigned long long;
    if (!in32 && !out32)
        hipLaunchKernelGGL((k_os_pass<T, u64, u64>), g, b, 0, s, (const u64*)kin, vin, count_dev, n, shift, drop, dp,
                           dnext, st, sn, tk, fl, (u64*)ko, vout);
    else if (!in32)
        hipLaunchKernelGGL((k_os_pass<T, u64, uint32_t>), g, b, 0, s, (const u64*)kin, vin, count_dev, n, shift, drop,
                           dp, dnext, st, sn, tk, fl, (uint32_t*)ko, vout);
    else
        hipLaunchKernelGGL((k_os_pass<T, uint32_t, uint32_t>), g, b, 0, s, (const uint32_t*)kin, vin, count_dev, n,
                           shift, drop, dp, dnext, st, sn, tk, fl, (uint32_t*)ko, vout);
}

// LSD radix sort of (key, value) pairs on the low `bits` key bits (8-bit
// digits, stable), the first pass's digit counts already in o.dig.  After
// each pass the low key bits already sorted are dropped from the keys
// written, up to `droppable` of them (bits the caller does not read), and
// once the rest fits 32 bits the keys travel as 32-bit words; with droppable
// >= bits the last pass writes no keys.
static SortedPairs onesweep_passes(unsigned long long* keys_a, unsigned long long* keys_b, uint32_t* vals_a,
                                   uint32_t* vals_b, const OnesweepScratch& o, const uint32_t* count_dev, uint32_t n,
                                   int bits, int droppable, hipStream_t s) {
    const int passes = std::min(kOsMaxPasses, (bits + 7) / 8);
    void* kin = keys_a;
    void* kout = keys_b;
    uint32_t* vin = vals_a;
    uint32_t* vout = vals_b;
    int dropped = 0;
    bool in32 = false;
    using u64 = unsigned long long;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p - dropped;
        const int nd = std::min(8 * (p + 1), std::max(droppable, dropped));
        const bool out32 = bits - nd <= 32;
        const int drop = nd - dropped;
        void* ko = (p == passes - 1 && droppable >= bits) ? nullptr : kout;
        uint32_t* dnext = p + 1 < passes ? o.dig + kOsCopies * 256 * (p + 1) : nullptr;
        const uint32_t* dp = o.dig + kOsCopies * 256 * p;
        u64* st = o.status[p & 1];
        u64* sn = o.status[(p + 1) & 1];
        uint32_t* tk = o.tickets + p;
        uint32_t* fl = o.tickets + kOsMaxPasses;
        const int threads = onesweep_threads(n);
        const dim3 g((unsigned)((n + threads * kOsItems - 1) / (threads * kOsItems))), b(threads);
        if (threads == 1024)
            launch_os_pass<1024>(in32, out32, g, b, s, kin, vin, count_dev, n, shift, drop, dp, dnext, st, sn, tk, fl,
                                 ko, vout);
        else if (threads == 512)
            launch_os_pass<512>(in32, out32, g, b, s, kin, vin, count_dev, n, shift, drop, dp, dnext, st, sn, tk, fl,
                                ko, vout);
        else
            launch_os_pass<256>(in32, out32, g, b, s, kin, vin, count_dev, n, shift, drop, dp, dnext, st, sn, tk, fl,
                                ko, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
        dropped = nd;
        in32 = out32;
    }
    return SortedPairs{kin, vin, in32, dropped};
}

SortedPairs radix_sort_pairs_drop(unsigned long long* keys_a, unsigned long long* keys_b, uint32_t* vals_a,
                                  uint32_t* vals_b, uint32_t* hist, const uint32_t* count_dev, uint32_t n, int bits,
                                  int droppable, hipStream_t s) {
    if (n == 0 || bits <= 0) return SortedPairs{keys_a, vals_a, false, 0};
    const OnesweepScratch o = onesweep_scratch(hist, n);
    onesweep_clear(o, s);
    const unsigned hg = (unsigned)std::min<size_t>(512, ((size_t)n + 8191) / 8192);
    hipLaunchKernelGGL(k_os_hist, dim3(hg), dim3(256), 0, s, keys_a, count_dev, n, o.dig);
    return onesweep_passes(keys_a, keys_b, vals_a, vals_b, o, count_dev, n, bits, droppable, s);
}

void radix_sort_pairs(unsigned long long* keys_a, unsigned long long* keys_b, uint32_t* vals_a, uint32_t* vals_b,
                      uint32_t* hist, uint32_t*, const uint32_t* count_dev, uint32_t n, int bits,
                      unsigned long long** keys_out, uint32_t** vals_out, hipStream_t s) {
    const SortedPairs r = radix_sort_pairs_drop(keys_a, keys_b, vals_a, vals_b, hist, count_dev, n, bits, 0, s);
    *keys_out = static_cast<unsigned long long*>(r.keys);
    *vals_out = r.vals;
}



__global__ void k_query_keys(CameraPassArgs a, uint32_t n, unsigned long long* keys, uint32_t* vals) {
    n = live_count(a.path_counter + 1, min(n, a.q.capacity));
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = a.q.pos[i];
    int scene = 0;
    PhotonMapView pm = a.pm;
    if (a.batch.n > 0) {
        scene = __float_as_int(a.q.gn[i].w);
        pm = a.batch.pm[scene];
    } else if (a.dyn) {
        pm.r_query = a.dyn->r_query;
        pm.inv_cell = a.dyn->inv_cell;
        pm.origin[0] = a.dyn->origin[0];
        pm.origin[1] = a.dyn->origin[1];
        pm.origin[2] = a.dyn->origin[2];
    }
    // the query's cell window: Morton code of its low corner (cell coordinates
    // modulo 1024, so neighbouring windows keep neighbouring codes) above the
    // window extents, so queries with the same window are adjacent
    int lo[3], hi[3];
    query_window(pm, f3(p.x, p.y, p.z), lo, hi);
    uint32_t code = (spread10((uint32_t)lo[0]) << 2) | (spread10((uint32_t)lo[1]) << 1) | spread10((uint32_t)lo[2]);
    uint32_t ext = (uint32_t)((hi[0] - lo[0]) & 3) | (uint32_t)((hi[1] - lo[1]) & 3) << 2 |
                   (uint32_t)((hi[2] - lo[2]) & 3) << 4;
    keys[i] = ((unsigned long long)(uint32_t)scene << 36) | ((unsigned long long)code << 6) | ext;
    vals[i] = i;
}

// 7 low bits of v spread to every fourth bit (4-D Morton codes).
__device__ __forceinline__ uint32_t spread7x4(uint32_t v) {
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) r |= ((v >> k) & 1u) << (4 * k);
    return r;
}

__global__ void k_light_keys(LightPassArgs a, unsigned long long* keys, uint32_t* vals) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_paths) return;
    PathRng rng{a.k0, a.k1, p, a.dyn ? a.dyn->iteration : a.iteration, STREAM_LIGHT};
    uint32_t scene = 0;
    if (a.batch.n > 0) {
        scene = (uint32_t)batch_scene_of(a.batch.light_starts, a.batch.n, p);
        const BatchScene& bs = a.batch.bs[scene];
        rng.k0 = bs.k0;
        rng.k1 = bs.k1;
        rng.iter = bs.iteration;
        rng.path = p - bs.light_base;
    }
    // the plain emission words: light pick, position (2), direction (2)
    U4 w0 = rng.block(0);
    uint32_t wy = rng.block(1).a;
    uint32_t code = (spread7x4(w0.b >> 25) << 3) | (spread7x4(w0.c >> 25) << 2) | (spread7x4(w0.d >> 25) << 1) |
                    spread7x4(wy >> 25);
    keys[p] = ((unsigned long long)scene << 32) | ((unsigned long long)(w0.a >> 28) << 28) | code;
    vals[p] = p;
}

const uint32_t* sort_light_paths(const LightPassArgs& a, const QuerySortBuffers& b, hipStream_t s) {
    const uint32_t n = a.n_paths;
    if (n == 0) return nullptr;
    int bits = 32;
    if (a.batch.n > 1) {
        int sb = 0;
        while ((1 << sb) < a.batch.n) ++sb;
        bits = 32 + sb;
    }
    hipLaunchKernelGGL(k_light_keys, dim3((n + 255) / 256), dim3(256), 0, s, a, b.keys_a, b.vals_a);
    return radix_sort_pairs_drop(b.keys_a, b.keys_b, b.vals_a, b.vals_b, b.hist, nullptr, n, bits, bits, s).vals;
}

const uint32_t* sort_queries(const CameraPassArgs& a, const QuerySortBuffers& b, uint32_t n, hipStream_t s) {
    if (n == 0) return nullptr;
    int bits = 36;
    if (a.batch.n > 1) {
        int sb = 0;
        while ((1 << sb) < a.batch.n) ++sb;
        bits = 36 + sb;
    }
    hipLaunchKernelGGL(k_query_keys, dim3((n + 255) / 256), dim3(256), 0, s, a, n, b.keys_a, b.vals_a);
    return radix_sort_pairs_drop(b.keys_a, b.keys_b, b.vals_a, b.vals_b, b.hist, a.path_counter + 1, n, bits, bits, s)
        .vals;
}

static void sort_permute_cells(const MapBuildArgs& a, const OnesweepScratch& o, hipStream_t s) {
    uint32_t n = a.count;
    unsigned g = (n + 255) / 256;
    // the light-path bits order photons within a bucket; the ranges below read only the bucket bits
    const SortedPairs r = onesweep_passes(a.keys_a, a.keys_b, a.vals_a, a.vals_b, o, a.count_dev, n,
                                          a.bucket_bits + a.uid_bits, a.uid_bits, s);
    hipLaunchKernelGGL(k_map_permute, dim3(g), dim3(256), 0, s, a.in_rec, r.vals, a.count_dev, n, a.out_pos,
                       a.out_dir, a.out_tp, a.out_index);
    const int shift = a.uid_bits - r.dropped;
    if (r.keys32)
        hipLaunchKernelGGL(k_map_cells<uint32_t>, dim3(g), dim3(256), 0, s, static_cast<const uint32_t*>(r.keys),
                           a.count_dev, n, shift, a.cells);
    else
        hipLaunchKernelGGL(k_map_cells<unsigned long long>, dim3(g), dim3(256), 0, s,
                           static_cast<const unsigned long long*>(r.keys), a.count_dev, n, shift, a.cells);
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

// ---------------------------------------------------------------------
// Radius sets without device allocation.  Every array beyond the caller's
// inputs and outputs is a slice of `scratch`, one file-scope device array
// per GPU, held until the process ends.  Slices (256-byte aligned):
// photons {pos, 0} and their indices as uids, the sort's keys and values
// (two of each), its histograms and scan partials, the sorted photons and
// their input indices, the bucket ranges, and the per-query counts.  The
// permute's direction and throughput inputs alias the position slice (the
// sets read positions only).  Calls that overlap on one GPU race on it.
// ---------------------------------------------------------------------
namespace {
constexpr size_t slice(size_t bytes) { return (bytes + 255) & ~size_t(255); }
constexpr size_t kNaP = kNoAllocMaxPhotons;
constexpr size_t kNaHist = onesweep_scratch_u32(kNaP);
constexpr size_t kNaSums = 1;
constexpr size_t kNaInRec = 0;
constexpr size_t kNaKeysA = kNaInRec + slice(16 * kPhotonRecord * kNaP);
constexpr size_t kNaKeysB = kNaKeysA + slice(8 * kNaP);
constexpr size_t kNaValsA = kNaKeysB + slice(8 * kNaP);
constexpr size_t kNaValsB = kNaValsA + slice(4 * kNaP);
constexpr size_t kNaHistAt = kNaValsB + slice(4 * kNaP);
constexpr size_t kNaSumsAt = kNaHistAt + slice(4 * kNaHist);
constexpr size_t kNaOutPos = kNaSumsAt + slice(4 * kNaSums);
constexpr size_t kNaOutDir = kNaOutPos + slice(16 * kNaP);
constexpr size_t kNaOutTp = kNaOutDir + slice(16 * kNaP);
constexpr size_t kNaOutIdx = kNaOutTp + slice(8 * kNaP);
constexpr size_t kNaCells = kNaOutIdx + slice(4 * kNaP);
constexpr size_t kNaCounts = kNaCells + slice(8 * (size_t)kNoAllocMaxBuckets);
constexpr size_t kNaBytes = kNaCounts + slice(4 * ((size_t)kNoAllocMaxQueries + 1));
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

size_t noalloc_scratch_bytes() { return kNaBytes; }

// Device address of `scratch` on the current device, looked up once per
// device (the host-side name is only a stand-in for it).
static unsigned char* scratch_address(const char** err) {
    static std::mutex mu;
    static unsigned char* addr[64] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        *err = "no HIP device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            *err = "hipGetSymbolAddress of the scratch array failed";
            return nullptr;
        }
        addr[dev] = static_cast<unsigned char*>(p);
    }
    return addr[dev];
}

__global__ void k_na_photons(const float* p, uint32_t n, float4* rec) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4* r = rec + kPhotonRecord * (size_t)i;
    r[0] = make_float4(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2], 0.0f);
    r[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r[2] = make_float4(0.0f, 0.0f, __uint_as_float(i), 0.0f);
}

// offsets[i] = counts[0] + ... + counts[i - 1] for i <= n (one workgroup:
// each thread sums a contiguous run, a scan over the runs in LDS, then each
// thread writes its run's offsets).
constexpr int kNaScanThreads = 1024;
__global__ __launch_bounds__(kNaScanThreads) void k_na_offsets(const uint32_t* counts, uint32_t n,
                                                               unsigned long long* offsets) {
    __shared__ unsigned long long run[kNaScanThreads];
    const uint32_t per = (n + kNaScanThreads - 1) / kNaScanThreads;
    const uint32_t b = threadIdx.x * per, e = min(n, b + per);
    unsigned long long sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += counts[i];
    run[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kNaScanThreads; off <<= 1) {
        unsigned long long t = threadIdx.x >= off ? run[threadIdx.x - off] : 0ull;
        __syncthreads();
        run[threadIdx.x] += t;
        __syncthreads();
    }
    unsigned long long at = run[threadIdx.x] - sum;
    for (uint32_t i = b; i < e; ++i) {
        offsets[i] = at;
        at += counts[i];
    }
    if (threadIdx.x == kNaScanThreads - 1) offsets[n] = run[kNaScanThreads - 1];
}

long long gather_sets_noalloc(const NoAllocGather& a, hipStream_t s, const char** err) {
    if (a.count > kNoAllocMaxPhotons || a.nq > kNoAllocMaxQueries || a.pm.mask + 1 > kNoAllocMaxBuckets) {
        *err = "rc_gather_sets_noalloc: size beyond the scratch array";
        return -1;
    }
    unsigned char* base = scratch_address(err);
    if (!base) return -1;
    float4* in_rec = reinterpret_cast<float4*>(base + kNaInRec);
    MapBuildArgs m{};
    m.in_rec = in_rec;
    m.count = a.count;
    for (int k = 0; k < 3; ++k) m.origin[k] = a.pm.origin[k];
    m.inv_cell = a.pm.inv_cell;
    m.mask = a.pm.mask;
    m.bucket_bits = a.bucket_bits;
    m.uid_bits = a.uid_bits;
    m.keys_a = reinterpret_cast<unsigned long long*>(base + kNaKeysA);
    m.keys_b = reinterpret_cast<unsigned long long*>(base + kNaKeysB);
    m.vals_a = reinterpret_cast<uint32_t*>(base + kNaValsA);
    m.vals_b = reinterpret_cast<uint32_t*>(base + kNaValsB);
    m.hist = reinterpret_cast<uint32_t*>(base + kNaHistAt);
    m.hist_sums = reinterpret_cast<uint32_t*>(base + kNaSumsAt);
    m.out_pos = reinterpret_cast<float4*>(base + kNaOutPos);
    m.out_dir = reinterpret_cast<float4*>(base + kNaOutDir);
    m.out_tp = reinterpret_cast<float2*>(base + kNaOutTp);
    m.out_index = reinterpret_cast<uint32_t*>(base + kNaOutIdx);
    m.cells = reinterpret_cast<uint2*>(base + kNaCells);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + kNaCounts);
    if (a.count) hipLaunchKernelGGL(k_na_photons, dim3((a.count + 255) / 256), dim3(256), 0, s, a.positions, a.count,
                                    in_rec);
    build_photon_map(m, s);  // clears every bucket range with a fill kernel first
    PhotonMapView pm = a.pm;
    pm.pos = m.out_pos;
    pm.dir_tr = m.out_dir;
    pm.tp_gb = m.out_tp;
    pm.cells = m.cells;
    launch_gather_count(pm, a.queries, (int)a.nq, m.out_index, counts, s);
    hipLaunchKernelGGL(k_na_offsets, dim3(1), dim3(kNaScanThreads), 0, s, counts, a.nq, a.offsets);
    unsigned long long total = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&total, a.offsets + a.nq, sizeof total, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        *err = "rc_gather_sets_noalloc: HIP error";
        return -1;
    }
    if ((long long)total <= a.cap && total > 0) {
        launch_gather_fill(pm, a.queries, (int)a.nq, m.out_index, a.offsets, a.indices, s);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
            *err = "rc_gather_sets_noalloc: HIP error";
            return -1;
        }
    }
    return (long long)total;
}

}  // namespace rc
