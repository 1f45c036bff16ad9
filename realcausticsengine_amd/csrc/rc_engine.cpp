// rc_engine.cpp -- host side of libRealCaustics.so: the exported C-ABI
// (init / main of src/main.cpp:89-223 plus the rc_* additions), Blender
// struct ingest, the binned-SAH BVH build, light set-up, the per-iteration
// schedule of the HIP kernels, and the EXR writer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <zlib.h>

#include "../../include/real_caustics.h"
#include "rc_engine.h"
#include "rc_guide_model.h"
#include "rc_host.h"
#include "rc_vec.h"

using namespace rc;

namespace rc {
std::string g_error;
std::mutex g_error_mutex;  // batch preparation runs scenes on several host threads
void set_error(const std::string& e) {
    std::lock_guard<std::mutex> lock(g_error_mutex);
    if (g_error.empty()) g_error = e;
}
bool has_error() {
    std::lock_guard<std::mutex> lock(g_error_mutex);
    return !g_error.empty();
}
}  // namespace rc

namespace {

std::atomic<int> g_cancel{0};

struct HostLight {
    char type;
    float color[3], position[3], width, height, spread, rotation[3], power, angle, radius;
};

struct Box {
    float lo[3], hi[3];
};

// ---------------------------------------------------------------------
// engine state (one process-wide scene, as src/main.cpp:59)
// ---------------------------------------------------------------------
struct Engine {
    bool initialised = false, prepared = false;
    std::string init_error;  // what init() rejected; reported again by main()
    int samples = 1;
    float base_radius = 0.01f, radius_alpha = 0.75f;
    int debug_test = 1;
    float debug_ior = 1.5f;  // debug_stuff.specular: the IOR of debug mode 4 (src/Debug_scene.cpp:182)
    // geometry in world space, global triangle order (mesh, then triangle)
    std::vector<float> wv;        // 3 floats per vertex (all meshes)
    std::vector<float> wn;        // 3 floats per vertex
    std::vector<uint32_t> tris;   // v0 v1 v2 material per triangle (global vertex ids)
    std::vector<uint32_t> mesh_tri_offset;
    std::vector<DevMaterial> mats;
    std::vector<HostLight> py_lights;
    int W = 0, H = 0;
    F3 cam_o, cam_bl, cam_h, cam_v;
    Box bbox;
    float scene_radius = 0.0f;
    // options
    uint64_t seed = 0;
    long long photons = 0;
    int write_exr = 1, compat_rows = 1, compat_point = 1, compat_pdf = 1, max_path_length = 6;
    int compat_glass = 1;  // 0: glass of roughness >= 0.01 is a GGX dielectric (BK_RGLASS)
    int light_emit = 0;    // 1: the light samples of every path in a coherent kernel before the light pass
    DevBuf<float4> emitted;
    long long map_ids = 0;    // > 0: photon ids of the next map are indices below this (caller photon lists)
    std::vector<rc_python_material> py_mats;  // as given to init; converted again at rc_prepare
    double photon_capacity = 0, cell_scale = 0.67;
    int persistent_blocks = 0;
    // launch shapes and queue / tile sizes (options, or chosen by rc_autotune)
    int light_blocks_per_cu = 8, camera_blocks_per_cu = 8, gather_blocks_per_cu = 8;
    uint32_t path_chunk = 256;
    uint32_t light_refill = 1, camera_refill = 1;  // idle lanes a tracing wave waits for before it refills
    int count_events = 0;
    int exact_merge = 0;
    std::string output_path = "D:\\caustics.exr";
    // device scene
    DevBuf<float4> d_tri, d_shade, d_nodes, d_nodes4;
    DevBuf<int> d_tri_id;
    DevBuf<DevMaterial> d_mats;
    DevBuf<DevLight> d_lights;
    DevBuf<float> d_cdf;
    std::vector<DevLight> lights;
    float integral = 0.0f;
    int root = 0, n_nodes = 0;
    int root4 = 0, n_nodes4 = 0;
    int bvh_width = 4;  // 4: traverse the 4-wide BVH, 2: the binary one
    int bvh_build = 0;  // 0: binned SAH on the host, 1: linear BVH on the device (binary nodes)
    int device_ingest = 1;  // with bvh_build 1: init converts the meshes on the GPU (d_tri, d_shade, d_tri_id, bbox)
    bool ingested_on_device = false;
    size_t n_tris = 0;      // triangles of the scene (either ingest)
    unsigned long long* merge_sets = nullptr;  // rc_merge_sets: per-slot set fingerprints (counting launches)
    int bvh_depth = 0;   // internal ancestors of the deepest leaf of the binary tree
    int bvh_depth4 = 0;  // depth of the deepest internal 4-wide node (root 0)
    double bvh_build_ms = 0.0;
    // per-iteration buffers
    DevBuf<float4> ph_rec, s_pos, s_dir;  // ph_rec: kPhotonRecord float4 per unsorted photon
    DevBuf<float2> s_tp;
    DevBuf<uint32_t> s_index, vals_a, vals_b, hist, hist_sums, counters;
    DevBuf<unsigned long long> keys_a, keys_b;
    DevBuf<uint2> cells;
    DevBuf<float> image;
    DevBuf<float4> q_pos, q_dir, q_sn, q_gn, q_tpn;
    DevBuf<int2> path_info;
    DevBuf<float> contrib;
    uint32_t q_capacity = 0;
    DevBuf<Counters> stats_dev;  // [0] light pass, [1] camera pass
    hipEvent_t ev[9] = {};  // light, map, camera (trace..resolve), trace end, gather end, query sort begin
    int query_order = 1;    // 1: the gather takes the merge queries in spatial (Morton) order
    int light_order = 0;    // 1: the light pass traces its paths sorted by emission position and direction
    int camera_qmc = 0;     // 1: camera jitter and first-bounce choice stratified across iterations
    // the k-th iteration rendered samples iteration index iteration_offset + k * iteration_stride (Philox
    // counter, radius schedule, stratified sample index): K frames of one scene with offsets 0..K-1 and
    // stride K together render iterations 0, 1, 2, ... of one estimator
    int iteration_offset = 0, iteration_stride = 1;
    DevBuf<unsigned long long> lk_a, lk_b;  // light-path sort buffers (one entry per path)
    DevBuf<uint32_t> lv_a, lv_b, l_hist, l_hist_sums;
    int group_queries = 8;  // queries per wave of the grouped merge (4, 8, 16), 0: one query per wave, 1: per lane
    uint32_t gather_tile = 1024;  // grouped merge: queries per XCD tile of the order (0: one segment per XCD)
    int gather_block = 2;  // grouped merge: 1 / 2 / 4 = rolled query loop over 64 x this many photons, 0 = unrolled
    // adaptive MLT light pass (option "mlt"): chains, their explicit states and counts
    int mlt = 0;
    uint32_t mlt_chains = 262144;
    DevBuf<uint4> mlt_state;
    DevBuf<uint32_t> mlt_counts;
    uint32_t mlt_last[3] = {0, 0, 0};  // visible uniform paths, mutations, accepted mutations
    DevBuf<unsigned long long> qk_a, qk_b;  // query sort buffers (q_capacity each)
    DevBuf<uint32_t> qv_a, qv_b, q_hist, q_hist_sums;
    int sort_passes = 0;
    uint32_t capacity = 0;
    int iterations_done = 0;
    double stats[RC_STAT_COUNT] = {0};
    hipStream_t stream = nullptr;
    // overlap = 1: the camera trace runs on stream2 (its own path counters)
    // while the light pass and the photon map run on stream; the merge waits
    // for both
    int overlap = 1;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_cam = nullptr;   // the overlapped trace is done (stream waits on it)
    hipEvent_t ev_join = nullptr;   // the query sort's start
    hipEvent_t ev_qdone = nullptr;  // the query sort's end (stream waits on it before the merge)
    hipEvent_t ev_fork = nullptr;  // stream2 starts after the work already queued on stream
    hipEvent_t ev_gjoin = nullptr;  // inside a captured iteration: stream's merge waits for stream2's query sort
    DevBuf<uint32_t> cam_counters;
    uint32_t* h_nq = nullptr;  // pinned: the overlapped trace's query count
    uint32_t* h_count = nullptr;  // pinned: the light pass's photon count
    DevBuf<float> film_prev;         // whole-iteration graph: the film before the replay
    DevBuf<uint32_t> graph_count;    // whole-iteration graph: the light pass's photon count
    int device_cus = 0;
    uint32_t last_queries = 0;
    GuideModel guide;  // learned emission proposal (option "guide")
    // captured iteration graph (option "graph"): photon map + camera pass
    int graph = 0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraphExec_t graph_exec2 = nullptr;  // with overlap: merge and resolve (graph_exec: the map)
    hipGraphExec_t graph_exec3 = nullptr;  // with overlap: the query sort, replayed on stream2
    std::vector<uintptr_t> graph_key;  // buffers and sizes the graph was captured with
    DevBuf<IterParams> d_iter;
    IterParams* h_iter = nullptr;      // pinned staging of the per-iteration values
    int graph_captures = 0;
    hipEvent_t ev_guide[4] = {};  // guide work begin / end, network step begin / end
};

Engine* g = nullptr;

// ---------------------------------------------------------------------
// Scene ingest (Scene::init_meshes, Scene.cpp:14-43; Mesh::Mesh, mesh.cpp:12-26)
// ---------------------------------------------------------------------
void inverse3x3_of4(const float* M, float inv[3][3]) {
    // Inverse (matrix.cpp:4-63) in double; only the upper 3x3 is used by ApplyToNormal.
    double a[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) a[r][c] = M[4 * r + c];
    double A2323 = a[2][2] * a[3][3] - a[2][3] * a[3][2], A1323 = a[2][1] * a[3][3] - a[2][3] * a[3][1];
    double A1223 = a[2][1] * a[3][2] - a[2][2] * a[3][1], A0323 = a[2][0] * a[3][3] - a[2][3] * a[3][0];
    double A0223 = a[2][0] * a[3][2] - a[2][2] * a[3][0], A0123 = a[2][0] * a[3][1] - a[2][1] * a[3][0];
    double A2313 = a[1][2] * a[3][3] - a[1][3] * a[3][2], A1313 = a[1][1] * a[3][3] - a[1][3] * a[3][1];
    double A1213 = a[1][1] * a[3][2] - a[1][2] * a[3][1], A0313 = a[1][0] * a[3][3] - a[1][3] * a[3][0];
    double A0213 = a[1][0] * a[3][2] - a[1][2] * a[3][0], A0113 = a[1][0] * a[3][1] - a[1][1] * a[3][0];
    double det = a[0][0] * (a[1][1] * A2323 - a[1][2] * A1323 + a[1][3] * A1223) -
                 a[0][1] * (a[1][0] * A2323 - a[1][2] * A0323 + a[1][3] * A0223) +
                 a[0][2] * (a[1][0] * A1323 - a[1][1] * A0323 + a[1][3] * A0123) -
                 a[0][3] * (a[1][0] * A1223 - a[1][1] * A0223 + a[1][2] * A0123);
    det = 1.0 / det;
    inv[0][0] = (float)(det * (a[1][1] * A2323 - a[1][2] * A1323 + a[1][3] * A1223));
    inv[0][1] = (float)(det * -(a[0][1] * A2323 - a[0][2] * A1323 + a[0][3] * A1223));
    inv[0][2] = (float)(det * (a[0][1] * A2313 - a[0][2] * A1313 + a[0][3] * A1213));
    inv[1][0] = (float)(det * -(a[1][0] * A2323 - a[1][2] * A0323 + a[1][3] * A0223));
    inv[1][1] = (float)(det * (a[0][0] * A2323 - a[0][2] * A0323 + a[0][3] * A0223));
    inv[1][2] = (float)(det * -(a[0][0] * A2313 - a[0][2] * A0313 + a[0][3] * A0213));
    inv[2][0] = (float)(det * (a[1][0] * A1323 - a[1][1] * A0323 + a[1][3] * A0123));
    inv[2][1] = (float)(det * -(a[0][0] * A1323 - a[0][1] * A0323 + a[0][3] * A0123));
    inv[2][2] = (float)(det * (a[0][0] * A1313 - a[0][1] * A0313 + a[0][3] * A0113));
}

DevMaterial convert_material(const rc_python_material& pm, int compat_glass) {
    // Scene::init_materials (Scene.cpp:45-68) and materials.h:59-146; with
    // compat_glass = 0 glass of roughness >= 0.01 is a GGX dielectric of
    // a2 = roughness^2 (the metal's mapping, materials.h:127-137)
    DevMaterial m{};
    F3 c = f3(pm.color[0], pm.color[1], pm.color[2]);
    char t = pm.type;
    if (t != 'P' && t != 'G' && t != 'M') {
        t = 'P';
        c = f3(1.0f, 1.0f, 1.0f);
    }
    auto cont_of = [](F3 v) { return smax(0.05f, smax(smax(v.x, v.y), v.z)); };
    m.R[0] = c.x;
    m.R[1] = c.y;
    m.R[2] = c.z;
    m.ior = 1.0f;
    if (t == 'P') {
        float rough = (pm.type == 'P') ? pm.roughness * pm.roughness : 1.0f;
        float spec = (pm.type == 'P') ? pm.specular * 0.08f : 0.0f * 0.08f;
        m.s = spec;
        m.cont = 1.0f;
        if (rough < 0.01f) {
            m.kind = BK_FSMG;
        } else {
            m.kind = BK_FMG;
            m.a2 = rough * rough;
        }
    } else if (t == 'G') {
        m.kind = (!compat_glass && pm.roughness >= 0.01f) ? BK_RGLASS : BK_FRESNEL_SPEC;
        if (m.kind == BK_RGLASS) m.a2 = pm.roughness * pm.roughness;
        m.ior = pm.ior;
        m.cont = cont_of(c);
    } else {
        if (pm.roughness < 0.01f) {
            m.kind = BK_SPEC_REFL;
        } else {
            m.kind = BK_MRG;
            m.a2 = pm.roughness * pm.roughness;
        }
        m.cont = cont_of(c);
    }
    return m;
}

// ---------------------------------------------------------------------
// Lights (Scene::init_lights, Scene.cpp:70-108; Lights.cpp)
// ---------------------------------------------------------------------
struct M3 {
    F3 i, j, k;
};
M3 rot_matrix(F3 r) {  // matrix_3x3(vec3), matrix.cpp:77-89 (double cos/sin on this platform)
    double x = r.x, y = r.y, z = r.z;
    M3 m;
    m.i = f3((float)(cos(z) * cos(y)), (float)(sin(z) * cos(y)), (float)(-sin(y)));
    m.j = f3((float)(cos(z) * sin(y) * sin(x) - sin(z) * cos(x)), (float)(sin(z) * sin(y) * sin(x) + cos(z) * cos(x)),
             (float)(cos(y) * sin(x)));
    m.k = f3((float)(cos(z) * sin(y) * cos(x) + sin(z) * sin(x)), (float)(sin(z) * sin(y) * cos(x) - cos(z) * sin(x)),
             (float)(cos(y) * cos(x)));
    return m;
}
F3 mul_vm(F3 v, M3 m) { return (m.i * v.x + m.j * v.y) + m.k * v.z; }
F3 rotate_deg(F3 v, F3 p, F3 rot) {
    const float PI = 3.14159265359f;
    M3 m = rot_matrix(f3(rot.x * PI / 180.f, rot.y * PI / 180.f, rot.z * PI / 180.f));
    return mul_vm(v - p, m) + p;
}
F3 project_onto_plane(F3 point, F3 center, F3 normal) {
    F3 tp = point - center;
    return (tp - normal * (dot(tp, normal) / len2(normal))) + center;
}
void put3(float* d, F3 v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}

void build_lights(Engine& e) {
    e.lights.clear();
    for (const HostLight& p : e.py_lights) {
        DevLight L{};
        F3 c = f3(p.color[0], p.color[1], p.color[2]);
        L.power = p.power;
        L.radius = p.radius;
        put3(L.pos, f3(p.position[0], p.position[1], p.position[2]));
        if (p.type == 'P') {
            L.kind = L_POINT;
            put3(L.color, c);
        } else if (p.type == 'S') {
            L.kind = L_SUN;
            put3(L.color, c * p.power);
            M3 m = rot_matrix(f3(p.rotation[0], p.rotation[1], p.rotation[2]));
            F3 normal = mul_vm(f3(0.0f, 0.0f, -1.0f) - f3s(0.0f), m) + f3s(0.0f);
            M3 inv{f3(m.i.x, m.j.x, m.k.x), f3(m.i.y, m.j.y, m.k.y), f3(m.i.z, m.j.z, m.k.z)};
            F3 mn = f3(e.bbox.lo[0], e.bbox.lo[1], e.bbox.lo[2]), mx = f3(e.bbox.hi[0], e.bbox.hi[1], e.bbox.hi[2]);
            F3 pos = (-normal) * sqrtf(len2(mx - mn)) + (mx + mn) * 0.5f;
            F3 pts[8] = {mn, f3(mx.x, mn.y, mn.z), f3(mn.x, mx.y, mn.z), f3(mx.x, mx.y, mn.z),
                         f3(mn.x, mn.y, mx.z), f3(mx.x, mn.y, mx.z), f3(mn.x, mx.y, mx.z), mx};
            F3 hi = f3(-INFINITY, -INFINITY, 0.0f), lo = f3(INFINITY, INFINITY, 0.0f);
            for (int k = 0; k < 8; ++k) {
                F3 loc = mul_vm(project_onto_plane(pts[k], pos, normal) - pos, inv);
                if (loc.x > hi.x) hi.x = loc.x;
                if (loc.y > hi.y) hi.y = loc.y;
                if (loc.x < lo.x) lo.x = loc.x;
                if (loc.y < lo.y) lo.y = loc.y;
            }
            F3 bl = mul_vm(lo, m) + pos;
            F3 brc = mul_vm(f3(hi.x, lo.y, 0.0f), m) + pos, tlc = mul_vm(f3(lo.x, hi.y, 0.0f), m) + pos;
            put3(L.bl, bl);
            put3(L.h, brc - bl);
            put3(L.v, tlc - bl);
            put3(L.normal, normal);
            put3(L.pos, pos);
        } else {
            L.kind = L_AREA;
            put3(L.color, c * p.power);
            F3 pos = f3(p.position[0], p.position[1], p.position[2]);
            F3 rr = f3(p.rotation[0], p.rotation[1], p.rotation[2]);
            F3 bl = pos - f3(p.width / 2.f, p.height / 2.f, 0.f);
            F3 hh = f3(p.width, 0.0f, 0.0f), vv = f3(0.0f, p.height, 0.0f), n = f3(0.0f, 0.0f, -1.0f);
            F3 tl = bl + vv, br = bl + hh;
            bl = rotate_deg(bl, pos, rr);
            tl = rotate_deg(tl, pos, rr);
            br = rotate_deg(br, pos, rr);
            n = rotate_deg(n, f3(0.0f, 0.0f, 0.0f), rr);
            put3(L.bl, bl);
            put3(L.h, br - bl);
            put3(L.v, tl - bl);
            put3(L.normal, n);
            Frame fr = frame_from_z(n);
            put3(L.fx, fr.x);
            put3(L.fy, fr.y);
            put3(L.fz, fr.z);
        }
        e.lights.push_back(L);
    }
}

// ---------------------------------------------------------------------
// BVH build: binned SAH over all world-space triangles, binary nodes that
// hold both child boxes, leaves of at most kLeafMax triangles stored
// contiguously (the triangle arrays are reordered to leaf order).
// ---------------------------------------------------------------------
struct BuildItem {
    float lo[3], hi[3], c[3];
    uint32_t id;
};

struct Builder {
    std::vector<BuildItem> items;
    std::vector<float4> nodes;  // 4 per node
    std::vector<uint32_t> order;
    int max_depth = 0;

    static void grow(Box& b, const float* lo, const float* hi) {
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = std::min(b.lo[k], lo[k]);
            b.hi[k] = std::max(b.hi[k], hi[k]);
        }
    }
    static Box empty() {
        Box b;
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = INFINITY;
            b.hi[k] = -INFINITY;
        }
        return b;
    }
    static float area(const Box& b) {
        float d[3];
        for (int k = 0; k < 3; ++k) d[k] = std::max(0.0f, b.hi[k] - b.lo[k]);
        return 2.0f * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
    }
    static Box padded(Box b) {
        // outward padding so the kernel's slab test never rejects a box
        // containing a hit because of rounding
        for (int k = 0; k < 3; ++k) {
            float m = std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k]));
            float pad = m * 9.5367431640625e-07f + 1e-30f;  // 2^-20 relative
            b.lo[k] -= pad;
            b.hi[k] += pad;
        }
        return b;
    }

    int make_leaf(int start, int end) {
        int first = (int)order.size();
        for (int i = start; i < end; ++i) order.push_back(items[i].id);
        return ~((first << 3) | (end - start - 1));
    }

    // returns child ref; writes bounds
    int build(int start, int end, int depth, Box& out) {
        max_depth = std::max(max_depth, depth);
        Box b = empty(), cb = empty();
        for (int i = start; i < end; ++i) {
            grow(b, items[i].lo, items[i].hi);
            grow(cb, items[i].c, items[i].c);
        }
        out = b;
        int n = end - start;
        if (n <= 2 && n <= kLeafMax) return make_leaf(start, end);
        int mid = -1;
        constexpr int kBins = 16;
        if (depth < kMaxDepth - 8) {
            float best_cost = INFINITY;
            int best_axis = -1, best_split = -1;
            for (int ax = 0; ax < 3; ++ax) {
                float ext = cb.hi[ax] - cb.lo[ax];
                if (!(ext > 0.0f)) continue;
                float scale = kBins / ext;
                Box bb[kBins];
                int cnt[kBins] = {0};
                for (int k = 0; k < kBins; ++k) bb[k] = empty();
                for (int i = start; i < end; ++i) {
                    int bi = std::min(kBins - 1, (int)((items[i].c[ax] - cb.lo[ax]) * scale));
                    cnt[bi]++;
                    grow(bb[bi], items[i].lo, items[i].hi);
                }
                float left_area[kBins];
                int left_cnt[kBins];
                Box acc = empty();
                int c = 0;
                for (int k = 0; k < kBins; ++k) {
                    if (cnt[k]) grow(acc, bb[k].lo, bb[k].hi);
                    c += cnt[k];
                    left_area[k] = c ? area(acc) : 0.0f;
                    left_cnt[k] = c;
                }
                acc = empty();
                c = 0;
                for (int k = kBins - 1; k >= 1; --k) {
                    if (cnt[k]) grow(acc, bb[k].lo, bb[k].hi);
                    c += cnt[k];
                    if (c == 0 || left_cnt[k - 1] == 0) continue;
                    float cost = left_area[k - 1] * left_cnt[k - 1] + area(acc) * c;
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = ax;
                        best_split = k;
                    }
                }
            }
            float parent = area(b);
            float leaf_cost = (float)n;
            float split_cost = 0.5f + (parent > 0.0f ? best_cost / parent : INFINITY);
            if (best_axis >= 0 && (n > kLeafMax || split_cost < leaf_cost)) {
                float scale = kBins / (cb.hi[best_axis] - cb.lo[best_axis]);
                auto pred = [&](const BuildItem& it) {
                    return std::min(kBins - 1, (int)((it.c[best_axis] - cb.lo[best_axis]) * scale)) < best_split;
                };
                mid = (int)(std::partition(items.begin() + start, items.begin() + end, pred) - items.begin());
                if (mid == start || mid == end) mid = -1;
            } else if (best_axis < 0 && n <= kLeafMax) {
                return make_leaf(start, end);
            } else if (n <= kLeafMax) {
                return make_leaf(start, end);
            }
        }
        if (mid < 0) {
            if (n <= kLeafMax) return make_leaf(start, end);
            // median split on the widest centroid axis
            int ax = 0;
            float best = -1.0f;
            for (int k = 0; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > best) {
                    best = cb.hi[k] - cb.lo[k];
                    ax = k;
                }
            mid = (start + end) / 2;
            std::nth_element(items.begin() + start, items.begin() + mid, items.begin() + end,
                             [&](const BuildItem& a, const BuildItem& b2) { return a.c[ax] < b2.c[ax]; });
        }
        int id = (int)(nodes.size() / 4);
        nodes.resize(nodes.size() + 4);
        Box lb, rb;
        int l = build(start, mid, depth + 1, lb);
        int r = build(mid, end, depth + 1, rb);
        lb = padded(lb);
        rb = padded(rb);
        float4* nd = &nodes[4 * (size_t)id];
        nd[0] = make_float4(lb.lo[0], lb.hi[0], lb.lo[1], lb.hi[1]);
        nd[1] = make_float4(lb.lo[2], lb.hi[2], rb.lo[0], rb.hi[0]);
        nd[2] = make_float4(rb.lo[1], rb.hi[1], rb.lo[2], rb.hi[2]);
        int refs[2] = {l, r};
        float fl, fr;
        memcpy(&fl, &refs[0], 4);
        memcpy(&fr, &refs[1], 4);
        nd[3] = make_float4(fl, fr, 0.0f, 0.0f);
        return id;
    }
};

// ---------------------------------------------------------------------
// 4-wide BVH: the binary tree collapsed top-down; a node takes its binary
// node's two children and keeps opening the inner child of largest surface
// area until it has four children or only leaves.  Node layout (8 float4):
// x lo, x hi, y lo, y hi, z lo, z hi of the four children, then their refs
// (a ref >= 0 is a node, < 0 the leaf encoding of the binary tree,
// kEmptyChild for a missing child), then padding.
// ---------------------------------------------------------------------
struct Child4 {
    float lo[3], hi[3];
    int ref;
};

void binary_children(const std::vector<float4>& bn, int ref, Child4 out[2]) {
    const float4* n = &bn[4 * (size_t)ref];
    out[0] = Child4{{n[0].x, n[0].z, n[1].x}, {n[0].y, n[0].w, n[1].y}, __builtin_bit_cast(int, n[3].x)};
    out[1] = Child4{{n[1].z, n[2].x, n[2].z}, {n[1].w, n[2].y, n[2].w}, __builtin_bit_cast(int, n[3].y)};
}

float child_area(const Child4& c) {
    float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

int collapse4(const std::vector<float4>& bn, int bref, std::vector<float4>& out, int depth, int& max_depth) {
    max_depth = std::max(max_depth, depth);
    Child4 pair[2];
    binary_children(bn, bref, pair);
    std::vector<Child4> c = {pair[0], pair[1]};
    while (c.size() < 4) {
        int pick = -1;
        float area = -1.0f;
        for (size_t i = 0; i < c.size(); ++i)
            if (c[i].ref >= 0 && child_area(c[i]) > area) {
                area = child_area(c[i]);
                pick = (int)i;
            }
        if (pick < 0) break;
        int r = c[pick].ref;
        c.erase(c.begin() + pick);
        binary_children(bn, r, pair);
        c.push_back(pair[0]);
        c.push_back(pair[1]);
    }
    const size_t idx = out.size() / 8;
    out.resize(out.size() + 8);
    int refs[4] = {kEmptyChild, kEmptyChild, kEmptyChild, kEmptyChild};
    for (size_t i = 0; i < c.size(); ++i)
        refs[i] = c[i].ref >= 0 ? collapse4(bn, c[i].ref, out, depth + 1, max_depth) : c[i].ref;
    float lo[3][4], hi[3][4];
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 4; ++i) {
            lo[k][i] = i < (int)c.size() ? c[i].lo[k] : 0.0f;
            hi[k][i] = i < (int)c.size() ? c[i].hi[k] : 0.0f;
        }
    float4* n = &out[8 * idx];
    for (int k = 0; k < 3; ++k) {
        n[2 * k] = make_float4(lo[k][0], lo[k][1], lo[k][2], lo[k][3]);
        n[2 * k + 1] = make_float4(hi[k][0], hi[k][1], hi[k][2], hi[k][3]);
    }
    n[6] = make_float4(__builtin_bit_cast(float, refs[0]), __builtin_bit_cast(float, refs[1]),
                       __builtin_bit_cast(float, refs[2]), __builtin_bit_cast(float, refs[3]));
    n[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return (int)idx;
}

template <typename T>
bool upload(DevBuf<T>& b, const std::vector<T>& v) {
    if (!b.alloc(v.size())) return false;
    if (!v.empty()) HIP_OK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return true;
}

bool uses_nodes4(const Engine& e) { return e.bvh_width == 4 && e.n_nodes4 > 0; }

// Materials converted again with the options in force at rc_prepare (compat_glass).
bool reconvert_materials(Engine& e) {
    for (size_t i = 0; i < e.py_mats.size() && i < e.mats.size(); ++i)
        e.mats[i] = convert_material(e.py_mats[i], e.compat_glass);
    return true;
}

SceneView scene_view(const Engine& e) {
    SceneView sc{};
    sc.tri = e.d_tri.p;
    sc.shade = e.d_shade.p;
    sc.tri_id = e.d_tri_id.p;
    sc.nodes = e.d_nodes.p;
    sc.root = e.root;
    sc.nodes4 = uses_nodes4(e) ? e.d_nodes4.p : nullptr;
    sc.root4 = e.root4;
    sc.mats = e.d_mats.p;
    sc.lights = e.d_lights.p;
    sc.cdf = e.d_cdf.p;
    sc.nlights = (int)e.lights.size();
    sc.integral = e.integral;
    put3(sc.cam_o, e.cam_o);
    put3(sc.cam_bl, e.cam_bl);
    put3(sc.cam_h, e.cam_h);
    put3(sc.cam_v, e.cam_v);
    sc.W = e.W;
    sc.H = e.H;
    sc.max_path_length = e.max_path_length;
    sc.compat_point = e.compat_point;
    sc.compat_pdf = e.compat_pdf;
    sc.compat_rows = e.compat_rows;
    return sc;
}

bool ensure_device(Engine& e) {
    if (e.stream) return true;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("no HIP device available: libRealCaustics needs an MI355X (gfx950) GPU");
        return false;
    }
    HIP_OK(hipSetDevice(0));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, 0));
    e.device_cus = prop.multiProcessorCount;
    HIP_OK(hipStreamCreate(&e.stream));
    for (auto& ev : e.ev) HIP_OK(hipEventCreate(&ev));
    for (auto& ev : e.ev_guide) HIP_OK(hipEventCreate(&ev));
    return true;
}

// The second stream, its join event and counters (overlapped camera trace), on first use.
bool ensure_stream2(Engine& e) {
    if (e.stream2) return true;  // created last: everything else is in place
    if (!e.ev_cam) HIP_OK(hipEventCreateWithFlags(&e.ev_cam, hipEventDisableTiming));
    if (!e.ev_join) HIP_OK(hipEventCreate(&e.ev_join));
    if (!e.ev_qdone) HIP_OK(hipEventCreate(&e.ev_qdone));
    if (!e.ev_fork) HIP_OK(hipEventCreateWithFlags(&e.ev_fork, hipEventDisableTiming));
    if (!e.ev_gjoin) HIP_OK(hipEventCreateWithFlags(&e.ev_gjoin, hipEventDisableTiming));
    if (!e.h_nq) HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&e.h_nq), sizeof(uint32_t), hipHostMallocDefault));
    if (!e.cam_counters.alloc(kPathCounters)) return false;
    HIP_OK(hipStreamCreate(&e.stream2));
    return true;
}

// Linear BVH built on the device over the triangles uploaded in input order
// (option bvh_build = 1); binary nodes only, root 0.
bool gpu_bvh(Engine& e, uint32_t n) {
    DevBuf<float4> tmp4;
    DevBuf<int> tmp_id, parent;
    DevBuf<unsigned long long> ka, kb;
    DevBuf<uint32_t> va, vb, hist, sums, flags, depth;
    DevBuf<float> boxes;
    bool ok = e.d_nodes.alloc(4 * (size_t)(n - 1)) && tmp4.alloc(6 * (size_t)n) && tmp_id.alloc(n) &&
              parent.alloc(2 * (size_t)n) && ka.alloc(n) && kb.alloc(n) && va.alloc(n) && vb.alloc(n) &&
              hist.alloc(photon_map_hist_entries(n)) && sums.alloc(photon_map_sum_entries(n)) && flags.alloc(n) &&
              depth.alloc(1) && boxes.alloc(12 * (size_t)n);
    int d = -1;
    if (ok) {
        LbvhBuffers b{e.d_tri.p, e.d_shade.p, e.d_tri_id.p, tmp4.p, tmp_id.p, e.d_nodes.p, ka.p, kb.p, va.p, vb.p,
                      hist.p, sums.p, parent.p, flags.p, boxes.p, depth.p};
        float ms = 0.0f;
        ok = hipEventRecord(e.ev[0], e.stream) == hipSuccess;
        d = ok ? build_lbvh(b, n, e.bbox.lo, e.bbox.hi, e.stream) : -1;
        // bvh_width 4: the binary tree collapsed to 4-wide nodes on the GPU
        e.n_nodes4 = 0;
        e.root4 = 0;
        bool collapsed = true;
        // the collapse keeps the internal nodes of even depth: the deepest
        // internal node (depth d - 1) lies under (d - 1) / 2 4-wide ancestors
        if (ok && d >= 1 && traversal_width(e.bvh_width, d, (d - 1) / 2) == 4) {
            const int n4 = lbvh_wide_count(b, n, e.stream);
            collapsed = n4 > 0 && e.d_nodes4.alloc(8 * (size_t)n4) && lbvh_collapse4(b, n, e.d_nodes4.p, e.stream);
            if (collapsed) e.n_nodes4 = n4;
            ok = collapsed;
        }
        ok = ok && hipEventRecord(e.ev[1], e.stream) == hipSuccess && hipEventSynchronize(e.ev[1]) == hipSuccess &&
             hipEventElapsedTime(&ms, e.ev[0], e.ev[1]) == hipSuccess;
        e.bvh_build_ms = ms;
        if (!ok) set_error(collapsed ? "GPU BVH build: HIP event error" : "GPU BVH build: 4-wide collapse failed");
    }
    tmp4.release();
    tmp_id.release();
    parent.release();
    ka.release();
    kb.release();
    va.release();
    vb.release();
    hist.release();
    sums.release();
    flags.release();
    depth.release();
    boxes.release();
    if (!ok) return false;
    if (d < 0) {
        set_error("GPU BVH build failed");
        return false;
    }
    if (traversal_width(2, d, 0) == 0) {
        set_error("BVH deeper than the traversal stack");
        return false;
    }
    e.bvh_depth = d;
    e.bvh_depth4 = d >= 1 ? (d - 1) / 2 : 0;
    if (e.n_nodes4 == 0) e.d_nodes4.release();
    return true;
}

bool prepare(Engine& e) {
    if (e.prepared) return true;
    if (!e.initialised) {
        set_error("main/rc_prepare called before init");
        return false;
    }
    if (has_error()) return false;
    if (!ensure_device(e)) return false;
    const bool dev = e.ingested_on_device;
    size_t ntri = dev ? e.n_tris : e.tris.size() / 4;
    e.n_tris = ntri;
    if (ntri == 0) {
        set_error("scene has no triangles");
        return false;
    }
    if (dev && e.bvh_build != 1) {
        set_error("the meshes were ingested on the GPU (option device_ingest); set bvh_build = 1 or device_ingest = 0 before init");
        return false;
    }
    // scene bounds (union of triangle boxes, as the root of BVH_world.cpp); the device ingest made them
    Builder bld;
    if (!dev) bld.items.resize(ntri);
    Box sb = dev ? e.bbox : Builder::empty();
    for (size_t t = 0; t < (dev ? 0 : ntri); ++t) {
        BuildItem& it = bld.items[t];
        it.id = (uint32_t)t;
        for (int k = 0; k < 3; ++k) {
            float a = e.wv[3 * e.tris[4 * t] + k], b = e.wv[3 * e.tris[4 * t + 1] + k], c = e.wv[3 * e.tris[4 * t + 2] + k];
            it.lo[k] = smin(smin(a, b), c);
            it.hi[k] = smax(smax(a, b), c);
            it.c[k] = (it.lo[k] + it.hi[k]) * 0.5f;
        }
        for (int k = 0; k < 3; ++k) {
            sb.lo[k] = smin(sb.lo[k], it.lo[k]);
            sb.hi[k] = smax(sb.hi[k], it.hi[k]);
        }
    }
    e.bbox = sb;
    const bool gpu_tree = e.bvh_build == 1 && ntri >= 2;
    if (gpu_tree && ntri >= (1u << 30)) {
        set_error("bvh_build=1 takes fewer than 2^30 triangles");
        return false;
    }
    Box rootb;
    int root = 0;
    auto t_build = std::chrono::steady_clock::now();
    if (gpu_tree) {
        if (!dev) {
            bld.order.resize(ntri);  // triangles go up in input order; the GPU build reorders them
            for (size_t t = 0; t < ntri; ++t) bld.order[t] = (uint32_t)t;
        }
    } else {
        bld.order.reserve(ntri);
        root = bld.build(0, (int)ntri, 0, rootb);
    }
    if (!gpu_tree && root < 0) {
        // whole scene in one leaf: wrap it in a node whose second child is an empty box
        bld.nodes.resize(4);
        Box lb = Builder::padded(rootb);
        bld.nodes[0] = make_float4(lb.lo[0], lb.hi[0], lb.lo[1], lb.hi[1]);
        bld.nodes[1] = make_float4(lb.lo[2], lb.hi[2], INFINITY, -INFINITY);
        bld.nodes[2] = make_float4(INFINITY, -INFINITY, INFINITY, -INFINITY);
        float fl, fr;
        int refs[2] = {root, root};
        memcpy(&fl, &refs[0], 4);
        memcpy(&fr, &refs[1], 4);
        bld.nodes[3] = make_float4(fl, fr, 0.0f, 0.0f);
        root = 0;
    }
    e.root = root;
    e.n_nodes = gpu_tree ? (int)ntri - 1 : (int)(bld.nodes.size() / 4);
    std::vector<float4> nodes4;
    e.root4 = 0;
    if (!gpu_tree) {
        nodes4.reserve(bld.nodes.size());
        int depth4 = 0;
        e.root4 = collapse4(bld.nodes, root, nodes4, 0, depth4);
        e.bvh_depth = std::max(1, bld.max_depth);
        e.bvh_depth4 = depth4;
        // a 4-wide tree whose traversal stack could exceed kMaxDepth entries is
        // dropped for the binary traversal (one push per level)
        const int width = traversal_width(e.bvh_width, e.bvh_depth, depth4);
        if (width == 0) {
            set_error("BVH deeper than the traversal stack");
            return false;
        }
        if (width != 4) {
            nodes4.clear();
            e.root4 = 0;
        }
        e.bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_build).count();
    }
    e.n_nodes4 = (int)(nodes4.size() / 8);
    std::vector<float4> tri(dev ? 0 : 3 * ntri), shade(dev ? 0 : 3 * ntri);
    std::vector<int> tri_id(dev ? 0 : ntri);
    for (size_t i = 0; i < (dev ? 0 : ntri); ++i) {
        uint32_t t = bld.order[i];
        const uint32_t* T = &e.tris[4 * (size_t)t];
        F3 p0 = f3(e.wv[3 * T[0]], e.wv[3 * T[0] + 1], e.wv[3 * T[0] + 2]);
        F3 p1 = f3(e.wv[3 * T[1]], e.wv[3 * T[1] + 1], e.wv[3 * T[1] + 2]);
        F3 p2 = f3(e.wv[3 * T[2]], e.wv[3 * T[2] + 1], e.wv[3 * T[2] + 2]);
        F3 e1 = p1 - p0, e2 = p2 - p0;
        float matf;
        int mat = (int)T[3];
        memcpy(&matf, &mat, 4);
        tri[3 * i] = make_float4(p0.x, p0.y, p0.z, matf);
        tri[3 * i + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
        tri[3 * i + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
        for (int k = 0; k < 3; ++k)
            shade[3 * i + k] = make_float4(e.wn[3 * T[k]], e.wn[3 * T[k] + 1], e.wn[3 * T[k] + 2], 0.0f);
        tri_id[i] = (int)t;
    }
    build_lights(e);
    if (e.lights.empty()) {
        set_error("scene has no lights");
        return false;
    }
    // Lights_list::update_pdf_distribution (Light_list.cpp:32-44)
    std::vector<float> cdf(e.lights.size() + 1);
    cdf[0] = 0.0f;
    for (size_t i = 1; i < cdf.size(); ++i) cdf[i] = cdf[i - 1] + e.lights[i - 1].power;
    e.integral = cdf.back();
    if (!(e.integral > 0.0f)) {
        set_error("total light power must be positive");
        return false;
    }
    for (size_t i = 1; i < cdf.size(); ++i) cdf[i] /= e.integral;
    // main.cpp:166-168
    F3 mn = f3(sb.lo[0], sb.lo[1], sb.lo[2]), mx = f3(sb.hi[0], sb.hi[1], sb.hi[2]);
    e.scene_radius = sqrtf(len2(mn - (mn + mx) * 0.5f));
    e.base_radius *= e.scene_radius;
    if ((!dev && (!upload(e.d_tri, tri) || !upload(e.d_shade, shade) || !upload(e.d_tri_id, tri_id))) ||
        !upload(e.d_nodes, bld.nodes) || !upload(e.d_nodes4, nodes4) || !reconvert_materials(e) ||
        !upload(e.d_mats, e.mats) ||
        !upload(e.d_lights, e.lights) ||
        !upload(e.d_cdf, cdf))
        return false;
    if (gpu_tree && !gpu_bvh(e, (uint32_t)ntri)) return false;
    if (!e.image.alloc((size_t)e.W * e.H * 3)) return false;
    HIP_OK(hipMemset(e.image.p, 0, (size_t)e.W * e.H * 3 * sizeof(float)));
    if (!e.counters.alloc(kPathCounters) || !e.stats_dev.alloc(2)) return false;
    if (!e.guide.setup_scene(e.lights, e.compat_point, e.bbox.lo, e.bbox.hi)) return false;
    e.iterations_done = 0;
    e.prepared = true;
    return true;
}

long long light_paths(const Engine& e) { return e.photons > 0 ? e.photons : (long long)e.W * e.H; }

// Iteration index of the k-th iteration a scene renders (options iteration_offset / iteration_stride).
int sample_iteration(const Engine& e, int k) { return e.iteration_offset + k * e.iteration_stride; }

// Visibility normalisation of the adaptive MLT light pass (Vertex_Merging.cpp:278-282):
// every photon's throughput times (visible uniform paths) / N, applied to the merge.
float mlt_factor(const Engine& e) { return e.mlt ? (float)e.mlt_last[0] / (float)light_paths(e) : 1.0f; }

// Vertex_Merging.cpp:201-212, in float with the C library powf.
void iteration_params(const Engine& e, int iteration, float& radius, float& r2, float& norm) {
    float count = (float)light_paths(e);
    radius = e.base_radius;
    radius /= powf((float)(iteration + 1), 0.5f * (1.0f - e.radius_alpha));
    radius = smax(radius, 1e-7f);
    r2 = radius * radius;
    const float PI = 3.14159265359f;
    norm = 1.0f / (PI * r2 * count);
}

int log2_ceil(uint64_t v) {
    int b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

int light_blocks(const Engine& e) {
    if (e.persistent_blocks > 0) return e.persistent_blocks;
    return std::max(1, e.device_cus) * e.light_blocks_per_cu;
}
int camera_blocks(const Engine& e) {
    if (e.persistent_blocks > 0) return e.persistent_blocks;
    return std::max(1, e.device_cus) * e.camera_blocks_per_cu;
}
// a multiple of the 8 XCDs (the gather assigns contiguous query segments per XCD)
int gather_blocks(const Engine& e) {
    int b = std::max(8, e.device_cus) * e.gather_blocks_per_cu;
    return (b + 7) / 8 * 8;
}

bool ensure_photon_buffers(Engine& e, uint32_t cap) {
    if (cap <= e.capacity && e.ph_rec.p) return true;
    e.capacity = cap;
    return e.ph_rec.alloc(kPhotonRecord * (size_t)cap) &&
           e.s_pos.alloc(cap) && e.s_dir.alloc(cap) && e.s_tp.alloc(cap) && e.s_index.alloc(cap) &&
           e.keys_a.alloc(cap) && e.keys_b.alloc(cap) && e.vals_a.alloc(cap) && e.vals_b.alloc(cap) &&
           e.hist.alloc(photon_map_hist_entries(cap)) && e.hist_sums.alloc(photon_map_sum_entries(cap));
}

// Training paths of one iteration: the first guide_batch paths, a multiple of 64.
int guide_rows(const Engine& e) {
    long long n = std::min<long long>(e.guide.batch, light_paths(e));
    return (int)(n - n % 64);
}

bool guide_active(const Engine& e) { return e.guide.enabled && e.guide.slots > 0; }

bool guide_trains(const Engine& e, int iteration) {
    const GuideModel& gd = e.guide;
    return guide_active(e) && gd.train && (gd.step < gd.warmup || iteration % std::max(1, gd.train_every) == 0);
}

GuideView guide_light_view(Engine& e, int iteration) {
    if (!guide_active(e)) return plain_guide_view();
    bool rec = guide_trains(e, iteration);
    GuideView v = e.guide.view(rec, e.guide.step >= e.guide.warmup);
    v.train_paths = rec ? (uint32_t)guide_rows(e) : 0u;
    return v;
}

// s: the stream the proposal's set-up runs on (the scene's own by default)
bool ensure_guide_buffers(Engine& e, hipStream_t s = nullptr) {
    GuideModel& gd = e.guide;
    if (!guide_active(e)) return true;
    if (!s) s = e.stream;
    if (gd.cells_built != gd.cells &&
        !gd.setup_scene(e.lights, e.compat_point, e.bbox.lo, e.bbox.hi))
        return false;
    int n = std::max(64, guide_rows(e));
    if (!gd.params_ready && !gd.init_params(s)) return false;
    if (!gd.table_ready && gd.step >= gd.warmup && !gd.build_table(s)) return false;
    return gd.train_x.alloc((size_t)n * kGuideRowFloats) && gd.train_bin.alloc(n) && gd.train_w.alloc(n);
}

// The light pass's arguments for one iteration (single scene, no event counts).
LightPassArgs light_pass_args(Engine& e, int iteration) {
    LightPassArgs a{};
    a.sc = scene_view(e);
    a.k0 = (uint32_t)e.seed;
    a.k1 = (uint32_t)(e.seed >> 32);
    a.iteration = (uint32_t)iteration;
    a.stream = STREAM_LIGHT;
    a.n_paths = (uint32_t)light_paths(e);
    a.path_chunk = e.path_chunk;
    a.refill_min = e.light_refill;
    a.path_counter = e.counters.p;
    a.photon_counter = e.counters.p + 1;
    a.capacity = e.capacity;
    a.out_rec = e.ph_rec.p;
    a.guide = guide_light_view(e, iteration);
    return a;
}

// Coherent order of this launch's light paths (option light_order), or null.
bool light_order_buffers(Engine& e, size_t n) {
    return e.lk_a.alloc(n) && e.lk_b.alloc(n) && e.lv_a.alloc(n) && e.lv_b.alloc(n) &&
           e.l_hist.alloc(photon_map_hist_entries((uint32_t)n)) && e.l_hist_sums.alloc(photon_map_sum_entries((uint32_t)n));
}
const uint32_t* light_path_order(Engine& e, const LightPassArgs& a, hipStream_t s) {
    if (!e.light_order || a.mlt_state) return nullptr;
    if (!light_order_buffers(e, a.n_paths)) return nullptr;
    QuerySortBuffers b{e.lk_a.p, e.lk_b.p, e.lv_a.p, e.lv_b.p, e.l_hist.p, e.l_hist_sums.p};
    return sort_light_paths(a, b, s);
}

struct MapInfo {
    PhotonMapView pm;
};

// The photon map's grid for this iteration (host arithmetic, no launch):
// key widths, bucket table, cell size and origin, and the view the merge and
// the query sort read.
struct MapGrid {
    int uid_bits = 0, bucket_bits = 0;
    uint32_t buckets = 0;
    float cell = 0.0f;
    PhotonMapView pm{};
};

// The photon-map grid for N light paths and a radius, over the scene's
// bounds, for photon ids below `ids` (the light path of each photon, or the
// index of a caller's photon); allocates nothing, sets no array of g.pm.
bool map_grid_params(const Engine& e, long long N, long long ids, float radius, float r2, MapGrid& g) {
    // keys sort by (bucket, photon id): a path's photons come from one lane in bounce order,
    // so the stable sort keeps them in that order
    g.uid_bits = log2_ceil((uint64_t)std::max<long long>(1, ids));
    uint64_t want = std::max<uint64_t>(4096, std::min<uint64_t>(1ull << 24, (uint64_t)N / 2));
    g.bucket_bits = log2_ceil(want);
    if (g.bucket_bits + g.uid_bits > 64) {
        set_error("photon map key does not fit 64 bits");
        return false;
    }
    g.buckets = 1u << g.bucket_bits;
    float maxc = 0.0f;
    for (int k = 0; k < 3; ++k) maxc = std::max(maxc, std::max(std::fabs(e.bbox.lo[k]), std::fabs(e.bbox.hi[k])));
    float rq = radius + std::max(radius * 1e-3f, 8.0f * maxc * 1.1920929e-07f);
    // cells of at least 2/3 of the query radius: a query touches at most 4 per axis
    g.cell = rq * (float)std::max(0.67, e.cell_scale);
    PhotonMapView& pm = g.pm;
    pm.mask = g.buckets - 1;
    for (int k = 0; k < 3; ++k) pm.origin[k] = e.bbox.lo[k] - g.cell;
    pm.inv_cell = 1.0f / g.cell;
    pm.r2 = r2;
    pm.r_query = rq;
    return true;
}

bool map_grid(Engine& e, float radius, float r2, MapGrid& g) {
    if (!map_grid_params(e, light_paths(e), e.map_ids > 0 ? e.map_ids : light_paths(e), radius, r2, g)) return false;
    if (!e.cells.alloc(g.buckets)) return false;
    PhotonMapView& pm = g.pm;
    pm.pos = e.s_pos.p;
    pm.dir_tr = e.s_dir.p;
    pm.tp_gb = e.s_tp.p;
    pm.cells = e.cells.p;
    return true;
}

// Light pass into the unsorted photon buffers; returns the photon count.
// after_launch (if set) queues other work once the first light pass is queued
// and before the host waits for its photon count.
// The light pass makes the photon map's keys as it writes the photons (the
// iteration's grid is known before any photon exists): MapKeys of grid g into
// the engine's sort buffers, laid out for the photon buffer's capacity.
MapKeys map_keys_of(Engine& e, const MapGrid& g) {
    MapKeys k{};
    k.keys = e.keys_a.p;
    k.vals = e.vals_a.p;
    k.dig = map_sort_digits(e.hist.p, e.capacity);
    for (int c = 0; c < 3; ++c) k.origin[c] = g.pm.origin[c];
    k.inv_cell = g.pm.inv_cell;
    k.mask = g.pm.mask;
    k.uid_bits = g.uid_bits;
    return k;
}

bool run_light_pass(Engine& e, int iteration, uint32_t& count, const std::function<bool()>& after_launch = nullptr) {
    long long N = light_paths(e);
    if (N <= 0 || N > (1ll << 31)) {
        set_error("light path count out of range (1 .. 2^31)");
        return false;
    }
    float radius, r2, norm;
    iteration_params(e, iteration, radius, r2, norm);
    MapGrid grid;
    if (!map_grid(e, radius, r2, grid)) return false;
    uint32_t cap = e.photon_capacity > 0 ? (uint32_t)e.photon_capacity : (uint32_t)std::max<long long>(N, 1 << 16);
    if (!ensure_photon_buffers(e, std::max(cap, e.capacity))) return false;
    if (!ensure_guide_buffers(e)) return false;
    if (e.mlt) {
        if (guide_active(e)) {
            set_error("options mlt and guide are exclusive");
            return false;
        }
        if (!e.mlt_state.alloc((size_t)kMltBlocks * e.mlt_chains) || !e.mlt_counts.alloc(4)) return false;
    }
    if (!e.h_count) HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&e.h_count), sizeof(uint32_t), hipHostMallocDefault));
    for (int attempt = 0; attempt < 2; ++attempt) {
        launch_fill_u32(e.counters.p, 0u, 4, e.stream);
        if (e.mlt) launch_fill_u32(e.mlt_counts.p, 0u, 4, e.stream);
        if (e.count_events) launch_fill_u32(reinterpret_cast<uint32_t*>(e.stats_dev.p), 0u, 2 * kCounterFields * 2, e.stream);
        LightPassArgs a = light_pass_args(e, iteration);
        a.counters = e.count_events ? e.stats_dev.p : nullptr;
        clear_map_sort(e.hist.p, e.capacity, e.stream);
        a.map = map_keys_of(e, grid);
        if (e.mlt) {
            a.mlt_state = e.mlt_state.p;
            // at most mlt_chains chains of at least 64 indices each (the mutation size adapts along a chain)
            a.mlt_chains = (uint32_t)std::max<long long>(1, std::min<long long>(e.mlt_chains, N / 64));
            a.mlt_counts = e.mlt_counts.p;
        }
        if (e.light_emit && !e.mlt) {
            if (!e.emitted.alloc(3 * (size_t)N)) return false;
            a.emitted = e.emitted.p;
        }
        HIP_OK(hipEventRecord(e.ev[0], e.stream));
        a.path_order = light_path_order(e, a, e.stream);  // inside the light pass's time
        if (e.light_order && !a.path_order && !e.mlt) return false;
        launch_light_pass(a, light_blocks(e), e.stream);
        HIP_OK(hipEventRecord(e.ev[1], e.stream));
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(e.h_count, e.counters.p + 1, 4, hipMemcpyDeviceToHost, e.stream));
        if (attempt == 0 && after_launch && !after_launch()) return false;
        if (e.mlt) HIP_OK(hipMemcpyAsync(e.mlt_last, e.mlt_counts.p, sizeof e.mlt_last, hipMemcpyDeviceToHost, e.stream));
        HIP_OK(hipStreamSynchronize(e.stream));
        const uint32_t c = *e.h_count;
        if (c <= e.capacity) {
            count = c;
            return true;
        }
        if (!ensure_photon_buffers(e, (uint32_t)std::min<double>((double)c * 1.25 + 1024, 4.0e9))) return false;
    }
    set_error("photon buffer overflow");
    return false;
}

// keys_ready: the light pass just made the keys (run_light_pass), else the
// photons are a caller's and the map build makes them.
bool build_map(Engine& e, uint32_t count, float radius, float r2, MapInfo& mi, bool keys_ready = false) {
    MapGrid g;
    if (!map_grid(e, radius, r2, g)) return false;
    const int uid_bits = g.uid_bits, bucket_bits = g.bucket_bits;
    MapBuildArgs a{};
    a.in_rec = e.ph_rec.p;
    a.count = count;
    for (int k = 0; k < 3; ++k) a.origin[k] = g.pm.origin[k];
    a.inv_cell = g.pm.inv_cell;
    a.mask = g.pm.mask;
    a.bucket_bits = bucket_bits;
    a.uid_bits = uid_bits;
    a.keys_a = e.keys_a.p;
    a.keys_b = e.keys_b.p;
    a.vals_a = e.vals_a.p;
    a.vals_b = e.vals_b.p;
    a.hist = e.hist.p;
    a.hist_sums = e.hist_sums.p;
    a.out_pos = e.s_pos.p;
    a.out_dir = e.s_dir.p;
    a.out_tp = e.s_tp.p;
    a.out_index = e.s_index.p;
    a.cells = e.cells.p;
    a.keys_ready = keys_ready;
    a.scratch_n = keys_ready ? e.capacity : 0u;
    e.sort_passes = (bucket_bits + uid_bits + 7) / 8;
    HIP_OK(hipEventRecord(e.ev[2], e.stream));
    build_photon_map(a, e.stream);
    HIP_OK(hipEventRecord(e.ev[3], e.stream));
    HIP_OK(hipGetLastError());
    mi.pm = g.pm;
    return true;
}

bool ensure_query_buffers(Engine& e, uint32_t cap, uint32_t n_paths) {
    if (!e.path_info.alloc(n_paths) || !e.contrib.alloc((size_t)n_paths * 8 * 3)) return false;
    if (e.query_order && !(e.qk_a.alloc(std::max(cap, e.q_capacity)) && e.qk_b.alloc(std::max(cap, e.q_capacity)) &&
                           e.qv_a.alloc(std::max(cap, e.q_capacity)) && e.qv_b.alloc(std::max(cap, e.q_capacity)) &&
                           e.q_hist.alloc(photon_map_hist_entries(std::max(cap, e.q_capacity))) &&
                           e.q_hist_sums.alloc(photon_map_sum_entries(std::max(cap, e.q_capacity)))))
        return false;
    if (cap <= e.q_capacity && e.q_pos.p) return true;
    e.q_capacity = cap;
    return e.q_pos.alloc(cap) && e.q_dir.alloc(cap) && e.q_sn.alloc(cap) && e.q_gn.alloc(cap) && e.q_tpn.alloc(cap);
}

// Spatial order of this pass's queries (option query_order), or null.
const uint32_t* query_order(Engine& e, const CameraPassArgs& a, uint32_t n, hipStream_t s) {
    if (!e.query_order) return nullptr;
    QuerySortBuffers qb{e.qk_a.p, e.qk_b.p, e.qv_a.p, e.qv_b.p, e.q_hist.p, e.q_hist_sums.p};
    return sort_queries(a, qb, std::min(n, e.q_capacity), s);
}

// Camera pass: trace (queries), merge, resolve into `image`.
CameraPassArgs camera_args(Engine& e, int iteration, float norm, float* image) {
    uint32_t n_paths = (uint32_t)((long long)e.W * e.H);
    CameraPassArgs a{};
    a.sc = scene_view(e);
    a.k0 = (uint32_t)e.seed;
    a.k1 = (uint32_t)(e.seed >> 32);
    a.iteration = (uint32_t)iteration;
    a.stream = STREAM_CAMERA;
    a.n_paths = n_paths;
    a.path_chunk = e.path_chunk;
    a.refill_min = e.camera_refill;
    a.path_counter = e.counters.p;
    a.vm_norm = norm;
    a.path_info = e.path_info.p;
    a.contrib = e.contrib.p;
    a.image = image;
    a.counters = e.count_events ? e.stats_dev.p + 1 : nullptr;
    a.exact_merge = e.exact_merge;
    a.group_queries = e.group_queries;
    a.gather_tile = e.gather_tile;
    a.gather_block = e.gather_block;
    a.camera_qmc = e.camera_qmc;
    a.q = QueryBuffers{e.q_pos.p, e.q_dir.p, e.q_sn.p, e.q_gn.p, e.q_tpn.p, e.q_capacity};
    return a;
}

// Overlapped camera pass, part 1: the camera trace on stream2 with its own
// path counters, its query count copied to pinned memory.  It starts after
// the work queued on stream when ev_fork was recorded: here, or (record_fork
// false) by the caller before it queued the light pass (overlap = 2: the
// light pass reaches the GPU first and the trace fills the compute units its
// tail and the photon-map build leave free).
bool camera_trace_async(Engine& e, int iteration, float norm, float* image, CameraPassArgs& a,
                        bool record_fork = true) {
    uint32_t n_paths = (uint32_t)((long long)e.W * e.H);
    uint32_t cap = std::max(e.q_capacity, std::max<uint32_t>(2 * n_paths, 1024));
    if (!ensure_query_buffers(e, cap, n_paths) || !ensure_stream2(e)) return false;
    a = camera_args(e, iteration, norm, image);
    a.path_counter = e.cam_counters.p;
    if (record_fork) HIP_OK(hipEventRecord(e.ev_fork, e.stream));
    HIP_OK(hipStreamWaitEvent(e.stream2, e.ev_fork, 0));
    launch_fill_u32(e.cam_counters.p, 0u, kPathCounters, e.stream2);
    HIP_OK(hipEventRecord(e.ev[4], e.stream2));
    launch_camera_trace(a, camera_blocks(e), e.stream2);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(e.h_nq, e.cam_counters.p + 1, 4, hipMemcpyDeviceToHost, e.stream2));
    HIP_OK(hipEventRecord(e.ev[8], e.stream2));
    HIP_OK(hipEventRecord(e.ev_cam, e.stream2));
    return true;
}

// Part 1b, queued once the light pass is: the query sort after the trace on
// stream2, sized for the buffer and reading the count from device memory.
// Its keys need only this iteration's map grid, which the host knows before
// any photon exists.
bool queue_query_sort(Engine& e, CameraPassArgs& a, float radius, float r2) {
    MapGrid g;
    if (!map_grid(e, radius, r2, g)) return false;
    a.pm = g.pm;
    HIP_OK(hipEventRecord(e.ev_join, e.stream2));
    a.q_order = query_order(e, a, e.q_capacity, e.stream2);
    HIP_OK(hipEventRecord(e.ev_qdone, e.stream2));
    return true;
}

// Part 2, once the photon map is queued on stream: the query count (a trace
// that overflowed the query buffer is traced again, serially, into larger
// buffers, and its queries sorted on stream), then stream waits for the query
// sort and runs the merge and the resolve.
bool camera_finish(Engine& e, int iteration, const MapInfo& mi, float norm, float* image, CameraPassArgs a) {
    HIP_OK(hipEventSynchronize(e.ev_cam));
    uint32_t nq = *e.h_nq;
    const bool overflow = nq > e.q_capacity;
    if (overflow) {
        HIP_OK(hipStreamSynchronize(e.stream2));  // the sort of the overflowed trace, before its buffers move
        uint32_t n_paths = (uint32_t)((long long)e.W * e.H);
        if (!ensure_query_buffers(e, (uint32_t)std::min<double>(nq * 1.25 + 1024, 4.0e9), n_paths)) {
            set_error("merge query buffer overflow");
            return false;
        }
        a = camera_args(e, iteration, norm, image);
        a.path_counter = e.cam_counters.p;
        launch_fill_u32(e.cam_counters.p, 0u, kPathCounters, e.stream);
        HIP_OK(hipEventRecord(e.ev[4], e.stream));
        launch_camera_trace(a, camera_blocks(e), e.stream);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(&nq, e.cam_counters.p + 1, 4, hipMemcpyDeviceToHost, e.stream));
        HIP_OK(hipStreamSynchronize(e.stream));
        HIP_OK(hipEventRecord(e.ev[8], e.stream));
        if (nq > e.q_capacity) {
            set_error("merge query buffer overflow");
            return false;
        }
    }
    e.stats[RC_STAT_GATHERS] = nq;
    e.last_queries = nq;
    a.pm = mi.pm;
    if (overflow) {
        HIP_OK(hipEventRecord(e.ev_join, e.stream));
        a.q_order = query_order(e, a, e.last_queries, e.stream);
        HIP_OK(hipEventRecord(e.ev_qdone, e.stream));
    }
    HIP_OK(hipStreamWaitEvent(e.stream, e.ev_qdone, 0));  // after the trace too (stream2's order)
    HIP_OK(hipEventRecord(e.ev[6], e.stream));
    launch_gather(a, gather_blocks(e), e.stream);
    HIP_OK(hipEventRecord(e.ev[7], e.stream));
    launch_resolve(a, e.stream);
    HIP_OK(hipEventRecord(e.ev[5], e.stream));
    HIP_OK(hipGetLastError());
    return true;
}

bool run_camera_pass(Engine& e, int iteration, const MapInfo& mi, float norm, float* image) {
    uint32_t n_paths = (uint32_t)((long long)e.W * e.H);
    uint32_t cap = std::max(e.q_capacity, std::max<uint32_t>(2 * n_paths, 1024));
    if (!ensure_query_buffers(e, cap, n_paths)) return false;
    CameraPassArgs a{};
    a.sc = scene_view(e);
    a.pm = mi.pm;
    a.k0 = (uint32_t)e.seed;
    a.k1 = (uint32_t)(e.seed >> 32);
    a.iteration = (uint32_t)iteration;
    a.stream = STREAM_CAMERA;
    a.n_paths = n_paths;
    a.path_chunk = e.path_chunk;
    a.refill_min = e.camera_refill;
    a.path_counter = e.counters.p;
    a.vm_norm = norm;
    a.path_info = e.path_info.p;
    a.contrib = e.contrib.p;
    a.image = image;
    a.counters = e.count_events ? e.stats_dev.p + 1 : nullptr;
    a.exact_merge = e.exact_merge;
    a.group_queries = e.group_queries;
    a.gather_tile = e.gather_tile;
    a.gather_block = e.gather_block;
    a.camera_qmc = e.camera_qmc;
    a.merge_sets = e.count_events ? e.merge_sets : nullptr;
    a.map_index = e.s_index.p;
    for (int attempt = 0; attempt < 2; ++attempt) {
        a.q = QueryBuffers{e.q_pos.p, e.q_dir.p, e.q_sn.p, e.q_gn.p, e.q_tpn.p, e.q_capacity};
        launch_fill_u32(e.counters.p, 0u, kPathCounters, e.stream);
        HIP_OK(hipEventRecord(e.ev[4], e.stream));
        launch_camera_trace(a, camera_blocks(e), e.stream);
        HIP_OK(hipGetLastError());
        uint32_t nq = 0;
        HIP_OK(hipMemcpyAsync(&nq, e.counters.p + 1, 4, hipMemcpyDeviceToHost, e.stream));
        HIP_OK(hipStreamSynchronize(e.stream));
        e.stats[RC_STAT_GATHERS] = nq;
        e.last_queries = nq;
        if (nq <= e.q_capacity) break;
        if (attempt == 1 || !ensure_query_buffers(e, (uint32_t)std::min<double>(nq * 1.25 + 1024, 4.0e9), n_paths)) {
            set_error("merge query buffer overflow");
            return false;
        }
        if (e.count_events) launch_fill_u32(reinterpret_cast<uint32_t*>(e.stats_dev.p + 1), 0u, kCounterFields * 2, e.stream);
    }
    HIP_OK(hipEventRecord(e.ev[8], e.stream));
    a.q_order = query_order(e, a, e.last_queries, e.stream);
    HIP_OK(hipEventRecord(e.ev[6], e.stream));
    launch_gather(a, gather_blocks(e), e.stream);
    HIP_OK(hipEventRecord(e.ev[7], e.stream));
    launch_resolve(a, e.stream);
    HIP_OK(hipEventRecord(e.ev[5], e.stream));
    HIP_OK(hipGetLastError());
    return true;
}

// One training step of the emission proposal on this iteration's first
// paths: camera-visible merge cells from the queries, per-path visible
// luminance, forward, backward, Adam, then the table for the next iteration.
bool guide_iteration(Engine& e, uint32_t photon_count, const MapInfo& mi) {
    GuideModel& gd = e.guide;
    const int n = guide_rows(e);
    if (n <= 0) return true;
    hipStream_t s = e.stream;
    const PhotonMapView& pm = mi.pm;
    size_t buckets = (size_t)pm.mask + 1;
    if (!gd.vis.alloc(buckets + 3)) return false;
    launch_fill_u32(reinterpret_cast<uint32_t*>(gd.vis.p), 0u, (buckets + 3) / 4, s);
    launch_mark_visible(e.q_pos.p, std::min(e.last_queries, e.q_capacity), pm.origin[0], pm.origin[1], pm.origin[2],
                        pm.inv_cell, pm.mask, gd.vis.p, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(gd.train_w.p), 0u, (size_t)n, s);
    launch_guide_targets(e.ph_rec.p, photon_count, pm.origin[0], pm.origin[1],
                         pm.origin[2], pm.inv_cell, pm.mask, gd.vis.p, (uint32_t)n, gd.train_w.p, s);
    HIP_OK(hipEventRecord(e.ev_guide[2], s));
    bool ok = gd.forward(gd.train_x.p, n, s, false) && gd.backward(gd.train_bin.p, gd.train_w.p, n, s) &&
              gd.adam(s);
    HIP_OK(hipEventRecord(e.ev_guide[3], s));
    return ok && gd.build_table(s);
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// The launches `body` queues on st, captured and instantiated into `out`.
bool capture_graph(hipStream_t st, const std::function<void()>& body, hipGraphExec_t& out) {
    hipGraph_t graph = nullptr;
    HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    body();
    hipError_t cap = hipStreamEndCapture(st, &graph);
    if (cap != hipSuccess || !graph) {
        set_error(std::string("hipGraph capture failed: ") + hipGetErrorString(cap));
        return false;
    }
    hipError_t inst = hipGraphInstantiate(&out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (inst != hipSuccess) {
        out = nullptr;
        set_error(std::string("hipGraphInstantiate failed: ") + hipGetErrorString(inst));
        return false;
    }
    return true;
}

// The iteration replayed from captured hipGraphs.  Inside a graph every count
// comes from device memory and the per-iteration values (iteration, radius,
// grid, VM normalisation) from d_iter, copied in before each replay; the
// query buffer holds the most queries a camera pass can make (8 per path).
// Plain light pass, no overlap: the whole iteration is one graph -- a copy of
// the film, the light pass, the photon map, the camera pass and the resolve
// -- and the host waits once, after it, for the photon count; when the
// photons overflowed their buffer the film is restored from the copy, the
// buffers grow (a new capture) and the iteration runs again.  Otherwise the
// light pass stays a direct launch (MLT and the learned proposal set
// per-iteration host values), its photon count is read back and checked
// before the replay of the photon map and the camera pass.  With option
// overlap the camera trace is a direct launch on stream2 beside the light
// pass, a graph of the query sort replays behind it on stream2, and on stream
// a graph of the photon map replays after the light pass, then, once stream
// has waited for the sort, a graph of the merge and the resolve.
bool graph_iteration(Engine& e, int iteration, uint32_t& count, MapInfo& mi) {
    hipStream_t st = e.stream;
    float radius, r2, norm;
    iteration_params(e, iteration, radius, r2, norm);
    const uint32_t n_paths = (uint32_t)((long long)e.W * e.H);
    if (!ensure_query_buffers(e, (uint32_t)std::min<uint64_t>(8ull * n_paths + 1024, 1ull << 31), n_paths)) return false;
    long long N = light_paths(e);
    // the grid of this iteration, as build_map
    int uid_bits = log2_ceil((uint64_t)N);  // (bucket, light path) keys, as map_grid_params
    uint64_t want = std::max<uint64_t>(4096, std::min<uint64_t>(1ull << 24, (uint64_t)N / 2));
    int bucket_bits = log2_ceil(want);
    uint32_t buckets = 1u << bucket_bits;
    if (!e.cells.alloc(buckets)) return false;
    float maxc = 0.0f;
    for (int k = 0; k < 3; ++k) maxc = std::max(maxc, std::max(std::fabs(e.bbox.lo[k]), std::fabs(e.bbox.hi[k])));
    float rq = radius + std::max(radius * 1e-3f, 8.0f * maxc * 1.1920929e-07f);
    float cell = rq * (float)std::max(0.67, e.cell_scale);
    if (!e.d_iter.alloc(1)) return false;
    if (!e.h_iter) HIP_OK(hipHostMalloc((void**)&e.h_iter, sizeof(IterParams), hipHostMallocDefault));
    IterParams& ip = *e.h_iter;
    ip.iteration = (uint32_t)iteration;
    ip.r2 = r2;
    ip.r_query = rq;
    ip.inv_cell = 1.0f / cell;
    for (int k = 0; k < 3; ++k) ip.origin[k] = e.bbox.lo[k] - cell;
    // copied in before the overlapped trace forks (the query sort behind it reads the grid), else after the
    // light pass (the MLT light pass sets the normalisation)
    auto copy_iter = [&]() -> bool {
        ip.vm_norm = norm * mlt_factor(e);
        HIP_OK(hipMemcpyAsync(e.d_iter.p, e.h_iter, sizeof(IterParams), hipMemcpyHostToDevice, st));
        return true;
    };
    const bool overlapped = e.overlap && !e.mlt;
    const bool whole = !e.mlt && !guide_active(e);
    CameraPassArgs ca{};
    if (overlapped && !whole && !(copy_iter() && camera_trace_async(e, iteration, norm, e.image.p, ca))) return false;
    if (overlapped && whole && !ensure_stream2(e)) return false;
    const size_t film = (size_t)e.W * e.H * 3;
    if (whole) {  // the buffers the captured light pass writes, sized as run_light_pass sizes them
        if (N <= 0 || N > (1ll << 31)) {
            set_error("light path count out of range (1 .. 2^31)");
            return false;
        }
        const uint32_t cap = e.photon_capacity > 0 ? (uint32_t)e.photon_capacity : (uint32_t)std::max<long long>(N, 1 << 16);
        if (!ensure_photon_buffers(e, std::max(cap, e.capacity)) || !e.film_prev.alloc(film) ||
            !e.graph_count.alloc(1) || (e.light_order && !light_order_buffers(e, (size_t)N)) ||
            (e.light_emit && !e.emitted.alloc(3 * (size_t)N)))
            return false;
        if (!e.h_count) HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&e.h_count), sizeof(uint32_t), hipHostMallocDefault));
        if (!copy_iter()) return false;
    } else {
        if (!run_light_pass(e, iteration, count)) return false;
        if (!overlapped && !copy_iter()) return false;
    }
    for (int attempt = 0;; ++attempt) {
    PhotonMapView& pm = mi.pm;
    pm.pos = e.s_pos.p;
    pm.dir_tr = e.s_dir.p;
    pm.tp_gb = e.s_tp.p;
    pm.cells = e.cells.p;
    pm.mask = buckets - 1;
    for (int k = 0; k < 3; ++k) pm.origin[k] = ip.origin[k];
    pm.inv_cell = ip.inv_cell;
    pm.r2 = r2;
    pm.r_query = rq;
    std::vector<uintptr_t> key = {(uintptr_t)e.ph_rec.p, e.capacity, (uintptr_t)e.q_pos.p, e.q_capacity,
                                  (uintptr_t)e.cells.p, buckets, (uintptr_t)uid_bits, (uintptr_t)e.image.p,
                                  (uintptr_t)e.path_info.p, (uintptr_t)e.contrib.p, (uintptr_t)e.path_chunk,
                                  (uintptr_t)gather_blocks(e), (uintptr_t)camera_blocks(e),
                                  (uintptr_t)e.exact_merge, (uintptr_t)e.query_order, (uintptr_t)e.group_queries,
                                  (uintptr_t)e.camera_refill, (uintptr_t)e.camera_qmc, (uintptr_t)e.gather_tile,
                                  (uintptr_t)e.gather_block, (uintptr_t)(overlapped ? e.overlap : 0), (uintptr_t)whole,
                                  (uintptr_t)light_blocks(e), (uintptr_t)e.light_refill, (uintptr_t)e.light_order,
                                  (uintptr_t)e.light_emit, (uintptr_t)e.emitted.p, (uintptr_t)e.film_prev.p,
                                  (uintptr_t)e.lk_a.p};
    if (!e.graph_exec || key != e.graph_key) {
        if (e.graph_exec) HIP_OK(hipGraphExecDestroy(e.graph_exec));
        if (e.graph_exec2) HIP_OK(hipGraphExecDestroy(e.graph_exec2));
        if (e.graph_exec3) HIP_OK(hipGraphExecDestroy(e.graph_exec3));
        e.graph_exec = e.graph_exec2 = e.graph_exec3 = nullptr;
        MapBuildArgs m{};
        m.in_rec = e.ph_rec.p;
        m.count = e.capacity;
        m.count_dev = e.counters.p + 1;
        m.dyn = e.d_iter.p;
        m.mask = buckets - 1;
        m.bucket_bits = bucket_bits;
        m.uid_bits = uid_bits;
        m.keys_a = e.keys_a.p;
        m.keys_b = e.keys_b.p;
        m.vals_a = e.vals_a.p;
        m.vals_b = e.vals_b.p;
        m.hist = e.hist.p;
        m.hist_sums = e.hist_sums.p;
        m.out_pos = e.s_pos.p;
        m.out_dir = e.s_dir.p;
        m.out_tp = e.s_tp.p;
        m.out_index = e.s_index.p;
        m.cells = e.cells.p;
        m.keys_ready = true;  // by the light pass (in the graph, or run_light_pass before it)
        m.scratch_n = e.capacity;
        CameraPassArgs a{};
        a.sc = scene_view(e);
        a.pm = pm;
        a.k0 = (uint32_t)e.seed;
        a.k1 = (uint32_t)(e.seed >> 32);
        a.stream = STREAM_CAMERA;
        a.n_paths = n_paths;
        a.path_chunk = e.path_chunk;
        a.refill_min = e.camera_refill;
        a.path_counter = overlapped ? e.cam_counters.p : e.counters.p;
        a.path_info = e.path_info.p;
        a.contrib = e.contrib.p;
        a.image = e.image.p;
        a.exact_merge = e.exact_merge;
        a.group_queries = e.group_queries;
        a.gather_tile = e.gather_tile;
        a.gather_block = e.gather_block;
        a.camera_qmc = e.camera_qmc;
        a.dyn = e.d_iter.p;
        a.q = QueryBuffers{e.q_pos.p, e.q_dir.p, e.q_sn.p, e.q_gn.p, e.q_tpn.p, e.q_capacity};
        auto sort = [&]() { a.q_order = query_order(e, a, e.q_capacity, st); };  // count read from the device
        auto merge = [&]() {
            launch_gather(a, gather_blocks(e), st);
            launch_resolve(a, st);
        };
        auto camera = [&]() {
            sort();
            merge();
        };
        bool ok;
        bool forked = true;  // the stream2 branch of a captured overlapped iteration was joined
        if (overlapped && !whole) {
            ok = capture_graph(st, [&]() { build_photon_map(m, st); }, e.graph_exec) &&
                 capture_graph(st, sort, e.graph_exec3) && capture_graph(st, merge, e.graph_exec2);
        } else if (whole) {
            ok = capture_graph(st,
                               [&]() {
                                   launch_copy_u32(reinterpret_cast<uint32_t*>(e.film_prev.p),
                                                   reinterpret_cast<const uint32_t*>(e.image.p), film, st);
                                   // the camera trace and the query sort on stream2, beside the light pass
                                   // and the photon map (a fork and a join inside the one graph); with
                                   // overlap 2 the branch is captured after the light pass, so the graph
                                   // holds its nodes after the light pass's
                                   auto branch = [&]() {
                                       forked = hipStreamWaitEvent(e.stream2, e.ev_fork, 0) == hipSuccess;
                                       if (forked) {
                                           launch_fill_u32(e.cam_counters.p, 0u, kPathCounters, e.stream2);
                                           launch_camera_trace(a, camera_blocks(e), e.stream2);
                                           a.q_order = query_order(e, a, e.q_capacity, e.stream2);
                                           forked = hipEventRecord(e.ev_gjoin, e.stream2) == hipSuccess;
                                       }
                                   };
                                   if (overlapped) {
                                       forked = hipEventRecord(e.ev_fork, st) == hipSuccess;
                                       if (forked && e.overlap != 2) branch();
                                   }
                                   launch_fill_u32(e.counters.p, 0u, 4, st);
                                   clear_map_sort(e.hist.p, e.capacity, st);
                                   LightPassArgs la = light_pass_args(e, iteration);
                                   la.dyn = e.d_iter.p;
                                   MapGrid grid;
                                   grid.pm.mask = buckets - 1;
                                   grid.uid_bits = uid_bits;
                                   la.map = map_keys_of(e, grid);  // origin and inv_cell from dyn
                                   if (e.light_emit) la.emitted = e.emitted.p;
                                   la.path_order = light_path_order(e, la, st);
                                   launch_light_pass(la, light_blocks(e), st);
                                   if (overlapped && forked && e.overlap == 2) branch();
                                   launch_copy_u32(e.graph_count.p, e.counters.p + 1, 1, st);
                                   build_photon_map(m, st);
                                   if (overlapped) {
                                       if (forked) forked = hipStreamWaitEvent(st, e.ev_gjoin, 0) == hipSuccess;
                                       merge();
                                   } else {
                                       launch_fill_u32(e.counters.p, 0u, kPathCounters, st);
                                       launch_camera_trace(a, camera_blocks(e), st);
                                       camera();
                                   }
                               },
                               e.graph_exec) &&
                 forked;
        } else {
            ok = capture_graph(st,
                               [&]() {
                                   build_photon_map(m, st);
                                   launch_fill_u32(e.counters.p, 0u, kPathCounters, st);
                                   launch_camera_trace(a, camera_blocks(e), st);
                                   camera();
                               },
                               e.graph_exec);
        }
        if (!ok) {
            if (e.graph_exec) (void)hipGraphExecDestroy(e.graph_exec);
            if (e.graph_exec3) (void)hipGraphExecDestroy(e.graph_exec3);
            e.graph_exec = e.graph_exec3 = nullptr;
            return false;
        }
        e.graph_key = key;
        e.graph_captures++;
    }
    if (overlapped && !whole) {  // the query sort behind the trace
        HIP_OK(hipGraphLaunch(e.graph_exec3, e.stream2));
        HIP_OK(hipEventRecord(e.ev_qdone, e.stream2));
    }
    if (whole) {  // no separate light pass: its span is empty
        HIP_OK(hipEventRecord(e.ev[0], st));
        HIP_OK(hipEventRecord(e.ev[1], st));
    }
    HIP_OK(hipEventRecord(e.ev[2], st));
    HIP_OK(hipGraphLaunch(e.graph_exec, st));
    if (overlapped && !whole) {
        HIP_OK(hipStreamWaitEvent(st, e.ev_qdone, 0));
        HIP_OK(hipGraphLaunch(e.graph_exec2, st));
    }
    HIP_OK(hipEventRecord(e.ev[5], st));
    HIP_OK(hipMemcpyAsync(&e.last_queries, (overlapped ? e.cam_counters.p : e.counters.p) + 1, 4,
                          hipMemcpyDeviceToHost, st));
    if (whole) HIP_OK(hipMemcpyAsync(e.h_count, e.graph_count.p, 4, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    e.stats[RC_STAT_GATHERS] = e.last_queries;
    if (!whole || *e.h_count <= e.capacity) break;
    if (attempt > 0) {
        set_error("photon buffer overflow");
        return false;
    }
    // the photons overflowed their buffer: this iteration's film contribution is undone and the
    // iteration runs again with larger buffers (a new capture)
    launch_copy_u32(reinterpret_cast<uint32_t*>(e.image.p), reinterpret_cast<const uint32_t*>(e.film_prev.p), film,
                    st);
    if (!ensure_photon_buffers(e, (uint32_t)std::min<double>((double)*e.h_count * 1.25 + 1024, 4.0e9))) return false;
    }
    if (whole) count = *e.h_count;
    return true;
}

bool render_iteration(Engine& e, int iteration) {
    float radius, r2, norm;
    iteration_params(e, iteration, radius, r2, norm);
    auto t0 = std::chrono::steady_clock::now();
    uint32_t count = 0;
    MapInfo mi;
    const bool graphed = e.graph && !e.count_events;
    bool overlapped = false;
    if (graphed) {
        if (!graph_iteration(e, iteration, count, mi)) return false;
    } else {
        // the camera trace needs neither the photons nor the light pass's results
        // (plain estimator), so it can run beside them
        overlapped = e.overlap && !e.count_events && !e.mlt;
        CameraPassArgs ca{};
        const bool light_first = overlapped && e.overlap == 2;
        if (light_first) {  // the trace's fork point, before the light pass is queued
            if (!ensure_stream2(e)) return false;
            HIP_OK(hipEventRecord(e.ev_fork, e.stream));
        }
        if (overlapped && !light_first && !camera_trace_async(e, iteration, norm, e.image.p, ca)) return false;
        auto sort_queries_next = [&]() {
            if (light_first && !camera_trace_async(e, iteration, norm, e.image.p, ca, false)) return false;
            return queue_query_sort(e, ca, radius, r2);
        };
        if (!run_light_pass(e, iteration, count, overlapped ? sort_queries_next : std::function<bool()>()))
            return false;
        if (!build_map(e, count, radius, r2, mi, true)) return false;
        if (overlapped) {
            if (!camera_finish(e, iteration, mi, norm, e.image.p, ca)) return false;
        } else if (!run_camera_pass(e, iteration, mi, norm * mlt_factor(e), e.image.p)) {
            return false;
        }
    }
    const bool training = guide_trains(e, iteration);
    e.stats[RC_STAT_MS_GUIDE] = e.stats[RC_STAT_MS_GUIDE_STEP] = e.stats[RC_STAT_GUIDE_ROWS] = 0;
    if (training) {
        HIP_OK(hipEventRecord(e.ev_guide[0], e.stream));
        if (!guide_iteration(e, count, mi)) return false;
        HIP_OK(hipEventRecord(e.ev_guide[1], e.stream));
        HIP_OK(hipMemcpyAsync(e.guide.last, e.guide.scal.p, sizeof e.guide.last, hipMemcpyDeviceToHost, e.stream));
    }
    Counters c[2] = {};
    if (e.count_events) HIP_OK(hipMemcpyAsync(c, e.stats_dev.p, sizeof c, hipMemcpyDeviceToHost, e.stream));
    HIP_OK(hipStreamSynchronize(e.stream));
    float ms[3] = {0, 0, 0};
    float ms_trace = 0, ms_gather = 0, ms_resolve = 0, ms_sort = 0;
    if (graphed) {  // map + camera pass are one replay: its time goes to ms_camera
        HIP_OK(hipEventElapsedTime(&ms[0], e.ev[0], e.ev[1]));
        HIP_OK(hipEventElapsedTime(&ms[2], e.ev[2], e.ev[5]));
    } else {
        for (int k = 0; k < 3; ++k) HIP_OK(hipEventElapsedTime(&ms[k], e.ev[2 * k], e.ev[2 * k + 1]));
        HIP_OK(hipEventElapsedTime(&ms_trace, e.ev[4], e.ev[8]));
        if (overlapped) HIP_OK(hipEventElapsedTime(&ms_sort, e.ev_join, e.ev_qdone));
        else HIP_OK(hipEventElapsedTime(&ms_sort, e.ev[8], e.ev[6]));
        HIP_OK(hipEventElapsedTime(&ms_gather, e.ev[6], e.ev[7]));
        HIP_OK(hipEventElapsedTime(&ms_resolve, e.ev[7], e.ev[5]));
    }
    e.stats[RC_STAT_MS_CAMERA_TRACE] = ms_trace;
    e.stats[RC_STAT_MS_GATHER] = ms_gather;
    e.stats[RC_STAT_MS_RESOLVE] = ms_resolve;
    e.stats[RC_STAT_MS_QUERY_SORT] = ms_sort;
    // overlapped: the span from the trace's start to the resolve's end covers
    // the light pass and the map build too, so the camera pass is its own spans
    if (overlapped && !graphed) ms[2] = ms_trace + ms_sort + ms_gather + ms_resolve;
    if (training) {
        float ms_guide = 0;
        HIP_OK(hipEventElapsedTime(&ms_guide, e.ev_guide[0], e.ev_guide[1]));
        int n = std::max(1, guide_rows(e));
        e.stats[RC_STAT_MS_GUIDE] = ms_guide;
        float ms_step = 0;
        HIP_OK(hipEventElapsedTime(&ms_step, e.ev_guide[2], e.ev_guide[3]));
        e.stats[RC_STAT_MS_GUIDE_STEP] = ms_step;
        e.stats[RC_STAT_GUIDE_ROWS] = guide_rows(e);
        e.stats[RC_STAT_GUIDE_VISIBLE] = e.guide.last[0] / n;
        e.stats[RC_STAT_GUIDE_HIT_FRACTION] = e.guide.last[1] / n;
        e.stats[RC_STAT_GUIDE_LOSS] = e.guide.last[2];
    }
    e.stats[RC_STAT_LIGHT_PATHS] = (double)light_paths(e);
    e.stats[RC_STAT_MLT_VISIBLE] = e.mlt ? e.mlt_last[0] : 0.0;
    e.stats[RC_STAT_MLT_MUTATIONS] = e.mlt ? e.mlt_last[1] : 0.0;
    e.stats[RC_STAT_MLT_ACCEPTED] = e.mlt ? e.mlt_last[2] : 0.0;
    e.stats[RC_STAT_PHOTONS] = count;
    e.stats[RC_STAT_CAMERA_PATHS] = (double)e.W * e.H;
    e.stats[RC_STAT_MS_LIGHT] = ms[0];
    e.stats[RC_STAT_MS_MAP] = ms[1];
    e.stats[RC_STAT_MS_CAMERA] = ms[2];
    e.stats[RC_STAT_MS_TOTAL] = ms_since(t0);
    e.stats[RC_STAT_RADIUS] = radius;
    e.stats[RC_STAT_NODES] = e.n_nodes;
    e.stats[RC_STAT_TRIS] = (double)e.n_tris;
    e.stats[RC_STAT_SORT_PASSES] = e.sort_passes;
    e.stats[RC_STAT_GRAPH_CAPTURES] = e.graph_captures;
    if (e.count_events) {
        e.stats[RC_STAT_LIGHT_NODE_VISITS] = (double)c[0].nodes;
        e.stats[RC_STAT_LIGHT_TRI_TESTS] = (double)c[0].tris;
        e.stats[RC_STAT_LIGHT_HITS] = (double)c[0].hits;
        e.stats[RC_STAT_CAMERA_NODE_VISITS] = (double)c[1].nodes;
        e.stats[RC_STAT_CAMERA_TRI_TESTS] = (double)c[1].tris;
        e.stats[RC_STAT_CAMERA_HITS] = (double)c[1].hits;
        e.stats[RC_STAT_CELL_VISITS] = (double)c[1].cells;
        e.stats[RC_STAT_GATHERS] = (double)c[1].gathers;
        e.stats[RC_STAT_PHOTONS_TESTED] = (double)c[1].photons_tested;
        e.stats[RC_STAT_PHOTONS_ACCEPTED] = (double)c[1].photons_accepted;
        e.stats[RC_STAT_PHOTONS_STREAMED] = (double)c[1].photons_streamed;
    }
    return true;
}

// ---------------------------------------------------------------------
// Autotuner of the launch shapes and queue / tile sizes: every candidate
// renders the same iteration (same photons, same queries) into a scratch
// film; each knob keeps the value with the smallest time of the kernels it
// affects (best of `repeats`), one knob after the other.  The last knob,
// overlap, is judged by the iteration's wall time: the camera trace beside
// the light pass shortens the iteration only while the light pass leaves
// compute units idle, and it is kept only when it saves at least 2 %, more
// than the run-to-run spread of the iteration's wall time (beside the trace
// the light pass's own time no longer measures that kernel alone).
// ---------------------------------------------------------------------
enum { T_LIGHT = 0, T_MAP, T_TRACE, T_GATHER, T_WALL, T_PARTS };

struct TuneKnob {
    const char* name;
    std::vector<double> candidates;
    std::function<void(double)> set;
    std::function<double()> get;
    std::function<double(const double*)> cost;
    double margin = 0.0;  // a later candidate must beat the best so far by this fraction
};

// overlap_candidates: {0, 1} or {0, 2, 1}, tried after the other knobs (timed with the camera trace serial);
// a later candidate must beat the best so far by 2 %, so the order breaks near-ties (2 before 1: the light
// pass first wins by more when the light pass is long).
std::vector<TuneKnob> tune_knobs(std::function<Engine&()> launch_owner, std::function<void(double)> set_cell_scale,
                                 std::function<double()> get_cell_scale, std::vector<double> overlap_candidates) {
    auto L = [](const double* t) { return t[T_LIGHT]; };
    auto T = [](const double* t) { return t[T_TRACE]; };
    auto G = [](const double* t) { return t[T_GATHER]; };
    return {
        {"path_chunk", {64, 128, 256, 512, 1024},
         [=](double v) { launch_owner().path_chunk = (uint32_t)v; },
         [=]() { return (double)launch_owner().path_chunk; },
         [](const double* t) { return t[T_LIGHT] + t[T_TRACE]; }},
        {"light_blocks_per_cu", {4, 6, 8, 12, 16},
         [=](double v) { launch_owner().light_blocks_per_cu = (int)v; },
         [=]() { return (double)launch_owner().light_blocks_per_cu; }, L},
        {"camera_blocks_per_cu", {4, 6, 8, 12, 16},
         [=](double v) { launch_owner().camera_blocks_per_cu = (int)v; },
         [=]() { return (double)launch_owner().camera_blocks_per_cu; }, T},
        {"cell_scale", {0.67, 0.8, 1.0, 1.25}, set_cell_scale, get_cell_scale,
         [](const double* t) { return t[T_MAP] + t[T_GATHER]; }},
        {"light_refill", {1, 8, 16, 32},
         [=](double v) { launch_owner().light_refill = (uint32_t)v; },
         [=]() { return (double)launch_owner().light_refill; }, L},
        {"light_order", {0, 1},
         [=](double v) { launch_owner().light_order = (int)v; },
         [=]() { return (double)launch_owner().light_order; }, L},
        {"camera_refill", {1, 8, 16, 32},
         [=](double v) { launch_owner().camera_refill = (uint32_t)v; },
         [=]() { return (double)launch_owner().camera_refill; }, T},
        {"group_queries", {4, 8, 16, 1},
         [=](double v) { launch_owner().group_queries = (int)v; },
         [=]() { return (double)launch_owner().group_queries; }, G},
        {"gather_tile", {1024, 256, 0},
         [=](double v) { launch_owner().gather_tile = (uint32_t)v; },
         [=]() { return (double)launch_owner().gather_tile; }, G},
        {"gather_block", {2, 4, 1, 0},
         [=](double v) { launch_owner().gather_block = (int)v; },
         [=]() { return (double)launch_owner().gather_block; }, G},
        {"gather_blocks_per_cu", {5, 8, 12, 16, 24},
         [=](double v) { launch_owner().gather_blocks_per_cu = (int)v; },
         [=]() { return (double)launch_owner().gather_blocks_per_cu; }, G},
        {"query_order", {1, 0},
         [=](double v) { launch_owner().query_order = (int)v; },
         [=]() { return (double)launch_owner().query_order; },
         [](const double* t) { return t[T_TRACE] + t[T_GATHER]; }},
        {"overlap", overlap_candidates,
         [=](double v) { launch_owner().overlap = (int)v; },
         [=]() { return (double)launch_owner().overlap; },
         [](const double* t) { return t[T_WALL]; }, 0.02},
    };
}

// report: [wall ms of the iteration before, after, then the chosen value of each knob]
bool run_autotune(std::vector<TuneKnob> knobs, std::function<bool(double*)> timed, int repeats, double* report) {
    repeats = std::max(1, repeats);
    double t[T_PARTS];
    if (!timed(t)) return false;  // warm-up: buffers, code objects
    auto measure = [&](double* best) -> bool {
        for (int k = 0; k < T_PARTS; ++k) best[k] = std::numeric_limits<double>::infinity();
        for (int r = 0; r < repeats; ++r) {
            if (!timed(t)) return false;
            for (int k = 0; k < T_PARTS; ++k) best[k] = std::min(best[k], t[k]);
        }
        return true;
    };
    double before[T_PARTS];
    if (!measure(before)) return false;
    for (TuneKnob& kb : knobs) {
        double keep = kb.get(), best_cost = std::numeric_limits<double>::infinity(), best_v = keep;
        for (double v : kb.candidates) {
            kb.set(v);
            double m[T_PARTS];
            if (!measure(m)) return false;
            double c = kb.cost(m);
            if (c < best_cost * (1.0 - kb.margin)) {
                best_cost = c;
                best_v = v;
            }
        }
        kb.set(best_v);
    }
    double after[T_PARTS];
    if (!measure(after)) return false;
    report[0] = before[T_WALL];
    report[1] = after[T_WALL];
    for (size_t i = 0; i < knobs.size(); ++i) report[2 + i] = knobs[i].get();
    return true;
}

// One iteration `it` into `film`, without training; kernel times and the
// host's wall time in t.
bool timed_iteration(Engine& e, int it, float* film, double* t) {
    float radius, r2, norm;
    iteration_params(e, it, radius, r2, norm);
    auto w0 = std::chrono::steady_clock::now();
    const bool overlapped = e.overlap && !e.mlt;
    CameraPassArgs ca{};
    const bool light_first = overlapped && e.overlap == 2;
    if (light_first) {
        if (!ensure_stream2(e)) return false;
        HIP_OK(hipEventRecord(e.ev_fork, e.stream));
    }
    if (overlapped && !light_first && !camera_trace_async(e, it, norm, film, ca)) return false;
    uint32_t count = 0;
    auto sort_queries_next = [&]() {
        if (light_first && !camera_trace_async(e, it, norm, film, ca, false)) return false;
        return queue_query_sort(e, ca, radius, r2);
    };
    if (!run_light_pass(e, it, count, overlapped ? sort_queries_next : std::function<bool()>())) return false;
    MapInfo mi;
    if (!build_map(e, count, radius, r2, mi, true)) return false;
    if (overlapped) {
        if (!camera_finish(e, it, mi, norm, film, ca)) return false;
    } else if (!run_camera_pass(e, it, mi, norm * mlt_factor(e), film)) {
        return false;
    }
    HIP_OK(hipStreamSynchronize(e.stream));
    t[T_WALL] = ms_since(w0);
    float ms[4], sort_ms = 0;
    HIP_OK(hipEventElapsedTime(&ms[T_LIGHT], e.ev[0], e.ev[1]));
    HIP_OK(hipEventElapsedTime(&ms[T_MAP], e.ev[2], e.ev[3]));
    if (overlapped) {  // the trace, then the query sort
        HIP_OK(hipEventElapsedTime(&ms[T_TRACE], e.ev[4], e.ev[8]));
        HIP_OK(hipEventElapsedTime(&sort_ms, e.ev_join, e.ev_qdone));
        ms[T_TRACE] += sort_ms;
    } else {
        HIP_OK(hipEventElapsedTime(&ms[T_TRACE], e.ev[4], e.ev[6]));
    }
    HIP_OK(hipEventElapsedTime(&ms[T_GATHER], e.ev[6], e.ev[7]));
    for (int k = 0; k < 4; ++k) t[k] = ms[k];
    return true;
}

bool autotune(Engine& e, int repeats, double* report) {
    if (!prepare(e)) return false;
    DevBuf<float> film;
    if (!film.alloc((size_t)e.W * e.H * 3)) return false;
    const int it = sample_iteration(e, e.iterations_done), counting = e.count_events, overlap = e.overlap;
    e.count_events = 0;
    e.overlap = 0;  // the other knobs are timed with the camera trace serial
    Engine* pe = &e;
    auto knobs = tune_knobs([pe]() -> Engine& { return *pe; }, [pe](double v) { pe->cell_scale = v; },
                            [pe]() { return pe->cell_scale; }, {0, 2, 1});
    bool ok = run_autotune(knobs, [&](double* t) { return timed_iteration(e, it, film.p, t); }, repeats, report);
    e.count_events = counting;
    if (!ok) e.overlap = overlap;
    film.release();
    return ok;
}

// ---------------------------------------------------------------------
// OpenEXR writer with the layout tinyexr's SaveEXR gives the reference's
// image (src/main.cpp:192-207): single-part scanline, FLOAT channels B G R,
// ZIP compression in blocks of 16 scanlines (no compression when both sides
// are below 16 pixels).  A ZIP block is the block's lines (each line: its B,
// G and R pixels) with the bytes split into even and odd halves and then
// delta-coded (OpenEXR's ZIP predictor), deflated into a zlib stream; a block
// whose stream would not be smaller is stored raw (its size then equals the
// raw size, which tells readers it is not compressed).
// ---------------------------------------------------------------------
static void zip_block(const unsigned char* src, size_t n, std::vector<unsigned char>& out) {
    std::vector<unsigned char> t(n);
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) t[(i & 1) ? b++ : a++] = src[i];
    for (size_t i = n; i-- > 1;) t[i] = (unsigned char)(t[i] - t[i - 1] + 128);
    uLongf cap = compressBound((uLong)n);
    out.resize(cap);
    if (compress(out.data(), &cap, t.data(), (uLong)n) != Z_OK || cap >= n) {
        out.assign(src, src + n);
        return;
    }
    out.resize(cap);
}

bool write_exr(const char* path, const float* rgb, int W, int H) {
    std::vector<unsigned char> out;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        out.insert(out.end(), c, c + n);
    };
    auto put_i32 = [&](int32_t v) { put(&v, 4); };
    auto put_u32 = [&](uint32_t v) { put(&v, 4); };
    auto put_f32 = [&](float v) { put(&v, 4); };
    auto put_str = [&](const char* s) { put(s, strlen(s) + 1); };
    auto attr = [&](const char* name, const char* type, uint32_t size) {
        put_str(name);
        put_str(type);
        put_u32(size);
    };
    const bool zip = !(W < 16 && H < 16);
    const int lines = zip ? 16 : 1;  // scanlines per block
    put_u32(20000630u);  // magic
    put_u32(2u);         // version 2, scanline
    attr("channels", "chlist", 3 * (2 + 16) + 1);
    for (const char* ch : {"B", "G", "R"}) {
        put_str(ch);
        put_i32(2);  // FLOAT
        unsigned char z[4] = {0, 0, 0, 0};
        put(z, 4);  // pLinear + reserved
        put_i32(1);
        put_i32(1);  // sampling
    }
    out.push_back(0);
    attr("compression", "compression", 1);
    out.push_back(zip ? 3 : 0);  // ZIP_COMPRESSION or NO_COMPRESSION
    attr("dataWindow", "box2i", 16);
    put_i32(0);
    put_i32(0);
    put_i32(W - 1);
    put_i32(H - 1);
    attr("displayWindow", "box2i", 16);
    put_i32(0);
    put_i32(0);
    put_i32(W - 1);
    put_i32(H - 1);
    attr("lineOrder", "lineOrder", 1);
    out.push_back(0);
    attr("pixelAspectRatio", "float", 4);
    put_f32(1.0f);
    attr("screenWindowCenter", "v2f", 8);
    put_f32(0.0f);
    put_f32(0.0f);
    attr("screenWindowWidth", "float", 4);
    put_f32(1.0f);
    out.push_back(0);  // end of header
    const int blocks = (H + lines - 1) / lines;
    size_t table = out.size();
    out.resize(out.size() + 8 * (size_t)blocks);
    std::vector<unsigned char> raw, packed;
    for (int k = 0; k < blocks; ++k) {
        const int y0 = k * lines, y1 = std::min(H, y0 + lines);
        raw.clear();
        for (int y = y0; y < y1; ++y)
            for (int ch : {2, 1, 0})
                for (int x = 0; x < W; ++x) {
                    const float v = rgb[3 * ((size_t)y * W + x) + ch];
                    const unsigned char* c = (const unsigned char*)&v;
                    raw.insert(raw.end(), c, c + 4);
                }
        if (zip) zip_block(raw.data(), raw.size(), packed);
        const std::vector<unsigned char>& data = zip ? packed : raw;
        uint64_t off = out.size();
        memcpy(&out[table + 8 * (size_t)k], &off, 8);
        put_i32(y0);
        put_u32((uint32_t)data.size());
        put(data.data(), data.size());
    }
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_error(std::string("cannot open ") + path);
        return false;
    }
    size_t w = fwrite(out.data(), 1, out.size(), f);
    fclose(f);
    if (w != out.size()) {
        set_error(std::string("short write to ") + path);
        return false;
    }
    return true;
}

void release(Engine*& e) {
    if (!e) return;
    DevBuf<float4>* f4[] = {&e->d_tri, &e->d_shade, &e->d_nodes, &e->d_nodes4, &e->ph_rec, &e->s_pos, &e->s_dir};
    for (auto* b : f4) b->release();
    e->s_tp.release();
    DevBuf<float4>* q4[] = {&e->q_pos, &e->q_dir, &e->q_sn, &e->q_gn, &e->q_tpn};
    for (auto* b : q4) b->release();
    e->path_info.release();
    e->contrib.release();
    DevBuf<uint32_t>* u[] = {&e->s_index, &e->vals_a, &e->vals_b, &e->hist, &e->hist_sums, &e->counters};
    for (auto* b : u) b->release();
    e->keys_a.release();
    e->keys_b.release();
    for (auto* d : {&e->qk_a, &e->qk_b, &e->lk_a, &e->lk_b}) d->release();
    for (auto* d : {&e->lv_a, &e->lv_b, &e->l_hist, &e->l_hist_sums}) d->release();
    e->mlt_state.release();
    e->emitted.release();
    e->mlt_counts.release();
    for (auto* d : {&e->qv_a, &e->qv_b, &e->q_hist, &e->q_hist_sums}) d->release();
    e->cells.release();
    e->image.release();
    e->stats_dev.release();
    e->d_tri_id.release();
    e->d_mats.release();
    e->d_lights.release();
    e->d_cdf.release();
    for (auto& ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : e->ev_guide)
        if (ev) (void)hipEventDestroy(ev);
    e->guide.release();
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph_exec2) (void)hipGraphExecDestroy(e->graph_exec2);
    if (e->graph_exec3) (void)hipGraphExecDestroy(e->graph_exec3);
    e->d_iter.release();
    if (e->h_iter) (void)hipHostFree(e->h_iter);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    e->cam_counters.release();
    if (e->h_nq) (void)hipHostFree(e->h_nq);
    if (e->h_count) (void)hipHostFree(e->h_count);
    e->film_prev.release();
    e->graph_count.release();
    if (e->ev_cam) (void)hipEventDestroy(e->ev_cam);
    if (e->ev_join) (void)hipEventDestroy(e->ev_join);
    if (e->ev_qdone) (void)hipEventDestroy(e->ev_qdone);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->ev_gjoin) (void)hipEventDestroy(e->ev_gjoin);
    if (e->stream2) (void)hipStreamDestroy(e->stream2);
    delete e;
    e = nullptr;
}

// The user's options (rc_set_option) of one Engine carried into a new one.
void copy_options(const Engine& from, Engine& to) {
    to.seed = from.seed;
    to.photons = from.photons;
    to.write_exr = from.write_exr;
    to.compat_rows = from.compat_rows;
    to.compat_point = from.compat_point;
    to.compat_pdf = from.compat_pdf;
    to.compat_glass = from.compat_glass;
    to.light_emit = from.light_emit;
    to.max_path_length = from.max_path_length;
    to.photon_capacity = from.photon_capacity;
    to.cell_scale = from.cell_scale;
    to.persistent_blocks = from.persistent_blocks;
    to.light_blocks_per_cu = from.light_blocks_per_cu;
    to.camera_blocks_per_cu = from.camera_blocks_per_cu;
    to.gather_blocks_per_cu = from.gather_blocks_per_cu;
    to.path_chunk = from.path_chunk;
    to.light_refill = from.light_refill;
    to.camera_refill = from.camera_refill;
    to.count_events = from.count_events;
    to.exact_merge = from.exact_merge;
    to.graph = from.graph;
    to.bvh_width = from.bvh_width;
    to.bvh_build = from.bvh_build;
    to.device_ingest = from.device_ingest;
    to.query_order = from.query_order;
    to.light_order = from.light_order;
    to.camera_qmc = from.camera_qmc;
    to.iteration_offset = from.iteration_offset;
    to.iteration_stride = from.iteration_stride;
    to.group_queries = from.group_queries;
    to.gather_tile = from.gather_tile;
    to.gather_block = from.gather_block;
    to.overlap = from.overlap;
    to.mlt = from.mlt;
    to.mlt_chains = from.mlt_chains;
    to.output_path = from.output_path;
    const GuideModel& pg = from.guide;  // the proposal's options, not its state
    GuideModel& ng = to.guide;
    ng.enabled = pg.enabled;
    ng.train = pg.train;
    ng.alpha = pg.alpha;
    ng.lr = pg.lr;
    ng.batch = pg.batch;
    ng.cells = pg.cells;
    ng.warmup = pg.warmup;
    ng.train_every = pg.train_every;
    ng.seed = pg.seed;
}

// Debug render modes 2-4 into the film (replacing it); one pass, so the
// film counts as one iteration.
bool render_debug(Engine& e, int mode) {
    if (mode < 2 || mode > 4) {
        set_error("debug mode must be 2 (normals), 3 (facing) or 4 (materials)");
        return false;
    }
    if (!prepare(e)) return false;
    launch_fill_u32(reinterpret_cast<uint32_t*>(e.image.p), 0u, (size_t)e.W * e.H * 3, e.stream);
    launch_debug_aov(scene_view(e), mode, e.debug_ior, e.image.p, e.stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(e.stream));
    e.iterations_done = 1;
    return true;
}

bool get_image(Engine& e, float* rgb, bool scaled) {
    size_t n = (size_t)e.W * e.H * 3;
    HIP_OK(hipMemcpy(rgb, e.image.p, n * sizeof(float), hipMemcpyDeviceToHost));
    if (scaled && e.iterations_done > 0) {
        // Image_Buffer::scale (Image_Buffer.h:38-44)
        float s = 1.0f / (float)e.iterations_done;
        for (size_t i = 0; i < n; ++i) rgb[i] *= s;
    }
    return true;
}

// ---------------------------------------------------------------------
// Batches: several scenes (each an Engine with its own BVH, lights, camera,
// film, options and iteration count) rendered by one launch per pass.  The
// work Engine holds the shared photon, map and query buffers.
// ---------------------------------------------------------------------
struct Batch {
    std::vector<Engine*> scenes;
    Engine work;
    bool prepared = false;
    // captured batch iteration (scene 0's option "graph"): photon map + camera pass
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uintptr_t> graph_key;
    int graph_captures = 0;
    DevBuf<SceneView> d_sc;
    DevBuf<PhotonMapView> d_pm;
    DevBuf<BatchScene> d_bs;
    DevBuf<uint32_t> d_starts;  // light starts (n + 1), then camera starts (n + 1)
    DevBuf<float> scratch;      // films of the autotuner's iterations
    DevBuf<GuideView> d_guides;     // per scene proposal views of the light pass
    DevBuf<GuideTargets> d_targets; // per scene training targets
    DevBuf<uint8_t> vis;            // camera-visible cells of every scene's grid
    double stats[RC_STAT_COUNT] = {0};
    int iterations = 0;
    double guide_ms = 0, guide_step_ms = 0, guide_rows = 0;  // last training of the scenes' proposals
};

Batch* g_batch = nullptr;

void release_batch(Batch*& b) {
    if (!b) return;
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
    for (Engine*& e : b->scenes) release(e);
    Engine& w = b->work;
    for (auto* d : {&w.ph_rec, &w.s_pos, &w.s_dir, &w.q_pos, &w.q_dir, &w.q_sn, &w.q_gn, &w.q_tpn})
        d->release();
    w.s_tp.release();
    w.emitted.release();
    for (auto* d : {&w.s_index, &w.vals_a, &w.vals_b, &w.hist, &w.hist_sums, &w.counters}) d->release();
    w.keys_a.release();
    w.keys_b.release();
    for (auto* d : {&w.qk_a, &w.qk_b, &w.lk_a, &w.lk_b}) d->release();
    for (auto* d : {&w.lv_a, &w.lv_b, &w.l_hist, &w.l_hist_sums}) d->release();
    for (auto* d : {&w.qv_a, &w.qv_b, &w.q_hist, &w.q_hist_sums}) d->release();
    w.cells.release();
    w.path_info.release();
    w.contrib.release();
    w.stats_dev.release();
    b->d_sc.release();
    b->d_pm.release();
    b->d_bs.release();
    b->d_starts.release();
    b->scratch.release();
    b->d_guides.release();
    b->d_targets.release();
    b->vis.release();
    for (auto& ev : w.ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : w.ev_guide)
        if (ev) (void)hipEventDestroy(ev);
    if (w.stream) (void)hipStreamDestroy(w.stream);
    delete b;
    b = nullptr;
}

bool batch_prepare(Batch& b) {
    if (b.prepared) return true;
    if (b.scenes.empty()) {
        set_error("batch has no scenes");
        return false;
    }
    for (Engine* e : b.scenes) {
        // one launch serves every scene: the merge arithmetic is chosen per launch
        if (e->exact_merge != b.scenes[0]->exact_merge) {
            set_error("batch scenes must agree on the option exact_merge");
            return false;
        }
        if (e->mlt) {
            set_error("the adaptive MLT light pass (option mlt) renders single scenes only");
            return false;
        }
    }
    // BVH builds and uploads of the scenes on parallel host threads
    std::atomic<size_t> next{0};
    std::atomic<bool> ok{true};
    auto worker = [&]() {
        for (size_t i; (i = next++) < b.scenes.size();)
            if (!prepare(*b.scenes[i])) ok = false;
    };
    size_t nt = std::max<size_t>(1, std::min<size_t>(b.scenes.size(), std::thread::hardware_concurrency()));
    nt = std::min<size_t>(nt, 16);
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    if (!ok) return false;
    if (!ensure_device(b.work)) return false;
    // one launch per pass serves every scene: its shapes are scene 0's options
    b.work.path_chunk = b.scenes[0]->path_chunk;
    b.work.light_blocks_per_cu = b.scenes[0]->light_blocks_per_cu;
    b.work.camera_blocks_per_cu = b.scenes[0]->camera_blocks_per_cu;
    b.work.gather_blocks_per_cu = b.scenes[0]->gather_blocks_per_cu;
    b.work.light_refill = b.scenes[0]->light_refill;
    b.work.camera_refill = b.scenes[0]->camera_refill;
    b.work.query_order = b.scenes[0]->query_order;
    b.work.light_order = b.scenes[0]->light_order;
    b.work.group_queries = b.scenes[0]->group_queries;
    b.work.gather_tile = b.scenes[0]->gather_tile;
    b.work.gather_block = b.scenes[0]->gather_block;
    b.work.overlap = b.scenes[0]->overlap;
    b.work.light_emit = b.scenes[0]->light_emit;
    if (!b.work.counters.alloc(kPathCounters) || !b.work.stats_dev.alloc(2)) return false;
    size_t n = b.scenes.size();
    if (!b.d_sc.alloc(n) || !b.d_pm.alloc(n) || !b.d_bs.alloc(n) || !b.d_starts.alloc(2 * (n + 1))) return false;
    std::vector<SceneView> sc(n);
    for (size_t i = 0; i < n; ++i) sc[i] = scene_view(*b.scenes[i]);
    HIP_OK(hipMemcpy(b.d_sc.p, sc.data(), n * sizeof(SceneView), hipMemcpyHostToDevice));
    b.prepared = true;
    return true;
}

// Training step of every scene whose proposal trains this iteration:
// camera-visible cells of each scene's grid (one launch over all queries),
// per-path visible luminance (one launch over all photons), then each
// scene's network step and table on the batch stream.
bool batch_guide_training(Batch& b, const BatchView& view, uint32_t count, uint32_t nq,
                          const std::vector<GuideTargets>& targets, hipStream_t st) {
    Engine& w = b.work;
    uint64_t buckets = w.cells.n;
    if (!b.vis.alloc(buckets + 3)) return false;
    HIP_OK(hipEventRecord(w.ev_guide[0], st));
    launch_fill_u32(reinterpret_cast<uint32_t*>(b.vis.p), 0u, (buckets + 3) / 4, st);
    launch_mark_visible_batch(w.q_pos.p, w.q_gn.p, std::min(nq, w.q_capacity), view, b.vis.p, st);
    for (const GuideTargets& t : targets)
        if (t.rows) launch_fill_u32(reinterpret_cast<uint32_t*>(t.w), 0u, t.rows, st);
    launch_guide_targets_batch(w.ph_rec.p, count, view, b.vis.p, b.d_targets.p,
                               st);
    HIP_OK(hipEventRecord(w.ev_guide[2], st));
    double rows = 0.0;
    for (size_t i = 0; i < b.scenes.size(); ++i) {
        if (!targets[i].rows) continue;
        GuideModel& gd = b.scenes[i]->guide;
        const int n = (int)targets[i].rows;
        if (!gd.forward(gd.train_x.p, n, st, false) || !gd.backward(gd.train_bin.p, gd.train_w.p, n, st) ||
            !gd.adam(st) ||
            !gd.build_table(st))
            return false;
        HIP_OK(hipMemcpyAsync(gd.last, gd.scal.p, sizeof gd.last, hipMemcpyDeviceToHost, st));
        rows += n;
    }
    HIP_OK(hipEventRecord(w.ev_guide[3], st));
    HIP_OK(hipEventRecord(w.ev_guide[1], st));
    HIP_OK(hipStreamSynchronize(st));
    float ms_guide = 0, ms_step = 0;
    HIP_OK(hipEventElapsedTime(&ms_guide, w.ev_guide[0], w.ev_guide[1]));
    HIP_OK(hipEventElapsedTime(&ms_step, w.ev_guide[2], w.ev_guide[3]));
    b.guide_ms = ms_guide;
    b.guide_step_ms = ms_step;
    b.guide_rows = rows;
    return true;
}

// The batch's photon map and camera pass replayed from one captured graph
// (scene 0's option graph).  The per-scene values live in the device tables
// the caller filled for this iteration; the photon and query counts are read
// from device memory, so the sort grids are sized for the buffers' capacity;
// the graph is captured again when a buffer, a table, the scene set or a
// launch shape changes.  nq = this pass's query count.
bool batch_graph_pass(Batch& b, const BatchView& view, CameraPassArgs c, uint32_t count, uint64_t buckets,
                      int bucket_bits, int uid_bits, uint32_t& nq) {
    Engine& w = b.work;
    hipStream_t st = w.stream;
    (void)count;
    c.path_counter = w.counters.p;
    c.q = QueryBuffers{w.q_pos.p, w.q_dir.p, w.q_sn.p, w.q_gn.p, w.q_tpn.p, w.q_capacity};
    std::vector<uintptr_t> key = {(uintptr_t)w.ph_rec.p, w.capacity, (uintptr_t)w.q_pos.p, w.q_capacity,
                                  (uintptr_t)w.cells.p, (uintptr_t)buckets, (uintptr_t)bucket_bits,
                                  (uintptr_t)uid_bits, (uintptr_t)w.path_info.p, (uintptr_t)w.contrib.p,
                                  (uintptr_t)view.n, (uintptr_t)view.sc, (uintptr_t)view.pm, (uintptr_t)view.bs,
                                  (uintptr_t)view.light_starts, (uintptr_t)view.guides, (uintptr_t)c.n_paths,
                                  (uintptr_t)w.path_chunk, (uintptr_t)gather_blocks(w), (uintptr_t)camera_blocks(w),
                                  (uintptr_t)c.exact_merge, (uintptr_t)w.query_order, (uintptr_t)w.group_queries,
                                  (uintptr_t)w.camera_refill, (uintptr_t)c.camera_qmc, (uintptr_t)w.gather_tile,
                                  (uintptr_t)w.gather_block};
    if (!b.graph_exec || key != b.graph_key) {
        if (b.graph_exec) HIP_OK(hipGraphExecDestroy(b.graph_exec));
        b.graph_exec = nullptr;
        MapBuildArgs m{};
        m.in_rec = w.ph_rec.p;
        m.count = w.capacity;
        m.count_dev = w.counters.p + 1;
        m.mask = (uint32_t)(buckets - 1);
        m.bucket_bits = bucket_bits;
        m.uid_bits = uid_bits;
        m.keys_a = w.keys_a.p;
        m.keys_b = w.keys_b.p;
        m.vals_a = w.vals_a.p;
        m.vals_b = w.vals_b.p;
        m.hist = w.hist.p;
        m.hist_sums = w.hist_sums.p;
        m.keys_ready = true;  // by the batch's light pass
        m.scratch_n = w.capacity;
        m.out_pos = w.s_pos.p;
        m.out_dir = w.s_dir.p;
        m.out_tp = w.s_tp.p;
        m.out_index = w.s_index.p;
        m.cells = w.cells.p;
        bool ok = capture_graph(st,
                                [&]() {
                                    build_photon_map_batch(m, view, st);
                                    launch_fill_u32(w.counters.p, 0u, kPathCounters, st);  // after the map read [1]
                                    launch_camera_trace(c, camera_blocks(w), st);
                                    c.q_order = query_order(w, c, w.q_capacity, st);  // count from the device
                                    launch_gather(c, gather_blocks(w), st);
                                    launch_resolve(c, st);
                                },
                                b.graph_exec);
        if (!ok) {
            if (b.graph_exec) (void)hipGraphExecDestroy(b.graph_exec);
            b.graph_exec = nullptr;
            return false;
        }
        b.graph_key = key;
        b.graph_captures++;
    }
    HIP_OK(hipEventRecord(w.ev[2], st));
    HIP_OK(hipGraphLaunch(b.graph_exec, st));
    HIP_OK(hipEventRecord(w.ev[5], st));
    HIP_OK(hipMemcpyAsync(&nq, w.counters.p + 1, 4, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (nq > w.q_capacity) {  // cannot happen: at most 8 queries per camera path
        set_error("merge query buffer overflow in the batch graph");
        return false;
    }
    return true;
}

// One iteration of every scene: light pass, photon map, camera trace,
// gather and resolve, each a single launch over all scenes.  With `scratch`
// the films and iteration counts stay untouched (autotuner runs).
bool batch_iteration(Batch& b, bool scratch = false) {
    Engine& w = b.work;
    hipStream_t st = w.stream;
    const size_t n = b.scenes.size();
    std::vector<BatchScene> bs(n);
    std::vector<PhotonMapView> pm(n);
    std::vector<uint32_t> starts(2 * (n + 1));
    uint64_t light = 0, camera = 0, buckets = 0;
    for (size_t i = 0; i < n; ++i) {
        Engine& e = *b.scenes[i];
        float radius, r2, norm;
        const int it = sample_iteration(e, e.iterations_done);
        iteration_params(e, it, radius, r2, norm);
        long long N = light_paths(e);
        BatchScene& s = bs[i];
        s.light_base = (uint32_t)light;
        s.light_paths = (uint32_t)N;
        s.camera_base = (uint32_t)camera;
        s.camera_paths = (uint32_t)((long long)e.W * e.H);
        s.k0 = (uint32_t)e.seed;
        s.k1 = (uint32_t)(e.seed >> 32);
        s.vm_norm = norm;
        s.bucket_base = (uint32_t)buckets;
        s.iteration = (uint32_t)it;
        s.image = e.image.p;
        starts[i] = (uint32_t)light;
        starts[n + 1 + i] = (uint32_t)camera;
        light += (uint64_t)N;
        camera += s.camera_paths;
        // the scene's grid, as build_map
        uint64_t want = std::max<uint64_t>(4096, std::min<uint64_t>(1ull << 24, (uint64_t)N / 2));
        uint64_t nb = 1ull << log2_ceil(want);
        float maxc = 0.0f;
        for (int k = 0; k < 3; ++k) maxc = std::max(maxc, std::max(std::fabs(e.bbox.lo[k]), std::fabs(e.bbox.hi[k])));
        float rq = radius + std::max(radius * 1e-3f, 8.0f * maxc * 1.1920929e-07f);
        float cell = rq * (float)std::max(0.67, e.cell_scale);
        PhotonMapView& v = pm[i];
        for (int k = 0; k < 3; ++k) v.origin[k] = e.bbox.lo[k] - cell;
        v.inv_cell = 1.0f / cell;
        v.mask = (uint32_t)(nb - 1);
        v.r2 = r2;
        v.r_query = rq;
        buckets += nb;
    }
    starts[n] = (uint32_t)light;
    starts[2 * n + 1] = (uint32_t)camera;
    if (scratch) {  // films of the autotuner's runs: one scratch buffer, scenes at their camera offsets
        if (!b.scratch.alloc(3 * camera)) return false;
        for (size_t i = 0; i < n; ++i) bs[i].image = b.scratch.p + 3 * (size_t)bs[i].camera_base;
    }
    // photon ids are global light path indices (32 bits); query slots are camera path * 8 + length
    if (light > (1ull << 31) || camera >= (1ull << 29) || buckets > (1ull << 31)) {
        set_error("batch too large for one launch (light paths > 2^31 or camera paths >= 2^29)");
        return false;
    }
    int uid_bits = log2_ceil(light);  // (bucket, global light path) keys, as map_grid_params
    int bucket_bits = log2_ceil(buckets);
    if (uid_bits + bucket_bits > 64) {
        set_error("photon map key does not fit 64 bits");
        return false;
    }
    // shared buffers; photons start at a quarter of the paths and grow on overflow
    uint32_t cap = std::max<uint32_t>(w.capacity, (uint32_t)std::max<uint64_t>(light / 4, 1 << 20));
    if (!ensure_photon_buffers(w, cap) || !w.cells.alloc(buckets)) return false;
    for (size_t i = 0; i < n; ++i) {
        pm[i].pos = w.s_pos.p;
        pm[i].dir_tr = w.s_dir.p;
        pm[i].tp_gb = w.s_tp.p;
        pm[i].cells = w.cells.p + bs[i].bucket_base;
    }
    // per-scene emission proposals: sampled in the light pass, trained after the camera pass
    std::vector<GuideView> gvs(n);
    std::vector<GuideTargets> targets(n);
    bool any_train = false;
    for (size_t i = 0; i < n; ++i) {
        Engine& e = *b.scenes[i];
        gvs[i] = plain_guide_view();
        targets[i] = GuideTargets{nullptr, 0u};
        if (!guide_active(e)) continue;
        if (!ensure_guide_buffers(e, st)) return false;
        gvs[i] = guide_light_view(e, e.iterations_done);
        if (scratch) gvs[i].train_paths = 0;
        if (gvs[i].train_paths > 0) {
            targets[i] = GuideTargets{e.guide.train_w.p, gvs[i].train_paths};
            any_train = true;
        }
    }
    if (!b.d_guides.alloc(n) || !b.d_targets.alloc(n)) return false;
    HIP_OK(hipMemcpyAsync(b.d_guides.p, gvs.data(), n * sizeof(GuideView), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(b.d_targets.p, targets.data(), n * sizeof(GuideTargets), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(b.d_bs.p, bs.data(), n * sizeof(BatchScene), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(b.d_pm.p, pm.data(), n * sizeof(PhotonMapView), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(b.d_starts.p, starts.data(), starts.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    BatchView view{};
    view.n = (int)n;
    view.sc = b.d_sc.p;
    view.pm = b.d_pm.p;
    view.bs = b.d_bs.p;
    view.light_starts = b.d_starts.p;
    view.camera_starts = b.d_starts.p + n + 1;
    view.guides = b.d_guides.p;
    const Engine& e0 = *b.scenes[0];
    const bool counting = e0.count_events != 0;  // event counters of the first scene's option
    // scene 0's option "graph": after the light pass (a direct launch whose photon count is checked),
    // the photon map and the camera pass replay one captured graph; every per-scene value comes from
    // the tables above, every count from device memory, and the query buffer holds the most queries
    // a camera pass can make (8 per path)
    const bool graphed = e0.graph && !counting;
    // camera trace and query sort on stream2 beside the light pass and the map (option overlap of the
    // batch's launch options; not while counting events or graphed): the trace reads only the scene tables
    const bool overlapped = w.overlap && !counting && !graphed;
    uint32_t qcap = std::max(w.q_capacity,
                             (uint32_t)std::min<uint64_t>((graphed ? 8 : 2) * camera + 1024, 1ull << 31));
    if (!ensure_query_buffers(w, qcap, (uint32_t)camera)) return false;
    CameraPassArgs c{};
    c.sc = scene_view(e0);
    c.iteration = (uint32_t)e0.iterations_done;
    c.stream = STREAM_CAMERA;
    c.n_paths = (uint32_t)camera;
    c.path_chunk = w.path_chunk;
    c.refill_min = w.camera_refill;
    c.path_counter = w.counters.p;
    c.path_info = w.path_info.p;
    c.contrib = w.contrib.p;
    c.exact_merge = e0.exact_merge;
    c.group_queries = w.group_queries;
    c.gather_tile = w.gather_tile;
    c.gather_block = w.gather_block;
    c.camera_qmc = e0.camera_qmc;
    c.batch = view;
    c.counters = counting ? w.stats_dev.p + 1 : nullptr;
    // overlap 2: the trace is queued after the light pass (forked before it), as for one scene
    const bool light_first = overlapped && w.overlap == 2;
    auto queue_trace = [&]() -> bool {
        hipStream_t s2 = w.stream2;
        HIP_OK(hipStreamWaitEvent(s2, w.ev_fork, 0));
        launch_fill_u32(w.cam_counters.p, 0u, kPathCounters, s2);
        HIP_OK(hipEventRecord(w.ev[4], s2));
        launch_camera_trace(c, camera_blocks(w), s2);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(w.h_nq, w.cam_counters.p + 1, 4, hipMemcpyDeviceToHost, s2));
        HIP_OK(hipEventRecord(w.ev[8], s2));
        HIP_OK(hipEventRecord(w.ev_cam, s2));
        HIP_OK(hipEventRecord(w.ev_join, s2));
        c.q_order = query_order(w, c, w.q_capacity, s2);  // count read from the device
        HIP_OK(hipEventRecord(w.ev_qdone, s2));
        return true;
    };
    if (overlapped) {
        if (!ensure_stream2(w)) return false;
        c.path_counter = w.cam_counters.p;
        c.q = QueryBuffers{w.q_pos.p, w.q_dir.p, w.q_sn.p, w.q_gn.p, w.q_tpn.p, w.q_capacity};
        HIP_OK(hipEventRecord(w.ev_fork, st));
        if (!light_first && !queue_trace()) return false;
    }

    // light pass (grow and repeat on overflow)
    uint32_t count = 0;
    for (int attempt = 0;; ++attempt) {
        launch_fill_u32(w.counters.p, 0u, 4, st);
        if (counting) launch_fill_u32(reinterpret_cast<uint32_t*>(w.stats_dev.p), 0u, 2 * kCounterFields * 2, st);
        LightPassArgs a{};
        a.sc = scene_view(e0);
        a.iteration = (uint32_t)e0.iterations_done;
        a.stream = STREAM_LIGHT;
        a.n_paths = (uint32_t)light;
        a.path_chunk = w.path_chunk;
        a.refill_min = w.light_refill;
        a.path_counter = w.counters.p;
        a.photon_counter = w.counters.p + 1;
        a.capacity = w.capacity;
        a.out_rec = w.ph_rec.p;
        a.guide = plain_guide_view();
        a.batch = view;
        a.counters = counting ? w.stats_dev.p : nullptr;
        if (w.light_emit) {
            if (!w.emitted.alloc(3 * (size_t)light)) return false;
            a.emitted = w.emitted.p;
        }
        // the photon map's keys with the photons (grid per scene from the views uploaded above)
        clear_map_sort(w.hist.p, w.capacity, st);
        a.map.keys = w.keys_a.p;
        a.map.vals = w.vals_a.p;
        a.map.dig = map_sort_digits(w.hist.p, w.capacity);
        a.map.uid_bits = uid_bits;
        HIP_OK(hipEventRecord(w.ev[0], st));
        a.path_order = light_path_order(w, a, st);
        if (w.light_order && !a.path_order) return false;
        launch_light_pass(a, light_blocks(w), st);
        if (light_first && attempt == 0 && !queue_trace()) return false;
        HIP_OK(hipEventRecord(w.ev[1], st));
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(&count, w.counters.p + 1, 4, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (count <= w.capacity) break;
        if (attempt == 1 || !ensure_photon_buffers(w, (uint32_t)std::min<double>(count * 1.25 + 1024, 4.0e9))) {
            set_error("photon buffer overflow");
            return false;
        }
    }
    if (pm[0].pos != w.s_pos.p || pm[0].dir_tr != w.s_dir.p || pm[0].tp_gb != w.s_tp.p) {
        // the overflow above reallocated the sorted photon arrays: the merge must read the new ones, not
        // the freed ones the scenes' photon-map views were uploaded with (the overlapped trace and query
        // sort read the same table on stream2, so they finish first)
        if (overlapped) HIP_OK(hipStreamSynchronize(w.stream2));
        for (size_t i = 0; i < n; ++i) {
            pm[i].pos = w.s_pos.p;
            pm[i].dir_tr = w.s_dir.p;
            pm[i].tp_gb = w.s_tp.p;
        }
        HIP_OK(hipMemcpyAsync(b.d_pm.p, pm.data(), n * sizeof(PhotonMapView), hipMemcpyHostToDevice, st));
    }
    uint32_t nq = 0;
    if (graphed) {
        if (!batch_graph_pass(b, view, c, count, buckets, bucket_bits, uid_bits, nq)) return false;
    } else {
        // one photon map over every scene
        MapBuildArgs m{};
        m.in_rec = w.ph_rec.p;
        m.count = count;
        m.mask = (uint32_t)(buckets - 1);
        m.bucket_bits = bucket_bits;
        m.uid_bits = uid_bits;
        m.keys_a = w.keys_a.p;
        m.keys_b = w.keys_b.p;
        m.vals_a = w.vals_a.p;
        m.vals_b = w.vals_b.p;
        m.hist = w.hist.p;
        m.hist_sums = w.hist_sums.p;
        m.out_pos = w.s_pos.p;
        m.out_dir = w.s_dir.p;
        m.out_tp = w.s_tp.p;
        m.out_index = w.s_index.p;
        m.cells = w.cells.p;
        m.keys_ready = true;  // by the light pass
        m.scratch_n = w.capacity;
        HIP_OK(hipEventRecord(w.ev[2], st));
        build_photon_map_batch(m, view, st);
        HIP_OK(hipEventRecord(w.ev[3], st));
        HIP_OK(hipGetLastError());
        // camera pass: the overlapped trace's count, or (serial, or after an overflow) the trace here
        bool serial = !overlapped;
        if (overlapped) {
            HIP_OK(hipEventSynchronize(w.ev_cam));
            nq = *w.h_nq;
            if (nq > w.q_capacity) {  // trace again, serially, into larger buffers
                HIP_OK(hipStreamSynchronize(w.stream2));
                if (!ensure_query_buffers(w, (uint32_t)std::min<double>(nq * 1.25 + 1024, 4.0e9), (uint32_t)camera)) {
                    set_error("merge query buffer overflow");
                    return false;
                }
                c.path_counter = w.counters.p;
                c.q_order = nullptr;
                serial = true;
            }
        }
        for (int attempt = 0; serial; ++attempt) {
            c.q = QueryBuffers{w.q_pos.p, w.q_dir.p, w.q_sn.p, w.q_gn.p, w.q_tpn.p, w.q_capacity};
            launch_fill_u32(w.counters.p, 0u, kPathCounters, st);
            if (counting && attempt > 0)
                launch_fill_u32(reinterpret_cast<uint32_t*>(w.stats_dev.p + 1), 0u, kCounterFields * 2, st);
            HIP_OK(hipEventRecord(w.ev[4], st));
            launch_camera_trace(c, camera_blocks(w), st);
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpyAsync(&nq, w.counters.p + 1, 4, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            if (nq <= w.q_capacity) break;
            if (attempt == 1 ||
                !ensure_query_buffers(w, (uint32_t)std::min<double>(nq * 1.25 + 1024, 4.0e9), (uint32_t)camera)) {
                set_error("merge query buffer overflow");
                return false;
            }
        }
        if (serial) {
            HIP_OK(hipEventRecord(w.ev[8], st));
            if (overlapped) HIP_OK(hipEventRecord(w.ev_join, st));
            c.q_order = query_order(w, c, nq, st);
            if (overlapped) HIP_OK(hipEventRecord(w.ev_qdone, st));
        } else {
            HIP_OK(hipStreamWaitEvent(st, w.ev_qdone, 0));  // the trace and the query sort on stream2
        }
        HIP_OK(hipEventRecord(w.ev[6], st));
        launch_gather(c, gather_blocks(w), st);
        HIP_OK(hipEventRecord(w.ev[7], st));
        launch_resolve(c, st);
        HIP_OK(hipEventRecord(w.ev[5], st));
        HIP_OK(hipGetLastError());
    }
    if (any_train && !batch_guide_training(b, view, count, nq, targets, st)) return false;
    Counters cnt[2] = {};
    if (counting) HIP_OK(hipMemcpyAsync(cnt, w.stats_dev.p, sizeof cnt, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    float ms[8] = {0};
    HIP_OK(hipEventElapsedTime(&ms[0], w.ev[0], w.ev[1]));
    if (graphed) {  // map + camera pass are one replay: its time goes to ms_camera
        HIP_OK(hipEventElapsedTime(&ms[2], w.ev[2], w.ev[5]));
    } else {
        HIP_OK(hipEventElapsedTime(&ms[1], w.ev[2], w.ev[3]));
        HIP_OK(hipEventElapsedTime(&ms[2], w.ev[4], w.ev[5]));
        HIP_OK(hipEventElapsedTime(&ms[3], w.ev[4], w.ev[8]));
        if (overlapped) HIP_OK(hipEventElapsedTime(&ms[6], w.ev_join, w.ev_qdone));
        else HIP_OK(hipEventElapsedTime(&ms[6], w.ev[8], w.ev[6]));
        HIP_OK(hipEventElapsedTime(&ms[4], w.ev[6], w.ev[7]));
        HIP_OK(hipEventElapsedTime(&ms[5], w.ev[7], w.ev[5]));
        if (overlapped) ms[2] = ms[3] + ms[6] + ms[4] + ms[5];  // the camera pass's own spans (see run_iteration)
    }
    std::fill(b.stats, b.stats + RC_STAT_COUNT, 0.0);
    b.stats[RC_STAT_LIGHT_PATHS] = (double)light;
    b.stats[RC_STAT_PHOTONS] = count;
    b.stats[RC_STAT_CAMERA_PATHS] = (double)camera;
    b.stats[RC_STAT_GATHERS] = nq;
    b.stats[RC_STAT_MS_LIGHT] = ms[0];
    b.stats[RC_STAT_MS_MAP] = ms[1];
    b.stats[RC_STAT_MS_CAMERA] = ms[2];
    b.stats[RC_STAT_MS_CAMERA_TRACE] = ms[3];
    b.stats[RC_STAT_MS_GATHER] = ms[4];
    b.stats[RC_STAT_MS_RESOLVE] = ms[5];
    b.stats[RC_STAT_MS_QUERY_SORT] = ms[6];
    b.stats[RC_STAT_SORT_PASSES] = (bucket_bits + uid_bits + 7) / 8;
    if (counting) {
        b.stats[RC_STAT_LIGHT_NODE_VISITS] = (double)cnt[0].nodes;
        b.stats[RC_STAT_LIGHT_TRI_TESTS] = (double)cnt[0].tris;
        b.stats[RC_STAT_LIGHT_HITS] = (double)cnt[0].hits;
        b.stats[RC_STAT_CAMERA_NODE_VISITS] = (double)cnt[1].nodes;
        b.stats[RC_STAT_CAMERA_TRI_TESTS] = (double)cnt[1].tris;
        b.stats[RC_STAT_CAMERA_HITS] = (double)cnt[1].hits;
        b.stats[RC_STAT_CELL_VISITS] = (double)cnt[1].cells;
        b.stats[RC_STAT_PHOTONS_TESTED] = (double)cnt[1].photons_tested;
        b.stats[RC_STAT_PHOTONS_ACCEPTED] = (double)cnt[1].photons_accepted;
        b.stats[RC_STAT_PHOTONS_STREAMED] = (double)cnt[1].photons_streamed;
    }
    if (any_train) {
        b.stats[RC_STAT_MS_GUIDE] = b.guide_ms;
        b.stats[RC_STAT_MS_GUIDE_STEP] = b.guide_step_ms;
        b.stats[RC_STAT_GUIDE_ROWS] = b.guide_rows;
    }
    b.stats[RC_STAT_GRAPH_CAPTURES] = b.graph_captures;
    if (scratch) return true;
    for (Engine* e : b.scenes) e->iterations_done++;
    b.iterations++;
    return true;
}

bool batch_autotune(Batch& b, int repeats, double* report) {
    if (!batch_prepare(b)) return false;
    Batch* pb = &b;
    auto knobs = tune_knobs([pb]() -> Engine& { return pb->work; },
                            [pb](double v) {
                                for (Engine* e : pb->scenes) e->cell_scale = v;
                            },
                            [pb]() { return pb->scenes[0]->cell_scale; }, {0, 2, 1});
    const int overlap = b.work.overlap;
    b.work.overlap = 0;  // the other knobs are timed with the camera trace serial
    const int graph = b.scenes[0]->graph;
    b.scenes[0]->graph = 0;  // per-kernel times need direct launches
    struct Restore {
        Engine* e;
        int g;
        ~Restore() { e->graph = g; }
    } restore{b.scenes[0], graph};
    auto timed = [pb](double* t) {
        auto w0 = std::chrono::steady_clock::now();
        if (!batch_iteration(*pb, true)) return false;
        t[T_WALL] = ms_since(w0);
        t[T_LIGHT] = pb->stats[RC_STAT_MS_LIGHT];
        t[T_MAP] = pb->stats[RC_STAT_MS_MAP];
        t[T_TRACE] = pb->stats[RC_STAT_MS_CAMERA_TRACE];
        t[T_GATHER] = pb->stats[RC_STAT_MS_GATHER];
        return true;
    };
    bool ok = run_autotune(knobs, timed, repeats, report);
    if (!ok) b.work.overlap = overlap;
    return ok;
}


// Blender DNA offsets of the Mesh struct's mpoly, mloop and mvert pointers.
constexpr size_t kMeshMpoly = 200, kMeshMloop = 208, kMeshMvert = 256;

// Scene ingest on the GPU (option device_ingest with bvh_build = 1): the
// host checks each polygon (three loops at least, a used material slot that
// maps into the material array) while it finds how many loops and material
// slots the polygons reach, copies the mesh's raw MVert, MPoly and MLoop
// bytes and its slot map to the device, and the ingest kernels make d_tri,
// d_shade, d_tri_id and the bounds there.  Non-finite coordinates and loop
// vertex indices beyond the mesh come back as flags.  Sets the error and
// returns false on a rejected scene.
bool ingest_on_device(Engine& e, const long long* meshes_pointers, unsigned n_meshes, const rc_matrix4x4* matrices,
                      const unsigned* n_verts, const unsigned* n_tris, int* const* materials_indices,
                      unsigned n_materials) {
    if (!ensure_device(e)) return false;
    struct Piece {
        const void* src;
        size_t bytes, at;
    };
    std::vector<Piece> pieces;
    std::vector<std::vector<int>> maps(n_meshes);
    std::vector<IngestMesh> meshes(n_meshes);
    size_t total = 0;
    auto place = [&](const void* src, size_t bytes) {
        size_t at = total;
        pieces.push_back(Piece{src, bytes, at});
        total += (bytes + 255) / 256 * 256;
        return at;
    };
    std::vector<size_t> at_vert(n_meshes), at_poly(n_meshes), at_loop(n_meshes), at_map(n_meshes);
    uint64_t vbase = 0, tbase = 0;
    for (unsigned i = 0; i < n_meshes; ++i) {
        const unsigned char* mb = reinterpret_cast<const unsigned char*>((intptr_t)meshes_pointers[i]);
        const unsigned nv = n_verts[i], nt = n_tris[i];
        e.mesh_tri_offset.push_back((uint32_t)tbase);
        if (!mb) {
            set_error("null Blender mesh pointer");
            return false;
        }
        const unsigned char *mpoly, *mloop, *mvert;
        memcpy(&mpoly, mb + kMeshMpoly, sizeof(void*));
        memcpy(&mloop, mb + kMeshMloop, sizeof(void*));
        memcpy(&mvert, mb + kMeshMvert, sizeof(void*));
        if ((nt > 0 && (!mpoly || !mloop || !mvert)) || (nv > 0 && !mvert)) {
            set_error("Blender mesh without mpoly/mloop/mvert");
            return false;
        }
        if (!materials_indices[i]) {
            set_error("null materials_indices entry");
            return false;
        }
        // polygons: the loops they reach and the material slots they use
        int64_t loops = 0;
        std::vector<uint8_t> used(1 << 15, 0);
        int max_slot = 0;
        for (unsigned t = 0; t < nt; ++t) {
            int32_t loopstart, totloop;
            int16_t mat_nr;
            memcpy(&loopstart, mpoly + 12 * (size_t)t, 4);
            memcpy(&totloop, mpoly + 12 * (size_t)t + 4, 4);
            memcpy(&mat_nr, mpoly + 12 * (size_t)t + 8, 2);
            if (totloop < 3 || loopstart < 0) {
                set_error("polygon with fewer than 3 loops");
                return false;
            }
            loops = std::max<int64_t>(loops, (int64_t)loopstart + 3);
            const int slot = mat_nr < 0 ? 0 : mat_nr;
            used[slot] = 1;
            max_slot = std::max(max_slot, slot);
        }
        std::vector<int>& map = maps[i];
        map.assign(max_slot + 1, 0);
        for (int k = 0; k <= max_slot; ++k) {
            if (!used[k]) continue;
            const int mat = materials_indices[i][k];
            if (mat < 0 || (unsigned)mat >= n_materials) {
                set_error("material index out of range");
                return false;
            }
            map[k] = mat;
        }
        at_vert[i] = place(mvert, 20 * (size_t)nv);
        at_poly[i] = place(mpoly, 12 * (size_t)nt);
        at_loop[i] = place(mloop, 8 * (size_t)loops);
        at_map[i] = place(map.data(), 4 * map.size());
        IngestMesh& m = meshes[i];
        m.nv = nv;
        m.nt = nt;
        m.vbase = (uint32_t)vbase;
        m.tbase = (uint32_t)tbase;
        const float* M = matrices[i].t;
        for (int k = 0; k < 12; ++k) m.M[k] = M[k];
        float inv[3][3];
        inverse3x3_of4(M, inv);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) m.inv[3 * r + c] = inv[r][c];
        vbase += nv;
        tbase += nt;
        if (vbase >= (1ull << 32) || tbase >= (1ull << 30)) {
            set_error("device ingest takes fewer than 2^32 vertices and 2^30 triangles");
            return false;
        }
    }
    const uint32_t nv = (uint32_t)vbase, nt = (uint32_t)tbase;
    e.n_tris = nt;
    if (nt == 0) return true;  // prepare reports the empty scene
    DevBuf<unsigned char> raw;
    DevBuf<IngestMesh> table;
    DevBuf<float4> wv, wn;
    DevBuf<int> small;  // bounds[6], flags
    if (!raw.alloc(std::max<size_t>(total, 1)) || !table.alloc(n_meshes) || !wv.alloc(std::max(nv, 1u)) ||
        !wn.alloc(std::max(nv, 1u)) || !small.alloc(8) || !e.d_tri.alloc(3 * (size_t)nt) ||
        !e.d_shade.alloc(3 * (size_t)nt) || !e.d_tri_id.alloc(nt))
        return false;
    for (const Piece& p : pieces)
        if (p.bytes) HIP_OK(hipMemcpyAsync(raw.p + p.at, p.src, p.bytes, hipMemcpyHostToDevice, e.stream));
    for (unsigned i = 0; i < n_meshes; ++i) {
        meshes[i].mvert = raw.p + at_vert[i];
        meshes[i].mpoly = raw.p + at_poly[i];
        meshes[i].mloop = raw.p + at_loop[i];
        meshes[i].mat_map = reinterpret_cast<const int*>(raw.p + at_map[i]);
    }
    const int init_small[8] = {ordered_float_bits(0x7f800000), ordered_float_bits(0x7f800000),
                               ordered_float_bits(0x7f800000), ordered_float_bits((int)0xff800000u),
                               ordered_float_bits((int)0xff800000u), ordered_float_bits((int)0xff800000u), 0, 0};
    HIP_OK(hipMemcpyAsync(table.p, meshes.data(), n_meshes * sizeof(IngestMesh), hipMemcpyHostToDevice, e.stream));
    HIP_OK(hipMemcpyAsync(small.p, init_small, sizeof init_small, hipMemcpyHostToDevice, e.stream));
    launch_ingest(table.p, (int)n_meshes, nv, nt, wv.p, wn.p, e.d_tri.p, e.d_shade.p, e.d_tri_id.p, small.p,
                  reinterpret_cast<uint32_t*>(small.p + 6), e.stream);
    HIP_OK(hipGetLastError());
    int out[8];
    HIP_OK(hipMemcpyAsync(out, small.p, sizeof out, hipMemcpyDeviceToHost, e.stream));
    HIP_OK(hipStreamSynchronize(e.stream));
    if (out[6] & (int)kIngestNonFinite) {
        set_error("non-finite vertex coordinate");
        return false;
    }
    if (out[6] & (int)kIngestLoopVertex) {
        set_error("loop vertex index out of range");
        return false;
    }
    for (int k = 0; k < 3; ++k) {  // +0.0f: a zero bound is +0
        int lo = ordered_float_bits(out[k]), hi = ordered_float_bits(out[3 + k]);
        float l, h;
        memcpy(&l, &lo, 4);
        memcpy(&h, &hi, 4);
        e.bbox.lo[k] = l + 0.0f;
        e.bbox.hi[k] = h + 0.0f;
    }
    e.ingested_on_device = true;
    return true;
}
}  // namespace

// =====================================================================
// exported C-ABI
// =====================================================================
extern "C" {

static void init_impl(int samples, float base_radius, float radius_reduction,
                                                 long long* meshes_pointers, unsigned int number_of_meshes,
                                                 rc_matrix4x4* mesh_matrices, unsigned int* meshes_number_of_verts,
                                                 unsigned int* meshes_number_of_tris, unsigned int camera_x,
                                                 unsigned int camera_y, rc_vec3 camera_position, rc_vec3 camera_corner0,
                                                 rc_vec3 camera_corner1, rc_vec3 camera_corner2, rc_vec3 camera_corner3,
                                                 rc_python_light* lights, unsigned int number_of_lights,
                                                 rc_python_material* materials, unsigned int number_of_materials,
                                                 int** materials_indices, int debug_test, rc_debug_stuff debug_stuff) {
    const float debug_ior = debug_stuff.specular;
    // options survive a re-init; geometry does not
    Engine* prev = g;
    Engine* e = new Engine();
    if (prev) {
        copy_options(*prev, *e);
        release(prev);
    }
    g = e;
    g_error.clear();
    g_cancel = 0;
    e->samples = samples;
    e->base_radius = base_radius;
    e->radius_alpha = radius_reduction;
    e->debug_test = debug_test;
    e->debug_ior = debug_ior;
    e->initialised = true;
    if (samples < 1) set_error("samples must be >= 1");
    if (camera_x == 0 || camera_y == 0) set_error("camera resolution must be positive");
    if (number_of_meshes > 0 && (!meshes_pointers || !mesh_matrices || !meshes_number_of_verts ||
                                 !meshes_number_of_tris || !materials_indices))
        set_error("null mesh arrays");
    if (number_of_materials > 0 && !materials) set_error("null material array");
    if (number_of_lights > 0 && !lights) set_error("null light array");
    if (!g_error.empty()) return;
    for (unsigned i = 0; i < number_of_materials; ++i) {
        e->py_mats.push_back(materials[i]);
        e->mats.push_back(convert_material(materials[i], e->compat_glass));
    }
    for (unsigned i = 0; i < number_of_lights; ++i) {
        HostLight h;
        memcpy(&h, &lights[i], sizeof h);
        e->py_lights.push_back(h);
    }
    if (e->bvh_build == 1 && e->device_ingest) {
        if (!ingest_on_device(*e, meshes_pointers, number_of_meshes, mesh_matrices, meshes_number_of_verts,
                              meshes_number_of_tris, materials_indices, number_of_materials))
            return;
    }
    uint32_t vbase = 0;
    for (unsigned i = 0; i < (e->ingested_on_device ? 0u : number_of_meshes); ++i) {
        const unsigned char* mb = reinterpret_cast<const unsigned char*>((intptr_t)meshes_pointers[i]);
        unsigned nv = meshes_number_of_verts[i], nt = meshes_number_of_tris[i];
        e->mesh_tri_offset.push_back((uint32_t)(e->tris.size() / 4));
        if (!mb) {
            set_error("null Blender mesh pointer");
            return;
        }
        const unsigned char *mpoly, *mloop, *mvert;
        memcpy(&mpoly, mb + kMeshMpoly, sizeof(void*));
        memcpy(&mloop, mb + kMeshMloop, sizeof(void*));
        memcpy(&mvert, mb + kMeshMvert, sizeof(void*));
        if (nt > 0 && (!mpoly || !mloop || !mvert)) {
            set_error("Blender mesh without mpoly/mloop/mvert");
            return;
        }
        if (!materials_indices[i]) {
            set_error("null materials_indices entry");
            return;
        }
        const float* M = mesh_matrices[i].t;
        float inv[3][3];
        inverse3x3_of4(M, inv);
        for (unsigned v = 0; v < nv; ++v) {
            float co[3];
            int16_t no[3];
            memcpy(co, mvert + 20 * (size_t)v, 12);
            memcpy(no, mvert + 20 * (size_t)v + 12, 6);
            if (!std::isfinite(co[0]) || !std::isfinite(co[1]) || !std::isfinite(co[2])) {
                set_error("non-finite vertex coordinate");
                return;
            }
            // Transform::ApplyToPoint (Transform.cpp:3-11)
            float x = co[0], y = co[1], z = co[2];
            e->wv.push_back(M[0] * x + M[1] * y + M[2] * z + M[3]);
            e->wv.push_back(M[4] * x + M[5] * y + M[6] * z + M[7]);
            e->wv.push_back(M[8] * x + M[9] * y + M[10] * z + M[11]);
            // normal_short_to_float (Triangle.cpp:11-16), ApplyToNormal (Transform.cpp:23-31)
            const float s = 1.0f / 32767.0f;
            float nx = no[0] * s, ny = no[1] * s, nz = no[2] * s;
            F3 n = f3(inv[0][0] * nx + inv[1][0] * ny + inv[2][0] * nz, inv[0][1] * nx + inv[1][1] * ny + inv[2][1] * nz,
                      inv[0][2] * nx + inv[1][2] * ny + inv[2][2] * nz);
            n = normalize(n);
            e->wn.push_back(n.x);
            e->wn.push_back(n.y);
            e->wn.push_back(n.z);
        }
        for (unsigned t = 0; t < nt; ++t) {
            int32_t loopstart, totloop;
            int16_t mat_nr;
            memcpy(&loopstart, mpoly + 12 * (size_t)t, 4);
            memcpy(&totloop, mpoly + 12 * (size_t)t + 4, 4);
            memcpy(&mat_nr, mpoly + 12 * (size_t)t + 8, 2);
            if (totloop < 3 || loopstart < 0) {
                set_error("polygon with fewer than 3 loops");
                return;
            }
            uint32_t vid[3];
            for (int k = 0; k < 3; ++k) {
                memcpy(&vid[k], mloop + 8 * ((size_t)loopstart + k), 4);
                if (vid[k] >= nv) {
                    set_error("loop vertex index out of range");
                    return;
                }
                e->tris.push_back(vbase + vid[k]);
            }
            int mat = materials_indices[i][mat_nr < 0 ? 0 : mat_nr];
            if (mat < 0 || (unsigned)mat >= number_of_materials) {
                set_error("material index out of range");
                return;
            }
            e->tris.push_back((uint32_t)mat);
        }
        vbase += nv;
    }
    // Camera (Camera.h:21-27): corners top right, bottom right, bottom left, top left
    e->W = (int)camera_x;
    e->H = (int)camera_y;
    e->cam_o = f3(camera_position.x, camera_position.y, camera_position.z);
    F3 br = f3(camera_corner1.x, camera_corner1.y, camera_corner1.z);
    F3 bl = f3(camera_corner2.x, camera_corner2.y, camera_corner2.z);
    F3 tl = f3(camera_corner3.x, camera_corner3.y, camera_corner3.z);
    (void)camera_corner0;
    e->cam_bl = bl;
    e->cam_h = br - bl;
    e->cam_v = tl - bl;
}

__attribute__((visibility("default"))) void init(int samples, float base_radius, float radius_reduction,
                                                 long long* meshes_pointers, unsigned int number_of_meshes,
                                                 rc_matrix4x4* mesh_matrices, unsigned int* meshes_number_of_verts,
                                                 unsigned int* meshes_number_of_tris, unsigned int camera_x,
                                                 unsigned int camera_y, rc_vec3 camera_position, rc_vec3 camera_corner0,
                                                 rc_vec3 camera_corner1, rc_vec3 camera_corner2, rc_vec3 camera_corner3,
                                                 rc_python_light* lights, unsigned int number_of_lights,
                                                 rc_python_material* materials, unsigned int number_of_materials,
                                                 int** materials_indices, int debug_test, rc_debug_stuff debug_stuff) {
    init_impl(samples, base_radius, radius_reduction, meshes_pointers, number_of_meshes, mesh_matrices,
              meshes_number_of_verts, meshes_number_of_tris, camera_x, camera_y, camera_position, camera_corner0,
              camera_corner1, camera_corner2, camera_corner3, lights, number_of_lights, materials, number_of_materials,
              materials_indices, debug_test, debug_stuff);
    // keep the first problem init found so that main() reports it
    if (g && !g_error.empty()) g->init_error = g_error;
}

__attribute__((visibility("default"))) int main(void) {
    g_error.clear();
    if (!g || !g->initialised) {
        set_error("main called before init");
        return 1;
    }
    if (!g->init_error.empty()) {
        set_error(g->init_error);
        return 1;
    }
    if (!prepare(*g)) return 2;
    Engine& e = *g;
    if (e.debug_test == 1) {
        for (int it = e.iterations_done; it < e.samples; ++it) {
            if (g_cancel.load()) break;
            if (!render_iteration(e, sample_iteration(e, it))) return 3;
            e.iterations_done = it + 1;
        }
    } else if (e.debug_test >= 2 && e.debug_test <= 4) {
        // src/main.cpp:172-180: one debug pass instead of the integrator
        if (!render_debug(e, e.debug_test)) return 3;
    }
    int rc = 0;
    if (e.debug_test != 5 && e.write_exr) {
        std::vector<float> img((size_t)e.W * e.H * 3, 0.0f);
        if (e.debug_test >= 1 && e.debug_test <= 4 && !get_image(e, img.data(), true)) rc = 4;
        if (rc == 0 && !write_exr(e.output_path.c_str(), img.data(), e.W, e.H)) rc = 5;
    }
    // src/main.cpp:220-221: the scene is deleted after rendering; options survive
    Engine* keep = new Engine();
    copy_options(e, *keep);
    release(g);
    g = keep;
    return rc;
}

__attribute__((visibility("default"))) int rc_set_option(const char* key, double v) {
    if (!g) g = new Engine();
    std::string k(key ? key : "");
    if (k == "seed") g->seed = (uint64_t)v;
    else if (k == "photons") g->photons = (long long)v;
    else if (k == "write_exr") g->write_exr = (int)v;
    else if (k == "compat_rows") g->compat_rows = (int)v;
    else if (k == "compat_point_light") g->compat_point = (int)v;
    else if (k == "compat_glass") g->compat_glass = (int)v;
    else if (k == "light_emit") g->light_emit = (int)v;
    else if (k == "compat_light_pdf") g->compat_pdf = (int)v;
    else if (k == "max_path_length") g->max_path_length = std::max(0, std::min(7, (int)v));
    else if (k == "photon_capacity") g->photon_capacity = v;
    else if (k == "cell_scale") g->cell_scale = v;
    else if (k == "persistent_blocks") g->persistent_blocks = (int)v;
    else if (k == "graph") g->graph = (int)v;
    else if (k == "bvh_width") g->bvh_width = (int)v == 2 ? 2 : 4;
    else if (k == "bvh_build") g->bvh_build = (int)v == 1 ? 1 : 0;
    else if (k == "device_ingest") g->device_ingest = v != 0.0 ? 1 : 0;
    else if (k == "query_order") g->query_order = (int)v ? 1 : 0;
    else if (k == "light_order") g->light_order = (int)v ? 1 : 0;
    else if (k == "camera_qmc") g->camera_qmc = (int)v ? 1 : 0;
    else if (k == "iteration_offset") g->iteration_offset = std::max(0, std::min((int)v, 1 << 30));
    else if (k == "iteration_stride") g->iteration_stride = std::max(1, std::min((int)v, 1 << 16));
    else if (k == "mlt") g->mlt = (int)v ? 1 : 0;
    else if (k == "mlt_chains") g->mlt_chains = (uint32_t)std::max(1.0, std::min(v, 16777216.0));
    else if (k == "group_queries") g->group_queries = (int)v >= 12 ? 16 : (int)v >= 6 ? 8 : (int)v >= 3 ? 4 : (int)v > 0 ? 1 : 0;
    else if (k == "overlap") g->overlap = (int)v >= 2 ? 2 : (int)v ? 1 : 0;
    else if (k == "gather_block") g->gather_block = (int)v >= 4 ? 4 : (int)v >= 2 ? 2 : (int)v >= 1 ? 1 : 0;
    else if (k == "gather_tile") g->gather_tile = (uint32_t)std::max(0.0, std::min(v, 1048576.0));
    else if (k == "light_blocks_per_cu") g->light_blocks_per_cu = std::max(1, std::min(64, (int)v));
    else if (k == "camera_blocks_per_cu") g->camera_blocks_per_cu = std::max(1, std::min(64, (int)v));
    else if (k == "gather_blocks_per_cu") g->gather_blocks_per_cu = std::max(1, std::min(128, (int)v));
    else if (k == "light_refill") g->light_refill = (uint32_t)std::max(1, std::min(64, (int)v));
    else if (k == "camera_refill") g->camera_refill = (uint32_t)std::max(1, std::min(64, (int)v));
    else if (k == "path_chunk") g->path_chunk = (uint32_t)std::max(1, std::min(1024, (int)v / 64)) * 64u;
    else if (k == "count_events") g->count_events = (int)v;
    else if (k == "exact_merge") g->exact_merge = (int)v;
    else if (k == "next_iteration") g->iterations_done = std::max(0, (int)v);
    else if (k == "guide") g->guide.enabled = (int)v;
    else if (k == "guide_alpha") g->guide.alpha = (float)std::min(1.0, std::max(1e-3, v));
    else if (k == "guide_lr") g->guide.lr = (float)v;
    else if (k == "guide_batch") g->guide.batch = (int)std::max(0.0, std::min(v, 1048576.0));
    else if (k == "guide_cells") g->guide.cells = (int)v;
    else if (k == "guide_warmup") g->guide.warmup = std::max(0, (int)v);
    else if (k == "guide_train") g->guide.train = (int)v;
    else if (k == "guide_train_every") g->guide.train_every = std::max(1, (int)v);
    else if (k == "guide_seed") g->guide.seed = (uint64_t)v;
    else if (k == "guide_reset") {
        g->guide.params_ready = g->guide.table_ready = false;
        g->guide.step = 0;
    } else return -1;
    return 0;
}

__attribute__((visibility("default"))) int rc_set_option_str(const char* key, const char* value) {
    if (!g) g = new Engine();
    if (key && value && std::string(key) == "output_path") {
        g->output_path = value;
        return 0;
    }
    return -1;
}

__attribute__((visibility("default"))) double rc_get_option(const char* key) {
    if (!g || !key) return NAN;
    std::string k(key);
    if (k == "seed") return (double)g->seed;
    if (k == "photons") return (double)light_paths(*g);
    if (k == "scene_radius") return g->scene_radius;
    if (k == "base_radius") return g->base_radius;
    if (k == "iterations") return g->iterations_done;
    if (k == "max_path_length") return g->max_path_length;
    if (k == "nodes") return g->n_nodes;
    if (k == "device_cus") return g->device_cus;
    if (k == "capacity") return g->capacity;
    if (k == "graph") return g->graph;
    if (k == "bvh_width") return g->bvh_width;
    if (k == "node_bytes") return uses_nodes4(*g) ? 128 : 64;
    if (k == "bvh_build") return g->bvh_build;
    if (k == "device_ingest") return g->device_ingest;
    if (k == "query_order") return g->query_order;
    if (k == "light_order") return g->light_order;
    if (k == "camera_qmc") return g->camera_qmc;
    if (k == "iteration_offset") return g->iteration_offset;
    if (k == "iteration_stride") return g->iteration_stride;
    if (k == "group_queries") return g->group_queries;
    if (k == "gather_tile") return g->gather_tile;
    if (k == "gather_block") return g->gather_block;
    if (k == "overlap") return g->overlap;
    if (k == "mlt") return g->mlt;
    if (k == "mlt_chains") return g->mlt_chains;
    if (k == "mlt_visible") return g->mlt_last[0];  // counts of the last MLT light pass
    if (k == "mlt_mutations") return g->mlt_last[1];
    if (k == "mlt_accepted") return g->mlt_last[2];
    if (k == "compat_glass") return g->compat_glass;
    if (k == "light_emit") return g->light_emit;
    if (k == "bvh_depth") return g->bvh_depth;
    if (k == "bvh_depth4") return g->bvh_depth4;
    if (k == "bvh_traversal_width") return uses_nodes4(*g) ? 4 : 2;
    if (k == "bvh_build_ms") return g->bvh_build_ms;
    if (k == "nodes4") return g->n_nodes4;
    if (k == "graph_captures") return g->graph_captures;
    if (k == "light_blocks_per_cu") return g->light_blocks_per_cu;
    if (k == "camera_blocks_per_cu") return g->camera_blocks_per_cu;
    if (k == "gather_blocks_per_cu") return g->gather_blocks_per_cu;
    if (k == "path_chunk") return g->path_chunk;
    if (k == "light_refill") return g->light_refill;
    if (k == "camera_refill") return g->camera_refill;
    if (k == "cell_scale") return g->cell_scale;
    if (k == "guide") return g->guide.enabled;
    if (k == "guide_step") return g->guide.step;
    if (k == "guide_slots") return g->guide.slots;
    if (k == "guide_table_rows") return g->guide.table_rows;
    if (k == "guide_alpha") return g->guide.alpha;
    return NAN;
}

__attribute__((visibility("default"))) int rc_prepare(void) {
    g_error.clear();
    if (g && !g->init_error.empty()) {
        set_error(g->init_error);
        return -1;
    }
    if (!g) {
        set_error("rc_prepare called before init");
        return -1;
    }
    return prepare(*g) ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_render_iteration(void) {
    if (rc_prepare() != 0) return -1;
    if (g_cancel.load()) {
        set_error("cancelled");
        return -2;
    }
    int it = g->iterations_done;
    if (!render_iteration(*g, sample_iteration(*g, it))) return -1;
    g->iterations_done = it + 1;
    return it;
}

__attribute__((visibility("default"))) int rc_render_debug(int mode) {
    g_error.clear();
    if (!g || !g->initialised) {
        set_error("rc_render_debug: no scene");
        return -1;
    }
    return render_debug(*g, mode) ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_get_image(float* rgb, int scaled) {
    if (!g || !g->prepared || !rgb) {
        set_error("no rendered image");
        return -1;
    }
    return get_image(*g, rgb, scaled != 0) ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_write_exr_image(const char* path, const float* rgb, int width,
                                                              int height) {
    if (!path || !rgb || width <= 0 || height <= 0) {
        set_error("rc_write_exr_image: bad arguments");
        return -1;
    }
    return write_exr(path, rgb, width, height) ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_write_exr(const char* path) {
    if (!g || !g->prepared || !path) {
        set_error("no rendered image");
        return -1;
    }
    std::vector<float> img((size_t)g->W * g->H * 3);
    if (!get_image(*g, img.data(), true)) return -1;
    return write_exr(path, img.data(), g->W, g->H) ? 0 : -1;
}

__attribute__((visibility("default"))) void rc_release(void) {
    release(g);
    g_error.clear();
}

__attribute__((visibility("default"))) const char* rc_last_error(void) { return g_error.c_str(); }

__attribute__((visibility("default"))) void rc_cancel(void) { g_cancel = 1; }

__attribute__((visibility("default"))) int rc_stats(double* out, int n) {
    if (!g || !out) return 0;
    int m = std::min(n, (int)RC_STAT_COUNT);
    for (int i = 0; i < m; ++i) out[i] = g->stats[i];
    return m;
}

__attribute__((visibility("default"))) int rc_bvh_shape(const float* tris, long long n, int requested_width,
                                                        int* out4) {
    // the host build of rc_prepare (binned SAH, then the 4-wide collapse) on
    // the given triangles; host memory only, no device call
    if (!tris || n <= 0 || n >= (1LL << 28) || !out4) return -1;
    Builder bld;
    bld.items.resize((size_t)n);
    for (long long t = 0; t < n; ++t) {
        BuildItem& it = bld.items[(size_t)t];
        it.id = (uint32_t)t;
        for (int k = 0; k < 3; ++k) {
            const float a = tris[9 * t + k], b = tris[9 * t + 3 + k], c = tris[9 * t + 6 + k];
            it.lo[k] = smin(smin(a, b), c);
            it.hi[k] = smax(smax(a, b), c);
            it.c[k] = (it.lo[k] + it.hi[k]) * 0.5f;
        }
    }
    bld.order.reserve((size_t)n);
    Box rootb;
    const int root = bld.build(0, (int)n, 0, rootb);
    int depth4 = 0;
    if (root >= 0) {
        std::vector<float4> nodes4;
        collapse4(bld.nodes, root, nodes4, 0, depth4);
    }
    const int depth = std::max(1, bld.max_depth);
    const int width = traversal_width(requested_width, depth, depth4);
    out4[0] = depth;
    out4[1] = depth4;
    out4[2] = width;
    out4[3] = width == 4 ? stack_need_wide(depth4) : stack_need_binary(depth);
    return width == 0 ? -1 : 0;
}

__attribute__((visibility("default"))) int rc_first_hit(const float* rays, long long n, float tmin, int* tri, float* t) {
    if (!g || !prepare(*g)) return -1;
    if (n <= 0) return 0;
    DevBuf<float> dr, dt;
    DevBuf<int> di;
    if (!dr.alloc((size_t)n * 6) || !dt.alloc((size_t)n) || !di.alloc((size_t)n)) return -1;
    bool ok = hipMemcpy(dr.p, rays, (size_t)n * 6 * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        launch_first_hit(scene_view(*g), dr.p, (int)n, tmin, di.p, dt.p, g->stream);
        ok = hipStreamSynchronize(g->stream) == hipSuccess &&
             hipMemcpy(tri, di.p, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(t, dt.p, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess;
    }
    dr.release();
    dt.release();
    di.release();
    if (!ok) set_error("rc_first_hit failed");
    return ok ? 0 : -1;
}

__attribute__((visibility("default"))) long long rc_trace_photons(int iteration, float* photons, unsigned int* uid,
                                                                  long long cap) {
    if (!g || !prepare(*g)) return -1;
    if (uid && light_paths(*g) > (1ll << 29)) {
        set_error("uid = path * 8 + length needs fewer than 2^29 light paths");
        return -1;
    }
    uint32_t count = 0;
    if (!run_light_pass(*g, iteration, count)) return -1;
    std::vector<float4> rec(kPhotonRecord * (size_t)count);
    if (count && hipMemcpy(rec.data(), g->ph_rec.p, rec.size() * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("photon readback failed");
        return -1;
    }
    std::vector<uint32_t> u(count);
    for (uint32_t i = 0; i < count; ++i) {
        const float4 r2 = rec[kPhotonRecord * (size_t)i + 2];
        uint32_t path, len;
        memcpy(&path, &r2.z, 4);
        memcpy(&len, &r2.w, 4);
        u[i] = path * 8u + len;
    }
    // path order (the order of the reference's sequential light loop)
    std::vector<uint32_t> ord(count);
    for (uint32_t i = 0; i < count; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return u[a] < u[b]; });
    long long m = std::min<long long>(cap, count);
    for (long long k = 0; k < m; ++k) {
        const float4* p = &rec[kPhotonRecord * (size_t)ord[k]];
        float* r = photons + 10 * k;
        r[0] = p[0].x;
        r[1] = p[0].y;
        r[2] = p[0].z;
        r[3] = p[1].x;
        r[4] = p[1].y;
        r[5] = p[1].z;
        r[6] = p[1].w;
        r[7] = p[2].x;
        r[8] = p[2].y;
        r[9] = p[0].w;
        if (uid) uid[k] = u[ord[k]];
    }
    return count;
}

// Loads a caller photon list (10 floats each) into the unsorted photon buffers.
static bool load_photons(const float* photons, long long count) {
    Engine& e = *g;
    if (!ensure_photon_buffers(e, (uint32_t)std::max<long long>(count, std::max<uint32_t>(e.capacity, 1024)))) return false;
    std::vector<float4> rec(kPhotonRecord * (size_t)count);
    for (long long i = 0; i < count; ++i) {  // photon id = input index, path length 0
        const float* r = photons + 10 * i;
        float id;
        const uint32_t u = (uint32_t)i;
        memcpy(&id, &u, 4);
        rec[kPhotonRecord * i] = make_float4(r[0], r[1], r[2], r[9]);
        rec[kPhotonRecord * i + 1] = make_float4(r[3], r[4], r[5], r[6]);
        rec[kPhotonRecord * i + 2] = make_float4(r[7], r[8], id, 0.0f);
    }
    if (count) HIP_OK(hipMemcpy(e.ph_rec.p, rec.data(), rec.size() * sizeof(float4), hipMemcpyHostToDevice));
    return true;
}

__attribute__((visibility("default"))) long long rc_gather_sets(const float* positions, long long count,
                                                                const float* queries, long long nq, float r2,
                                                                long long* offsets, int* indices, long long cap) {
    if (!g || !prepare(*g)) return -1;
    Engine& e = *g;
    std::vector<float> ph((size_t)count * 10, 0.0f);
    for (long long i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) ph[10 * i + k] = positions[3 * i + k];
    if (!load_photons(ph.data(), count)) return -1;
    // a photon map whose id bits cover the input indices
    long long saved = e.photons;
    e.photons = std::max<long long>(1, (count + 7) / 8 + 1);
    e.map_ids = count;
    float radius = sqrtf(r2);
    MapInfo mi;
    bool ok = build_map(e, (uint32_t)count, radius, r2, mi);
    e.photons = saved;
    e.map_ids = 0;
    if (!ok) return -1;
    mi.pm.r2 = r2;
    DevBuf<float> dq;
    DevBuf<uint32_t> dc;
    DevBuf<unsigned long long> doff;
    DevBuf<int> dout;
    if (!dq.alloc((size_t)nq * 3 + 1) || !dc.alloc((size_t)nq + 1) || !doff.alloc((size_t)nq + 1)) return -1;
    long long total = -1;
    do {
        if (hipMemcpy(dq.p, queries, (size_t)nq * 12, hipMemcpyHostToDevice) != hipSuccess) break;
        launch_gather_count(mi.pm, dq.p, (int)nq, e.s_index.p, dc.p, e.stream);
        std::vector<uint32_t> cnt(nq);
        if (hipStreamSynchronize(e.stream) != hipSuccess) break;
        if (nq && hipMemcpy(cnt.data(), dc.p, (size_t)nq * 4, hipMemcpyDeviceToHost) != hipSuccess) break;
        std::vector<unsigned long long> off(nq + 1, 0);
        for (long long i = 0; i < nq; ++i) off[i + 1] = off[i] + cnt[i];
        total = (long long)off[nq];
        for (long long i = 0; i <= nq; ++i) offsets[i] = (long long)off[i];
        if (total > cap) break;
        if (hipMemcpy(doff.p, off.data(), (size_t)(nq + 1) * 8, hipMemcpyHostToDevice) != hipSuccess) {
            total = -1;
            break;
        }
        if (!dout.alloc((size_t)total + 1)) {
            total = -1;
            break;
        }
        launch_gather_fill(mi.pm, dq.p, (int)nq, e.s_index.p, doff.p, dout.p, e.stream);
        if (hipStreamSynchronize(e.stream) != hipSuccess ||
            (total && hipMemcpy(indices, dout.p, (size_t)total * 4, hipMemcpyDeviceToHost) != hipSuccess)) {
            total = -1;
            break;
        }
    } while (0);
    dq.release();
    dc.release();
    doff.release();
    dout.release();
    if (total < 0) set_error("rc_gather_sets failed");
    return total;
}

__attribute__((visibility("default"))) long long rc_gather_sets_noalloc(const float* positions, long long count,
                                                                        const float* queries, long long nq, float r2,
                                                                        long long* offsets, int* indices,
                                                                        long long cap) {
    // no device allocation: the scene must be prepared already, and the
    // photon map and counts go into the library's fixed scratch array
    if (!g || !g->prepared) {
        set_error("rc_gather_sets_noalloc: the scene is not prepared (call rc_prepare first)");
        return -1;
    }
    if (count < 0 || nq < 0 || count > (long long)kNoAllocMaxPhotons || nq > (long long)kNoAllocMaxQueries) {
        set_error("rc_gather_sets_noalloc: at most 262144 photons and 262144 queries");
        return -1;
    }
    if ((count > 0 && !positions) || (nq > 0 && !queries) || !offsets || (cap > 0 && !indices) || cap < 0) {
        set_error("rc_gather_sets_noalloc: null buffer");
        return -1;
    }
    Engine& e = *g;
    MapGrid gr;
    // the grid rc_gather_sets builds for the same input (uid bits covering the indices)
    if (!map_grid_params(e, std::max<long long>(1, (count + 7) / 8 + 1), count, sqrtf(r2), r2, gr)) return -1;
    NoAllocGather a{};
    a.positions = positions;
    a.count = (uint32_t)count;
    a.queries = queries;
    a.nq = (uint32_t)nq;
    a.pm = gr.pm;
    a.bucket_bits = gr.bucket_bits;
    a.uid_bits = gr.uid_bits;
    a.offsets = reinterpret_cast<unsigned long long*>(offsets);
    a.indices = indices;
    a.cap = cap;
    const char* err = nullptr;
    long long total = gather_sets_noalloc(a, e.stream, &err);
    if (total < 0) set_error(err ? err : "rc_gather_sets_noalloc failed");
    return total;
}

__attribute__((visibility("default"))) long long rc_merge_sets(int iteration, const float* photons, long long count,
                                                               float* query_pos, long long* query_slot,
                                                               unsigned long long* set_count,
                                                               unsigned long long* set_hash, long long cap,
                                                               float* r2_out) {
    if (!g || !prepare(*g)) return -1;
    Engine& e = *g;
    if (!load_photons(photons, count)) return -1;
    float radius, r2, norm;
    iteration_params(e, iteration, radius, r2, norm);
    e.map_ids = count;  // photon ids are the input indices
    MapInfo mi;
    bool ok = build_map(e, (uint32_t)count, radius, r2, mi);
    e.map_ids = 0;
    if (!ok) return -1;
    const size_t slots = (size_t)e.W * e.H * 8;
    DevBuf<unsigned long long> sets;
    DevBuf<float> img;
    if (!sets.alloc(2 * slots) || !img.alloc((size_t)e.W * e.H * 3)) return -1;
    launch_fill_u32(reinterpret_cast<uint32_t*>(sets.p), 0u, 4 * slots, e.stream);
    launch_fill_u32(reinterpret_cast<uint32_t*>(img.p), 0u, (size_t)e.W * e.H * 3, e.stream);
    const int saved_count = e.count_events;
    e.count_events = 1;
    e.merge_sets = sets.p;
    ok = run_camera_pass(e, iteration, mi, norm, img.p) && hipStreamSynchronize(e.stream) == hipSuccess;
    e.merge_sets = nullptr;
    e.count_events = saved_count;
    const long long nq = ok ? (long long)e.last_queries : -1;
    if (ok && nq > 0) {
        const long long n = std::min(nq, cap);
        std::vector<float4> qp((size_t)nq), qs((size_t)nq);
        std::vector<unsigned long long> fs(2 * slots);
        ok = hipMemcpy(qp.data(), e.q_pos.p, (size_t)nq * 16, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(qs.data(), e.q_sn.p, (size_t)nq * 16, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(fs.data(), sets.p, 2 * slots * 8, hipMemcpyDeviceToHost) == hipSuccess;
        for (long long i = 0; ok && i < n; ++i) {
            uint32_t slot;
            memcpy(&slot, &qs[(size_t)i].w, 4);
            query_pos[3 * i] = qp[(size_t)i].x;
            query_pos[3 * i + 1] = qp[(size_t)i].y;
            query_pos[3 * i + 2] = qp[(size_t)i].z;
            query_slot[i] = slot;
            set_count[i] = fs[2 * (size_t)slot];
            set_hash[i] = fs[2 * (size_t)slot + 1];
        }
    }
    sets.release();
    img.release();
    if (r2_out) *r2_out = mi.pm.r2;
    if (!ok) {
        set_error("rc_merge_sets failed");
        return -1;
    }
    return nq;
}

__attribute__((visibility("default"))) int rc_camera_pass_photons(int iteration, const float* photons, long long count,
                                                                 float* image) {
    if (!g || !prepare(*g)) return -1;
    Engine& e = *g;
    if (!load_photons(photons, count)) return -1;
    float radius, r2, norm;
    iteration_params(e, iteration, radius, r2, norm);
    e.map_ids = count;  // photon ids are the input indices
    MapInfo mi;
    bool ok = build_map(e, (uint32_t)count, radius, r2, mi);
    e.map_ids = 0;
    if (!ok) return -1;
    size_t n = (size_t)e.W * e.H * 3;
    DevBuf<float> img;
    if (!img.alloc(n)) return -1;
    ok = hipMemcpy(img.p, image, n * 4, hipMemcpyHostToDevice) == hipSuccess &&
         run_camera_pass(e, iteration, mi, norm, img.p) && hipStreamSynchronize(e.stream) == hipSuccess &&
         hipMemcpy(image, img.p, n * 4, hipMemcpyDeviceToHost) == hipSuccess;
    img.release();
    if (!ok) set_error("rc_camera_pass_photons failed");
    return ok ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_autotune(int repeats, double* report, int n) {
    g_error.clear();
    if (!g) {
        set_error("rc_autotune: no scene");
        return -1;
    }
    double r[RC_TUNE_REPORT] = {0};
    if (!autotune(*g, repeats, r)) return -1;
    if (report) std::copy(r, r + std::max(0, std::min(n, (int)RC_TUNE_REPORT)), report);
    return 0;
}

// ---- batches ----
__attribute__((visibility("default"))) int rc_batch_add(void) {
    g_error.clear();
    if (!g || !g->initialised) {
        set_error("rc_batch_add: no scene (call init first)");
        return -1;
    }
    if (!g->init_error.empty()) {
        set_error(g->init_error);
        return -1;
    }
    if (!g_batch) g_batch = new Batch();
    if (g_batch->prepared) {
        set_error("rc_batch_add: the batch is already prepared");
        return -1;
    }
    g_batch->scenes.push_back(g);
    g = nullptr;
    return (int)g_batch->scenes.size() - 1;
}

__attribute__((visibility("default"))) int rc_batch_set_option(int scene, const char* key, double value) {
    if (!g_batch || scene < 0 || scene >= (int)g_batch->scenes.size()) {
        set_error("rc_batch_set_option: no such scene");
        return -1;
    }
    Engine* saved = g;
    g = g_batch->scenes[scene];
    int rc = rc_set_option(key, value);
    g = saved;
    return rc;
}

__attribute__((visibility("default"))) int rc_batch_size(void) { return g_batch ? (int)g_batch->scenes.size() : 0; }

__attribute__((visibility("default"))) int rc_batch_prepare(void) {
    g_error.clear();
    if (!g_batch) {
        set_error("no batch");
        return -1;
    }
    return batch_prepare(*g_batch) ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_batch_render_iteration(void) {
    if (rc_batch_prepare() != 0) return -1;
    if (g_cancel.load()) {
        set_error("cancelled");
        return -2;
    }
    int it = g_batch->iterations;
    return batch_iteration(*g_batch) ? it : -1;
}

__attribute__((visibility("default"))) int rc_batch_render_scene(int scene) {
    if (rc_batch_prepare() != 0) return -1;
    if (scene < 0 || scene >= (int)g_batch->scenes.size()) {
        set_error("rc_batch_render_scene: no such scene");
        return -1;
    }
    Engine& e = *g_batch->scenes[scene];
    int it = e.iterations_done;
    if (!render_iteration(e, sample_iteration(e, it))) return -1;
    e.iterations_done = it + 1;
    return it;
}

__attribute__((visibility("default"))) int rc_batch_get_image(int scene, float* rgb, int scaled) {
    if (!g_batch || !g_batch->prepared || scene < 0 || scene >= (int)g_batch->scenes.size() || !rgb) {
        set_error("rc_batch_get_image: no such rendered scene");
        return -1;
    }
    return get_image(*g_batch->scenes[scene], rgb, scaled != 0) ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_batch_autotune(int repeats, double* report, int n) {
    g_error.clear();
    if (!g_batch) {
        set_error("no batch");
        return -1;
    }
    double r[RC_TUNE_REPORT] = {0};
    if (!batch_autotune(*g_batch, repeats, r)) return -1;
    if (report) std::copy(r, r + std::max(0, std::min(n, (int)RC_TUNE_REPORT)), report);
    return 0;
}

__attribute__((visibility("default"))) int rc_batch_stats(double* out, int n) {
    if (!g_batch || !out) return 0;
    int m = std::min(n, (int)RC_STAT_COUNT);
    std::copy(g_batch->stats, g_batch->stats + m, out);
    return m;
}

__attribute__((visibility("default"))) void rc_batch_release(void) { release_batch(g_batch); }

// ---- learned emission proposal ----
static GuideModel* guide_ready() {
    g_error.clear();
    if (!g || !prepare(*g)) return nullptr;
    GuideModel& gd = g->guide;
    if (!gd.params_ready && !gd.init_params(g->stream)) return nullptr;
    return &gd;
}

// Uploads n rows of 6 floats as 8-float records, padded with zero rows to a multiple of 64.
static bool upload_rows(GuideModel& gd, const float* x, long long n, int padded) {
    std::vector<float> rec((size_t)padded * kGuideRowFloats, 0.0f);
    for (long long i = 0; i < n; ++i)
        for (int k = 0; k < kGuideIn; ++k) rec[(size_t)i * kGuideRowFloats + k] = x[i * kGuideIn + k];
    if (!gd.train_x.alloc(rec.size())) return false;
    HIP_OK(hipMemcpy(gd.train_x.p, rec.data(), rec.size() * sizeof(float), hipMemcpyHostToDevice));
    return true;
}

__attribute__((visibility("default"))) int rc_guide_get_params(float* params, long long n) {
    GuideModel* gd = guide_ready();
    if (!gd) return -1;
    if (!params || n != kGuideParams) {
        set_error("rc_guide_get_params: n must be RC_GUIDE_PARAMS");
        return -1;
    }
    return gd->get_params(params, g->stream) ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_guide_set_params(const float* params, long long n) {
    GuideModel* gd = guide_ready();
    if (!gd) return -1;
    if (!params || n != kGuideParams) {
        set_error("rc_guide_set_params: n must be RC_GUIDE_PARAMS");
        return -1;
    }
    return gd->set_params(params, g->stream) ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_guide_forward(const float* x, long long n, float* logits) {
    GuideModel* gd = guide_ready();
    if (!gd) return -1;
    if (!x || !logits || n <= 0 || n > (1 << 20)) {
        set_error("rc_guide_forward: 1 <= n <= 2^20 rows and non-null buffers required");
        return -1;
    }
    int padded = (int)((n + 63) / 64 * 64);
    bool ok = upload_rows(*gd, x, n, padded) && gd->forward(gd->train_x.p, padded, g->stream) &&
              hipMemcpyAsync(logits, gd->Z.p, (size_t)n * kGuideBins * sizeof(float), hipMemcpyDeviceToHost,
                             g->stream) == hipSuccess &&
              hipStreamSynchronize(g->stream) == hipSuccess;
    gd->table_ready = false;  // the work buffers were reused
    if (!ok) set_error("rc_guide_forward failed");
    return ok ? 0 : -1;
}

__attribute__((visibility("default"))) int rc_guide_train_step(const float* x, const int* bins, const float* w,
                                                              long long n, float* grad, float* loss, int apply) {
    GuideModel* gd = guide_ready();
    if (!gd) return -1;
    if (!x || !bins || !w || n <= 0 || n > (1 << 20)) {
        set_error("rc_guide_train_step: 1 <= n <= 2^20 rows and non-null inputs required");
        return -1;
    }
    for (long long i = 0; i < n; ++i)
        if (bins[i] < -1 || bins[i] >= kGuideBins || !(w[i] >= 0.0f)) {
            set_error("rc_guide_train_step: bins must be in [-1, 1023] and weights >= 0");
            return -1;
        }
    int padded = (int)((n + 63) / 64 * 64);
    std::vector<int> b(padded, -1);
    std::vector<float> ww(padded, 0.0f);
    std::copy(bins, bins + n, b.begin());
    std::copy(w, w + n, ww.begin());
    bool ok = upload_rows(*gd, x, n, padded) && gd->train_bin.alloc(padded) && gd->train_w.alloc(padded) &&
              hipMemcpy(gd->train_bin.p, b.data(), padded * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(gd->train_w.p, ww.data(), padded * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
              gd->forward(gd->train_x.p, padded, g->stream, false) &&
              gd->backward(gd->train_bin.p, gd->train_w.p, padded, g->stream) && (!apply || gd->adam(g->stream));
    float sc[4] = {0, 0, 0, 0};
    ok = ok && hipMemcpyAsync(sc, gd->scal.p, sizeof sc, hipMemcpyDeviceToHost, g->stream) == hipSuccess;
    if (ok && grad)
        ok = hipMemcpyAsync(grad, gd->grad.p, sizeof(float) * kGuideParams, hipMemcpyDeviceToHost, g->stream) ==
             hipSuccess;
    ok = ok && hipStreamSynchronize(g->stream) == hipSuccess;
    gd->table_ready = false;
    if (ok && loss) *loss = sc[2];
    if (!ok) set_error("rc_guide_train_step failed");
    return ok ? 0 : -1;
}

__attribute__((visibility("default"))) long long rc_guide_table(float* cdf, float* inputs, long long max_rows) {
    GuideModel* gd = guide_ready();
    if (!gd) return -1;
    if (!gd->table_ready && !gd->build_table(g->stream)) return -1;
    long long rows = std::min<long long>(gd->table_rows, std::max(0ll, max_rows));
    std::vector<float> t((size_t)rows * kGuideCdfStride);
    if (rows > 0 && cdf) {
        if (hipMemcpy(t.data(), gd->cdf.p, t.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("rc_guide_table: copy failed");
            return -1;
        }
        for (long long r = 0; r < rows; ++r)
            std::copy(t.begin() + r * kGuideCdfStride, t.begin() + r * kGuideCdfStride + kGuideBins + 1,
                      cdf + r * (kGuideBins + 1));
    }
    if (inputs)
        for (long long r = 0; r < rows; ++r)
            for (int k = 0; k < kGuideIn; ++k) inputs[r * kGuideIn + k] = gd->table_inputs[r * kGuideRowFloats + k];
    return gd->table_rows;
}

__attribute__((visibility("default"))) long long rc_guide_records(float* x, int* bins, long long n) {
    GuideModel* gd = guide_ready();
    if (!gd) return -1;
    long long avail = std::min<long long>(n, std::min<long long>(gd->train_bin.n, guide_rows(*g)));
    if (avail <= 0 || !gd->train_x.p) return 0;
    std::vector<float> rec((size_t)avail * kGuideRowFloats);
    if (hipMemcpy(rec.data(), gd->train_x.p, rec.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
        (bins && hipMemcpy(bins, gd->train_bin.p, avail * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)) {
        set_error("rc_guide_records: copy failed");
        return -1;
    }
    if (x)
        for (long long i = 0; i < avail; ++i)
            for (int k = 0; k < kGuideIn; ++k) x[i * kGuideIn + k] = rec[i * kGuideRowFloats + k];
    return avail;
}

}  // extern "C"
