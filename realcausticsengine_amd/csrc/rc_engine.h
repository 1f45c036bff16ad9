// rc_engine.h -- device data layout of the vertex-merging engine and the
// launchers the host side calls.  Layout in HBM:
//   triangles   3 x float4 per triangle in BVH-leaf order: {v0.xyz, material},
//               {v1-v0, 0}, {v2-v0, 0}  (48 B; the three world-space points of
//               Triangle::hit, src/Triangle.cpp:20-88, pre-transformed once)
//   shading     3 x float4 per triangle: vertex normals (world, normalised)
//   nodes       4 x float4 per binary BVH node: both child boxes + child refs
//               (64 B; a ref >= 0 is a node, < 0 is ~(first << 3 | count - 1))
//   photons     SoA after the Morton/hash sort: float4 {pos, dVM},
//               float4 {dir, throughput.r}, float2 {throughput.g, .b}
//   cells       uint2 {start, end} per hash bucket
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rc_guide.h"

namespace rc {

// BSDF classes of the reference (materials.h:59-146), plus BK_RGLASS: glass of
// roughness >= 0.01 as a GGX microfacet dielectric (option compat_glass = 0;
// the reference's glass is always a delta interface).
enum { BK_SPEC_REFL = 0, BK_FRESNEL_SPEC = 1, BK_FMG = 2, BK_FSMG = 3, BK_MRG = 4, BK_RGLASS = 5 };
enum { EV_REFLECTION = 1, EV_TRANSMISSION = 2, EV_DIFFUSE = 4, EV_GLOSSY = 8, EV_SPECULAR = 16 };
enum { L_AREA = 0, L_SUN = 1, L_POINT = 2 };

constexpr int kLeafMax = 8;
constexpr int kBlock = 128;      // threads per block of the tracing kernels
constexpr int kLdsStack = 8;     // traversal stack entries ({node ref, entry distance}) held in LDS per lane
constexpr int kMaxDepth = 64;    // total stack capacity (LDS + private overflow)
constexpr int kEmptyChild = (int)0x80000000;  // missing child of a 4-wide BVH node

// Traversal stack entries a tree can need.  The binary traversal pushes at
// most one entry (the far child) per internal ancestor of the node it is in,
// so a tree whose deepest leaf has `leaf_depth` internal ancestors needs that
// many.  The 4-wide traversal pushes at most three entries per 4-wide node on
// the path, so a tree whose deepest internal 4-wide node sits at depth
// `depth4` (root 0) needs 3 (depth4 + 1).  The builders keep a tree only when
// the traversal that runs it fits kMaxDepth; the kernels do not check.
constexpr int stack_need_binary(int leaf_depth) { return leaf_depth; }
constexpr int stack_need_wide(int depth4) { return 3 * (depth4 + 1); }
// Traversal width for a tree: `requested` (4 or 2) when its stack fits,
// else 2 when the binary stack fits, else 0 (the tree cannot be traversed).
constexpr int traversal_width(int requested, int leaf_depth, int depth4) {
    return (requested == 4 && stack_need_wide(depth4) <= kMaxDepth)
               ? 4
               : (stack_need_binary(leaf_depth) <= kMaxDepth ? 2 : 0);
}

struct DevMaterial {
    int kind;
    float R[3];
    float a2, s, ior, cont;
};

struct DevLight {
    int kind;
    float power, radius, pad;
    float color[3], bl[3], h[3], v[3], normal[3], pos[3];
    float fx[3], fy[3], fz[3];
};

struct SceneView {
    const float4* tri;
    const float4* shade;
    const int* tri_id;
    const float4* nodes;
    int root;
    const float4* nodes4;  // 4-wide BVH (8 float4 per node), or null: traverse the binary nodes
    int root4;
    const DevMaterial* mats;
    const DevLight* lights;
    const float* cdf;
    int nlights;
    float integral;
    float cam_o[3], cam_bl[3], cam_h[3], cam_v[3];
    int W, H;
    int max_path_length;
    int compat_point, compat_pdf, compat_rows;
};

struct PhotonMapView {
    const float4* pos;     // sorted {pos, dVM}
    const float4* dir_tr;  // sorted {dir, tp.r}
    const float2* tp_gb;   // sorted {tp.g, tp.b}
    const uint2* cells;    // [start, end) per bucket
    uint32_t mask;         // buckets - 1
    float origin[3];
    float inv_cell;
    float r2;       // merge radius squared (strict d2 < r2)
    float r_query;  // radius with margin used to pick the cells
};

// Event counts of one pass (counting launches only): they define the
// algorithmic bytes of the roofline figures.
struct Counters {
    unsigned long long paths, nodes, tris, hits, photons_stored, gathers, cells, photons_tested, photons_accepted;
    unsigned long long photons_streamed;  // photon records read by the merge (once per shared window)
};
constexpr int kCounterFields = 10;

// Several scenes rendered by the same launches: global light and camera path
// indices are cut into per-scene ranges; each scene keeps its own device
// arrays (SceneView), Philox key, photon-map grid (PhotonMapView, with its
// own range of the shared bucket table) and film.
struct BatchScene {
    uint32_t light_base, light_paths;    // global light-path range
    uint32_t camera_base, camera_paths;  // global camera-path range
    uint32_t k0, k1;                     // Philox key (seed) of the scene
    float vm_norm;                       // 1 / (pi r^2 N) of this iteration
    uint32_t bucket_base;                // first entry of the scene in the shared bucket table
    uint32_t iteration;                  // the scene's iteration index (Philox counter)
    float* image;                        // the scene's film
};
struct BatchView {
    int n;                           // scenes (0: single-scene launch)
    const SceneView* sc;             // per scene
    const PhotonMapView* pm;         // per scene (cells already offset by bucket_base)
    const BatchScene* bs;            // per scene
    const uint32_t* light_starts;    // n + 1 prefix sums of light paths
    const uint32_t* camera_starts;   // n + 1 prefix sums of camera paths
    const GuideView* guides;         // per scene emission proposal (the plain view for unguided scenes)
};

// Training targets of one scene of a batch: w[local path] for local paths < rows.
struct GuideTargets {
    float* w;
    uint32_t rows;
};

// Scene of a global path index: the last s with starts[s] <= i.
__device__ __forceinline__ int batch_scene_of(const uint32_t* starts, int n, uint32_t i) {
    int lo = 0, hi = n;  // starts[lo] <= i < starts[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (starts[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// An unsorted photon as the light pass writes it: kPhotonRecord float4,
//   {position, dVM}  {direction towards the light, throughput.r}
//   {throughput.g, throughput.b, light path index (bits), path length (bits)}
//   (a fourth, unused: records start on 64-B boundaries)
// one record per photon, so that the map's permutation gathers one record per
// photon, and an aligned 64-B record lies in one 128-B cache line (a 48-B one
// spans two lines for 3 records in 8); the light path index is the map's sort
// key below the bucket.
constexpr int kPhotonRecord = 4;

struct IterParams;

// Photon-map keys made by the light pass as it writes its photons: the key
// (bucket, light path) of the photon in record slot i goes to keys[i], i to
// vals[i], and the count of the key's low 8-bit digit (the sort's first pass)
// into dig (8 copies of 256 counts, by workgroup index modulo 8).  The grid
// is origin / inv_cell / mask, or the light pass's dyn values, or per scene
// the batch's photon-map views (bucket = the scene's bucket_base + its own).
struct MapKeys {
    unsigned long long* keys;  // null: the light pass makes no keys
    uint32_t* vals;
    uint32_t* dig;
    float origin[3];
    float inv_cell;
    uint32_t mask;
    int uid_bits;
};

struct LightPassArgs {
    SceneView sc;
    uint32_t k0, k1, iteration, stream;
    uint32_t n_paths;
    uint32_t path_chunk;  // path indices a wave takes per refill (>= 64)
    uint32_t refill_min;  // a wave refills when at least this many lanes are idle (1..64) or none is live
    uint32_t* path_counter;
    uint32_t* photon_counter;
    uint32_t capacity;
    float4* out_rec;     // kPhotonRecord float4 per photon (layout below)
    Counters* counters;  // non-null selects the counting variant
    GuideView guide;     // learned emission proposal (all slots -1: plain)
    BatchView batch;     // batch.n > 0: several scenes (sc, k0, k1, proposal per scene)
    const uint32_t* path_order;  // non-null: slot s traces path path_order[s] (paths sorted for coherence)
    // non-null: every slot's emission (light sample and proposal record) is made by a coherent kernel
    // before the persistent pass: 3 float4 per slot {o, tp.x}, {d, tp.y}, {tp.z, emits, 0, 0}
    float4* emitted;
    // non-null: the iteration (Philox counter) is read here (a captured iteration graph)
    const IterParams* dyn;
    // adaptive MLT light pass (non-null mlt_state selects it; single scene, plain proposal)
    uint4* mlt_state;     // kMltBlocks * mlt_chains explicit chain states, block-major
    uint32_t mlt_chains;  // Markov chains; chain c takes the path indices c, c + chains, ...
    uint32_t* mlt_counts; // [0] visible uniform paths, [1] mutations, [2] accepted mutations
    MapKeys map;          // photon-map keys written with the photons (map.keys null: none)
};
// Philox blocks of a light path's primary sample (light sample 0-1, bounces 2-9).
constexpr int kMltBlocks = 10;

// Per-iteration values read from device memory by the kernels of a captured
// iteration graph (a graph replays fixed kernel arguments).
struct IterParams {
    uint32_t iteration;
    float r2, r_query, inv_cell;
    float origin[3];
    float vm_norm;
};

// One merge query per non-delta camera vertex (80 B, SoA):
//   pos  {position, camera dVM}          dir {ray direction, material}
//   sn   {shading normal, slot}          gn  {geometric normal, 0}
//   tpn  {throughput * VM normalisation, 0}
// slot = path * 8 + path_length: the merge result lands in contrib[slot] and
// the resolve adds a path's slots in vertex order, as the reference does.
struct QueryBuffers {
    float4* pos;
    float4* dir;
    float4* sn;
    float4* gn;
    float4* tpn;
    uint32_t capacity;
};

// Path counters of a pass: [0] paths taken, [1] queries written, [2..3] pass
// bookkeeping, [4..11] the grouped merge's per-XCD group counters.
constexpr int kPathCounters = 12;

struct CameraPassArgs {
    SceneView sc;
    PhotonMapView pm;
    uint32_t k0, k1, iteration, stream;
    uint32_t n_paths;
    uint32_t path_chunk;      // path indices a wave takes per refill (>= 64)
    uint32_t refill_min;      // a wave refills when at least this many lanes are idle (1..64) or none is live
    uint32_t* path_counter;   // [0] paths, [1] queries, [2] gather work counter
    float vm_norm;
    QueryBuffers q;
    int2* path_info;          // per path: {image buffer index or -1, mask of vertices with a query}
    float* contrib;           // 3 floats per slot
    float* image;
    Counters* counters;  // non-null selects the counting variant
    int exact_merge;     // 1: correctly rounded division/sqrt in the merge
    BatchView batch;     // batch.n > 0: several scenes (queries carry their scene in gn.w)
    const IterParams* dyn;  // non-null: iteration, vm_norm and the grid of pm come from here
    const uint32_t* q_order;  // non-null: the gather takes the queries in this order (spatially sorted)
    int camera_qmc;           // 1: jitter and first-bounce choice stratified across iterations
    int group_queries;        // 4, 8 or 16: grouped merge of that many queries per wave (queries with the
                              // same cell window share photon loads); 0: one query per wave; 1: one per lane
    uint32_t gather_tile;     // grouped merge: XCD x takes the tiles t = x mod 8 of this many queries of the
                              // order; 0: XCD x takes the x-th contiguous eighth
    int gather_block;         // grouped merge: 2 or 4 = photons taken 64 x this many at a time and the query
                              // loop rolled (sums in LDS); 0 = the query loop unrolled (sums in registers)
    // counting launches (counters non-null) with merge_sets non-null: per
    // contribution slot, {photons the merge accepted, sum of mix64 of their
    // input indices}; map_index = input index of each sorted photon
    unsigned long long* merge_sets;
    const uint32_t* map_index;
};

// SplitMix64 finaliser (Steele, Lea and Flood 2014): the per-photon term of
// the merge-set fingerprints.
__host__ __device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// ---- launchers (rc_kernels.hip) ----
void launch_light_pass(const LightPassArgs& a, int blocks, hipStream_t s);
void launch_camera_trace(const CameraPassArgs& a, int blocks, hipStream_t s);
void launch_gather(const CameraPassArgs& a, int blocks, hipStream_t s);
void launch_resolve(const CameraPassArgs& a, hipStream_t s);
// Spatial order of the merge queries: key = scene << 36 | Morton code of the
// low corner of the query's photon-map cell window (10 bits per axis) << 6 |
// window extents, one stable radix sort; the query
// count comes from a.path_counter[1] (at most n).  Returns the order array
// (one of vals_a / vals_b).
struct QuerySortBuffers {
    unsigned long long* keys_a;
    unsigned long long* keys_b;
    uint32_t* vals_a;
    uint32_t* vals_b;
    uint32_t* hist;
    uint32_t* hist_sums;
};
const uint32_t* sort_queries(const CameraPassArgs& a, const QuerySortBuffers& b, uint32_t n, hipStream_t s);
// Coherent order of the light paths: key = (scene, top bits of the light-pick word, 4-D Morton code of the
// emission position and direction words), one stable radix sort.  Returns the order (one of the vals buffers).
const uint32_t* sort_light_paths(const LightPassArgs& a, const QuerySortBuffers& b, hipStream_t s);
void launch_first_hit(const SceneView& sc, const float* rays, int n, float tmin, int* tri, float* t, hipStream_t s);
// Debug AOV of debug_test mode 2, 3 or 4 added into image (W * H * 3).
void launch_debug_aov(const SceneView& sc, int mode, float ior, float* image, hipStream_t s);

// Photon map build: keys, sort, permute, cell ranges.
struct MapBuildArgs {
    const float4* in_rec;     // unsorted photon records (kPhotonRecord float4 each)
    uint32_t count;           // photons (with count_dev: the capacity the grids are sized for)
    const uint32_t* count_dev;  // non-null: the photon count is read here (captured graphs)
    const IterParams* dyn;      // non-null: origin and inv_cell come from here
    float origin[3];
    float inv_cell;
    uint32_t mask;
    int bucket_bits, uid_bits;
    // scratch
    unsigned long long* keys_a;
    unsigned long long* keys_b;
    uint32_t* vals_a;
    uint32_t* vals_b;
    uint32_t* hist;       // 256 * max_blocks
    uint32_t* hist_sums;  // scan partials
    // outputs
    float4* out_pos;
    float4* out_dir;
    float2* out_tp;
    uint32_t* out_index;  // original index of each sorted photon
    uint2* cells;         // mask + 1 entries
    // true: the light pass made the keys (keys_a, vals_a) and their first digit counts, into the
    // sort scratch laid out for scratch_n keys and cleared before it (clear_map_sort)
    bool keys_ready;
    uint32_t scratch_n;  // keys the sort scratch is laid out for (0: count)
};
void build_photon_map(const MapBuildArgs& a, hipStream_t s);
// The sort scratch (hist) laid out for n keys: zeroes its digit counts, tickets and first status
// rows before a light pass makes keys into it; map_sort_digits: where those counts go.
void clear_map_sort(uint32_t* hist, uint32_t n, hipStream_t s);
uint32_t* map_sort_digits(uint32_t* hist, uint32_t n);
// Batch: the keys of every scene's photons in one sort (bucket = the scene's
// bucket_base + its own cell bucket, above the global light path index).
void build_photon_map_batch(const MapBuildArgs& a, const BatchView& b, hipStream_t s);
// Training targets of the scenes of a batch (rc_guide.hip): vis[bucket_base +
// bucket] = 1 for the cell of every query (its scene in gn.w), then
// t[scene].w[local path] += luminance of each photon of a training path in a
// visible cell of its scene.
void launch_mark_visible_batch(const float4* qpos, const float4* qgn, uint32_t nq, const BatchView& b, uint8_t* vis,
                               hipStream_t s);
void launch_guide_targets_batch(const float4* rec, uint32_t count, const BatchView& b, const uint8_t* vis,
                                const GuideTargets* t, hipStream_t s);
// Stable LSD radix sort (8-bit digits, one kernel per digit) of (key, value)
// pairs on the low `bits` key bits; the sorted pairs end in *keys_out /
// *vals_out (one of the two buffers).  hist: photon_map_hist_entries(n)
// entries of scratch; hist_sums is not used.
void radix_sort_pairs(unsigned long long* keys_a, unsigned long long* keys_b, uint32_t* vals_a, uint32_t* vals_b,
                      uint32_t* hist, uint32_t* hist_sums, const uint32_t* count_dev, uint32_t n, int bits,
                      unsigned long long** keys_out, uint32_t** vals_out, hipStream_t s);
// The same sort, dropping from the keys it writes up to `droppable` low bits
// once they are sorted (the caller reads only the bits above them); keys
// travel as 32-bit words once the rest fits.  The sorted keys are `keys`
// (32-bit words when keys32), shifted right by `dropped`; with droppable >=
// bits no keys are returned.
struct SortedPairs {
    void* keys;
    uint32_t* vals;
    bool keys32;
    int dropped;
};
SortedPairs radix_sort_pairs_drop(unsigned long long* keys_a, unsigned long long* keys_b, uint32_t* vals_a,
                                  uint32_t* vals_b, uint32_t* hist, const uint32_t* count_dev, uint32_t n, int bits,
                                  int droppable, hipStream_t s);
// GPU BVH build (rc_bvh.hip): linear BVH over Morton codes (one triangle per
// leaf) in the binary node layout; tri / shade / tri_id are reordered into
// leaf order.  Returns the tree depth (< 0 on a launch error).
struct LbvhBuffers {
    float4* tri;      // 3 per triangle, reordered in place (via tmp)
    float4* shade;
    int* tri_id;
    float4* tmp4;     // 6 * n float4 scratch
    int* tmp_id;      // n
    float4* nodes;    // 4 * max(1, n - 1) float4: the binary nodes, root 0
    unsigned long long* keys_a;
    unsigned long long* keys_b;
    uint32_t* vals_a;
    uint32_t* vals_b;
    uint32_t* hist;
    uint32_t* hist_sums;
    int* parent;      // 2 * n: parents of internal nodes, then of leaves
    uint32_t* flags;  // n
    float* boxes;     // 6 * n: unpadded boxes of the internal nodes
    uint32_t* depth;  // 1
};
int build_lbvh(const LbvhBuffers& b, uint32_t n, const float lo[3], const float hi[3], hipStream_t s);
// After build_lbvh (same buffers): the number of 4-wide nodes of the collapsed
// tree (-1 on error), then the collapse into nodes4 (8 float4 per node, root 0).
int lbvh_wide_count(const LbvhBuffers& b, uint32_t n, hipStream_t s);
bool lbvh_collapse4(const LbvhBuffers& b, uint32_t n, float4* nodes4, hipStream_t s);
size_t photon_map_hist_entries(uint32_t max_count);
size_t photon_map_sum_entries(uint32_t max_count);

// Radius sets through the photon map (count then fill).
// Radius sets without device allocation (rc_gather_sets_noalloc): the
// photon-map build's scratch and the per-query counts live in one
// file-scope device array sized for these maxima.
constexpr uint32_t kNoAllocMaxPhotons = 1u << 18;
constexpr uint32_t kNoAllocMaxQueries = 1u << 18;
constexpr uint32_t kNoAllocMaxBuckets = 1u << 15;
struct NoAllocGather {
    const float* positions;  // device, 3 floats per photon
    uint32_t count;
    const float* queries;  // device, 3 floats per query
    uint32_t nq;
    PhotonMapView pm;  // the grid (origin, inv_cell, mask, r2, r_query); the launcher sets its arrays
    int bucket_bits, uid_bits;
    unsigned long long* offsets;  // device, nq + 1
    int* indices;                 // device, cap
    long long cap;
};
// Returns the total set size (indices written only when it is <= cap), or -1
// with *err set.  Synchronous on `s`.
long long gather_sets_noalloc(const NoAllocGather& a, hipStream_t s, const char** err);
size_t noalloc_scratch_bytes();
void launch_gather_count(const PhotonMapView& pm, const float* q, int nq, const uint32_t* orig, uint32_t* counts,
                         hipStream_t s);
void launch_gather_fill(const PhotonMapView& pm, const float* q, int nq, const uint32_t* orig,
                        const unsigned long long* offsets, int* out, hipStream_t s);

// Scene ingest on the GPU: one mesh's Blender arrays, copied to the device as
// raw bytes (MVert 20 B: float co[3], short no[3], flag, bweight; MPoly 12 B:
// int loopstart, int totloop, short mat_nr, flag, pad; MLoop 8 B: unsigned v,
// unsigned e), with its object matrix.
struct IngestMesh {
    const unsigned char* mvert;
    const unsigned char* mpoly;
    const unsigned char* mloop;
    const int* mat_map;  // material slot (mat_nr, negative as 0) -> scene material
    uint32_t nv, nt;     // vertices, polygons (triangles: the first three loops of each)
    uint32_t vbase, tbase;  // first global vertex and triangle of the mesh
    float M[12];         // rows 0-2 of the row-major object matrix
    float inv[9];        // inverse of its upper 3x3, row-major (normals: inv^T n)
};
// Bits of the ingest's flag word.
constexpr uint32_t kIngestNonFinite = 1u;    // a vertex coordinate is not finite
constexpr uint32_t kIngestLoopVertex = 2u;   // a loop's vertex index is >= the mesh's vertex count
// World-space triangles {p0, material bits}, {e1, 0}, {e2, 0}, shading normals
// (3 per triangle), ids (input order) and the scene bounds (6 order-preserving
// int encodings of lo.xyz, hi.xyz; bounds[0..5] start at +inf / -inf) of
// n_meshes meshes holding nv vertices and nt triangles; wv and wn hold nv
// float4 each (scratch).  Asynchronous on s.
void launch_ingest(const IngestMesh* meshes, int n_meshes, uint32_t nv, uint32_t nt, float4* wv, float4* wn,
                   float4* tri, float4* shade, int* tri_id, int* bounds, uint32_t* flags, hipStream_t s);
// The order-preserving int of a float (the bounds' encoding; its own inverse).
__host__ __device__ inline int ordered_float_bits(int bits) { return bits >= 0 ? bits : bits ^ 0x7fffffff; }

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// dst[i] = src[i] for n 32-bit words (a kernel: graph captures hold no memory-copy nodes).
void launch_copy_u32(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s);
void launch_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* partials, hipStream_t s);

}  // namespace rc
