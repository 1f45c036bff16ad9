/*
 * real_caustics.h -- C-ABI of libRealCaustics.so, the MI355X (gfx950) drop-in
 * for the exports of "Real Caustics.dll".
 *
 * The first two entry points keep the reference's signatures, struct layouts
 * and call protocol exactly (init uploads a scene, main renders it and writes
 * the image file, then releases the scene).  Every other entry point is a
 * non-breaking addition (prefix rc_): per-iteration rendering, in-memory image
 * readback, options, error reporting, cancellation, and the diagnostic calls
 * the parity tests use.  All calls act on one process-wide scene, as the
 * reference does (Real Caustics/src/main.cpp:59); they are not re-entrant.
 */
#ifndef REAL_CAUSTICS_H
#define REAL_CAUSTICS_H

#ifdef __cplusplus
extern "C" {
#endif

/* Value types passed across the boundary (layouts of the reference). */
typedef struct { float x, y, z; } rc_vec3;                       /* src/vec3.h:13-19 (12 B)  */
typedef struct { float t[16]; } rc_matrix4x4;                     /* src/matrix.h:151-171, row-major t00..t33 (64 B) */
typedef struct {                                                  /* src/Lights.h:15-36 (64 B) */
    char type;        /* 'P' point, 'S' sun, anything else area */
    float color[3];
    float position[3];
    float width, height, spread;
    float rotation[3]; /* degrees for area lights, radians for the sun */
    float power, angle, radius;
} rc_python_light;
typedef struct {                                                  /* src/materials.h:21-32 (28 B) */
    char type;        /* 'P' plastic, 'G' glass, 'M' metal, else white plastic */
    float ior, roughness, specular;
    float color[3];
} rc_python_material;
typedef struct {                                                  /* src/Vertex_Merging.h:13-19 (24 B) */
    int number_of_points;
    rc_vec3 direction;
    float roughness, specular;
} rc_debug_stuff;

/* Replaces init (Real Caustics/src/DLL.h:15, defined at src/main.cpp:89-143).
 * meshes_pointers[i] is the address of a Blender Mesh (mpoly@200, mloop@208,
 * mvert@256, src/Blender_definitions.h:214-311); every poly is read as the
 * triangle of its first three loops (src/mesh.cpp:20-26).  Geometry,
 * materials, lights and camera are copied, so unlike the reference nothing
 * needs to outlive this call.  Invalid input records an error (rc_last_error)
 * that main() then reports. */
void init(int samples, float base_radius, float radius_reduction,
          long long *meshes_pointers, unsigned int number_of_meshes, rc_matrix4x4 *mesh_matrices,
          unsigned int *meshes_number_of_verts, unsigned int *meshes_number_of_tris,
          unsigned int camera_x, unsigned int camera_y, rc_vec3 camera_position,
          rc_vec3 camera_corner0, rc_vec3 camera_corner1, rc_vec3 camera_corner2, rc_vec3 camera_corner3,
          rc_python_light *lights, unsigned int number_of_lights,
          rc_python_material *materials, unsigned int number_of_materials,
          int **materials_indices, int debug_test, rc_debug_stuff debug_stuff);

/* Replaces main (src/DLL.h:14, defined at src/main.cpp:148-223): builds the
 * acceleration structure, renders `samples` iterations on the GPU, averages,
 * writes the fp32 RGB EXR (default path "D:\caustics.exr" as in
 * src/main.cpp:206; option "output_path") and releases the scene.
 * Returns 0 on success, nonzero on error (the reference always returns 0). */
int main(void);

/* ---- additions (no counterpart in the reference) ---- */

/* Options.  Numeric keys: "seed", "photons" (light paths per iteration,
 * 0 = width*height as the reference), "write_exr", "compat_rows",
 * "compat_point_light", "compat_light_pdf", "compat_glass" (1, the default:
 * glass is a delta interface whatever its roughness, as src/materials.h:104;
 * 0: glass of roughness >= 0.01 is a GGX microfacet dielectric of
 * a2 = roughness^2 -- reflection or refraction through a sampled microfacet,
 * non-delta, so photons are stored and merged on it), "max_path_length",
 * "photon_capacity", "cell_scale", "persistent_blocks", "count_events",
 * "exact_merge" (1 = correctly rounded division and sqrt in the merge),
 * "next_iteration" (index of the next iteration rc_render_iteration renders).
 * Learned emission proposal: "guide" (1 = sample area and sphere lights from
 * the trained proposal, mixed with plain emission, and train it every
 * iteration), "guide_alpha" (plain share, default 0.2), "guide_lr" (Adam,
 * 1e-3), "guide_batch" (training paths per iteration, 65536), "guide_cells"
 * (position cells per light side, multiple of 8, default 16), "guide_warmup"
 * (Adam steps before the proposal is sampled, default 2), "guide_train"
 * (0 = keep the network fixed), "guide_train_every" (after warm-up, train on
 * every n-th iteration, default 1), "guide_seed", "guide_reset".
 * Light-pass estimator: "mlt" (1 = the reference's adaptive MLT light pass,
 * Vertex_Merging.cpp:231-282: for each light-path index the uniform sample,
 * and when it stores no photon a mutation of a Markov chain's state, kept
 * when it stores one; the merge is scaled by visible uniform paths / N), with
 * min("mlt_chains", N / 64) independent chains (default at most 262144, so
 * a chain takes at least 64 indices; chain c takes c, c + chains, ...).
 * Single scenes, without the learned proposal.
 * Acceleration structure: "bvh_build" (0 = binned SAH on the host, the
 * default; 1 = linear BVH over Morton codes built on the GPU, for scenes of
 * millions of triangles), "bvh_width" (4 = the tree collapsed to 4-wide
 * nodes -- on the host for the SAH tree, on the GPU for the linear BVH --
 * the default; 2 = binary nodes), "device_ingest" (1 = the default: with
 * bvh_build 1, init copies the raw Blender arrays to the GPU and converts
 * them there -- the same triangles as the host ingest; init then needs the
 * GPU; 0 = convert on the host).
 * Iteration indices: "iteration_offset", "iteration_stride" (the k-th
 * iteration rendered uses iteration index offset + k * stride for its random
 * numbers, radius and stratified samples; defaults 0 and 1.  K copies of a
 * scene in a batch with offsets 0 .. K-1 and stride K render consecutive
 * iterations of one estimator, K per batched launch).
 * Camera sampling: "camera_qmc" (1 = pixel jitter and the first bounce's
 * reflect / refract choice follow a randomly shifted low-discrepancy sequence
 * over iterations -- iteration i takes point i -- instead of independent
 * draws; every iteration keeps its expectation, the running image converges
 * faster).
 * Merge order: "query_order" (1 = the merge takes the queries sorted by the
 * Morton code of their photon-map cell window, the default; 0 = in trace
 * order); the film is the same either way.  "group_queries" (8 = the default):
 * queries per wave of the grouped merge, in which queries with the same cell
 * window share the photon loads (same terms as one query per wave, summed in
 * another order); 1 = one query per lane, each lane walking its own window
 * (for sparse photon maps; the same terms again in another order).
 * "gather_tile" (1024 = the default): the grouped merge deals
 * tiles of this many consecutive queries of the order to the eight XCDs in
 * turn, and each XCD's waves take its groups one at a time as they free up;
 * 0 = each XCD takes one contiguous eighth, dealt to its waves in a fixed
 * order.  The film is the same for every value.  "gather_block" (2 = the
 * default): the grouped merge takes a sub-run's photons 64 x this many (1, 2
 * or 4) at a time into registers and runs a scalar loop over the queries,
 * whose per-lane sums wait in LDS; 0 = the query loop unrolled, sums in
 * registers.  Each query adds its terms in the same order either way: the
 * film is the same.
 * Read-only: "mlt_visible", "mlt_mutations", "mlt_accepted" (counts of the
 * last MLT light pass), "bvh_depth", "bvh_build_ms" (build time of the last rc_prepare),
 * "nodes", "nodes4", "node_bytes" (bytes of the node layout traversed).
 * String keys: "output_path". Return 0, or -1 for an unknown key. */
int rc_set_option(const char *key, double value);
int rc_set_option_str(const char *key, const char *value);
double rc_get_option(const char *key);

/* Builds the BVH and uploads the scene (the prelude of main, src/main.cpp:157-168). */
int rc_prepare(void);
/* Renders the next iteration and accumulates it (one run_iteration,
 * src/Vertex_Merging.cpp:196-366).  Returns the iteration index or <0. */
int rc_render_iteration(void);
/* Renders debug_test mode 2 (shading normals), 3 (facing) or 4 (Fresnel of
 * the view direction, IOR = debug_stuff.specular) into the film, replacing it
 * (src/Debug_scene.cpp:21-205; main() does this for debug_test 2-4). */
int rc_render_debug(int mode);
/* Copies the accumulated image (width*height*3 floats, buffer row order of
 * src/Image_Buffer.h).  scaled != 0 divides by the iterations rendered. */
int rc_get_image(float *rgb, int scaled);
/* Writes the current image as the fp32 RGB OpenEXR file main() writes:
 * tinyexr's layout for the reference's SaveEXR call (src/main.cpp:192-207,
 * lib/tinyexr.h SaveEXR) -- single-part scanline, FLOAT channels B, G, R,
 * ZIP compression in 16-line blocks (none when both sides are below 16). */
int rc_write_exr(const char *path);
/* The same writer for a host image (width * height * 3 floats, buffer row
 * order); needs no scene and no GPU. */
int rc_write_exr_image(const char *path, const float *rgb, int width, int height);
/* Frees the scene and device memory (the end of main, src/main.cpp:220-221). */
void rc_release(void);
/* Last error message, "" when none. */
const char *rc_last_error(void);
/* Requests cancellation; the running render stops after the current iteration. */
void rc_cancel(void);
/* Statistics of the last iteration: see RC_STAT_* indices. Returns the count written. */
int rc_stats(double *out, int n);
enum {
    RC_STAT_LIGHT_PATHS = 0, RC_STAT_PHOTONS, RC_STAT_CAMERA_PATHS, RC_STAT_GATHERS,
    RC_STAT_PHOTONS_TESTED, RC_STAT_PHOTONS_ACCEPTED, RC_STAT_MS_LIGHT, RC_STAT_MS_MAP,
    RC_STAT_MS_CAMERA, RC_STAT_MS_TOTAL, RC_STAT_RADIUS, RC_STAT_NODES, RC_STAT_TRIS,
    /* event counts, filled only when the option "count_events" is 1 */
    RC_STAT_LIGHT_NODE_VISITS, RC_STAT_LIGHT_TRI_TESTS, RC_STAT_LIGHT_HITS,
    RC_STAT_CAMERA_NODE_VISITS, RC_STAT_CAMERA_TRI_TESTS, RC_STAT_CAMERA_HITS, RC_STAT_CELL_VISITS,
    RC_STAT_SORT_PASSES,
    /* camera pass split: trace (includes the query-count readback), merge, resolve */
    RC_STAT_MS_CAMERA_TRACE, RC_STAT_MS_GATHER, RC_STAT_MS_RESOLVE,
    /* learned proposal (option "guide"), on iterations that train: training +
     * table time, mean visible luminance per training path, share of training
     * paths with visible photons, weighted cross-entropy of the step, time of
     * the network step alone (forward, backward, Adam) and its row count */
    RC_STAT_MS_GUIDE, RC_STAT_GUIDE_VISIBLE, RC_STAT_GUIDE_HIT_FRACTION, RC_STAT_GUIDE_LOSS,
    RC_STAT_MS_GUIDE_STEP, RC_STAT_GUIDE_ROWS,
    /* sort of the merge queries into spatial order (option "query_order") */
    RC_STAT_MS_QUERY_SORT,
    /* photon records the merge streamed (count_events): once per query with
     * group_queries = 0 or 1, once per group of queries sharing a window otherwise */
    RC_STAT_PHOTONS_STREAMED,
    /* adaptive MLT light pass (option "mlt"): visible uniform paths, mutations, accepted mutations */
    RC_STAT_MLT_VISIBLE, RC_STAT_MLT_MUTATIONS, RC_STAT_MLT_ACCEPTED,
    /* graphs captured so far (option "graph"; for a batch, scene 0's option) */
    RC_STAT_GRAPH_CAPTURES,
    RC_STAT_COUNT
};

/* Launch shapes and queue / tile sizes are options too: "path_chunk" (path
 * indices a wave takes per queue refill, a multiple of 64), "light_blocks_per_cu",
 * "camera_blocks_per_cu", "gather_blocks_per_cu" (persistent workgroups per
 * compute unit), "cell_scale" (photon-map cell size over the merge radius),
 * "light_refill" / "camera_refill" (a persistent tracing wave takes new paths
 * once at least this many of its 64 lanes are idle, or none is live; 1 = at
 * once), "light_order" (1 = trace the light paths sorted by their emission
 * samples, for ray coherence; the film is the same), "query_order" and
 * "group_queries" (merge queries per wave: 4, 8 or 16; 0 = one per wave;
 * 1 = one per lane, for sparse photon maps),
 * "gather_tile" (0, 256 or 1024 queries per XCD tile of the merge),
 * "gather_block" (0, 1, 2 or 4 photon steps per block of the merge),
 * "overlap" (1, the default: the camera trace runs on a second stream beside
 * the light pass and the photon-map build, which it does not read; the merge
 * waits for both.  2: the same, with the trace queued after the light pass,
 * so that it fills what the light pass's tail and the photon-map build leave
 * free instead of taking compute units from the light pass first (direct
 * launches of a scene or a batch, and the whole-iteration graph, whose
 * trace branch is then captured after the light pass).  Direct launches
 * and graphs of one scene without "mlt", and batches (scene 0's option); not
 * with "count_events"; the film is the same).
 * rc_autotune chooses them for the current scene: every candidate renders the
 * scene's next iteration into a scratch film (the film and the iteration count
 * are unchanged); each knob keeps its fastest value (best of `repeats`) for the
 * kernels it affects, "overlap" last, by the iteration's wall time (kept only
 * when it saves at least 2 %; the other knobs are timed with the camera trace
 * serial).  report (up to RC_TUNE_REPORT
 * doubles): wall ms of one iteration before (camera trace serial) and after,
 * then path_chunk, light_blocks_per_cu, camera_blocks_per_cu, cell_scale,
 * light_refill, light_order, camera_refill, group_queries, gather_tile,
 * gather_block, gather_blocks_per_cu (after the merge's shape, which it
 * depends on), query_order, overlap. */
enum { RC_TUNE_REPORT = 15 };
/* Option "graph" = 1: with the plain light pass the whole iteration (light
 * pass, photon map, camera pass, resolve; with option overlap the camera
 * trace and the query sort as a branch on a second stream inside the graph)
 * replays one captured hipGraph and the host waits once, after it: photons
 * beyond the buffer restore the film from a copy taken in the graph, grow
 * the buffers and run the iteration again.  Otherwise (MLT, a learned
 * proposal) the light pass is a direct launch whose photon count is checked
 * on the host, then the photon-map build and the camera pass replay.
 * Graphs read counts and per-iteration values from device memory and are
 * captured again when buffers grow or the light-path count changes
 * (rc_get_option("graph_captures")).  The film is identical to the one
 * direct launches give.  In a batch, scene
 * 0's "graph" captures the batched map and camera pass of all scenes (the
 * per-scene values come from device tables refilled before each replay;
 * RC_STAT_GRAPH_CAPTURES counts the captures). */
int rc_autotune(int repeats, double *report, int n);

/* ---- diagnostics used by the parity tests ---- */

/* Shape of the BVH that rc_prepare builds on the host (binned SAH, then the
 * 4-wide collapse) over n world-space triangles of 9 floats each; host
 * memory only, no GPU call.  out4 = {internal ancestors of the deepest leaf
 * of the binary tree, depth of the deepest internal 4-wide node, traversal
 * width kept (requested_width 4 or 2; 2 when the 4-wide traversal's stack,
 * 3 entries per 4-wide level, could exceed its 64 entries; 0 when neither
 * fits), stack entries the kept traversal can need}.  Returns 0, or -1 when
 * the tree cannot be traversed (rc_prepare then fails the same way). */
int rc_bvh_shape(const float *tris, long long n, int requested_width, int *out4);

/* Closest hit for n rays {o.xyz, d.xyz}; tri = global triangle id (mesh order,
 * then triangle order) or -1; t = distance. */
int rc_first_hit(const float *rays, long long n, float tmin, int *tri, float *t);
/* Light pass of one iteration; photons in path order as 10 floats
 * {pos.xyz, dir.xyz, throughput.rgb, dVM} and uid = path*8 + path_length
 * (uid non-null needs fewer than 2^29 light paths; the engine's own photon
 * ids are the light path indices, up to 2^31 paths per launch).
 * Returns the photon count (records beyond cap are not written). */
long long rc_trace_photons(int iteration, float *photons, unsigned int *uid, long long cap);
/* Radius sets through the engine's photon map: for each query, the input
 * indices i with fl(fl(dx*dx+dy*dy)+dz*dz) < r2 (src/KD_Tree.cpp:316,349).
 * CSR output; returns the total (indices beyond cap are not written). */
long long rc_gather_sets(const float *positions, long long count, const float *queries, long long nq,
                         float r2, long long *offsets, int *indices, long long cap);
/* Radius sets of the production merge: the camera pass of `iteration`
 * against a given photon list (10 floats each, as rc_trace_photons) with the
 * merge kernel the options select, in its counting form.  Per merge query in
 * trace order (up to cap): position (3 floats), contribution slot (camera
 * path * 8 + vertex), the number of photons the merge accepted and the sum,
 * modulo 2^64, of the SplitMix64 finaliser of their input indices; r2_out =
 * the squared merge radius.  Returns the query count. */
long long rc_merge_sets(int iteration, const float *photons, long long count, float *query_pos, long long *query_slot,
                        unsigned long long *set_count, unsigned long long *set_hash, long long cap, float *r2_out);
/* rc_gather_sets without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture set-up, tight serving loops, harnesses that forbid
 * hipMalloc).  positions, queries, offsets (nq + 1 int64) and indices (cap)
 * are DEVICE pointers; the photon map and the per-query counts use a fixed
 * scratch array of about 36 MB that the library holds on each GPU until the
 * process ends, so count and nq are at most 262144 each (larger sizes are an
 * error).  The scene must be prepared (rc_prepare) first; the grid is the one
 * rc_gather_sets builds, so the sets are the same.  Returns the total set
 * size; indices are written only when it is <= cap.  Synchronous.  Calls that
 * overlap on one GPU share its scratch array and race. */
long long rc_gather_sets_noalloc(const float *positions, long long count, const float *queries, long long nq,
                                 float r2, long long *offsets, int *indices, long long cap);
/* Camera pass of one iteration against a given photon list (10 floats each,
 * as rc_trace_photons); adds into image (width*height*3). */
int rc_camera_pass_photons(int iteration, const float *photons, long long count, float *image);

/* ---- batches of scenes (no counterpart in the reference) ----
 * Several scenes render together, one launch per pass for all of them:
 * define a scene with rc_set_option(...) and init(...) as usual, then
 * rc_batch_add() moves it into the batch (the next init starts a new scene,
 * with default options).  Each scene keeps its own options, seed, radius
 * schedule and film; a scene in a batch renders the same image as alone.
 * "exact_merge" selects the merge kernel and must be equal in all scenes. */
int rc_batch_add(void);               /* returns the scene's index in the batch */
int rc_batch_size(void);
int rc_batch_prepare(void);           /* BVH build and upload of every scene */
int rc_batch_render_iteration(void);  /* one iteration of every scene; returns its index or < 0 */
int rc_batch_render_scene(int scene); /* one iteration of one scene alone (its own launches) */
int rc_batch_get_image(int scene, float *rgb, int scaled);
int rc_batch_stats(double *out, int n);  /* RC_STAT_* totals of the last batch iteration */
int rc_batch_autotune(int repeats, double *report, int n);  /* as rc_autotune, for the whole batch */
/* rc_set_option for one scene of the batch (e.g. "count_events" on scene 0,
 * which switches the event counters of the batched launches on). */
int rc_batch_set_option(int scene, const char *key, double value);
void rc_batch_release(void);

/* ---- learned emission proposal (no counterpart in the reference) ----
 * Parameters are one flat float vector of RC_GUIDE_PARAMS entries: W1
 * [192][64], b1[64], W2 [64][64], b2, W3 [64][64], b3, W4 [64][1024], b4[1024]
 * (weights row-major [in][out]).  An input row is 6 floats: the emission point
 * in scene-box coordinates and the emission axis.  Calls need rc_prepare. */
enum { RC_GUIDE_PARAMS = 87232, RC_GUIDE_BINS = 1024 };
/* Copies the FP32 master parameters out / in (setting resets Adam). */
int rc_guide_get_params(float *params, long long n);
int rc_guide_set_params(const float *params, long long n);
/* Logits (n x 1024) of n input rows through the BF16 matrix-core network. */
int rc_guide_forward(const float *x, long long n, float *logits);
/* One training step on n rows with target bins (-1 = none) and weights:
 * writes the loss sum_i w_i CE_i / sum_i w_i and the gradient (RC_GUIDE_PARAMS
 * floats, either may be null); apply != 0 also takes the Adam step. */
int rc_guide_train_step(const float *x, const int *bins, const float *w, long long n, float *grad, float *loss,
                        int apply);
/* The proposal table: rows = lights with a table x cells^2, each the CDF of
 * 1025 floats over the bins; inputs = the rows' 6-float input records.
 * Returns the row count (nothing beyond max_rows is written). */
long long rc_guide_table(float *cdf, float *inputs, long long max_rows);
/* Training records of the last light pass: 6 floats and the bin of each of the
 * first n paths (bin -1: no guidable emission).  Returns the count written. */
long long rc_guide_records(float *x, int *bins, long long n);

#ifdef __cplusplus
}
#endif
#endif
