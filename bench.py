"""Benchmark of the vertex-merging hot path on one MI355X.

Workload (water caustic): 512 x 512-vertex sine-wave heightfield (522,242
triangles, IOR 1.33) over a diffuse floor, 2 x 2 area light, 1024 x 1024
camera, 16,777,216 light paths per iteration, plain estimator, fp32.
A step is one iteration: light pass, photon-map build, camera pass + merge.
The scene is synthetic (seeded); nothing is loaded from disk.

Prints one JSON line with photon paths/s (whole iteration), the roofline of
the dominant kernel (algorithmic bytes from a separate counting launch over
its HIP-event duration), the CPU oracle timed on a bounded sample, and
time-to-RMSE <= 0.01 against the FP64 oracle's reference image
(tests/golden/water1024_fp64_s61.npz: fixed merge radius, 1/61 of the
pixels) for the plain and the guided estimator, each at 16,777,216 light
paths per iteration and at the light-path count the photon-budget tuner picks.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-ttr]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PHOTON_BYTES = 40  # Path_Vertex: position, direction, throughput, dVM (Path_Vertices.h:17-27)
PHOTON_RECORD_BYTES = 48  # the light pass's photon record: the 40 B + light path index and length


def sort_pass_bytes(light_paths: float) -> int:
    """Bytes per photon of the photon-map sort's passes: the key is (bucket, light path), bucket bits for
    max(4096, min(2^24, N / 2)) buckets above ceil(log2 N) light-path bits; after each 8-bit pass the sorted
    light-path bits are dropped, and once the rest fits 32 bits the keys are 4-B words."""
    n = max(1, int(light_paths))
    uid = max(0, math.ceil(math.log2(n)))
    bucket = math.ceil(math.log2(max(4096, min(1 << 24, n // 2))))
    bits = uid + bucket
    total, dropped, in32 = 0, 0, False
    for p in range(min(8, (bits + 7) // 8)):
        nd = min(8 * (p + 1), max(uid, dropped))
        out32 = bits - nd <= 32
        total += (4 if in32 else 8) + 4 + (4 if out32 else 8) + 4
        dropped, in32 = nd, out32
    return total


def algorithmic_bytes(st: dict, node_bytes: int = 128, grouped: bool = True) -> dict:
    """Bytes per launch from the event counts of one counting iteration.

    traversal: one node per visit (128 B: the four child boxes and refs of the
    4-wide BVH; 64 B for the binary one); 36 B per triangle
    test (three points); 40 B per closest hit (three normals + material).
    light pass: + 60 B per stored photon (its 48-B record, and its photon-map
    key and value).  photon map, per photon: the radix passes (the pair read and
    written: 8-B keys while the key's unsorted bits exceed 32, then 4-B keys;
    sort_pass_bytes), 96 B permutation (record and index read, sorted photon and
    index written), 8 B cell ranges.  camera trace:
    + 80 B per merge query written + 8 B per path record.  gather: 80 B per
    query read, 12 B per result, 8 B per cell lookup, and per photon: with one
    query per wave (grouped = False) 12 B per photon tested and 28 B per photon
    accepted; grouped, the 40 B record once per photon streamed (a group of
    queries sharing a window reads it once).  resolve: 8 B per path, 12 B per
    query, 24 B per film update."""
    light = (node_bytes * st["light_node_visits"] + 36 * st["light_tri_tests"] + 40 * st["light_hits"]
             + (PHOTON_RECORD_BYTES + 12) * st["photons"])
    trace = (node_bytes * st["camera_node_visits"] + 36 * st["camera_tri_tests"] + 40 * st["camera_hits"]
             + 80 * st["gathers"] + 8 * st["camera_paths"])
    if grouped:
        gather = 92 * st["gathers"] + 8 * st["cell_visits"] + PHOTON_BYTES * st["photons_streamed"]
    else:
        gather = (92 * st["gathers"] + 8 * st["cell_visits"] + 12 * st["photons_tested"]
                  + 28 * st["photons_accepted"])
    resolve = 32 * st["camera_paths"] + 12 * st["gathers"]
    photon_map = st["photons"] * (sort_pass_bytes(st["light_paths"]) + 96 + 8)
    return {"light_pass": light, "photon_map": photon_map, "camera_trace": trace, "gather": gather,
            "resolve": resolve}


TRAFFIC_FILE = os.path.join("profiles", "traffic_water.json")


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/), or None."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


# What limits each kernel beyond its bytes, from the SQ / TCC counter passes
# of profiles/water_sq_counters.md (VALU lanes active per instruction, VALU
# instructions per vector memory instruction, L2 hit rate).
LIMITER = {
    "light_pass": "instruction issue across divergent lanes (about 17 of 64 lanes active per VALU instruction, "
                  "27 VALU per VMEM instruction, 59 % L2 hits): algorithmic bytes mostly come from L2 / MALL",
    "gather": "instruction issue (distance tests and BSDF evaluations per photon-query pair)",
    "camera_trace": "instruction issue across divergent lanes",
}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, photons: int, stride: int) -> dict:
    """The CPU oracle (reference-mode fp32 restatement) on one iteration of the
    same scene: the full light pass, its photon map, and the camera pass over
    every `stride`-th pixel, extrapolated to all pixels -- on one core (the
    reference is single-threaded), then with its OpenMP passes on the host's
    cores (at most 16, one GPU's share of the host), whose photon lists and
    films equal the one-core run's."""
    from oracle import binding

    def one(threads):
        o = binding.OracleScene(args, "f32", photons=photons, seed=0, camera_stride=stride, threads=threads)
        o.prepare()
        t0 = time.perf_counter()
        ph = o.trace_light(0)
        t_light = time.perf_counter() - t0
        t1 = time.perf_counter()
        o.camera_pass(0, ph)
        t_cam = time.perf_counter() - t1
        return t_light, t_cam, t_light + stride * t_cam

    t_light, t_cam, total = one(1)
    threads = max(1, min(16, os.cpu_count() or 1))
    tl_n, tc_n, total_n = one(threads)
    return {
        "value": photons / total,
        "unit": "photon paths/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle f32, 1 iteration: {photons} light paths ({t_light:.1f} s), KD photon map + camera "
                   f"pass over 1/{stride} of the pixels ({t_cam:.1f} s, x{stride} extrapolated)"),
        "all_cores": {"value": photons / total_n, "cores": threads,
                      "sample": f"the same iteration on {threads} OpenMP threads: light pass {tl_n:.2f} s, "
                                f"camera pass over 1/{stride} of the pixels {tc_n:.2f} s"},
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
    }


def uneven_batch(a):
    """The uneven-scene batch (scenes.uneven_batch: seeded displaced glass blobs,
    1,200 .. 499,392 triangles log-uniform, GGX roughness U(0, 0.3) as the rough
    dielectric (compat_glass = 0), 256^2 .. 1024^2, 1 .. 8 x 2^20 light paths
    each): one batched iteration (one launch per pass for every scene) against
    the same scenes rendered one after another with their own launches, and
    the batched iteration with its photon map and camera pass replayed from a
    captured graph."""
    from realcausticsengine_amd import scenes
    from realcausticsengine_amd.engine import Batch

    t = time.perf_counter()
    cases = scenes.uneven_batch(a.batch_scenes)
    b = Batch([sc for sc, _ in cases], [dict(photons=p, seed=i, write_exr=0, compat_glass=0)
                                        for i, (_, p) in enumerate(cases)])
    b.prepare()
    prep_s = time.perf_counter() - t
    print(f"[bench] uneven batch: {len(cases)} scenes ready in {prep_s:.1f} s", file=sys.stderr, flush=True)
    tuned = None
    if not a.no_autotune:
        tuned = {k: round(v, 3) for k, v in b.autotune(repeats=1).items()}
        print(f"[bench]   batch autotune: {tuned}", file=sys.stderr, flush=True)
    b.render_iteration()  # warm-up: buffers of the batch and of every scene alone
    for i in range(len(cases)):
        b.render_scene(i)
    t0 = time.perf_counter()
    for _ in range(a.batch_steps):
        b.render_iteration()
    batched = (time.perf_counter() - t0) / a.batch_steps
    st = b.stats()
    t0 = time.perf_counter()
    for _ in range(a.batch_steps):
        for i in range(len(cases)):
            b.render_scene(i)
    sequential = (time.perf_counter() - t0) / a.batch_steps
    b.set(0, "graph", 1)
    b.render_iteration()  # capture
    t0 = time.perf_counter()
    for _ in range(a.batch_steps):
        b.render_iteration()
    graphed = (time.perf_counter() - t0) / a.batch_steps
    captures = int(b.stats()["graph_captures"])
    b.release()
    light = sum(p for _, p in cases)
    out = {
        "scenes": len(cases),
        "triangles": int(sum(sc.total_tris for sc, _ in cases)),
        "light_paths_per_iteration": int(light),
        "camera_paths_per_iteration": int(sum(sc.camera.width * sc.camera.height for sc, _ in cases)),
        "ms_per_iteration_batched": round(batched * 1e3, 2),
        "ms_per_iteration_one_scene_at_a_time": round(sequential * 1e3, 2),
        "ms_per_iteration_batched_graph": round(graphed * 1e3, 2),
        "graph_captures": captures,
        "speedup": round(sequential / batched, 3),
        "photon_paths_per_s_batched": round(light / batched, 1),
        "kernel_ms_batched": {k: round(st[k], 3) for k in ("ms_light", "ms_map", "ms_camera_trace", "ms_gather",
                                                            "ms_resolve")},
        "prepare_s": round(prep_s, 2),
        "autotune": tuned,
    }
    print(f"[bench]   {out}", file=sys.stderr, flush=True)
    return out


# Launch shapes the batch autotuner chose for the 16 at-scale scenes in one launch (the sample leg uses them).
AT_SCALE_LAUNCH = ("path_chunk=512,light_blocks_per_cu=12,camera_blocks_per_cu=16,cell_scale=0.67,"
                   "gather_blocks_per_cu=24,light_refill=16,light_order=1,camera_refill=32,group_queries=8,"
                   "gather_tile=0,gather_block=2,query_order=1,overlap=0")


def at_scale(a) -> int:
    """The at-scale batched + learned workload (scenes.c4_batch: displaced glass
    meshes and water heightfields of 2M-10M triangles, four lights each, 2048^2,
    64 x 2^20 light paths per scene and iteration, one learned proposal per
    scene trained every iteration, trees built on the GPU).  The scenes run in
    batches of --scale-group (default 16: all scenes in one launch per pass; a
    launch holds at most 2^31 light paths), one batch resident at a time; each batch's launch
    shapes are autotuned (untimed) unless --launch fixes them; prints one JSON
    line."""
    print(json.dumps({"at_scale": at_scale_run(a.at_scale, a.scale_group, a.scale_res, a.scale_photons,
                                               a.scale_steps, a.launch, a.no_autotune)}), flush=True)
    return 0


def at_scale_sample(a) -> dict:
    """The first four at-scale scenes at their full size (13.9 M triangles,
    2048^2, 64 x 2^20 light paths each, proposals trained every iteration) in
    one launch per pass, with the launch shapes tuned for the whole set (no
    tuning here): a sample of the at-scale workload that fits the default run."""
    out = at_scale_run(4, 4, 2048, 64 << 20, 2, AT_SCALE_LAUNCH, True)
    out["sample"] = "scenes 0-3 of the 16 at-scale scenes, launch shapes fixed (bench.py --at-scale 16 runs all)"
    return out


def at_scale_run(n, group, res, photons, steps, launch, no_autotune) -> dict:
    from realcausticsengine_amd import scenes
    from realcausticsengine_amd.engine import GUIDE_STEP_FLOPS_PER_ROW, Batch

    opts = dict(write_exr=0, bvh_build=1, guide=1, guide_batch=65536, guide_warmup=1, compat_point_light=0,
                compat_light_pdf=0, light_order=1)  # coherent light rays: -22 % light pass on these scenes
    fixed = {}
    for kv in filter(None, launch.split(",")):  # fixed launch shapes (every scene; scene 0's serve the batch)
        k, v = kv.split("=")
        fixed[k.strip()] = float(v)
    opts.update(fixed)
    groups, totals = [], dict(light=0, camera=0, tris=0, s_iter=0.0, ms_guide=0.0, guide_rows=0.0, prep=0.0)
    for g0 in range(0, n, group):
        t = time.perf_counter()
        cases = [scenes.c4_scene(s, res, photons) for s in range(g0, min(n, g0 + group))]
        gen_s = time.perf_counter() - t
        t = time.perf_counter()
        b = Batch([sc for sc, _ in cases], [dict(opts, photons=p, seed=g0 + i) for i, (_, p) in enumerate(cases)])
        b.prepare()
        prep_s = time.perf_counter() - t
        tris = int(sum(sc.total_tris for sc, _ in cases))
        print(f"[bench] at scale: scenes {g0}..{g0 + len(cases) - 1}, {tris} triangles, generated in {gen_s:.1f} s, "
              f"prepared in {prep_s:.1f} s", file=sys.stderr, flush=True)
        b.render_iteration()  # warm-up (buffers; the proposals' first step)
        tuned = {"fixed": launch} if fixed else None
        if not fixed and not no_autotune:  # launch shapes for this batch (untimed; renders into scratch films)
            t = time.perf_counter()
            tuned = {k: round(v, 3) for k, v in b.autotune(repeats=1).items()}
            print(f"[bench]   batch autotune ({time.perf_counter() - t:.1f} s): {tuned}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        ms_guide = rows = 0.0
        for _ in range(steps):
            b.render_iteration()
            st = b.stats()
            ms_guide += st["ms_guide"]
            rows += st["guide_rows"]
        per_it = (time.perf_counter() - t0) / steps
        # one counting iteration after the timed ones (its radius is a step smaller, so its counts are a
        # slight underestimate of the timed iterations'): algorithmic bytes over the last timed kernel times
        b.set(0, "count_events", 1)
        b.render_iteration()
        cst = b.stats()
        b.release()
        cst["sort_passes"] = st["sort_passes"]
        alg = algorithmic_bytes(cst, 64)  # binary nodes of the GPU-built trees
        light = sum(p for _, p in cases)
        camera = sum(sc.camera.width * sc.camera.height for sc, _ in cases)
        rec = {"scenes": [g0, g0 + len(cases) - 1], "triangles": tris, "light_paths": int(light),
               "camera_paths": int(camera), "photons": int(st["photons"]), "ms_per_iteration": round(per_it * 1e3, 1),
               "kernel_ms": {k: round(st[k], 2) for k in ("ms_light", "ms_map", "ms_camera_trace", "ms_gather",
                                                           "ms_resolve")},
               "gather": {"queries": int(cst["gathers"]), "photons_tested_per_query":
                          round(cst["photons_tested"] / max(1, cst["gathers"]), 1),
                          "algorithmic_gb_per_s": round(alg["gather"] / (st["ms_gather"] * 1e-3) / 1e9, 1)},
               "light_pass_algorithmic_gb_per_s": round(alg["light_pass"] / (st["ms_light"] * 1e-3) / 1e9, 1),
               "ms_guide_per_iteration": round(ms_guide / steps, 3), "prepare_s": round(prep_s, 2),
               "autotune": tuned}
        print(f"[bench]   {rec}", file=sys.stderr, flush=True)
        groups.append(rec)
        totals["light"] += light
        totals["camera"] += camera
        totals["tris"] += tris
        totals["s_iter"] += per_it
        totals["ms_guide"] += ms_guide / steps
        totals["guide_rows"] += rows / steps
        totals["prep"] += prep_s
    gflop = GUIDE_STEP_FLOPS_PER_ROW * totals["guide_rows"] / 1e9
    return {
        "scenes": n, "scenes_per_launch": group, "triangles": totals["tris"], "resolution": res,
        "light_paths_per_iteration": totals["light"], "camera_paths_per_iteration": totals["camera"],
        "s_per_iteration_all_scenes": round(totals["s_iter"], 3),
        "photon_paths_per_s": round(totals["light"] / totals["s_iter"], 1),
        "ms_guide_per_iteration_all_scenes": round(totals["ms_guide"], 2),
        "guide_gflop_per_iteration": round(gflop, 2),
        "prepare_s_total": round(totals["prep"], 1), "steps": steps, "groups": groups}


def guide_steps(a) -> int:
    """Only the learned proposal's training step: `--guide-steps` steps of 65,536
    synthetic rows (forward, backward, Adam) through rc_guide_train_step, timed
    with the engine's network-step events; prints one JSON line."""
    import numpy as np

    from realcausticsengine_amd import scenes
    from realcausticsengine_amd.engine import GUIDE_STEP_FLOPS_PER_ROW, Engine

    eng = Engine(scenes.water_scene(32, 16, 16, samples=1), photons=65536, seed=0, write_exr=0)
    eng.prepare()
    rng = np.random.default_rng(0)
    n = 65536
    x = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    bins = rng.integers(0, 1024, n).astype(np.int32)
    w = rng.exponential(1.0, n).astype(np.float32)
    eng.guide_train_step(x, bins, w, apply=True)  # warm-up
    t0 = time.perf_counter()
    for _ in range(a.guide_steps):
        eng.guide_train_step(x, bins, w, apply=True)
    wall = (time.perf_counter() - t0) / a.guide_steps
    eng.release()
    flops = GUIDE_STEP_FLOPS_PER_ROW * n
    print(json.dumps({"guide_training_step": {"rows": n, "steps": a.guide_steps, "gflop_per_step": flops / 1e9,
                                              "ms_per_step_wall_incl_upload": round(wall * 1e3, 3)}}), flush=True)
    return 0


def graph_leg(a):
    """Iteration time of a small scene (64 x 64 pixels, 65,536 light paths,
    5,120-triangle glass sphere) with direct launches and with the captured
    iteration graph, without and with the overlapped camera trace (overlap 1:
    trace queued first; 2: after the light pass).  The whole iteration is one
    graph either way; with overlap the trace and query sort are a second-stream
    branch of it, joined before the merge.  At this size the iteration
    is bounded by the latency of the longest light and camera paths (paths
    caught in the glass), not by launches."""
    from realcausticsengine_amd import scenes
    from realcausticsengine_amd.engine import Engine

    out = {"workload": "glass icosphere (5,120 tris) over a plane, 64x64, 65,536 light paths/iteration"}
    for overlap in (0, 1, 2):
        for graph in (0, 1):
            eng = Engine(scenes.c0_scene(64, 64, subdiv=4, samples=1), photons=65536, seed=0, write_exr=0,
                         graph=graph, overlap=overlap)
            eng.prepare()
            for _ in range(5):
                eng.render_iteration()
            n = 200
            t0 = time.perf_counter()
            for _ in range(n):
                eng.render_iteration()
            key = f"ms_per_iteration_{'graph' if graph else 'direct'}_overlap{overlap}"
            out[key] = round((time.perf_counter() - t0) * 1e3 / n, 4)
            if graph:
                out[f"graph_captures_overlap{overlap}"] = int(eng.get("graph_captures"))
            eng.release()
        out[f"speedup_overlap{overlap}"] = round(out[f"ms_per_iteration_direct_overlap{overlap}"] /
                                                out[f"ms_per_iteration_graph_overlap{overlap}"], 3)
    print(f"[bench] graph: {out}", file=sys.stderr, flush=True)
    return out


def time_to_rmse(a, cpu_s_per_iteration, cpu_all_s_per_iteration=None):
    """Wall time of rc_render_iteration calls until the running image reaches
    RMSE <= target against the FP64 reference (readback and scoring untimed)."""
    from realcausticsengine_amd import convergence, scenes
    from realcausticsengine_amd.engine import GUIDE_STEP_FLOPS_PER_ROW, Engine

    ref = convergence.SubsetReference.load(os.path.join(ROOT, "tests", "golden", "water1024_fp64_s61.npz"))
    target = a.ttr_target
    legs = {}

    def engine(guided, photons, qmc=False):
        args = scenes.water_scene(a.grid, 1024, 1024, alpha=1.0, samples=1, base_radius=0.01)
        opts = dict(photons=photons, seed=0, write_exr=0, compat_rows=0, camera_qmc=int(qmc))
        if guided:
            opts.update(guide=1, guide_warmup=2, guide_train_every=a.guide_train_every)
        e = Engine(args, **opts)
        e.prepare()
        return e

    # (guided, stratified camera samples): the 16M legs only for independent draws
    for guided, qmc in ((False, False), (True, False), (False, True), (True, True)):
        name = ("guided" if guided else "plain") + ("_qmc" if qmc else "")
        if not qmc:
            print(f"[bench] time-to-RMSE: {name}, 16,777,216 light paths/iteration", file=sys.stderr, flush=True)
            eng = engine(guided, a.photons)
            run = convergence.run_to_target(eng, ref, target, a.ttr_max_iterations, label=f"{name} 16M",
                                            max_seconds=a.ttr_max_seconds)
            eng.release()
            legs[f"{name}_16M"] = run.summary(target)
            if guided:
                legs[f"{name}_16M"]["guide_hit_fraction_last"] = round(run.stats[-1].get("guide_hit_fraction", 0), 4)
            print(f"[bench]   {legs[f'{name}_16M']}", file=sys.stderr, flush=True)
        print(f"[bench] time-to-RMSE: {name}, tuned light-path count", file=sys.stderr, flush=True)
        eng = engine(guided, a.photons, qmc)
        tuner = convergence.PhotonBudgetTuner(4 << 20, 1 << 20, pilots=6)
        run = convergence.run_tuned(eng, ref, target, a.ttr_max_iterations, tuner, label=f"{name} tuned",
                                    max_seconds=a.ttr_max_seconds, autotune=not a.no_autotune,
                                    online=not a.offline_tune)
        mfma = None
        if guided and not qmc:
            eng.set("guide_train_every", 1)
            eng.render_iteration()  # one training iteration to time the network step
            st = eng.stats()
            if st["ms_guide_step"] > 0:
                tf = GUIDE_STEP_FLOPS_PER_ROW * st["guide_rows"] / (st["ms_guide_step"] * 1e-3) / 1e12
                mfma = {"rows": int(st["guide_rows"]), "ms_step": round(st["ms_guide_step"], 4),
                        "ms_guide_total": round(st["ms_guide"], 4), "tflops": round(tf, 2),
                        "bf16_peak_tflops": 2500.0, "frac": round(tf / 2500.0, 5)}
        eng.release()
        f = tuner.fit
        legs[f"{name}_tuned"] = run.summary(target)
        if f:
            legs[f"{name}_tuned"]["fit"] = {"C": round(f.C, 5), "P_per_path": f.P, "T0_ms": round(f.T0, 3),
                                            "ms_per_Mpath": round(f.t_per_path * 1e6, 4), "chosen": f.best}
        if mfma:
            legs["guide_training_step"] = mfma
        print(f"[bench]   {legs[f'{name}_tuned']}", file=sys.stderr, flush=True)
    # batched frames: the tuned plain estimator with stratified camera samples, K consecutive iterations of
    # the scene per batched launch (convergence.FrameBatch)
    for guided in (False,):
        name = ("guided" if guided else "plain") + "_qmc_batched"
        print(f"[bench] time-to-RMSE: {name}, {a.ttr_frames} frames per launch, tuned light-path count",
              file=sys.stderr, flush=True)
        args = scenes.water_scene(a.grid, 1024, 1024, alpha=1.0, samples=1, base_radius=0.01)
        opts = dict(photons=1 << 20, seed=0, write_exr=0, compat_rows=0, camera_qmc=1)
        if guided:
            opts.update(guide=1, guide_warmup=2, guide_train_every=a.guide_train_every)
        fb = convergence.FrameBatch(args, a.ttr_frames, opts)
        tuner = convergence.PhotonBudgetTuner(1 << 20, 1 << 18, pilots=6)
        run = convergence.run_tuned(fb, ref, target, a.ttr_max_iterations, tuner, label=f"{name}",
                                    max_seconds=a.ttr_max_seconds, autotune=not a.no_autotune,
                                    online=not a.offline_tune)
        fb.release()
        f = tuner.fit
        leg = run.summary(target)
        leg["frames_per_launch"] = a.ttr_frames
        if leg["iterations"]:
            leg["frame_iterations"] = leg["iterations"] * a.ttr_frames
        if f:
            leg["fit"] = {"C_per_launch": round(f.C, 6), "P_per_path": f.P, "T0_ms": round(f.T0, 3),
                          "ms_per_Mpath": round(f.t_per_path * 1e6, 4), "chosen_per_frame": f.best}
        legs[name] = leg
        print(f"[bench]   {leg}", file=sys.stderr, flush=True)
    plain = legs["plain_16M"]
    it_plain = plain["iterations"] or plain.get("extrapolated_iterations")
    out = {"target": target, "definition": "RMSE over reference pixels and channels of (img - ref) / mean(ref)",
           "reference": {"file": "tests/golden/water1024_fp64_s61.npz", "estimator": "FP64 oracle, fixed radius",
                         "pixels": int(len(ref.rows)), "own_rmse": round(ref.noise, 5)},
           "legs": legs}
    if cpu_s_per_iteration and it_plain:
        out["cpu_reference_mode_estimate_s"] = round(cpu_s_per_iteration * it_plain, 1)
    if cpu_all_s_per_iteration and it_plain:
        out["cpu_all_cores_estimate_s"] = round(cpu_all_s_per_iteration * it_plain, 1)

    def secs(leg):
        return leg["seconds"] if leg["seconds"] is not None else leg.get("extrapolated_seconds")

    base = secs(plain)
    if base:
        out["speedup_vs_plain_16M"] = {k: round(base / secs(v), 3) for k, v in legs.items()
                                       if isinstance(v, dict) and "iterations_run" in v and secs(v)}
        if "cpu_reference_mode_estimate_s" in out:
            best = min(secs(v) for v in legs.values() if isinstance(v, dict) and "iterations_run" in v and secs(v))
            out["speedup_best_vs_cpu"] = round(out["cpu_reference_mode_estimate_s"] / best, 1)
            if "cpu_all_cores_estimate_s" in out:
                out["speedup_best_vs_cpu_all_cores"] = round(out["cpu_all_cores_estimate_s"] / best, 1)
    return out


def rough_glass_leg(a, launch: dict) -> dict:
    """The same water workload with the water as a rough GGX dielectric
    (roughness 0.02, option compat_glass = 0; the reference renders its glass
    as delta whatever the roughness): photon paths/s and kernel times with the
    main leg's launch shapes (photons are now also stored on the water and
    camera paths merge there)."""
    from realcausticsengine_amd import scenes
    from realcausticsengine_amd.engine import Engine

    args = scenes.water_scene(a.grid, a.res, a.res, roughness=0.02, samples=64)
    eng = Engine(args, photons=a.photons, seed=0, write_exr=0, compat_glass=0)
    eng.prepare()
    for k, v in launch.items():
        eng.set(k, v)
    for _ in range(2):
        eng.render_iteration()
    n = max(3, a.steps // 2)
    kern = {"ms_light": 0.0, "ms_map": 0.0, "ms_camera_trace": 0.0, "ms_gather": 0.0}
    photons = gathers = 0.0
    t0 = time.perf_counter()
    for _ in range(n):
        eng.render_iteration()
        st = eng.stats()
        for k in kern:
            kern[k] += st[k] / n
        photons += st["photons"] / n
        gathers += st["gathers"] / n
    ms = (time.perf_counter() - t0) * 1e3 / n
    eng.release()
    out = {"workload": "the water workload with roughness 0.02 as a GGX dielectric (compat_glass = 0)",
           "ms_per_step": round(ms, 3), "photon_paths_per_s": round(a.photons / (ms * 1e-3), 1),
           "photons_per_iteration": int(photons), "queries_per_iteration": int(gathers),
           "kernel_ms": {k: round(v, 3) for k, v in kern.items()}}
    print(f"[bench] rough glass: {out}", file=sys.stderr, flush=True)
    return out


def copy_calibration() -> int:
    """Three copies of 1 GiB between device buffers: what FETCH_SIZE and
    WRITE_SIZE read for a known number of bytes moved."""
    import torch

    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copies": 3, "bytes_per_copy": n * 4}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photons", type=int, default=16_777_216)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ttr", action="store_true", help="skip the time-to-RMSE legs")
    ap.add_argument("--ttr-target", type=float, default=0.01)
    ap.add_argument("--ttr-max-iterations", type=int, default=3000)
    ap.add_argument("--ttr-max-seconds", type=float, default=75.0)
    ap.add_argument("--guide-train-every", type=int, default=8)
    ap.add_argument("--ttr-frames", type=int, default=8,
                    help="independent frames of the scene per batched launch in the batched time-to-RMSE legs")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-scene batch leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--offline-tune", action="store_true",
                    help="time-to-RMSE legs: tune launch shapes with rc_autotune (scratch renders, timed) instead of "
                         "on the run's own iterations")
    ap.add_argument("--launch", default="", help="fixed launch shapes instead of autotuning, "
                    "e.g. path_chunk=512,light_blocks_per_cu=8,gather_blocks_per_cu=32")
    ap.add_argument("--no-graph", action="store_true", help="skip the iteration-graph leg")
    ap.add_argument("--no-rough", action="store_true", help="skip the rough-glass leg")
    ap.add_argument("--no-scale", action="store_true", help="skip the at-scale sample leg (four at-scale scenes)")
    ap.add_argument("--guide-steps", type=int, default=0,
                    help="only time this many training steps of the learned proposal (65,536 rows each)")
    ap.add_argument("--at-scale", type=int, default=0,
                    help="only run the at-scale batched + learned workload over this many scenes (16 = the full set)")
    ap.add_argument("--scale-group", type=int, default=16)
    ap.add_argument("--scale-steps", type=int, default=3)
    ap.add_argument("--scale-res", type=int, default=2048)
    ap.add_argument("--scale-photons", type=int, default=64 << 20)
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 1 GiB device-to-device copies (the reference point of the HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the engine renders on one GPU; run with --gpus 1", file=sys.stderr)
        return 2
    if a.copy_calibration:
        return copy_calibration()
    if a.guide_steps > 0:
        return guide_steps(a)
    if a.at_scale > 0:
        return at_scale(a)

    from realcausticsengine_amd import scenes
    from realcausticsengine_amd.engine import Engine

    # BASELINE's water is GGX roughness 0.02; with compat_glass = 1 (the default) glass is a delta interface
    # whatever its roughness, as in the reference (src/materials.h:104), so this is the reference's rendering
    args = scenes.water_scene(a.grid, a.res, a.res, roughness=0.02, samples=a.warmup + 2 * a.steps + 4)
    eng = Engine(args, photons=a.photons, seed=0, write_exr=0)
    t = time.perf_counter()
    eng.prepare()
    prep_s = time.perf_counter() - t
    print(f"[bench] scene ready: {args.total_tris} triangles, BVH {eng.get('nodes'):.0f} nodes, {prep_s:.2f} s",
          file=sys.stderr, flush=True)
    tuned = None
    if a.launch:
        for kv in a.launch.split(","):
            k, v = kv.split("=")
            eng.set(k.strip(), float(v))
        tuned = {"fixed": a.launch}
    elif not a.no_autotune:  # launch shapes and queue / tile sizes for this scene (untimed; renders nothing)
        tuned = {k: round(v, 3) for k, v in eng.autotune(repeats=2).items()}
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        eng.render_iteration()
    keys = {"light_pass": "ms_light", "photon_map": "ms_map", "camera_trace": "ms_camera_trace",
            "gather": "ms_gather", "resolve": "ms_resolve"}
    kern = {k: 0.0 for k in keys}
    t0 = time.perf_counter()
    for i in range(a.steps):
        eng.render_iteration()
        st = eng.stats()
        for k, sk in keys.items():
            kern[k] += st[sk]
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1000.0 / a.steps
    kern = {k: v / a.steps for k, v in kern.items()}
    overlap = int(eng.get("overlap"))
    print(f"[bench] {a.steps} steps, {ms_step:.2f} ms/step, kernel ms {kern} (overlap {overlap})",
          file=sys.stderr, flush=True)
    kern_timed = dict(kern)
    if overlap:
        # beside the light pass the camera trace shares the GPU: the per-kernel times of the roofline
        # come from serial iterations (untimed for the step rate) of the same iteration indices, as the
        # merge radius shrinks from one iteration to the next
        eng.set("overlap", 0)
        eng.set("next_iteration", a.warmup)
        kern = {k: 0.0 for k in keys}
        for _ in range(a.steps):
            eng.render_iteration()
            st = eng.stats()
            for k, sk in keys.items():
                kern[k] += st[sk] / a.steps
        eng.set("overlap", 1)
        print(f"[bench] serial kernel ms {kern}", file=sys.stderr, flush=True)
    launch = {k: eng.get(k) for k in ("path_chunk", "light_blocks_per_cu", "camera_blocks_per_cu", "cell_scale",
                                      "gather_blocks_per_cu", "light_refill", "light_order", "camera_refill",
                                      "group_queries", "gather_tile", "gather_block", "query_order", "overlap")}
    # counting launch (not timed) of the middle timed iteration: its event
    # counts give the algorithmic bytes (the merge radius shrinks per iteration)
    eng.set("next_iteration", a.warmup + a.steps // 2)
    eng.set("count_events", 1)
    eng.render_iteration()
    cst = eng.stats()
    eng.set("count_events", 0)
    by = algorithmic_bytes(cst, int(eng.get("node_bytes")), eng.get("group_queries") > 1)
    dom = max(kern, key=kern.get)
    achieved = by[dom] / (kern[dom] * 1e-3) / 1e9
    traffic = profiled_traffic(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": f"{TRAFFIC_FILE}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel on this "
                                  "workload, scaled by a 1 GiB copy (profiles/water_kernels.md)",
                "traffic_frac": (round(traffic / (kern[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None),
                "kernel_ms": round(kern[dom], 4), "timing": "serial iterations (overlap 0), HIP events on the "
                                                           "engine's stream",
                "limiter": LIMITER.get(dom)}
    extra = {
        "autotune": tuned,
        "kernel_ms": {k: round(v, 3) for k, v in kern.items()},
        "kernel_ms_in_timed_steps": {k: round(v, 3) for k, v in kern_timed.items()},
        "overlap_in_timed_steps": overlap,
        "kernel_achieved_GBs": {k: round(by[k] / (kern[k] * 1e-3) / 1e9, 1) for k in kern if kern[k] > 0},
        "algorithmic_bytes": {k: int(v) for k, v in by.items()},
        "photons_per_iteration": int(cst["photons"]),
        "gathers_per_s": cst["gathers"] / (ms_step * 1e-3),
        "photons_tested_per_gather": cst["photons_tested"] / max(1.0, cst["gathers"]),
        "counts": {k: int(cst[k]) for k in ("light_node_visits", "light_tri_tests", "light_hits",
                                             "camera_node_visits", "camera_tri_tests", "camera_hits",
                                             "gathers", "cell_visits", "photons_tested", "photons_accepted",
                                             "photons_streamed")},
    }
    eng.release()
    cpu = None
    if not a.no_cpu_baseline:
        print("[bench] timing the CPU oracle sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args, a.photons, a.cpu_stride)
    batch = None
    if not a.no_batch:
        batch = uneven_batch(a)
    graph = None if a.no_graph else graph_leg(a)
    rough = None if a.no_rough else rough_glass_leg(a, launch)
    scale = None if a.no_scale else at_scale_sample(a)
    ttr = None
    if not a.no_ttr:
        ttr = time_to_rmse(a, a.photons / cpu["value"] if cpu else None,
                           a.photons / cpu["all_cores"]["value"] if cpu else None)
    value = a.photons / (ms_step * 1e-3)
    line = {
        "metric": "time-to-RMSE<=0.01 vs FP64 (guided vs plain) + photon paths/s, 1 MI355X",
        "value": round(value, 1),
        "unit": "photon paths/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "water caustic: 512x512-vertex sine heightfield (522,242 tris, IOR 1.33, GGX "
                               "roughness 0.02) over diffuse floor, 2x2 area light, 1024x1024, 16,777,216 light "
                               "paths/iteration",
                   "glass": "delta interface (compat_glass = 1): the reference ignores glass roughness "
                            "(src/materials.h:104), so this is its rendering of the configuration; the rough "
                            "dielectric (compat_glass = 0) is the rough_glass leg",
                   "estimator": "plain vertex merging", "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "time_to_rmse": ttr,
        "uneven_batch": batch,
        "iteration_graph": graph,
        "rough_glass": rough,
        "at_scale_sample": scale,
        "detail": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
