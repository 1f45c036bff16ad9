"""Renders the FP64 reference used for time-to-RMSE (writes water1024_fp64_s61.npz).

Scene: the water-caustic workload (512 x 512-vertex seeded sine heightfield,
IOR 1.33, over a diffuse floor, 2 x 2 area light) with its 1024 x 1024
camera, a fixed merge radius (radius_reduction = 1, base radius 0.01 of the
scene half-diagonal) and image rows without the reference's one-row offset
(compat_rows = 0).  With a fixed radius every iteration has the same
expectation whatever the number of light paths per iteration, so the
reference may use fewer paths per iteration than the runs it judges.

Only every 61st pixel (row-major pixel index p with p % 61 == 0, 17,190
pixels scattered over the whole frame) is rendered; RMSE over these pixels is
an unbiased estimate of the RMSE over the frame.  The double-precision oracle
renders `--iterations` iterations of `--photons` light paths each, with Philox
seed 7 (the measured runs use seed 0, so the two share no random numbers), in
`--procs` processes.  The file holds the mean, the difference of the two
half-means (the reference's own noise: RMSE_ref = rms(diff) / 2 / mean) and
the pixel indices.

    python tests/golden/make_water_reference.py --procs 6 --iterations 7200
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

RES = 1024
STRIDE = 61
SEED = 7
PHOTONS = 131072


def pixel_rows_cols():
    """Image (row, col) of each rendered pixel; compat_rows = 0 puts camera row y at image row H-1-y."""
    p = np.arange(0, RES * RES, STRIDE, dtype=np.int64)
    return RES - 1 - p // RES, p % RES


def scene():
    from realcausticsengine_amd import scenes

    return scenes.water_scene(512, RES, RES, alpha=1.0, samples=1, base_radius=0.01)


def _worker(job):
    first, count, photons = job
    from oracle import binding

    o = binding.OracleScene(scene(), "f64", photons=photons, seed=SEED, camera_stride=STRIDE, compat_rows=0)
    o.prepare()
    rows, cols = pixel_rows_cols()
    acc = np.zeros((rows.size, 3))
    for it in range(first, first + count):
        ph = o.trace_light(it)
        acc += o.camera_pass(it, ph)[rows, cols]
    return first, count, acc


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=6)
    ap.add_argument("--iterations", type=int, default=7200)
    ap.add_argument("--photons", type=int, default=PHOTONS)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "water1024_fp64_s61.npz"))
    a = ap.parse_args()
    chunk = 40
    jobs = [(s, min(chunk, a.iterations - s), a.photons) for s in range(0, a.iterations, chunk)]
    rows, cols = pixel_rows_cols()
    t0 = time.time()
    halves = [np.zeros((rows.size, 3)), np.zeros((rows.size, 3))]
    counts = [0, 0]
    with mp.Pool(a.procs) as pool:
        for n_done, (first, count, acc) in enumerate(pool.imap_unordered(_worker, jobs), 1):
            h = (first // chunk) % 2
            halves[h] += acc
            counts[h] += count
            if n_done % 10 == 0:
                print(f"{n_done}/{len(jobs)} chunks, {time.time() - t0:.0f} s", flush=True)
    mean = (halves[0] + halves[1]) / (counts[0] + counts[1])
    diff = halves[0] / counts[0] - halves[1] / counts[1]
    noise = float(np.sqrt(np.mean(diff ** 2)) / 2.0 / mean.mean())
    np.savez_compressed(a.out, mean=mean.astype(np.float32), half_diff=diff.astype(np.float32),
                        rows=rows.astype(np.int32), cols=cols.astype(np.int32), iterations=a.iterations,
                        photons=a.photons, resolution=RES, stride=STRIDE, seed=SEED, base_radius=0.01,
                        radius_reduction=1.0, rmse_noise=noise)
    print(f"wrote {a.out}: {a.iterations} iterations x {a.photons} paths, own RMSE noise {noise:.4f}, "
          f"{time.time() - t0:.0f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
