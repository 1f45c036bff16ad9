"""Convergence measurement and the photon-budget tuner.

RMSE follows one definition throughout: over all reference pixels and the
three channels of (image - reference) / mean(reference).  With a fixed merge
radius (radius_reduction = 1) every iteration of the vertex-merging estimator
has the same expectation, whatever the number of light paths it traces, so

  * iterations with different light-path counts may be averaged, and
  * the per-iteration relative variance splits as  V(N) = C + P / N:
    C from the camera side (pixel jitter, Fresnel choices, camera bounces),
    P / N from the photon density at the merge vertices.

With an iteration time T(N) = T0 + t1 N, the wall time to a target error is
proportional to V(N) T(N), smallest at N* = sqrt(P T0 / (C t1)).  The tuner
measures V at two light-path counts from the per-pixel variance across pilot
iterations (no reference image needed) and T from the engine's own iteration
times, then renders on at N*.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .engine import Engine


def relative_rmse(values: np.ndarray, reference: np.ndarray) -> float:
    ref = np.asarray(reference, dtype=np.float64)
    v = np.asarray(values, dtype=np.float64)
    return float(np.sqrt(np.mean((v - ref) ** 2)) / ref.mean())


@dataclass
class SubsetReference:
    """Reference values at a subset of pixels (image rows/cols, 3 channels each)."""
    rows: np.ndarray
    cols: np.ndarray
    mean: np.ndarray
    noise: float = 0.0  # the reference's own RMSE estimate

    @classmethod
    def load(cls, path: str) -> "SubsetReference":
        d = np.load(path)
        return cls(d["rows"], d["cols"], d["mean"].astype(np.float64), float(d["rmse_noise"]))

    def rmse(self, accumulated: np.ndarray, iterations: int) -> float:
        return relative_rmse(accumulated[self.rows, self.cols] / float(iterations), self.mean)


@dataclass
class ConvergenceRun:
    label: str
    photons: list = field(default_factory=list)   # light paths of each iteration
    seconds: list = field(default_factory=list)   # cumulative wall time after each iteration
    rmse: list = field(default_factory=list)      # running-image RMSE after each iteration
    reached_at: int | None = None                  # 1-based iteration that first reached the target
    stats: list = field(default_factory=list)
    autotune: dict | None = None                   # launch shapes chosen inside the run (run_tuned)

    def summary(self, target: float) -> dict:
        it = self.reached_at
        out = {"label": self.label, "target": target, "iterations": it,
               "seconds": round(self.seconds[it - 1], 4) if it else None,
               "final_rmse": round(self.rmse[-1], 5) if self.rmse else None,
               "iterations_run": len(self.rmse), "light_paths_per_iteration": sorted(set(self.photons))}
        if self.autotune:
            out["autotune"] = {k: round(v, 3) for k, v in self.autotune.items()}
        if not it and len(self.rmse) >= 8:
            # not reached: extrapolate rmse^2 = a / k from the last quarter of the run
            k = np.arange(1, len(self.rmse) + 1, dtype=np.float64)
            tail = slice(3 * len(k) // 4, None)
            a = float(np.mean(np.asarray(self.rmse)[tail] ** 2 * k[tail]))
            need = a / target ** 2
            out["extrapolated_iterations"] = int(math.ceil(need))
            out["extrapolated_seconds"] = round(self.seconds[-1] / len(k) * need, 3)
        return out


def run_to_target(eng: Engine, ref: SubsetReference, target: float, max_iterations: int,
                  schedule=None, label: str = "", max_seconds: float = 1e9) -> ConvergenceRun:
    """Renders until the running image reaches `target` RMSE against `ref`.

    `schedule(i)` may return the light-path count of iteration i (None keeps
    the current one).  Only rc_render_iteration calls are timed; the image
    readback and the RMSE evaluation after each iteration are not."""
    run = ConvergenceRun(label)
    spent = 0.0
    for i in range(max_iterations):
        if schedule is not None:
            n = schedule(i)
            if n is not None:
                eng.set("photons", int(n))
        t0 = time.perf_counter()
        eng.render_iteration()
        spent += time.perf_counter() - t0
        img = eng.image(scaled=False)
        e = ref.rmse(img, i + 1)
        run.photons.append(int(eng.get("photons")))
        run.seconds.append(spent)
        run.rmse.append(e)
        st = eng.stats()
        run.stats.append({k: st[k] for k in ("ms_total", "ms_guide", "guide_hit_fraction") if k in st})
        if e <= target and run.reached_at is None:
            run.reached_at = i + 1
            break
        if spent > max_seconds:
            break
    return run


@dataclass
class BudgetFit:
    n1: int
    n2: int
    v1: float
    v2: float
    t1_ms: float
    t2_ms: float
    C: float = 0.0
    P: float = 0.0
    T0: float = 0.0
    t_per_path: float = 0.0
    best: int = 0

    def solve(self, n_min: int, n_max: int) -> "BudgetFit":
        # V(N) = C + P / N through the two pilot points; T(N) = T0 + t N likewise
        self.P = max(0.0, (self.v2 - self.v1) / (1.0 / self.n2 - 1.0 / self.n1))
        self.C = max(1e-12, self.v1 - self.P / self.n1)
        self.t_per_path = max(1e-12, (self.t1_ms - self.t2_ms) / (self.n1 - self.n2))
        self.T0 = max(1e-6, self.t1_ms - self.t_per_path * self.n1)
        n = math.sqrt(self.P * self.T0 / (self.C * self.t_per_path)) if self.P > 0 else n_min
        # keep N a multiple of 65536 (whole training batches and sort tiles)
        self.best = int(min(n_max, max(n_min, round(n / 65536) * 65536)))
        return self

    def predicted_cost(self, n: int) -> float:
        """Relative variance x milliseconds per iteration at n light paths."""
        return (self.C + self.P / n) * (self.T0 + self.t_per_path * n)


def pixel_variance(deltas: list) -> float:
    """Mean over pixels of the per-pixel variance across iterations, relative to the squared mean."""
    a = np.stack(deltas).astype(np.float64)
    m = a.mean()
    return float(a.var(axis=0, ddof=1).mean() / (m * m)) if m > 0 else 0.0


class PhotonBudgetTuner:
    """Pilot iterations at two light-path counts, then the cost-optimal count.

    The pilot iterations are ordinary iterations of the render (they count in
    its image and its wall time)."""

    def __init__(self, n1: int, n2: int, pilots: int = 6, n_min: int = 65536, n_max: int = 1 << 26):
        self.n1, self.n2, self.pilots = n1, n2, pilots
        self.n_min, self.n_max = n_min, n_max
        self.fit: BudgetFit | None = None
        self._deltas = {n1: [], n2: []}
        self._ms = {n1: [], n2: []}
        self._prev = None

    def schedule(self, i: int):
        if i < self.pilots:
            return self.n1
        if i < 2 * self.pilots:
            return self.n2
        return self.fit.best if self.fit else None

    def observe(self, eng: Engine, i: int, accumulated: np.ndarray) -> None:
        """Called after iteration i with the accumulated (unscaled) image."""
        prev = self._prev if self._prev is not None else np.zeros_like(accumulated)
        delta = accumulated - prev
        self._prev = accumulated.copy()
        if i < 2 * self.pilots:
            n = self.n1 if i < self.pilots else self.n2
            # the first pilot of each count also carries one-time costs (buffers, warm-up)
            if i not in (0, self.pilots):
                self._ms[n].append(eng.stats()["ms_total"])
            self._deltas[n].append(delta)
        if i == 2 * self.pilots - 1:
            self.fit = BudgetFit(self.n1, self.n2, pixel_variance(self._deltas[self.n1]),
                                 pixel_variance(self._deltas[self.n2]), float(np.median(self._ms[self.n1])),
                                 float(np.median(self._ms[self.n2]))).solve(self.n_min, self.n_max)
            self._deltas = {}


# The launch knobs rc_autotune tries, in its order (rc_engine.cpp, tune_knobs), each with its candidates and the
# per-iteration times (stats, ms) that judge it.  The merge's time is scaled by (r_first / r)^2: with a shrinking merge
# radius the merge's work shrinks from one iteration to the next.
ONLINE_KNOBS = (
    ("path_chunk", (64, 128, 256, 512, 1024), ("ms_light", "ms_camera_trace")),
    ("light_blocks_per_cu", (4, 6, 8, 12, 16), ("ms_light",)),
    ("camera_blocks_per_cu", (4, 6, 8, 12, 16), ("ms_camera_trace",)),
    ("cell_scale", (0.67, 0.8, 1.0, 1.25), ("ms_map", "ms_gather")),
    ("light_refill", (1, 8, 16, 32), ("ms_light",)),
    ("light_order", (0, 1), ("ms_light",)),
    ("camera_refill", (1, 8, 16, 32), ("ms_camera_trace",)),
    ("group_queries", (4, 8, 16, 1), ("ms_gather",)),
    ("gather_tile", (1024, 256, 0), ("ms_gather",)),
    ("gather_block", (2, 4, 1, 0), ("ms_gather",)),
    ("gather_blocks_per_cu", (5, 8, 12, 16, 24), ("ms_gather",)),
    ("query_order", (1, 0), ("ms_camera_trace", "ms_query_sort", "ms_gather")),
    ("overlap", (0, 2, 1), ("ms_total",)),
)
OVERLAP_MARGIN = 0.02  # overlap is kept only when it saves this fraction of the iteration's wall time


class OnlineTuner:
    """Launch-shape tuning on the run's own iterations.

    rc_autotune renders one iteration again and again into a scratch film,
    once per candidate; here each candidate value of each knob renders the
    next iteration of the run instead, which counts in the image and in the
    wall time like any other (launch shapes change how the work is spread
    over the GPU and the order of floating-point sums, not what is
    estimated).  Each knob keeps its cheapest candidate, one knob after the
    other in rc_autotune's order, judged by the same kernel times; the camera
    trace stays serial until the last knob, overlap, which is judged by the
    iteration's wall time and kept only when it saves OVERLAP_MARGIN.  One
    iteration per candidate."""

    def __init__(self, eng, knobs=ONLINE_KNOBS):
        self.eng = eng
        self.knobs = list(knobs)
        self.k = 0      # knob being tuned
        self.c = 0      # its candidate rendering next
        self.best = None
        self.r0 = None
        self.chosen: dict = {}
        self.iterations = 0
        self.walls: list = []

    @property
    def done(self) -> bool:
        return self.k >= len(self.knobs)

    def before_iteration(self) -> None:
        if self.done:
            return
        if self.k == 0 and self.c == 0:
            self.eng.set("overlap", 0)  # the kernel knobs are judged with the camera trace serial
        name, cands, _ = self.knobs[self.k]
        self.eng.set(name, cands[self.c])

    def after_iteration(self, st: dict) -> None:
        if self.done:
            return
        name, cands, parts = self.knobs[self.k]
        r = float(st.get("radius", 0.0))
        if self.r0 is None:
            self.r0 = r
        scale = (self.r0 / r) ** 2 if r > 0 and self.r0 > 0 else 1.0
        cost = sum(float(st[p]) * (scale if p == "ms_gather" else 1.0) for p in parts)
        margin = OVERLAP_MARGIN if name == "overlap" else 0.0
        if self.best is None or cost < self.best[0] * (1.0 - margin):
            self.best = (cost, cands[self.c])
        self.iterations += 1
        self.walls.append(float(st["ms_total"]))
        self.c += 1
        if self.c == len(cands):
            self.eng.set(name, self.best[1])
            self.chosen[name] = self.best[1]
            self.k += 1
            self.c = 0
            self.best = None

    def report(self) -> dict:
        """The chosen values, the tuning iterations and the wall time of the first and last of them (ms)."""
        out = dict(self.chosen)
        out["tuning_iterations"] = self.iterations
        if self.walls:
            out["ms_first"] = self.walls[0]
            out["ms_last"] = self.walls[-1]
        return out


class FrameBatch:
    """K frames of one scene rendered by the same launches.

    The frames are copies of the scene in one batch (rc_batch_*) with the
    same Philox key; frame k renders the iteration indices k, k + K, k + 2K,
    ... (options iteration_offset = k, iteration_stride = K), so the first B
    batched iterations render iterations 0 .. K B - 1 of one estimator -- the
    same samples, and with camera_qmc the same stratified sequence, as K B
    iterations of the scene alone -- with one light-pass, photon-map,
    camera-trace and merge launch per pass for all K.  With a fixed merge
    radius every iteration has the same expectation, so the running image is
    the mean of the K accumulated films.  It offers the calls run_tuned and
    PhotonBudgetTuner make of an Engine: `photons` is the light-path count of
    each frame, `ms_total` the wall time of the last batched iteration, and
    image() the mean of the frames' accumulated films (one batched iteration
    then counts as one iteration of that mean)."""

    def __init__(self, args, frames: int, options: dict):
        from .engine import Batch

        self.frames = frames
        opts = [dict(options, iteration_offset=k, iteration_stride=frames) for k in range(frames)]
        self.batch = Batch([args] * frames, opts)
        self.batch.prepare()
        self.photons = int(options.get("photons", 0))
        self.last_ms = 0.0
        self.W, self.H = args.camera.width, args.camera.height

    def set(self, key: str, value) -> None:
        for k in range(self.frames):
            self.batch.set(k, key, value)
        if key == "photons":
            self.photons = int(value)

    def get(self, key: str) -> float:
        if key == "photons":
            return float(self.photons)
        return self.batch.lib.rc_get_option(key.encode())

    def render_iteration(self) -> int:
        t0 = time.perf_counter()
        rc = self.batch.render_iteration()
        self.last_ms = (time.perf_counter() - t0) * 1e3
        return rc

    def image(self, scaled: bool = False) -> np.ndarray:
        acc = np.zeros((self.H, self.W, 3), dtype=np.float64)
        for k in range(self.frames):
            acc += self.batch.image(k, scaled=scaled)
        return (acc / self.frames).astype(np.float32)

    def stats(self) -> dict:
        st = self.batch.stats()
        st["ms_total"] = self.last_ms
        return st

    def autotune(self, repeats: int = 2) -> dict:
        return self.batch.autotune(repeats=repeats)

    def release(self) -> None:
        self.batch.release()


def run_tuned(eng: Engine, ref: SubsetReference, target: float, max_iterations: int, tuner: PhotonBudgetTuner,
              label: str = "", max_seconds: float = 1e9, autotune: bool = False,
              online: bool = False) -> ConvergenceRun:
    """run_to_target with the tuner choosing the light-path count.  With
    `autotune`, the launch shapes are tuned for the chosen count once the
    pilots are done: by rc_autotune, whose wall time counts in the run's time
    (its renders go to a scratch film and add nothing to the image), or, with
    `online`, by an OnlineTuner on the run's next iterations."""
    run = ConvergenceRun(label)
    spent = 0.0
    tuned = False
    live = None
    for i in range(max_iterations):
        n = tuner.schedule(i)
        if n is not None:
            eng.set("photons", int(n))
        if autotune and not tuned and tuner.fit is not None:
            tuned = True
            if online:
                live = OnlineTuner(eng)
            else:
                t0 = time.perf_counter()
                run.autotune = eng.autotune(repeats=1)
                spent += time.perf_counter() - t0
        if live is not None:
            live.before_iteration()
        t0 = time.perf_counter()
        eng.render_iteration()
        spent += time.perf_counter() - t0
        if live is not None and not live.done:
            live.after_iteration(eng.stats())
            if live.done:
                run.autotune = live.report()
        img = eng.image(scaled=False)
        tuner.observe(eng, i, img)
        e = ref.rmse(img, i + 1)
        run.photons.append(int(eng.get("photons")))
        run.seconds.append(spent)
        run.rmse.append(e)
        if e <= target and run.reached_at is None:
            run.reached_at = i + 1
            break
        if spent > max_seconds:
            break
    return run
