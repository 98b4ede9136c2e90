"""Learned tile scheduler for deep-zoom fields with a long-tailed cost.

A field is cut into 32x32 tiles.  A quarter-resolution preview of the same
view (1/16 of the pixels) is rendered first; its per-pixel work map (loop
passes, returned by the renderer) gives each tile an 8x8 patch.  A small
network (65 -> 64 -> 64 -> 1, bf16 weights, trained on the GPU by
sched_mlp.hip) predicts log(1 + cost) of each tile from its patch and the
zoom, and the tiles are launched heaviest first, so the expensive tiles do
not start last and leave the GPU idle at the end (longest-processing-time
order).  The schedule never changes a pixel: every order gives the same
field, and the plain grid is the fallback when no model is trained.

Cost of a tile, the training target: the sum over its sixteen 8x8 pixel
squares (one wavefront each) of the largest per-pixel work in the square,
divided by 16 -- a wavefront runs as long as its slowest lane.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

import fav_hip

TILE = 32
PREVIEW_SCALE = 4
PATCH = TILE // PREVIEW_SCALE  # 8x8 preview pixels per tile
WAVE = 8                       # 8x8 pixels per wavefront
INPUTS = PATCH * PATCH + 1
FEATURE_SCALE = 1.0 / 8.0      # log(1 + work) is at most ~12.2 for a 200000 cap


def tile_grid(width: int, height: int) -> tuple[int, int]:
    return (int(width) + TILE - 1) // TILE, (int(height) + TILE - 1) // TILE


def preview_size(width: int, height: int) -> tuple[int, int]:
    return (int(width) + PREVIEW_SCALE - 1) // PREVIEW_SCALE, (int(height) + PREVIEW_SCALE - 1) // PREVIEW_SCALE


def tile_features(preview_work: np.ndarray, width: int, height: int, log10_zoom: float) -> np.ndarray:
    """(tiles, 65) float32 inputs in row-major tile order: the tile's 8x8
    patch of log(1 + preview work) * FEATURE_SCALE (edge-padded past the
    preview's border) and log10(zoom) / 100."""
    tx, ty = tile_grid(width, height)
    work = np.log1p(np.maximum(preview_work.astype(np.float64), 0.0)) * FEATURE_SCALE
    ph, pw = work.shape
    padded = np.pad(work, ((0, ty * PATCH - ph), (0, tx * PATCH - pw)), mode="edge") if (
        ty * PATCH > ph or tx * PATCH > pw) else work[:ty * PATCH, :tx * PATCH]
    patches = padded.reshape(ty, PATCH, tx, PATCH).transpose(0, 2, 1, 3).reshape(tx * ty, PATCH * PATCH)
    zoom_column = np.full((tx * ty, 1), float(log10_zoom) / 100.0)
    return np.ascontiguousarray(np.hstack([patches, zoom_column]), dtype=np.float32)


def tile_costs(work: np.ndarray, width: int, height: int) -> np.ndarray:
    """(tiles,) float64 measured cost per tile (row-major): mean over the
    tile's sixteen 8x8 wavefront squares of the square's largest work."""
    tx, ty = tile_grid(width, height)
    h, w = work.shape
    padded = np.zeros((ty * TILE, tx * TILE), dtype=np.float64)
    padded[:h, :w] = work
    squares = padded.reshape(ty, TILE // WAVE, WAVE, tx, TILE // WAVE, WAVE).max(axis=(2, 5))
    return squares.sum(axis=(1, 3)).reshape(-1) / float((TILE // WAVE) ** 2)


def cost_targets(costs: np.ndarray) -> np.ndarray:
    return np.log1p(costs).astype(np.float32)


def heaviest_first(predicted: np.ndarray) -> np.ndarray:
    """Tile order by decreasing predicted cost (ties in row-major order)."""
    return np.argsort(-np.asarray(predicted, dtype=np.float64), kind="stable").astype(np.uint32)


def log10_of_zoom(zoom_text: str) -> float:
    """log10 of a zoom given as decimal text ("1e100", "2.5e37")."""
    mantissa, _, exponent = zoom_text.lower().partition("e")
    return math.log10(float(mantissa)) + (int(exponent) if exponent else 0)


class CostModel:
    """The network's fp32 master weights (rounded to bf16 on use) and its
    Adam state, trained and evaluated by the HIP kernels."""

    def __init__(self, seed: int = 0, library=None):
        self.lib = library or fav_hip.load_library()
        self.params = self.lib.fractal_hip_cost_model_parameter_count()
        self.weights = self.initial_weights(seed)
        self.adam_m = np.zeros(self.params, dtype=np.float32)
        self.adam_v = np.zeros(self.params, dtype=np.float32)
        self.steps_done = 0
        self.hip = fav_hip.HipRuntime()

    @staticmethod
    def initial_weights(seed: int) -> np.ndarray:
        """He-normal weights, zero biases (layout W1[65][64], b1, W2[64][64],
        b2, W3[64], b3)."""
        rng = np.random.default_rng(seed)
        hidden = 64
        parts = [rng.normal(0.0, math.sqrt(2.0 / INPUTS), INPUTS * hidden), np.zeros(hidden),
                 rng.normal(0.0, math.sqrt(2.0 / hidden), hidden * hidden), np.zeros(hidden),
                 rng.normal(0.0, math.sqrt(1.0 / hidden), hidden), np.zeros(1)]
        return np.concatenate(parts).astype(np.float32)

    def _device(self, array: np.ndarray):
        ptr = self.hip.malloc(max(array.nbytes, 4))
        self.hip.to_device(ptr, array)
        return ptr

    def _check(self, status: int) -> None:
        if status != 0:
            raise RuntimeError(self.lib.fractal_last_error().decode())

    def predict(self, features: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(features, dtype=np.float32)
        n = x.shape[0]
        w_d, x_d, y_d = self._device(self.weights), self._device(x), self.hip.malloc(4 * n)
        try:
            self._check(self.lib.fractal_hip_cost_model_forward(w_d, x_d, n, y_d))
            y = np.empty(n, dtype=np.float32)
            self.hip.to_host(y, y_d)
            return y
        finally:
            for ptr in (w_d, x_d, y_d):
                self.hip.free(ptr)

    def gradient(self, features: np.ndarray, targets: np.ndarray) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(features, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        n = x.shape[0]
        w_d, x_d, t_d = self._device(self.weights), self._device(x), self._device(t)
        g_d, l_d = self.hip.malloc(4 * self.params), self.hip.malloc(4)
        try:
            self._check(self.lib.fractal_hip_cost_model_gradient(w_d, x_d, t_d, n, g_d, l_d))
            g = np.empty(self.params, dtype=np.float32)
            loss = np.empty(1, dtype=np.float32)
            self.hip.to_host(g, g_d)
            self.hip.to_host(loss, l_d)
            return g, float(loss[0])
        finally:
            for ptr in (w_d, x_d, t_d, g_d, l_d):
                self.hip.free(ptr)

    def train(self, features: np.ndarray, targets: np.ndarray, steps: int, learning_rate: float = 1e-3) -> np.ndarray:
        """`steps` full-batch Adam updates on the GPU; returns the loss before
        each update."""
        x = np.ascontiguousarray(features, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        n = x.shape[0]
        buffers = [self._device(self.weights), self._device(self.adam_m), self._device(self.adam_v),
                   self._device(x), self._device(t), self.hip.malloc(4 * steps)]
        w_d, m_d, v_d, x_d, t_d, h_d = buffers
        try:
            self._check(self.lib.fractal_hip_cost_model_train(w_d, m_d, v_d, x_d, t_d, n, int(steps),
                                                              float(learning_rate), self.steps_done + 1, h_d))
            history = np.empty(steps, dtype=np.float32)
            for host, ptr in ((self.weights, w_d), (self.adam_m, m_d), (self.adam_v, v_d), (history, h_d)):
                self.hip.to_host(host, ptr)
            self.steps_done += int(steps)
            return history
        finally:
            for ptr in buffers:
                self.hip.free(ptr)

    def state(self) -> dict:
        return {"weights": self.weights.copy(), "adam_m": self.adam_m.copy(), "adam_v": self.adam_v.copy(),
                "steps_done": self.steps_done}


@dataclass
class FieldSample:
    """One field's training pair: per-tile features and measured costs."""
    features: np.ndarray
    costs: np.ndarray
    preview_ms: float
    render_ms: float


PREVIEW_WORK_LIMIT = 128
# Iteration cap of a Mandelbrot preview: every loop pass of the field
# kernels advances at least one iteration, so the cap also bounds the
# preview's passes per pixel.  BLA blocks make a deep Mandelbrot pass cover
# many iterations, so the cap is larger than the alternate formulas' pass
# budget.
MANDELBROT_PREVIEW_CAP = 4096


def preview_limit(formula: int) -> int:
    """Preview budget: loop passes per pixel for the alternate formulas, the
    preview's iteration cap for the Mandelbrot map."""
    return PREVIEW_WORK_LIMIT if int(formula) else MANDELBROT_PREVIEW_CAP


def case_options(case: dict) -> fav_hip.NativeRenderOptions:
    return fav_hip.NativeRenderOptions(**case.get("options", {}))


def is_direct(case: dict) -> bool:
    """A Mandelbrot view below zoom 1e6: rendered by the direct renderer
    (no reference orbit, no preview entry; its work is its iterations)."""
    return int(case.get("formula", 0)) == 0 and log10_of_zoom(case["zoom"]) < 6.0


def _series(case: dict) -> tuple[int, int]:
    return int(case.get("series_order", 3)), int(case.get("series_block", 256))


def direct_work(reference: fav_hip.NativeReference, case: dict, width: int, height: int, max_iter: int) -> np.ndarray:
    """(height, width) iterations per pixel of a direct-renderer view (the
    escape index, max_iter inside): that renderer's loop passes."""
    _, planes = reference.render(width, height, case["zoom"], max_iter, *_series(case), planes=True,
                                 options=case_options(case))
    oi = planes["orbit_iteration"].reshape(height, width)
    return np.where(oi >= max_iter, max_iter, oi + 1).astype(np.int32)


def preview_features(reference: fav_hip.NativeReference, width: int, height: int, zoom: str, max_iter: int,
                     work_limit: int = PREVIEW_WORK_LIMIT, options=None, series=(3, 256)) -> tuple[np.ndarray, float]:
    """Tile features of one field from its quarter-size preview, rendered on
    the GPU with a per-pixel budget (`work_limit`: loop passes for the
    alternate formulas, the iteration cap for the Mandelbrot map; a preview
    has too few wavefronts to fill the GPU, so its time is its slowest
    wavefront's, and the budget bounds that; a pixel that reaches it marks a
    heavy region).  Returns the features and the preview's kernel ms."""
    features = reference.preview_features(width, height, zoom, max_iter, work_limit, *series, options=options)
    return features, fav_hip.last_timing()["kernel_ms"]


def field_features(reference: fav_hip.NativeReference, case: dict) -> tuple[np.ndarray, float]:
    """Cost-model inputs of the field `case` (any formula, any zoom) and the
    preview's kernel ms.  A direct-renderer view's preview is its
    quarter-size render with cap min(max_iter, MANDELBROT_PREVIEW_CAP),
    its features formed on the host from the iteration map."""
    w, h, z, n = case["width"], case["height"], case["zoom"], case["max_iter"]
    if is_direct(case):
        pw, ph = preview_size(w, h)
        work = direct_work(reference, case, pw, ph, min(n, MANDELBROT_PREVIEW_CAP))
        return tile_features(work, w, h, log10_of_zoom(z)), fav_hip.last_timing()["kernel_ms"]
    return preview_features(reference, w, h, z, n, preview_limit(case.get("formula", 0)), case_options(case),
                            _series(case))


def field_work(reference: fav_hip.NativeReference, case: dict) -> np.ndarray:
    """(height, width) measured loop passes per pixel of the field `case`."""
    w, h, z, n = case["width"], case["height"], case["zoom"], case["max_iter"]
    if is_direct(case):
        return direct_work(reference, case, w, h, n)
    _, work = reference.render_scheduled(w, h, z, n, work=True, series_order=_series(case)[0],
                                         series_block=_series(case)[1], options=case_options(case))
    return work


def sample_field(reference: fav_hip.NativeReference, width: int, height: int, zoom: str, max_iter: int,
                 options=None) -> FieldSample:
    """Renders the preview and the full field of an alternate-formula
    reference with work maps, and returns features and measured tile costs."""
    features, preview_ms = preview_features(reference, width, height, zoom, max_iter, options=options)
    _, work = reference.render_scheduled(width, height, zoom, max_iter, work=True, options=options)
    render_ms = fav_hip.last_timing()["kernel_ms"]
    return FieldSample(features, tile_costs(work, width, height), preview_ms, render_ms)


def spearman(a: np.ndarray, b: np.ndarray) -> float:
    """Rank correlation (average ranks for ties)."""
    def ranks(v):
        order = np.argsort(v, kind="stable")
        r = np.empty(len(v), dtype=np.float64)
        sv = np.asarray(v)[order]
        i = 0
        while i < len(v):
            j = i
            while j + 1 < len(v) and sv[j + 1] == sv[i]:
                j += 1
            r[order[i:j + 1]] = 0.5 * (i + j)
            i = j + 1
        return r
    ra, rb = ranks(a), ranks(b)
    ra -= ra.mean()
    rb -= rb.mean()
    denominator = math.sqrt(float((ra * ra).sum() * (rb * rb).sum()))
    return float((ra * rb).sum() / denominator) if denominator > 0 else 0.0


def prediction_report(predicted_log: np.ndarray, measured_costs: np.ndarray) -> dict:
    """Prediction error against measured tile costs: RMS and mean absolute
    error of log(1 + cost), R^2 in log space and rank correlation."""
    target = np.log1p(measured_costs)
    err = predicted_log.astype(np.float64) - target
    var = float(np.var(target))
    return {"rms_log_error": float(np.sqrt(np.mean(err * err))), "mean_abs_log_error": float(np.mean(np.abs(err))),
            "r2_log": 1.0 - float(np.mean(err * err)) / var if var > 0 else 0.0,
            "spearman": spearman(predicted_log, target), "tiles": int(len(target))}


def render_learned(reference: fav_hip.NativeReference, model: CostModel | None, width: int, height: int, zoom: str,
                   max_iter: int, work_limit: int | None = None, options=None, series=(3, 256)):
    """One field with the learned schedule (preview, cost model and ordered
    render in one library call); without a trained model, the plain grid.
    Returns (field, tile order or None, timings dict)."""
    if work_limit is None:
        work_limit = preview_limit(reference.formula)
    if model is None or model.steps_done == 0:
        field = reference.render_scheduled(width, height, zoom, max_iter, series_order=series[0],
                                           series_block=series[1], options=options)
        return field, None, {"render_ms": fav_hip.last_timing()["kernel_ms"]}
    return reference.render_learned(width, height, zoom, max_iter, model.weights, work_limit, *series, options=options)


def training_fields(count: int = 12, width: int = 1920, height: int = 1080, max_iter: int = 20000) -> list[dict]:
    """The seeded alternate-formula training set: `count` Julia fields
    (seeds 10, 11, ...) viewed at their repelling fixed point, zooms cycling
    1e30 / 1e60 / 1e100, and the tricorn preperiodic fork and antler
    junction at the same three zooms.  Disjoint from the evaluation fields
    (Julia seeds 0-9, the tricorn branching junction)."""
    import fav_workloads
    zooms = ("1e30", "1e60", "1e100")
    julia = [fav_workloads.julia_field(seed=10 + i, width=width, height=height, zoom=zooms[i % 3],
                                       max_iter=max_iter) for i in range(count)]
    tricorn = [fav_workloads.tricorn_field(point, width, height, zoom, max_iter)
               for point in ("preperiodic-fork", "antler-junction") for zoom in zooms]
    return julia + tricorn


def mandelbrot_training_fields(c3_seeds=(21, 22), segment_seed: int = 3, segment_stride: int = 8) -> list[dict]:
    """Seeded Mandelbrot-map training fields of the four measured workload
    families: the 640x360 zoom-1e12 field; the 1920x1080 zoom-1e100 field;
    2560x1440 near-nucleus fields (Mandelbrot map and burning ship, zooms
    1e30 and 1e100, nucleus seeds `c3_seeds`); and every `segment_stride`-th
    keyframe of the 48-field atlas segment of seed `segment_seed` (direct
    views below zoom 1e6 included).  Disjoint from the evaluation fields
    (nucleus seeds 1-9, segment seed 7)."""
    import fav_batch
    import fav_workloads
    fields = [dict(fav_workloads.field_1e12_640x360(), formula=0), dict(fav_workloads.field_1e100_1080p(), formula=0)]
    for formula in (0, 2):
        for seed, exponent in zip(c3_seeds, (30, 100)):
            fields.append(fav_workloads.near_boundary_field(formula, seed, exponent))
    for i, f in enumerate(fav_batch.segment_fields(seed=segment_seed)[::segment_stride]):
        fields.append(segment_case(f, f"segment{segment_seed}_{i}"))
    return fields


def segment_case(field: dict, name: str) -> dict:
    """A training / evaluation case of one atlas keyframe (fav_batch's
    segment field dict: Kalles bailout and grid, image series off)."""
    return dict(name=name, formula=0, julia=("0", "0"), x=field["x"], y=field["y"], zoom=field["zoom"],
                width=field["width"], height=field["height"], max_iter=field["max_iter"], bits=field["bits"],
                image_series_order=8, series_order=3, series_block=256,
                options=dict(series_min_terms=32, escape_radius_mode=1, coordinate_mode=1))


def reference_for(case: dict, library=None) -> fav_hip.NativeReference:
    return fav_hip.NativeReference(case["x"], case["y"], case["zoom"], case["max_iter"], case["bits"],
                                   int(case.get("image_series_order", 8)), case_options(case), library=library,
                                   formula=int(case.get("formula", 0)), julia=tuple(case.get("julia", ("0", "0"))))


def field_samples(fields: list[dict], library=None) -> tuple[np.ndarray, np.ndarray, float]:
    """Features and log-cost targets of every field (rows in field order,
    tiles row-major) and the seconds spent building and rendering them."""
    import time
    features, targets = [], []
    start = time.perf_counter()
    for case in fields:
        ref = reference_for(case, library)
        try:
            f, _ = field_features(ref, case)
            work = field_work(ref, case)
        finally:
            ref.close()
        features.append(f)
        targets.append(cost_targets(tile_costs(work, case["width"], case["height"])))
    return np.vstack(features), np.concatenate(targets), time.perf_counter() - start


def train_model(fields: list[dict], seed: int = 0, schedule=((3e-3, 1500), (1e-3, 1500)),
                library=None) -> tuple[CostModel, dict]:
    """Renders each field's preview features and measured tile costs, then
    trains a cost model on all of them with full-batch Adam on the GPU.
    Returns the model and a summary (samples, loss before / after, times)."""
    import time
    x, t, data_s = field_samples(fields, library)
    model = CostModel(seed, library)
    losses = []
    start = time.perf_counter()
    for lr, steps in schedule:
        losses.append(model.train(x, t, steps, learning_rate=lr))
    history = np.concatenate(losses)
    return model, {"samples": int(x.shape[0]), "fields": len(fields), "steps": int(history.size),
                   "loss_first": float(history[0]), "loss_last": float(history[-1]), "data_s": data_s,
                   "train_s": time.perf_counter() - start}


# ---- schedule quality without timing noise ----------------------------------

BLOCK = 16  # the field kernels' 16x16 thread blocks (four 8x8 wavefronts)


def block_costs(work: np.ndarray, width: int, height: int) -> np.ndarray:
    """(blocks,) float64 cost per 16x16 block, row-major: the largest of its
    four 8x8 wavefronts' largest per-pixel work (a block holds its slot
    until its slowest wavefront ends; a wavefront runs as long as its
    slowest lane)."""
    bx, by = (int(width) + BLOCK - 1) // BLOCK, (int(height) + BLOCK - 1) // BLOCK
    padded = np.zeros((by * BLOCK, bx * BLOCK), dtype=np.float64)
    h, w = work.shape
    padded[:h, :w] = work
    return padded.reshape(by, BLOCK, bx, BLOCK).max(axis=(1, 3)).reshape(-1)


def blocks_of_tiles(tile_order: np.ndarray, width: int, height: int) -> np.ndarray:
    """Row-major 16x16 block indices in the launch order of a 32x32 tile
    order: each tile's blocks top-left, top-right, bottom-left,
    bottom-right (render_kernel.hip block_pixel); blocks past the field's
    edge are dropped."""
    tx, _ = tile_grid(width, height)
    bx, by = (int(width) + BLOCK - 1) // BLOCK, (int(height) + BLOCK - 1) // BLOCK
    t = np.asarray(tile_order, dtype=np.int64)
    x0, y0 = (t % tx) * 2, (t // tx) * 2
    xs = np.stack([x0, x0 + 1, x0, x0 + 1], axis=1).reshape(-1)
    ys = np.stack([y0, y0, y0 + 1, y0 + 1], axis=1).reshape(-1)
    keep = (xs < bx) & (ys < by)
    return (ys * bx + xs)[keep]


def makespan(costs: np.ndarray, order: np.ndarray, slots: int) -> float:
    """Greedy list schedule of jobs `costs[order]` in that order onto
    `slots` identical machines, each job on the machine that frees first
    (how the dispatcher hands blocks to free slots): the finishing time of
    the last job."""
    import heapq
    heap = [0.0] * int(slots)
    for job in np.asarray(order, dtype=np.int64):
        start = heapq.heappop(heap)
        heapq.heappush(heap, start + float(costs[job]))
    return max(heap)


def resident_blocks(compute_units: int = 256, blocks_per_cu: int = 8) -> int:
    """Blocks resident at once: 8 blocks of four wavefronts per CU fill its
    four SIMDs at 8 waves each (the plain pass's occupancy)."""
    return int(compute_units) * int(blocks_per_cu)


def kernel_slots(formula: int, compute_units: int = 256) -> int:
    """Resident 16x16 blocks of the kernel that renders `formula`: the
    alternate-formula kernel runs 5 waves per SIMD (5 blocks per CU), the
    Mandelbrot plain pass 8."""
    return resident_blocks(compute_units, 5 if int(formula) else 8)


def makespan_report(work: np.ndarray, predicted: np.ndarray, width: int, height: int, slots: int) -> dict:
    """Makespans (in loop passes) of one field's blocks under the plain grid,
    the learned heaviest-first tile order of `predicted`, and the order of
    the measured tile costs themselves (what a perfect predictor reaches),
    with the lower bound max(total / slots, largest block) and the learned
    order's gain over the plain grid in percent."""
    costs = block_costs(work, width, height)
    learned = blocks_of_tiles(heaviest_first(predicted), width, height)
    measured = blocks_of_tiles(heaviest_first(tile_costs(work, width, height)), width, height)
    r = {"plain": makespan(costs, np.arange(costs.size), slots), "learned": makespan(costs, learned, slots),
         "measured_order": makespan(costs, measured, slots),
         "bound": max(float(costs.sum()) / slots, float(costs.max())), "slots": int(slots)}
    r["gain_pct"] = 100.0 * (r["plain"] - r["learned"]) / r["plain"] if r["plain"] > 0 else 0.0
    return r
