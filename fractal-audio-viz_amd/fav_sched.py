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


def preview_features(reference: fav_hip.NativeReference, width: int, height: int, zoom: str, max_iter: int,
                     work_limit: int = PREVIEW_WORK_LIMIT, options=None) -> tuple[np.ndarray, float]:
    """Tile features of one field from its quarter-size preview, rendered on
    the GPU with at most `work_limit` loop passes per pixel (a preview has
    too few wavefronts to fill the GPU, so its time is its slowest
    wavefront's; the budget bounds that, and a pixel that reaches it marks
    a heavy region).  Returns the features and the preview's kernel ms."""
    features = reference.preview_features(width, height, zoom, max_iter, work_limit, options=options)
    return features, fav_hip.last_timing()["kernel_ms"]


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
                   max_iter: int, work_limit: int = PREVIEW_WORK_LIMIT, options=None):
    """One field with the learned schedule (preview, cost model and ordered
    render in one library call); without a trained model, the plain grid.
    Returns (field, tile order or None, timings dict)."""
    if model is None or model.steps_done == 0:
        field = reference.render_scheduled(width, height, zoom, max_iter, options=options)
        return field, None, {"render_ms": fav_hip.last_timing()["kernel_ms"]}
    return reference.render_learned(width, height, zoom, max_iter, model.weights, work_limit, options=options)


def training_fields(count: int = 12, width: int = 1920, height: int = 1080, max_iter: int = 20000) -> list[dict]:
    """The seeded training set: `count` Julia fields (seeds 10, 11, ...)
    viewed at their repelling fixed point, zooms cycling 1e30 / 1e60 / 1e100,
    and the tricorn preperiodic fork and antler junction at the same three
    zooms.  Disjoint from the evaluation fields (Julia seeds 0-9, the
    tricorn branching junction)."""
    import fav_workloads
    zooms = ("1e30", "1e60", "1e100")
    julia = [fav_workloads.julia_field(seed=10 + i, width=width, height=height, zoom=zooms[i % 3],
                                       max_iter=max_iter) for i in range(count)]
    tricorn = [fav_workloads.tricorn_field(point, width, height, zoom, max_iter)
               for point in ("preperiodic-fork", "antler-junction") for zoom in zooms]
    return julia + tricorn


def reference_for(case: dict, library=None) -> fav_hip.NativeReference:
    return fav_hip.NativeReference(case["x"], case["y"], case["zoom"], case["max_iter"], case["bits"],
                                   library=library, formula=case["formula"], julia=tuple(case["julia"]))


def train_model(fields: list[dict], seed: int = 0, schedule=((3e-3, 1500), (1e-3, 1500)),
                work_limit: int = PREVIEW_WORK_LIMIT, library=None) -> tuple[CostModel, dict]:
    """Renders each field's preview features and measured tile costs, then
    trains a cost model on all of them with full-batch Adam on the GPU.
    Returns the model and a summary (samples, loss before / after, times)."""
    import time
    features, targets = [], []
    start = time.perf_counter()
    for case in fields:
        ref = reference_for(case, library)
        try:
            f, _ = preview_features(ref, case["width"], case["height"], case["zoom"], case["max_iter"], work_limit)
            _, work = ref.render_scheduled(case["width"], case["height"], case["zoom"], case["max_iter"], work=True)
        finally:
            ref.close()
        features.append(f)
        targets.append(cost_targets(tile_costs(work, case["width"], case["height"])))
    data_s = time.perf_counter() - start
    x, t = np.vstack(features), np.concatenate(targets)
    model = CostModel(seed, library)
    losses = []
    start = time.perf_counter()
    for lr, steps in schedule:
        losses.append(model.train(x, t, steps, learning_rate=lr))
    history = np.concatenate(losses)
    return model, {"samples": int(x.shape[0]), "steps": int(history.size), "loss_first": float(history[0]),
                   "loss_last": float(history[-1]), "data_s": data_s, "train_s": time.perf_counter() - start}
