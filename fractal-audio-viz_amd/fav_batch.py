"""Batched atlas fields: one launch sequence over the keyframe fields of a
segment, plus the launch autotuner.

A segment's keyframes are nested fields at log-spaced zooms around one
centre, each with its own reference orbit, size and iteration cap.  The
batch renders them over one device tile queue (batch_kernel.hip); the
autotuner times the result-invariant launch choices (tile order and the
fast kernel's grid) on the batch itself and keeps the fastest.  The
per-field BLA settings are not tuned: they change the arithmetic, and
every field must stay bit-identical to its single-field render.
"""
import concurrent.futures
import ctypes
import math
import random
import time

import numpy as np

import fav_hip
import fav_workloads

SIZES = ((1280, 720), (1920, 1080), (3840, 2160))


def atlas_cap(zoom_log10: float) -> int:
    """Iteration cap of a keyframe: 2000 + 1000 per decade, 64..100000
    (visualizer.py:6976-6979 with base 2000)."""
    return int(min(100000, max(64, int(2000 + 1000 * zoom_log10))))


def precision_bits(zoom_log10: float) -> int:
    """MPFR bits for a reference at this zoom: the decimal digits of the
    zoom plus a guard, at least 128 (the reference's floor)."""
    return max(128, int(math.ceil((zoom_log10 + 20) * 3.33)) + 32)


def segment_fields(count: int = 48, seed: int = 7, decades: float = 100.0, min_log10: float = 0.0,
                   sizes=SIZES, cap_scale: float = 1.0):
    """`count` keyframes at zooms 10^(decades*k/(count-1)) around the 100-digit
    extended deep point, sizes drawn uniformly (seeded) from `sizes`.  Fields
    below zoom 1e6 are direct-renderer fields (the batch renders them from
    the same tile queue); fields below `min_log10` are dropped."""
    rng = random.Random(seed)
    x, y = fav_workloads.extended_centre()
    fields = []
    for k in range(count):
        z10 = decades * k / max(1, count - 1)
        w, h = sizes[rng.randrange(len(sizes))]
        if z10 < min_log10:
            continue
        whole = math.floor(z10 + 1e-12)
        zoom = f"{10.0 ** (z10 - whole):.12f}e{whole}"
        cap = max(64, int(atlas_cap(z10) * cap_scale))
        fields.append(dict(x=x, y=y, zoom=zoom, zoom_log10=z10, width=w, height=h, max_iter=cap,
                           bits=precision_bits(z10)))
    return fields


class Batch:
    """References, device outputs and the prepared batch of a field list."""

    def __init__(self, fields, options=None, library=None, hip=None, threads: int = 8):
        self.lib = library or fav_hip.load_library()
        self.hip = hip or fav_hip.HipRuntime()
        self.options = options or fav_hip.NativeRenderOptions(series_min_terms=32, escape_radius_mode=1,
                                                              coordinate_mode=1)
        self.fields = fields
        started = time.perf_counter()

        def build(f):
            return fav_hip.NativeReference(f["x"], f["y"], f["zoom"], f["max_iter"], f["bits"], 8, self.options,
                                           library=self.lib)

        with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as pool:
            self.references = list(pool.map(build, fields))
        self.reference_build_s = time.perf_counter() - started
        self.outputs = [self.hip.malloc(f["width"] * f["height"] * 4) for f in fields]
        entries = (fav_hip.NativeBatchField * len(fields))()
        self._zooms = [f["zoom"].encode() for f in fields]
        for i, f in enumerate(fields):
            entries[i].handle = self.references[i].handle
            entries[i].zoom_text = self._zooms[i]
            entries[i].device_output = self.outputs[i]
            entries[i].width, entries[i].height, entries[i].max_iter = f["width"], f["height"], f["max_iter"]
            entries[i].series_order, entries[i].series_block = 3, 256
        self.handle = self.lib.fractal_hip_batch_create(entries, len(fields), ctypes.byref(self.options))
        if not self.handle:
            raise RuntimeError(self.lib.fractal_last_error().decode())
        self.tile_count = self.lib.fractal_hip_batch_tile_count(self.handle)
        self.settings = dict(order=0, blocks_per_cu=5)

    def pixels(self) -> int:
        return sum(f["width"] * f["height"] for f in self.fields)

    def set_order(self, order: int, costs=None) -> None:
        cost = None
        if costs is not None:
            cost = np.ascontiguousarray(costs, dtype=np.float32)
        status = self.lib.fractal_hip_batch_set_order(self.handle, order,
                                                      cost.ctypes.data if cost is not None else None,
                                                      0 if cost is None else cost.size)
        if status != 0:
            raise RuntimeError(self.lib.fractal_last_error().decode())

    def set_precision(self, floatexp: bool) -> None:
        """FP64 launch sequence (False) or the floatexp pixel (True)."""
        if self.lib.fractal_hip_batch_set_precision(self.handle, 1 if floatexp else 0) != 0:
            raise RuntimeError(self.lib.fractal_last_error().decode())

    def render(self, blocks_per_cu: int = 0, graph: bool = True) -> float:
        """One batched render; returns the HIP-event time in ms."""
        status = self.lib.fractal_hip_batch_render(self.handle, blocks_per_cu, 1 if graph else 0)
        if status != 0:
            raise RuntimeError(self.lib.fractal_last_error().decode())
        return fav_hip.last_timing()["kernel_ms"]

    def render_per_field(self) -> float:
        """The same fields with one launch sequence each (the unbatched
        baseline); returns the summed HIP-event time in ms."""
        total = 0.0
        for i, f in enumerate(self.fields):
            status = self.lib.fractal_hip_render_reference_device(
                self.outputs[i], f["width"], f["height"], f["zoom"].encode(), self.references[i].handle,
                f["max_iter"], 3, 256, ctypes.byref(self.options), None)
            if status != 0:
                raise RuntimeError(self.lib.fractal_last_error().decode())
            total += fav_hip.last_timing()["kernel_ms"]
        return total

    def field(self, i: int) -> np.ndarray:
        f = self.fields[i]
        out = np.empty(f["width"] * f["height"], dtype=np.float32)
        self.hip.to_host(out, self.outputs[i])
        return out.reshape(f["height"], f["width"])

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.fractal_hip_batch_destroy(self.handle)
            self.handle = None
        for ptr in getattr(self, "outputs", []):
            self.hip.free(ptr)
        self.outputs = []
        for ref in getattr(self, "references", []):
            ref.close()


def expand_to_blocks(tile_values: np.ndarray, width: int, height: int) -> np.ndarray:
    """Per-32x32-tile values (row-major) repeated onto the field's row-major
    16x16 batch tiles: each batch tile takes the value of the 32x32 tile
    that contains it."""
    tx = (int(width) + 31) // 32
    bx, by = (int(width) + 15) // 16, (int(height) + 15) // 16
    ys, xs = np.divmod(np.arange(bx * by), bx)
    return np.asarray(tile_values)[(ys // 2) * tx + xs // 2]


def learned_tile_costs(batch: Batch, model) -> tuple[np.ndarray, dict]:
    """The cost model's prediction for every batch tile, in field-major tile
    order (the input of Batch.set_order(2, costs)): each field's preview
    features (fav_sched.field_features), the model on the GPU, and the
    32x32 predictions spread over the field's 16x16 batch tiles.  Returns
    the costs and the preview / model times."""
    import fav_sched
    costs, preview_ms, model_s = [], 0.0, 0.0
    for i, f in enumerate(batch.fields):
        case = fav_sched.segment_case(f, f"field{i}")
        features, ms = fav_sched.field_features(batch.references[i], case)
        preview_ms += ms
        started = time.perf_counter()
        predicted = model.predict(features)
        model_s += time.perf_counter() - started
        costs.append(expand_to_blocks(predicted, f["width"], f["height"]))
    return np.concatenate(costs).astype(np.float32), {"preview_ms": preview_ms, "model_wall_ms": model_s * 1e3}


def measured_block_costs(batch: Batch) -> np.ndarray:
    """Every batch tile's measured cost (fav_sched.block_costs of the
    field's per-pixel loop passes), in field-major tile order."""
    import fav_sched
    costs = []
    for i, f in enumerate(batch.fields):
        work = fav_sched.field_work(batch.references[i], fav_sched.segment_case(f, f"field{i}"))
        costs.append(fav_sched.block_costs(work, f["width"], f["height"]))
    return np.concatenate(costs)


def autotune(batch: Batch, costs=None, repeats: int = 2):
    """Times every (tile order, fast-kernel blocks per CU) candidate on the
    batch (graph launches, best of `repeats`) and keeps the fastest.
    Returns the trial table and the chosen setting."""
    orders = [0, 1] + ([2] if costs is not None else [])
    trials = []
    for order in orders:
        batch.set_order(order, costs if order == 2 else None)
        for blocks in (4, 5, 8):
            batch.render(blocks, graph=True)  # capture + warm
            ms = min(batch.render(blocks, graph=True) for _ in range(repeats))
            trials.append(dict(order=order, blocks_per_cu=blocks, ms=ms))
    best = min(trials, key=lambda t: t["ms"])
    batch.set_order(best["order"], costs if best["order"] == 2 else None)
    batch.render(best["blocks_per_cu"], graph=True)
    batch.settings = dict(order=best["order"], blocks_per_cu=best["blocks_per_cu"])
    return trials, best
