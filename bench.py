#!/usr/bin/env python3
"""Deep-zoom field throughput on one MI355X.

Workload (one step): the 1920x1080 field at zoom 1e100 with a 360-bit MPFR
reference orbit, cap 100000, BLA on, rebase on, Kalles bailout/grid, render
planes on, FP64 perturbation, then Kalles colouring of the field (method 7
field-difference distance, slopes at 45 degrees, seeded dither mask) with a
1024-entry cyclic palette of 12 seeded stops.  The reference orbit and BLA
tables are built on the host once and copied to HBM before timing; each
timed step is one field render and one colour pass into device-resident
buffers.

value = Gpixel-iterations/s = sum over pixels of n_p (escape index, or the
cap for interior / unresolved pixels) per second of wall time per step.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "fractal-audio-viz_amd"))

import fav_batch  # noqa: E402
import fav_hip  # noqa: E402
import fav_palette  # noqa: E402
import fav_sched  # noqa: E402
import fav_workloads  # noqa: E402

MI355X_FP64_VECTOR_PEAK_TFLOPS = 78.6
TRAFFIC_PROFILE = ROOT / "profiles" / "field_kernel_1920x1080_zoom1e100_traffic.md"


def measured_traffic_bytes():
    """HBM bytes per field launch from the committed FETCH_SIZE/WRITE_SIZE
    profile (copy-calibrated), or None when that summary is absent."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def escape_count_sum(orbit_iteration, field, cap):
    counts = orbit_iteration.astype(np.int64) + 1
    counts[orbit_iteration >= cap] = cap
    counts[np.isnan(field)] = cap
    return int(counts.sum())


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo (platform.processor()
    only names the architecture on Linux)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.machine()


def reference_build_flags():
    """Compiler flags oracle/_ref/mandelbrot.so was built with (recorded by
    oracle/Makefile next to the library)."""
    try:
        return (ROOT / "oracle" / "_ref" / "build_flags.txt").read_text().strip()
    except OSError:
        return "unknown"


def cpu_baseline(work, threads):
    """The reference's renderer.cpp (oracle/_ref, built from source) on the
    host cores, over a quarter-area sample (half width, half height) of the
    same workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    import native  # checker-side ctypes mirrors

    lib = native.load_reference(release=True)
    kind = "reference"
    sample_w, sample_h = work["width"] // 2, work["height"] // 2
    opts = native.RenderOptions.defaults(backend=0, **work["options"])
    build_s = None
    if lib is not None:
        t0 = time.perf_counter()
        ref = native.AbiReference(lib, work["x"], work["y"], work["zoom"], work["max_iter"],
                                  work["bits"], work["image_series_order"], opts)
        build_s = time.perf_counter() - t0
    else:
        kind = "port"
        lib = native.load_oracle()
        ref = native.OracleReference(lib, work["x"], work["y"], work["zoom"], work["max_iter"],
                                     work["bits"], work["image_series_order"],
                                     work["options"].get("escape_radius_mode", 0),
                                     work["options"].get("coordinate_mode", 0))
    started = time.perf_counter()
    field, planes, _ = ref.render(sample_w, sample_h, work["zoom"], work["max_iter"],
                                  work["series_order"], work["series_block"], opts,
                                  planes=work["planes"], threads=threads)
    seconds = time.perf_counter() - started
    total = escape_count_sum(planes.orbit_iteration, field.reshape(-1), work["max_iter"])
    return {"value": total / seconds / 1e9, "unit": "Gpixel-iter/s", "cores": threads, "kind": kind,
            "sample": f"{sample_w}x{sample_h} of the same field (1/4 of the pixels), "
                      f"{seconds:.1f} s on {threads} threads",
            "cpu_model": cpu_model(),
            "build_flags": reference_build_flags() if kind == "reference" else "oracle/Makefile"}, build_s


def copy_calibration(mib, copies):
    """Plain device-to-device copies of a known size: the reference point
    for the FETCH_SIZE / WRITE_SIZE counters (each copy reads and writes
    `mib` MiB)."""
    hip = fav_hip.HipRuntime()
    nbytes = mib << 20
    src, dst = hip.malloc(nbytes), hip.malloc(nbytes)
    for _ in range(copies):
        status = hip.lib.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)
        if status != 0:
            raise RuntimeError(f"hipMemcpy D2D failed with {status}")
    hip.synchronize()
    hip.free(src)
    hip.free(dst)
    print(json.dumps({"copy_calibration_mib": mib, "copies": copies}))
    return 0


def batch_leg(lib, hip, steps, model=None):
    """One atlas segment (48 keyframe fields from zoom 1 to 1e100, each with
    its own 100-digit-centre reference, sizes drawn from 720p/1080p/4K, caps
    2000 + 1000/decade up to 100000; the three below 1e6 are direct-renderer
    fields) rendered as one autotuned, graph-captured batch against one
    launch sequence per field."""
    fields = fav_batch.segment_fields()
    batch = fav_batch.Batch(fields, library=lib, hip=hip, threads=16)
    try:
        # pixel-iterations from the per-field statistics (logical iterations
        # plus the first one of every pixel)
        lib.fractal_set_stats_enabled(1)
        pixel_iterations = 0
        stats = np.zeros(39, dtype=np.uint64)
        for i, f in enumerate(fields):
            status = lib.fractal_hip_render_reference_device(
                batch.outputs[i], f["width"], f["height"], f["zoom"].encode(), batch.references[i].handle,
                f["max_iter"], 3, 256, ctypes.byref(batch.options), None)
            if status != 0:
                raise RuntimeError(lib.fractal_last_error().decode())
            if f["zoom_log10"] < 6.0:
                # the direct renderer keeps no statistics: its pixel
                # iterations are the escape indices from the iteration plane
                # (orbit_iteration + 1, the cap for interior pixels)
                _, planes = batch.references[i].render(f["width"], f["height"], f["zoom"], f["max_iter"],
                                                       planes=True)
                oi = planes["orbit_iteration"]
                pixel_iterations += int(np.where(oi >= f["max_iter"], f["max_iter"], oi + 1).sum())
                continue
            lib.fractal_get_last_stats_ex(stats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 39)
            pixel_iterations += int(stats[1]) + int(stats[0])
        lib.fractal_set_stats_enabled(0)
        per_field_ms = min(batch.render_per_field() for _ in range(max(1, steps)))
        learned = None
        costs = None
        if model is not None:
            # the cost model's order over the batch queue, and the makespan
            # model of the batch's blocks (measured work maps) under it
            costs, times = fav_batch.learned_tile_costs(batch, model)
            measured = fav_batch.measured_block_costs(batch)
            slots = fav_sched.resident_blocks(blocks_per_cu=5)
            order = np.argsort(-costs.astype(np.float64), kind="stable")
            learned = {"preview_ms": round(times["preview_ms"], 3),
                       "makespan_field_major": fav_sched.makespan(measured, np.arange(measured.size), slots),
                       "makespan_learned": fav_sched.makespan(measured, order, slots),
                       "makespan_measured_order": fav_sched.makespan(
                           measured, np.argsort(-measured, kind="stable"), slots),
                       "makespan_bound": round(max(float(measured.sum()) / slots, float(measured.max())), 1)}
        trials, best = fav_batch.autotune(batch, costs=costs, repeats=1)
        if learned is not None:
            learned["learned_order_ms"] = round(min(t["ms"] for t in trials if t["order"] == 2), 3)
            learned["field_major_ms"] = round(min(t["ms"] for t in trials if t["order"] == 0), 3)
        batched_ms = min(batch.render(0, graph=True) for _ in range(max(1, steps)))
        # end to end with the references built here: every reference first,
        # then the renders, against references built on host threads while
        # the GPU renders the ones already staged
        serial = fav_hip.render_segment(fields, batch.outputs, batch.options, threads=16, pipelined=False)
        overlapped = fav_hip.render_segment(fields, batch.outputs, batch.options, threads=16, pipelined=True)
        return {"workload": f"atlas segment: {len(fields)} fields, {batch.pixels() / 1e6:.1f} Mpx, zoom 1..1e100",
                "value": round(pixel_iterations / (batched_ms * 1e-3) / 1e9, 3), "unit": "Gpixel-iter/s",
                "batched_ms": round(batched_ms, 3), "per_field_launches_ms": round(per_field_ms, 3),
                "speedup_vs_per_field": round(per_field_ms / batched_ms, 3),
                "pixel_iterations": pixel_iterations, "reference_build_s": round(batch.reference_build_s, 3),
                "autotuned": best, "trials": len(trials), "learned_order": learned,
                "end_to_end": {"build_then_render_s": round(serial["wall_ms"] * 1e-3, 3),
                               "pipelined_s": round(overlapped["wall_ms"] * 1e-3, 3),
                               "speedup": round(serial["wall_ms"] / overlapped["wall_ms"], 3),
                               "host_build_s_summed": round(overlapped["build_ms"] * 1e-3, 3),
                               "last_reference_staged_s": round(overlapped["built_ms"] * 1e-3, 3),
                               "field_launches_s": round(overlapped["render_ms"] * 1e-3, 3),
                               "build_threads": 16}}
    finally:
        batch.close()


def learned_leg(lib, repeats=2):
    """The learned tile scheduler: one cost model trained on the GPU over the
    seeded training fields of both families (Mandelbrot map: the 640x360
    zoom-1e12 field, the 1920x1080 zoom-1e100 field, 2560x1440 near-nucleus
    fields and atlas keyframes; Julia and tricorn fields at 1920x1080, cap
    20000), then held-out fields rendered in the plain grid and in the
    learned heaviest-first order: the tricorn branching junction and Julia
    seeds 1, 2, 4 at 3840x2160, Julia seed 1 at 7680x4320 (zoom 1e100, cap
    200000), a near-nucleus 2560x1440 zoom-1e30 Mandelbrot field, the
    Misiurewicz 1920x1080 zoom-1e100 field and the 3840x2160 zoom-1e100
    interior field.  Net gain counts the preview, features and model time;
    prediction error and the makespan model (fav_sched.makespan_report) are
    against each field's measured work map.  Returns the leg's JSON and the
    model."""
    model, training = fav_sched.train_model(fav_sched.mandelbrot_training_fields() + fav_sched.training_fields(),
                                            library=lib)
    fields = []
    cases = [fav_workloads.tricorn_field("branching-junction", 3840, 2160, "1e100", 200000)] + [
        fav_workloads.julia_field(seed=seed, width=3840, height=2160, zoom="1e100", max_iter=200000)
        for seed in (1, 2, 4)] + [
        fav_workloads.julia_field(seed=1, width=7680, height=4320, zoom="1e100", max_iter=200000),
        fav_workloads.near_boundary_field(0, 1, 30), dict(fav_workloads.misiurewicz_field(), formula=0),
        dict(fav_workloads.field_1e100_1080p(), formula=0, width=3840, height=2160,
             name="field_3840x2160_zoom1e100_cap100000")]
    for case in cases:
        w, h, z, n = case["width"], case["height"], case["zoom"], case["max_iter"]
        opts, series = fav_sched.case_options(case), fav_sched._series(case)
        ref = fav_sched.reference_for(case, lib)
        try:
            work = fav_sched.field_work(ref, case)
            costs = fav_sched.tile_costs(work, w, h)
            features, _ = fav_sched.field_features(ref, case)
            predicted = model.predict(features)
            report = fav_sched.prediction_report(predicted, costs)
            schedule = fav_sched.makespan_report(work, predicted, w, h, fav_sched.kernel_slots(case.get("formula", 0)))
            plain = []
            for _ in range(repeats):
                ref.render_scheduled(w, h, z, n, series_order=series[0], series_block=series[1], options=opts)
                plain.append(fav_hip.last_timing()["kernel_ms"])
            learned = [fav_sched.render_learned(ref, model, w, h, z, n, options=opts, series=series)[2]
                       for _ in range(repeats)]
            best = min(learned, key=lambda t: t["preview_ms"] + t["model_ms"] + t["render_ms"])
            total = best["preview_ms"] + best["model_ms"] + best["render_ms"]
            fields.append({"field": case["name"], "plain_ms": round(min(plain), 3),
                           "learned_total_ms": round(total, 3), "preview_ms": round(best["preview_ms"], 3),
                           "model_ms": round(best["model_ms"], 3), "ordered_render_ms": round(best["render_ms"], 3),
                           "net_gain_pct": round(100.0 * (min(plain) - total) / min(plain), 2),
                           "makespan_model": {k: round(v, 2) if isinstance(v, float) else v
                                              for k, v in schedule.items()},
                           "prediction": {k: round(v, 4) if isinstance(v, float) else v for k, v in report.items()}})
        finally:
            ref.close()
    alternate = [f["net_gain_pct"] for f in fields if not f["field"].startswith(("mandelbrot", "misiurewicz", "field"))]
    return {"workload": "held-out fields: tricorn branching junction and Julia seeds 1 2 4 at 3840x2160, Julia "
                        "seed 1 at 7680x4320 (zoom 1e100, cap 200000); Mandelbrot near-nucleus 2560x1440 zoom "
                        "1e30, Misiurewicz 1920x1080 zoom 1e100, interior 3840x2160 zoom 1e100",
            "preview": f"quarter size, {fav_sched.PREVIEW_WORK_LIMIT} passes per pixel (alternate formulas), "
                       f"cap {fav_sched.MANDELBROT_PREVIEW_CAP} iterations (Mandelbrot map)",
            "training": {k: round(v, 4) if isinstance(v, float) else v for k, v in training.items()},
            "mean_net_gain_pct_alternate": round(float(np.mean(alternate)), 2),
            "mean_makespan_gain_pct": round(float(np.mean([f["makespan_model"]["gain_pct"] for f in fields])), 2),
            "fields": fields}, model


def atlas_leg(lib, repeats=5):
    """The raw KFP atlas colour pass of one video frame: 1920x1080 output
    from two 1920x1080 scalar tiles (parent zoomed 1.5x, child a third of
    the frame), Kalles-default colouring; HIP-event time of the compose and
    colour kernels with the tiles resident on the host (each call copies
    them in and the RGB out, as the reference's entry does)."""
    parent = fav_workloads.atlas_tile(1920, 1080, 20000, seed=41)
    child = fav_workloads.atlas_tile(1920, 1080, 24000, seed=42)
    opts = fav_hip.NativeKfpOptions.kalles_default()
    lut = fav_palette.kfp_palette_lut(fav_palette.seeded_stops(12, 2024), 1024)
    out = np.empty((1080, 1920, 3), dtype=np.uint8)
    fn = lib.fractal_atlas_colourise_kfp_raw
    times, walls = [], []
    for _ in range(repeats):
        t0 = time.perf_counter()
        status = fn(parent.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 1920, 1080, 20000,
                    child.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 1920, 1080, 24000,
                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), 1920, 1080, 1.5, 0.33, 1.6, 24000,
                    0.0, 0.0, 0.0, 0.0, ctypes.byref(opts), lut.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                    lut.shape[0], 0)
        walls.append((time.perf_counter() - t0) * 1e3)
        if status != 0:
            raise RuntimeError(lib.fractal_last_error().decode())
        times.append(fav_hip.last_timing()["kernel_ms"])
    kernel_ms = min(times)
    # the ordinary-palette atlas pass of the same frame (accented palette)
    accents = np.array([250, 40, 90, 30, 200, 120, 60, 80, 240], dtype=np.uint8)
    aurora_out = np.empty((1080, 1920, 3), dtype=np.uint8)
    aurora_times, aurora_walls = [], []
    for _ in range(repeats):
        t0 = time.perf_counter()
        status = lib.fractal_atlas_colourise_accents(
            parent.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 1920, 1080, 20000,
            child.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 1920, 1080, 24000,
            aurora_out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), 1920, 1080, 1.5, 0.33, 24000, 1.3, 0.7, 0.2,
            0.8, accents.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), 10, 20, 30, 0)
        aurora_walls.append((time.perf_counter() - t0) * 1e3)
        if status != 0:
            raise RuntimeError(lib.fractal_last_error().decode())
        aurora_times.append(fav_hip.last_timing()["kernel_ms"])
    aurora_ms = min(aurora_times)
    return {"workload": "1920x1080 frame from 1920x1080 parent (zoom 1.5) and child (1/3 of the frame, zoom 1.6)",
            "kernel_ms": round(kernel_ms, 4), "call_ms": round(min(walls), 3),
            "frames_per_s_kernel": round(1e3 / kernel_ms, 1), "rgb_checksum": int(out.astype(np.uint64).sum()),
            "aurora": {"entry": "fractal_atlas_colourise_accents", "kernel_ms": round(aurora_ms, 4),
                       "call_ms": round(min(aurora_walls), 3), "frames_per_s_kernel": round(1e3 / aurora_ms, 1),
                       "rgb_checksum": int(aurora_out.astype(np.uint64).sum())}}


def boundary_leg(lib, hip, kfp, lut, steps):
    """Boundary-rich Mandelbrot fields timed like the headline step (field
    render with planes, then the KFP colour pass, device-resident): the
    2560x1440 zoom-1e30 field around a Newton-refined period-121 nucleus
    (escapes 590..895; the 4% of pixels around the orbit's near-zero
    returns are glitch-rebased) and the 1920x1080 zoom-1e100 field at the
    1/3-limb Misiurewicz point (every pixel escapes, 806..881)."""
    nucleus = fav_workloads.nucleus_centre(0, 1, 30)
    fields = [fav_workloads.near_boundary_field(0, 1, 30, centre=nucleus), fav_workloads.misiurewicz_field()]
    out = []
    lut_p = lut.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    for work in fields:
        opts = fav_hip.NativeRenderOptions(**work["options"])
        ref = fav_hip.NativeReference(work["x"], work["y"], work["zoom"], work["max_iter"], work["bits"],
                                      work["image_series_order"], opts)
        ref.upload()
        w, h, cap = work["width"], work["height"], work["max_iter"]
        n = w * h
        d_field, d_planes, d_rgb = hip.malloc(n * 4), hip.malloc(n * 48), hip.malloc(n * 3)
        planes = fav_hip.NativeRenderPlanes()
        planes.struct_size = ctypes.sizeof(fav_hip.NativeRenderPlanes)
        planes.version = 1
        for i, name in enumerate(("orbit_iteration", "phase", "de_x", "de_y", "test1", "test2")):
            setattr(planes, name, ctypes.cast(ctypes.c_void_p(d_planes.value + i * n * 8), planes._fields_[2 + i][1]))

        def step():
            status = lib.fractal_hip_render_reference_device(
                d_field, w, h, work["zoom"].encode(), ref.handle, cap, work["series_order"], work["series_block"],
                ctypes.byref(opts), ctypes.byref(planes))
            if status != 0:
                raise RuntimeError(lib.fractal_last_error().decode())
            field_ms = fav_hip.last_timing()["kernel_ms"]
            status = lib.fractal_hip_colourise_kfp_device(d_field, d_rgb, w, h, cap, ctypes.byref(kfp), lut_p,
                                                          lut.shape[0])
            if status != 0:
                raise RuntimeError(lib.fractal_last_error().decode())
            return field_ms, fav_hip.last_timing()["kernel_ms"]

        lib.fractal_set_stats_enabled(1)
        step()
        stats = np.zeros(39, dtype=np.uint64)
        lib.fractal_get_last_stats_ex(stats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 39)
        lib.fractal_set_stats_enabled(0)
        field = np.empty(n, dtype=np.float32)
        orbit_iteration = np.empty(n, dtype=np.int64)
        hip.to_host(field, d_field)
        hip.to_host(orbit_iteration, d_planes)
        pixel_iterations = escape_count_sum(orbit_iteration, field, cap)
        step()
        hip.synchronize()
        field_ms, colour_ms = [], []
        started = time.perf_counter()
        for _ in range(steps):
            a, b = step()
            field_ms.append(a)
            colour_ms.append(b)
        hip.synchronize()
        elapsed = (time.perf_counter() - started) / steps
        rgb = np.empty(n * 3, dtype=np.uint8)
        hip.to_host(rgb, d_rgb)
        for ptr in (d_field, d_planes, d_rgb):
            hip.free(ptr)
        ref.close()
        escaped = np.isfinite(field) & (orbit_iteration < cap)
        out.append({"workload": work["name"], "value": round(pixel_iterations / elapsed / 1e9, 3),
                    "unit": "Gpixel-iter/s", "ms_per_step": round(elapsed * 1e3, 3),
                    "mean_kernel_ms": round(float(np.mean(field_ms)), 3),
                    "mean_colour_ms": round(float(np.mean(colour_ms)), 4),
                    "pixel_iterations": pixel_iterations, "escaped_pct": round(float(escaped.mean()) * 100, 2),
                    "nan_pixels": int(np.isnan(field).sum()), "rgb_checksum": int(rgb.astype(np.uint64).sum()),
                    "linear_blocks": int(stats[3]), "cubic_blocks": int(stats[4]), "exact_steps": int(stats[5]),
                    "tail_steps": int(stats[11]), "series_jumps": int(stats[17]), "glitches": int(stats[18])})
    return out


def floatexp_boundary_mismatch(lib):
    """The floatexp path where escape counts are sensitive: the 640x360
    zoom-1e12 field (cap 20000, every pixel near the boundary), against the
    FP64 kernel's counts on the same inputs."""
    work = fav_workloads.field_1e12_640x360()
    opts = fav_hip.NativeRenderOptions(**work["options"])
    ref = fav_hip.NativeReference(work["x"], work["y"], work["zoom"], work["max_iter"], work["bits"],
                                  work["image_series_order"], opts)
    w, h, cap = work["width"], work["height"], work["max_iter"]
    fp64_field, planes = ref.render(w, h, work["zoom"], cap, work["series_order"], work["series_block"],
                                    planes=True, options=opts)
    oi = planes["orbit_iteration"].reshape(-1)
    f64 = fp64_field.reshape(-1)
    reference = np.where(np.isnan(f64), -1, np.where(oi >= cap, cap, oi + 1))
    field = np.empty(w * h, dtype=np.float32)
    counts = np.empty(w * h, dtype=np.int32)
    status = lib.fractal_hip_render_reference_floatexp(
        field.ctypes.data, counts.ctypes.data, w, h, work["zoom"].encode(), ref.handle, cap, work["series_order"],
        work["series_block"], ctypes.byref(opts), 0)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    ref.close()
    agree = (counts == reference) & ~np.isnan(field) & ~np.isnan(f64)
    return {"workload": work["name"],
            "pixels_differing_from_fp64_pct": round(float(np.mean(counts != reference)) * 100.0, 3),
            "max_smooth_difference_where_counts_agree":
                float(np.max(np.abs(field[agree].astype(np.float64) - f64[agree]))) if agree.any() else 0.0}


def floatexp_precision_fields(lib, repeats=2):
    """The floatexp path on the 2560x1440 near-nucleus precision fields
    (Mandelbrot map and burning ship, zooms 1e30 and 1e100, cap 100000,
    glitch rebase on): the fraction of pixels whose escape count differs
    from the FP64 counts of the same handle (for the Mandelbrot map also
    over the pixels a strict FP64 render resolves without a glitch), and
    kernel time against the FP64 path."""
    out = []
    for formula in (0, 2):
        for exponent in (30, 100):
            f = fav_workloads.near_boundary_field(formula, 1, exponent)
            w, h, z, n = f["width"], f["height"], f["zoom"], f["max_iter"]
            series = (f["series_order"], f["series_block"])
            opts = fav_hip.NativeRenderOptions(**f["options"])
            ref = fav_hip.NativeReference(f["x"], f["y"], z, n, f["bits"], f["image_series_order"], opts,
                                          formula=formula, julia=tuple(f["julia"]))
            try:
                fp64_field, fp64 = ref.render_counts(w, h, z, n, *series, options=opts)
                fp64_ms = []
                for _ in range(repeats):
                    if formula == 0:
                        ref.render(w, h, z, n, *series, options=opts)
                    else:
                        ref.render_scheduled(w, h, z, n, series_order=series[0], series_block=series[1], options=opts)
                    fp64_ms.append(fav_hip.last_timing()["kernel_ms"])
                fx_ms = []
                for _ in range(repeats):
                    field, counts = ref.render_floatexp(w, h, z, n, *series, options=opts)
                    fx_ms.append(fav_hip.last_timing()["kernel_ms"])
                entry = {"workload": f["name"], "fp64_kernel_ms": round(min(fp64_ms), 3),
                         "floatexp_kernel_ms": round(min(fx_ms), 3),
                         "pixels_differing_from_fp64_pct": round(float(np.mean(counts != fp64)) * 100.0, 4)}
                if formula == 0:
                    strict_opts = fav_hip.NativeRenderOptions(**dict(f["options"], allow_recovery=0))
                    strict = ref.render(w, h, z, n, *series, options=strict_opts).reshape(h, w)
                    resolved = ~np.isnan(strict)
                    entry["resolved_without_glitch_pct"] = round(float(resolved.mean()) * 100.0, 2)
                    entry["resolved_pixels_differing_pct"] = (
                        round(float((counts != fp64)[resolved].mean()) * 100.0, 4) if resolved.any() else None)
                out.append(entry)
            finally:
                ref.close()
    return out


def combined_leg(lib, hip, model, steps=3):
    """The three additions together on the 3840x2160 zoom-1e100 cap-100000
    Mandelbrot field (the deep point's 100-digit extension): a batch of
    that field rendered by the floatexp pixel in the cost model's order,
    against the plain FP64 kernel (no planes) on the same field, and the
    fraction of pixels whose count differs from FP64.  The same for the
    48-field atlas segment: floatexp batch in learned order against the
    FP64 batch."""
    work = dict(fav_workloads.field_1e100_1080p(), width=3840, height=2160)
    w, h, z, n = work["width"], work["height"], work["zoom"], work["max_iter"]
    opts = fav_hip.NativeRenderOptions(**work["options"])
    field = dict(x=work["x"], y=work["y"], zoom=z, zoom_log10=100.0, width=w, height=h, max_iter=n, bits=work["bits"])
    batch = fav_batch.Batch([field], options=opts, library=lib, hip=hip)
    try:
        ref = batch.references[0]
        fp64_ms = []
        for _ in range(steps + 1):
            status = lib.fractal_hip_render_reference_device(batch.outputs[0], w, h, z.encode(), ref.handle, n,
                                                             work["series_order"], work["series_block"],
                                                             ctypes.byref(opts), None)
            if status != 0:
                raise RuntimeError(lib.fractal_last_error().decode())
            fp64_ms.append(fav_hip.last_timing()["kernel_ms"])
        fp64_field = batch.field(0).copy()
        costs, times = fav_batch.learned_tile_costs(batch, model) if model is not None else (None, {"preview_ms": 0})
        result = {"workload": "field_3840x2160_zoom1e100_cap100000", "plain_fp64_kernel_ms": round(min(fp64_ms[1:]), 3)}
        for name, floatexp, order in (("batched_fp64_learned", False, 2), ("batched_floatexp_plain_order", True, 0),
                                      ("batched_floatexp_learned", True, 2)):
            batch.set_precision(floatexp)
            batch.set_order(order, costs if order == 2 else None)
            batch.render(5)
            ms = min(batch.render(5) for _ in range(steps))
            result[name + "_ms"] = round(ms, 3)
        fx_field = batch.field(0)
        batch.set_precision(False)
        result["preview_and_model_ms"] = round(times["preview_ms"], 3)
        combined = result["batched_floatexp_learned_ms"] + times["preview_ms"]
        result["combined_vs_plain_fp64"] = round(result["plain_fp64_kernel_ms"] / combined, 3)
        result["pixels_differing_from_fp64_pct"] = round(
            float(np.mean(np.floor(fx_field) != np.floor(fp64_field))) * 100.0, 4)
    finally:
        batch.close()
    fields = fav_batch.segment_fields()
    seg = fav_batch.Batch(fields, library=lib, hip=hip, threads=16)
    try:
        seg.set_order(0)
        seg.render(5)
        fp64_ms = min(seg.render(5) for _ in range(2))
        fp64 = [seg.field(i).copy() for i in range(len(fields))]
        seg_costs, seg_times = fav_batch.learned_tile_costs(seg, model) if model is not None else (None, {})
        seg.set_precision(True)
        seg.set_order(2 if seg_costs is not None else 0, seg_costs)
        seg.render(5)
        fx_ms = min(seg.render(5) for _ in range(2))
        moved = sum(int(np.sum(np.floor(seg.field(i)) != np.floor(fp64[i]))) for i in range(len(fields)))
        result["segment"] = {"workload": "atlas segment: 48 fields, zoom 1..1e100",
                             "batched_fp64_ms": round(fp64_ms, 3), "batched_floatexp_learned_ms": round(fx_ms, 3),
                             "preview_and_model_ms": round(seg_times.get("preview_ms", 0.0), 3),
                             "ratio": round(fp64_ms / (fx_ms + seg_times.get("preview_ms", 0.0)), 3),
                             "pixels_whose_integer_part_differs_pct": round(100.0 * moved / seg.pixels(), 4)}
    finally:
        seg.close()
    return result


def bla_epsilon_leg(lib, hip, work, opts, d_field, planes, headline_ms, steps):
    """The headline field with the BLA validity radius eps = 2^-53 (a tuned
    reference, fractal_hip_create_reference_tuned) instead of the
    reference's 2^-38: field kernel time with planes, the BLA blocks per
    pixel, and whether every pixel's count equals the 2^-38 render's."""
    w, h, n = work["width"], work["height"], work["max_iter"]
    lib.fractal_hip_create_reference_tuned.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(fav_hip.NativeRenderOptions), ctypes.c_int]
    lib.fractal_hip_create_reference_tuned.restype = ctypes.c_void_p
    handle = lib.fractal_hip_create_reference_tuned(work["x"].encode(), work["y"].encode(), work["zoom"].encode(),
                                                     n, work["bits"], work["image_series_order"], 0, b"0", b"0",
                                                     ctypes.byref(opts), -53)
    if not handle:
        raise RuntimeError(lib.fractal_last_error().decode())
    try:
        def render():
            status = lib.fractal_hip_render_reference_device(d_field, w, h, work["zoom"].encode(),
                                                             ctypes.c_void_p(handle), n, work["series_order"],
                                                             work["series_block"], ctypes.byref(opts),
                                                             ctypes.byref(planes))
            if status != 0:
                raise RuntimeError(lib.fractal_last_error().decode())
            return fav_hip.last_timing()["kernel_ms"]

        lib.fractal_set_stats_enabled(1)
        render()
        stats = np.zeros(39, dtype=np.uint64)
        lib.fractal_get_last_stats_ex(stats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 39)
        lib.fractal_set_stats_enabled(0)
        render()
        ms = float(np.mean([render() for _ in range(max(1, steps))]))
        oi = np.empty(w * h, dtype=np.int64)
        hip.to_host(oi, ctypes.c_void_p(ctypes.cast(planes.orbit_iteration, ctypes.c_void_p).value))
    finally:
        lib.fractal_destroy_reference(ctypes.c_void_p(handle))
    return {"workload": work["name"], "bla_epsilon": "2^-53", "mean_kernel_ms": round(ms, 3),
            "headline_mean_kernel_ms_at_2^-38": round(headline_ms, 3),
            "value": round(float(w * h) * n / (ms * 1e-3) / 1e9, 3), "unit": "Gpixel-iter/s (kernel)",
            "linear_blocks_per_pixel": round(int(stats[3]) / (w * h), 1),
            "all_counts_at_cap": bool(np.all(oi >= n))}


def floatexp_leg(lib, hip, ref, work, opts, d_field, fp64_field, fp64_orbit_iteration, args):
    """The FP32 floatexp path on the same field: throughput over the same
    step count, and the fraction of pixels whose escape count differs from
    the FP64 kernel's (plus the largest smooth-value difference where the
    counts agree)."""
    w, h = work["width"], work["height"]
    n = w * h
    d_counts = hip.malloc(n * 4)

    def step():
        status = lib.fractal_hip_render_reference_floatexp(
            d_field, d_counts, w, h, work["zoom"].encode(), ref.handle, work["max_iter"], work["series_order"],
            work["series_block"], ctypes.byref(opts), 1)
        if status != 0:
            raise RuntimeError(lib.fractal_last_error().decode())
        return fav_hip.last_timing()["kernel_ms"]

    for _ in range(max(1, args.warmup)):
        step()
    hip.synchronize()
    started = time.perf_counter()
    kernel_ms = [step() for _ in range(args.steps)]
    hip.synchronize()
    elapsed = (time.perf_counter() - started) / args.steps
    field = np.empty(n, dtype=np.float32)
    counts = np.empty(n, dtype=np.int32)
    hip.to_host(field, d_field)
    hip.to_host(counts, d_counts)
    hip.free(d_counts)
    cap = work["max_iter"]
    reference = np.where(np.isnan(fp64_field), -1,
                         np.where(fp64_orbit_iteration >= cap, cap, fp64_orbit_iteration + 1))
    agree = (counts == reference) & ~np.isnan(field) & ~np.isnan(fp64_field)
    max_err = float(np.max(np.abs(field[agree].astype(np.float64) - fp64_field[agree]))) if agree.any() else 0.0
    pixel_iterations = int(np.where(counts < 0, cap, counts).astype(np.int64).sum())
    # the floatexp path writes no planes: time the FP64 path without planes too
    def fp64_step():
        status = lib.fractal_hip_render_reference_device(
            d_field, w, h, work["zoom"].encode(), ref.handle, work["max_iter"], work["series_order"],
            work["series_block"], ctypes.byref(opts), None)
        if status != 0:
            raise RuntimeError(lib.fractal_last_error().decode())

    fp64_step()
    hip.synchronize()
    started = time.perf_counter()
    for _ in range(args.steps):
        fp64_step()
    hip.synchronize()
    fp64_elapsed = (time.perf_counter() - started) / args.steps
    fp64_pixel_iterations = int(np.where(reference < 0, cap, reference).astype(np.int64).sum())
    boundary = floatexp_boundary_mismatch(lib)
    return {"value": round(pixel_iterations / elapsed / 1e9, 3), "unit": "Gpixel-iter/s",
            "planes": False,
            "boundary_field": boundary,
            "precision_fields": floatexp_precision_fields(lib),
            "fp64_without_planes_value": round(fp64_pixel_iterations / fp64_elapsed / 1e9, 3),
            "ms_per_step": round(elapsed * 1e3, 3), "mean_kernel_ms": round(float(np.mean(kernel_ms)), 3),
            "pixels_differing_from_fp64_pct": round(float(np.mean(counts != reference)) * 100.0, 4),
            "max_smooth_difference": max_err, "dtype": "f32 mantissa + i32 exponent deltas"}


def main():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=30)
    parser.add_argument("--warmup", type=int, default=10)
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-batch", action="store_true", help="skip the batched atlas-segment leg")
    parser.add_argument("--no-learned", action="store_true", help="skip the learned tile-scheduler leg")
    parser.add_argument("--learned-only", action="store_true",
                        help="run only the learned tile-scheduler leg and print its JSON")
    parser.add_argument("--batch-only", action="store_true",
                        help="run only the batched atlas-segment leg and print its JSON")
    parser.add_argument("--combined-only", action="store_true",
                        help="run only the floatexp precision fields and the combined leg (with a trained "
                             "cost model) and print their JSON")
    parser.add_argument("--boundary-only", action="store_true",
                        help="run only the boundary-rich Mandelbrot fields and print their JSON")
    parser.add_argument("--copy-calibration-mib", type=int, default=0,
                        help="only run N device-to-device copies of this many MiB (HBM counter calibration)")
    args = parser.parse_args()
    if args.copy_calibration_mib:
        return copy_calibration(args.copy_calibration_mib, max(1, args.steps))
    if args.gpus != 1:
        print("bench.py measures one GPU: this build renders on a single device (fields and tiles "
              "shard across GPUs without exchange, but no multi-GPU path is built); run with --gpus 1",
              file=sys.stderr)
        return 2

    work = fav_workloads.field_1e100_1080p()
    lib = fav_hip.load_library()
    if not lib.fractal_backend_capabilities() & 8:
        raise SystemExit(f"no gfx950 device: {lib.fractal_last_error()!r}")
    hip = fav_hip.HipRuntime()
    if args.learned_only:
        print(json.dumps({"learned_schedule": learned_leg(lib)[0]}))
        return 0
    if args.batch_only:
        print(json.dumps({"batched": batch_leg(lib, hip, 2)}))
        return 0
    if args.boundary_only:
        lut = fav_palette.kfp_palette_lut(fav_palette.seeded_stops(12, 2024), 1024)
        print(json.dumps({"boundary_fields": boundary_leg(lib, hip, fav_hip.NativeKfpOptions.kalles_default(),
                                                          lut, args.steps)}))
        return 0
    if args.combined_only:
        model, _ = fav_sched.train_model(fav_sched.mandelbrot_training_fields() + fav_sched.training_fields(),
                                         library=lib)
        print(json.dumps({"precision_fields": floatexp_precision_fields(lib),
                          "combined": combined_leg(lib, hip, model)}))
        return 0
    opts = fav_hip.NativeRenderOptions(**work["options"])
    t0 = time.perf_counter()
    ref = fav_hip.NativeReference(work["x"], work["y"], work["zoom"], work["max_iter"], work["bits"],
                                  work["image_series_order"], opts)
    build_s = time.perf_counter() - t0
    ref.upload()
    w, h = work["width"], work["height"]
    n = w * h
    d_field = hip.malloc(n * 4)
    d_planes_block = hip.malloc(n * 48)
    planes = fav_hip.NativeRenderPlanes()
    planes.struct_size = ctypes.sizeof(fav_hip.NativeRenderPlanes)
    planes.version = 1
    base = d_planes_block.value
    for i, name in enumerate(("orbit_iteration", "phase", "de_x", "de_y", "test1", "test2")):
        setattr(planes, name, ctypes.cast(ctypes.c_void_p(base + i * n * 8), planes._fields_[2 + i][1]))

    d_rgb = hip.malloc(n * 3)
    kfp = fav_hip.NativeKfpOptions.kalles_default()
    lut = fav_palette.kfp_palette_lut(fav_palette.seeded_stops(12, 2024), 1024)
    lut_p = lut.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))

    def step():
        status = lib.fractal_hip_render_reference_device(
            d_field, w, h, work["zoom"].encode(), ref.handle, work["max_iter"], work["series_order"],
            work["series_block"], ctypes.byref(opts), ctypes.byref(planes) if work["planes"] else None)
        if status != 0:
            raise RuntimeError(lib.fractal_last_error().decode())
        timing = fav_hip.last_timing()
        status = lib.fractal_hip_colourise_kfp_device(d_field, d_rgb, w, h, work["max_iter"], ctypes.byref(kfp),
                                                      lut_p, lut.shape[0])
        if status != 0:
            raise RuntimeError(lib.fractal_last_error().decode())
        return timing["kernel_ms"], timing["plain_ms"], fav_hip.last_timing()["kernel_ms"]

    # event counters for the roofline numerator (deterministic, untimed)
    lib.fractal_set_stats_enabled(1)
    step()
    stats = np.zeros(39, dtype=np.uint64)
    lib.fractal_get_last_stats_ex(stats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 39)
    lib.fractal_set_stats_enabled(0)
    field = np.empty(n, dtype=np.float32)
    orbit_iteration = np.empty(n, dtype=np.int64)
    hip.to_host(field, d_field)
    hip.to_host(orbit_iteration, ctypes.c_void_p(base))
    pixel_iterations = escape_count_sum(orbit_iteration, field, work["max_iter"])
    flops = fav_workloads.flops_from_stats(stats, work["planes"])

    for _ in range(args.warmup):
        step()
    hip.synchronize()
    kernel_ms, plain_ms, colour_ms = [], [], []
    started = time.perf_counter()
    for _ in range(args.steps):
        field_ms, pass_ms, kfp_ms = step()
        kernel_ms.append(field_ms)
        plain_ms.append(pass_ms)
        colour_ms.append(kfp_ms)
    hip.synchronize()
    elapsed = time.perf_counter() - started
    ms_per_step = elapsed / args.steps * 1e3
    value = pixel_iterations / (elapsed / args.steps) / 1e9
    mean_kernel_ms = float(np.mean(kernel_ms))
    # the dominant kernel's own launch time (HIP events around the plain pass
    # on its stream); every block of this field but the reference pixel's
    # 1,566 runs in it, so the field's FLOP count is its count to 1e-5
    mean_plain_ms = float(np.mean(plain_ms))
    achieved = flops / (mean_plain_ms * 1e-3) / 1e12
    rgb = np.empty(n * 3, dtype=np.uint8)
    hip.to_host(rgb, d_rgb)
    bla_epsilon = bla_epsilon_leg(lib, hip, work, opts, d_field, planes, mean_kernel_ms, args.steps)
    floatexp = floatexp_leg(lib, hip, ref, work, opts, d_field, field, orbit_iteration, args)
    boundary = boundary_leg(lib, hip, kfp, lut, args.steps)
    learned, model = (None, None) if args.no_learned else learned_leg(lib)
    batched = None if args.no_batch else batch_leg(lib, hip, 2, model)
    combined = None if args.no_batch else combined_leg(lib, hip, model)
    atlas = atlas_leg(lib)
    hip.free(d_field)
    hip.free(d_planes_block)
    hip.free(d_rgb)

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    baseline, reference_build_s = (None, None) if args.no_cpu_baseline else cpu_baseline(work, threads)
    # host setup of the same reference (MPFR orbit, image series, BLA tables):
    # this library against the reference's renderer.cpp on the same host
    host_setup = {"reference_build_s": round(build_s, 3),
                  "reference_build_s_renderer_cpp": None if reference_build_s is None else round(reference_build_s, 3),
                  "speedup": None if not reference_build_s else round(reference_build_s / build_s, 2),
                  "threads": threads}
    result = {
        "metric": "Gpixel-iterations/s at 1e100 zoom (FP64 vs floatexp/batched), % pixels ≠ FP64",
        "value": round(value, 3),
        "unit": "Gpixel-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded 100-digit extension of the reference's bundled deep point",
        "config": {"workload": work["name"], "width": w, "height": h, "zoom": work["zoom"],
                   "max_iter": work["max_iter"], "precision_bits": work["bits"],
                   "path": "plain FP64 HIP kernel", "planes": work["planes"],
                   "colouring": "KFP method 7, differences 3, slopes 45 deg, 12 seeded stops",
                   "pixel_iterations": pixel_iterations, "mean_kernel_ms": round(mean_kernel_ms, 3),
                   "mean_plain_kernel_ms": round(mean_plain_ms, 3),
                   "mean_colour_ms": round(float(np.mean(colour_ms)), 4),
                   "rgb_checksum": int(rgb.astype(np.uint64).sum()),
                   "reference_build_s": round(build_s, 3),
                   "linear_blocks": int(stats[3]), "cubic_blocks": int(stats[4]),
                   "exact_steps": int(stats[5]), "tail_steps": int(stats[11]),
                   "nan_pixels": int(np.isnan(field).sum())},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": MI355X_FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / MI355X_FP64_VECTOR_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic_bytes()},
        "cpu_baseline": baseline,
        "bla_epsilon": bla_epsilon,
        "floatexp": floatexp,
        "boundary_fields": boundary,
        "batched": batched,
        "learned_schedule": learned,
        "combined": combined,
        "atlas_colour": atlas,
        "host_setup": host_setup,
    }
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
