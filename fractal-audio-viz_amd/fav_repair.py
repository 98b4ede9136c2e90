"""Glitch repair over the point ABI: NaN pixels of a strict deep field are
re-rendered around secondary references centred on the failed regions
(visualizer.py:8215-9252, `_atlas_glitch_reference_field`).

A strict render marks a pixel whose perturbation lost its low bits as NaN
(renderer.cpp:12931-12988).  This module groups those pixels into
8-connected regions, creates an MPFR reference at each region's probe pixel
and renders only the region's box through `fractal_render_points`
(renderer.cpp:20019-20186) on the HIP device, with offsets measured from the
probe pixel (visualizer.py:6674-6790).  Regions that stay unresolved are
retried from other probe pixels, split into quadrants, and finally rendered
once more with BLA off, as the reference's queue does.

Kept from the reference: the region search and its order, the probe and the
local-view geometry in Decimal arithmetic, the point offsets, the split of
large boxes, the largest-first stack, the in-cell and edge alternate probes
of small boxes, the quadrant split of a box that stays whole, the final
BLA-off retry and the strict failure.  Not kept (they are optimisations or
draft-mode fallbacks of the Python pipeline, not parts of the device path):
the retargeted root clones tried before the regions
(visualizer.py:8652-8783), the exact MPFR pixel retry
(visualizer.py:8472-8585), the up to 64 halo probes outside a small box
(visualizer.py:8960-9050), per-call time budgets, plane merging, and the
parent-tile fallback of draft renders (spatial recovery is kept for
`allow_recovery`).
"""
from __future__ import annotations

import math
from collections import deque
from decimal import Decimal, localcontext
from typing import Callable, Optional

import numpy as np

import fav_hip

MAX_POINT_CELL_PIXELS = 8192     # ATLAS_LOCAL_REFERENCE_MAX_POINT_CELL_PIXELS, visualizer.py:628
MAX_DIVISIONS = 32               # ATLAS_LOCAL_REFERENCE_MAX_DIVISIONS, visualizer.py:611
MINIMUM_CELL = 8                 # visualizer.py:8797
IMAGE_SERIES_ORDER = 32          # _create_native_reference default, visualizer.py:6257


def fractional_decimal_places(value: str) -> int:
    """Decimal place of the last supplied digit (visualizer.py:4692-4708)."""
    try:
        decimal_value = Decimal(str(value).strip())
    except Exception:
        return 0
    if not decimal_value.is_finite():
        return 0
    return max(0, -int(decimal_value.as_tuple().exponent))


def decimal_precision(x_center: str, y_center: str, log10_zoom: float) -> int:
    """Working Decimal precision of a view (visualizer.py:5011-5022)."""
    return max(50, 32 + int(math.ceil(max(0.0, log10_zoom))),
               fractional_decimal_places(x_center), fractional_decimal_places(y_center))


def native_precision_bits(x_center: str, y_center: str, log10_zoom: float) -> int:
    """MPFR precision of a reference (visualizer.py:5025-5036)."""
    return max(256, int(decimal_precision(x_center, y_center, log10_zoom) * math.log2(10.0)) + 32)


def zoom_text(log10_zoom: float) -> str:
    """10**log10_zoom as decimal text (visualizer.py:5118-5124)."""
    exponent = math.floor(log10_zoom)
    mantissa = 10.0 ** (log10_zoom - exponent)
    return f"{mantissa:.17g}e{exponent:+d}"


def decimal_to_scaled(value: Decimal) -> tuple[float, int]:
    """(mantissa in [0.5, 1), binary exponent) of a Decimal (visualizer.py:6642-6671)."""
    if not value.is_finite():
        raise ValueError("scaled Decimal value must be finite")
    if value.is_zero():
        return 0.0, 0
    sign = -1.0 if value.is_signed() else 1.0
    magnitude = abs(value)
    decimal_exponent = magnitude.adjusted()
    decimal_mantissa = float(magnitude.scaleb(-decimal_exponent))
    log2_value = math.log2(decimal_mantissa) + float(decimal_exponent) * math.log2(10.0)
    binary_exponent = math.floor(log2_value)
    fractional = float(decimal_exponent) * math.log2(10.0) + math.log2(decimal_mantissa) - binary_exponent
    mantissa, shift = math.frexp(2.0 ** fractional)
    return sign * float(mantissa), int(binary_exponent + shift)


def unresolved_regions(mask) -> list[tuple[int, int, int, int, int, int, int]]:
    """8-connected regions of a mask as (x0, x1, y0, y1, count, probe_x,
    probe_y), largest first (visualizer.py:7966-8023).  The box carries a
    one-pixel halo clipped to the image; the probe is the rounded mean
    position; equal counts keep the raster order of their first pixel."""
    unresolved = np.asarray(mask, dtype=bool)
    if unresolved.ndim != 2:
        raise ValueError("unresolved mask must be two-dimensional")
    height, width = unresolved.shape
    visited = np.zeros_like(unresolved, dtype=bool)
    regions = []
    for seed_y, seed_x in zip(*np.nonzero(unresolved)):
        if visited[seed_y, seed_x]:
            continue
        stack = deque([(int(seed_y), int(seed_x))])
        visited[seed_y, seed_x] = True
        x0 = x1 = int(seed_x)
        y0 = y1 = int(seed_y)
        count = sum_x = sum_y = 0
        while stack:
            y, x = stack.pop()
            count += 1
            sum_x += x
            sum_y += y
            x0, x1, y0, y1 = min(x0, x), max(x1, x), min(y0, y), max(y1, y)
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    ny, nx = y + dy, x + dx
                    if (dy or dx) and 0 <= ny < height and 0 <= nx < width and unresolved[ny, nx] \
                            and not visited[ny, nx]:
                        visited[ny, nx] = True
                        stack.append((ny, nx))
        regions.append((max(0, x0 - 1), min(width, x1 + 2), max(0, y0 - 1), min(height, y1 + 2), count,
                        int(round(sum_x / max(count, 1))), int(round(sum_y / max(count, 1)))))
    regions.sort(key=lambda region: region[4], reverse=True)
    return regions


def _view(log10_zoom: float, width: int, height: int, coordinate_mode: int):
    """(view_width, view_height) as Decimals in the current context."""
    exponent = math.floor(log10_zoom)
    zoom_mantissa = Decimal(str(10.0 ** (log10_zoom - exponent)))
    inverse_zoom = Decimal(1).scaleb(-exponent) / zoom_mantissa
    view_height = (Decimal("4.0") if coordinate_mode == 1 else Decimal("2.8")) * inverse_zoom
    return view_height * Decimal(width) / Decimal(height), view_height


def local_reference_cell(width: int, height: int, log10_zoom: float, x_center: str, y_center: str,
                         cell: tuple[int, int, int, int], probe: Optional[tuple[int, int]] = None,
                         coordinate_mode: int = 0) -> tuple[int, int, int, int, str, str, float]:
    """Centre (decimal text) and zoom of a secondary reference for a cell
    (visualizer.py:7906-7963): the probe pixel's centre, or the cell's
    centre without a probe, and the zoom that fits the cell to the frame."""
    x0, x1, y0, y1 = cell
    if not (0 <= x0 < x1 <= width and 0 <= y0 < y1 <= height):
        raise ValueError("invalid local-reference cell")
    if coordinate_mode not in (0, 1):
        raise ValueError("coordinate mode must be project (0) or Kalles (1)")
    with localcontext() as context:
        context.prec = max(96, decimal_precision(x_center, y_center, log10_zoom) + 48)
        view_width, view_height = _view(log10_zoom, width, height, coordinate_mode)
        origin_x = Decimal(width if coordinate_mode else width - 1)
        origin_y = Decimal(height if coordinate_mode else height - 1)
        if probe is None:
            x_numerator = Decimal(x0 + x1 - 1) - origin_x
            y_numerator = origin_y - Decimal(y0 + y1 - 1)
        else:
            probe_x, probe_y = probe
            if not (0 <= probe_x < width and 0 <= probe_y < height):
                raise ValueError("glitch probe is outside the render")
            x_numerator = Decimal(2 * probe_x) - origin_x
            y_numerator = origin_y - Decimal(2 * probe_y)
        local_x = Decimal(x_center) + view_width * (x_numerator / (Decimal(2) * Decimal(width)))
        local_y = Decimal(y_center) + view_height * (y_numerator / (Decimal(2) * Decimal(height)))
    scale = max(width / float(x1 - x0), height / float(y1 - y0))
    return x0, x1, y0, y1, str(local_x), str(local_y), float(log10_zoom) + math.log10(scale)


def point_offsets(width: int, height: int, log10_zoom: float, cell: tuple[int, int, int, int],
                  probe: tuple[int, int], x_center: str, y_center: str, reference_x: str, reference_y: str,
                  coordinate_mode: int = 0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """dc of every pixel of a cell from the probe pixel, as mantissa pairs
    with a shared exponent (visualizer.py:6713-6755), row-major over the cell."""
    x0, x1, y0, y1 = cell
    probe_x, probe_y = probe
    if not (0 <= x0 < x1 <= width and 0 <= y0 < y1 <= height):
        raise ValueError("invalid point-render cell")
    if not (0 <= probe_x < width and 0 <= probe_y < height):
        raise ValueError("point-render probe is outside the render")
    with localcontext() as context:
        context.prec = max(96, decimal_precision(x_center, y_center, log10_zoom) + 48,
                           decimal_precision(reference_x, reference_y, log10_zoom) + 48)
        view_width, view_height = _view(log10_zoom, width, height, coordinate_mode)
        step_x_mantissa, step_x_exponent = decimal_to_scaled(view_width / Decimal(width))
        step_y_mantissa, step_y_exponent = decimal_to_scaled(view_height / Decimal(height))
    w, h = x1 - x0, y1 - y0
    x_delta = np.arange(x0, x1, dtype=np.float64) - float(probe_x)
    y_delta = float(probe_y) - np.arange(y0, y1, dtype=np.float64)
    real_values = np.broadcast_to(step_x_mantissa * x_delta[None, :], (h, w)).copy()
    imag_values = np.broadcast_to(step_y_mantissa * y_delta[:, None], (h, w)).copy()
    real_mantissa, real_shift = np.frexp(real_values)
    imag_mantissa, imag_shift = np.frexp(imag_values)
    real_exponent = real_shift.astype(np.int32) + int(step_x_exponent)
    imag_exponent = imag_shift.astype(np.int32) + int(step_y_exponent)
    common = np.maximum(real_exponent, imag_exponent)
    real_mantissa = np.ldexp(real_mantissa, real_exponent.astype(np.int64) - common.astype(np.int64))
    imag_mantissa = np.ldexp(imag_mantissa, imag_exponent.astype(np.int64) - common.astype(np.int64))
    return (np.ascontiguousarray(real_mantissa, dtype=np.float64).ravel(),
            np.ascontiguousarray(imag_mantissa, dtype=np.float64).ravel(),
            np.ascontiguousarray(common, dtype=np.int32).ravel())


def spatial_recover_field(local_field, fallback_field=None):
    """Fill NaN pixels from a fallback, then by averaging finite
    8-neighbours (visualizer.py:8026-8087); draft renders only."""
    result = np.asarray(local_field, dtype=np.float32).copy()
    missing_before = ~np.isfinite(result)
    if not missing_before.any():
        return result, 0
    if fallback_field is not None:
        fallback = np.asarray(fallback_field, dtype=np.float32)
        use = missing_before & np.isfinite(fallback)
        result[use] = fallback[use]
    shape = result.shape
    for _ in range(max(shape)):
        finite = np.isfinite(result)
        if finite.all():
            break
        padded = np.pad(result, 1, mode="constant", constant_values=np.nan)
        total = np.zeros(shape, dtype=np.float32)
        count = np.zeros(shape, dtype=np.float32)
        for dy in range(3):
            for dx in range(3):
                if dx == 1 and dy == 1:
                    continue
                neighbour = padded[dy:dy + shape[0], dx:dx + shape[1]]
                valid = np.isfinite(neighbour)
                total[valid] += neighbour[valid]
                count[valid] += 1.0
        fillable = ~finite & (count > 0.0)
        if not fillable.any():
            break
        result[fillable] = total[fillable] / count[fillable]
    if not np.isfinite(result).all():
        raise RuntimeError("adaptive local reference has no spatially valid recovery")
    return result, int(missing_before.sum())


def hip_reference_factory(formula: int = 0, julia: tuple[str, str] = ("0", "0"), library=None) -> Callable:
    """Secondary references on the HIP library: (x, y, zoom text, max_iter,
    bits, escape mode, coordinate mode) -> fav_hip.NativeReference."""
    def create(x, y, zoom, max_iter, bits, escape_radius_mode, coordinate_mode):
        options = fav_hip.NativeRenderOptions(escape_radius_mode=escape_radius_mode,
                                              coordinate_mode=coordinate_mode)
        return fav_hip.NativeReference(x, y, zoom, max_iter, bits, IMAGE_SERIES_ORDER, options,
                                       library=library, formula=formula, julia=julia)
    return create


def repair_glitches(field, *, log10_zoom: float, x_center: str, y_center: str, max_iter: int,
                    series_order: int = 3, series_block: int = 256, escape_radius_mode: int = 0,
                    coordinate_mode: int = 0, output_bias: float = 0.0, allow_recovery: bool = False,
                    reference_factory: Optional[Callable] = None, backend: int = 2,
                    diagnostics: Optional[dict] = None) -> np.ndarray:
    """The field with its NaN pixels re-rendered around secondary references
    (the region queue of visualizer.py:8788-9226).  Strict renders
    (allow_recovery False) raise if a pixel stays unresolved; draft renders
    fill the rest spatially."""
    field = np.array(field, dtype=np.float32, copy=True)
    if field.ndim != 2:
        raise ValueError("glitch repair needs a two-dimensional field")
    height, width = field.shape
    factory = reference_factory or hip_reference_factory()
    stats = {"initial_regions": 0, "secondary_references": 0, "refined_regions": 0, "recovered_pixels": 0}
    unresolved = ~np.isfinite(field)
    if not unresolved.any():
        if diagnostics is not None:
            diagnostics.update(stats, final_unresolved_pixels=0)
        return field
    regions = unresolved_regions(unresolved)
    stats["initial_regions"] = len(regions)
    pending = [((x0, x1, y0, y1), 0, count, (px, py)) for x0, x1, y0, y1, count, px, py in regions]
    max_depth = int(math.log2(MAX_DIVISIONS))
    zoom = zoom_text(log10_zoom)
    local_bla_log_zoom = max(12.0, log10_zoom - 1.0)  # a decade of BLA headroom, visualizer.py:8870-8878

    def options(final_retry):
        return fav_hip.NativeRenderOptions(strict=True, allow_recovery=False, disable_bla=final_retry,
                                           strict_cycle=True, backend=backend, output_bias=output_bias,
                                           escape_radius_mode=escape_radius_mode,
                                           coordinate_mode=coordinate_mode)

    def secondary(cell, probe):
        _, _, _, _, local_x, local_y, local_log_zoom = local_reference_cell(
            width, height, log10_zoom, x_center, y_center, cell, probe, coordinate_mode)
        stats["secondary_references"] += 1
        reference = factory(local_x, local_y, zoom_text(local_bla_log_zoom), max_iter,
                            native_precision_bits(local_x, local_y, local_log_zoom), escape_radius_mode,
                            coordinate_mode)
        return reference, local_x, local_y

    def render_cell(reference, cell, probe, local_x, local_y, final_retry=False):
        re, im, e = point_offsets(width, height, log10_zoom, cell, probe, x_center, y_center, local_x, local_y,
                                  coordinate_mode)
        values = reference.render_points(re, im, e, zoom, max_iter, series_order, series_block,
                                         options(final_retry))
        return np.asarray(values, dtype=np.float32).reshape(cell[3] - cell[2], cell[1] - cell[0])

    def recover(cell, local):
        if not allow_recovery:
            x0, x1, y0, y1 = cell
            raise RuntimeError("strict atlas render encountered unresolved pixels in a glitch-driven secondary "
                               f"reference at cell ({x0}:{x1}, {y0}:{y1}) with "
                               f"{int((~np.isfinite(local)).sum())} unresolved pixel(s)")
        x0, x1, y0, y1 = cell
        existing = field[y0:y1, x0:x1]
        fallback = existing if np.isfinite(existing).any() else None
        repaired, count = spatial_recover_field(local, fallback)
        field[y0:y1, x0:x1] = repaired
        stats["recovered_pixels"] += count

    while pending:
        cell, depth, _, probe = pending.pop()
        x0, x1, y0, y1 = cell
        if x1 <= x0 or y1 <= y0:
            continue
        cell_pixels = (x1 - x0) * (y1 - y0)
        if cell_pixels > MAX_POINT_CELL_PIXELS and depth < max_depth:  # visualizer.py:8821-8852
            stats["refined_regions"] += 1
            x_mid, y_mid = x0 + (x1 - x0) // 2, y0 + (y1 - y0) // 2
            for sx0, sx1, sy0, sy1 in reversed(((x0, x_mid, y0, y_mid), (x_mid, x1, y0, y_mid),
                                                (x0, x_mid, y_mid, y1), (x_mid, x1, y_mid, y1))):
                if sx1 > sx0 and sy1 > sy0:
                    pending.append(((sx0, sx1, sy0, sy1), depth + 1, cell_pixels,
                                    (sx0 + (sx1 - sx0) // 2, sy0 + (sy1 - sy0) // 2)))
            continue
        reference, local_x, local_y = secondary(cell, probe)
        try:
            local_field = render_cell(reference, cell, probe, local_x, local_y)
            local_view = field[y0:y1, x0:x1]
            local_finite = np.isfinite(local_field)
            local_view[local_finite] = local_field[local_finite]
            if local_finite.all():
                continue
            if max(x1 - x0, y1 - y0) <= MINIMUM_CELL:  # alternate probes, visualizer.py:8906-8955, 9051-9100
                alternates = []
                for ly, lx in sorted(np.argwhere(np.isfinite(local_view)).tolist(),
                                     key=lambda p: ((x0 + p[1] - probe[0]) ** 2 + (y0 + p[0] - probe[1]) ** 2,
                                                    p[0], p[1])):
                    candidate = (x0 + lx, y0 + ly)
                    if candidate != probe and candidate not in alternates:
                        alternates.append(candidate)
                    if len(alternates) >= 8:
                        break
                if len(alternates) < 8:
                    for candidate in ((x0, y0), (x1 - 1, y0), (x0, y1 - 1), (x1 - 1, y1 - 1),
                                      ((x0 + x1 - 1) // 2, y0), ((x0 + x1 - 1) // 2, y1 - 1),
                                      (x0, (y0 + y1 - 1) // 2), (x1 - 1, (y0 + y1 - 1) // 2)):
                        if candidate != probe and candidate not in alternates:
                            alternates.append(candidate)
                        if len(alternates) >= 8:
                            break
                for candidate in alternates:
                    alternate, alt_x, alt_y = secondary(cell, candidate)
                    try:
                        alternate_field = render_cell(alternate, cell, candidate, alt_x, alt_y)
                    finally:
                        alternate.close()
                    repair = ~np.isfinite(local_view) & np.isfinite(alternate_field)
                    local_view[repair] = alternate_field[repair]
                    if np.isfinite(local_view).all():
                        break
                if np.isfinite(local_view).all():
                    continue
                local_finite = np.isfinite(local_view)
            local_regions = unresolved_regions(~local_finite)
            if depth < max_depth and max(x1 - x0, y1 - y0) > MINIMUM_CELL:  # visualizer.py:9104-9178
                stats["refined_regions"] += 1
                next_pending = []
                for lx0, lx1, ly0, ly1, count, probe_x, probe_y in local_regions:
                    region = (x0 + lx0, x0 + lx1, y0 + ly0, y0 + ly1)
                    if region == cell:  # the whole box failed: split it so each child gets its own probe
                        x_mid, y_mid = x0 + (x1 - x0) // 2, y0 + (y1 - y0) // 2
                        for sx0, sx1, sy0, sy1 in ((x0, x_mid, y0, y_mid), (x_mid, x1, y0, y_mid),
                                                   (x0, x_mid, y_mid, y1), (x_mid, x1, y_mid, y1)):
                            if sx1 <= sx0 or sy1 <= sy0:
                                continue
                            sub_mask = ~local_finite[sy0 - y0:sy1 - y0, sx0 - x0:sx1 - x0]
                            if not sub_mask.any():
                                continue
                            positions = np.argwhere(sub_mask)
                            py, px = positions[len(positions) // 2]
                            next_pending.append(((sx0, sx1, sy0, sy1), depth + 1, int(sub_mask.sum()),
                                                 (sx0 + int(px), sy0 + int(py))))
                    else:
                        next_pending.append((region, depth + 1, count, (x0 + probe_x, y0 + probe_y)))
                pending.extend(reversed(next_pending))
                continue
            # the final retry with BLA off (visualizer.py:9180-9214)
            retry = render_cell(reference, cell, probe, local_x, local_y, final_retry=True)
            retry_finite = np.isfinite(retry)
            local_view[retry_finite] = retry[retry_finite]
            if not np.isfinite(local_view).all():
                recover(cell, local_view)
        finally:
            reference.close()
    if not np.isfinite(field).all():
        raise RuntimeError(f"strict atlas render left {int((~np.isfinite(field)).sum())} unresolved pixels "
                           "after glitch-driven references")
    if diagnostics is not None:
        diagnostics.update(stats, final_unresolved_pixels=0)
    return field
