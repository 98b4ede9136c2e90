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
    """(mantissa in [0.5, 1), binary exponent) of a Decimal, formed as the
    reference's point offsets are (visualizer.py:6642-6671): log2 of the
    magnitude in double from the Decimal's own mantissa and exponent, its
    floor as the exponent, 2 to the fractional rest as the mantissa.  The
    double logarithm leaves the mantissa exact only to about 1e-14, and
    the offsets must carry that same rounding to equal the reference's."""
    if not value.is_finite():
        raise ValueError("scaled Decimal value must be finite")
    if value.is_zero():
        return 0.0, 0
    digits_exponent = abs(value).adjusted()                   # |value| = d * 10^k, 1 <= d < 10
    leading = float(abs(value).scaleb(-digits_exponent))
    log_magnitude = math.log2(leading) + digits_exponent * math.log2(10.0)
    exponent = math.floor(log_magnitude)
    rest = digits_exponent * math.log2(10.0) + math.log2(leading) - exponent
    mantissa, carry = math.frexp(2.0 ** rest)
    if not 0.5 <= mantissa < 1.0:
        raise ValueError("scaled Decimal mantissa is not representable")
    return math.copysign(mantissa, -1.0 if value.is_signed() else 1.0), int(exponent + carry)


def unresolved_regions(mask) -> list[tuple[int, int, int, int, int, int, int]]:
    """8-connected components of `mask` as (x0, x1, y0, y1, count, probe_x,
    probe_y), the order and fields of the reference's region search
    (visualizer.py:7966-8023): the bounding box widened by one pixel and
    clipped to the image (x1, y1 exclusive), the pixel count, and the
    component's mean position rounded half to even; largest first, equal
    counts in the raster order of their first pixel.  scipy's labelling
    numbers components in that raster order, so a stable sort by count
    gives the reference's order."""
    from scipy import ndimage
    flags = np.asarray(mask, dtype=bool)
    if flags.ndim != 2:
        raise ValueError("unresolved mask must be two-dimensional")
    labels, components = ndimage.label(flags, structure=np.ones((3, 3), dtype=bool))
    if components == 0:
        return []
    height, width = flags.shape
    ys, xs = np.nonzero(labels)
    ids = labels[ys, xs]
    counts = np.bincount(ids, minlength=components + 1)[1:]
    mean_x = np.rint(np.bincount(ids, weights=xs, minlength=components + 1)[1:] / counts)
    mean_y = np.rint(np.bincount(ids, weights=ys, minlength=components + 1)[1:] / counts)
    out = []
    for k, (rows, cols) in enumerate(ndimage.find_objects(labels)):
        out.append((max(0, cols.start - 1), min(width, cols.stop + 1), max(0, rows.start - 1),
                    min(height, rows.stop + 1), int(counts[k]), int(mean_x[k]), int(mean_y[k])))
    order = np.argsort(-counts, kind="stable")
    return [out[k] for k in order]


class _View:
    """A frame's view in Decimal arithmetic: width x height pixels around
    (x_center, y_center) at 10^log10_zoom, view height 2.8 / zoom (project
    grid, pixel centres at (p - (n-1)/2) / n) or 4 / zoom (Kalles grid,
    (p - n/2) / n) (renderer.cpp:142-158).  Positions are given doubled,
    as integers: 2p for pixel p's centre, x0 + x1 - 1 for the middle of
    the pixel range [x0, x1)."""

    def __init__(self, width: int, height: int, log10_zoom: float, x_center: str, y_center: str,
                 coordinate_mode: int):
        if coordinate_mode not in (0, 1):
            raise ValueError("coordinate mode must be project (0) or Kalles (1)")
        self.width, self.height, self.mode = width, height, coordinate_mode
        self.x_center, self.y_center = Decimal(x_center), Decimal(y_center)
        whole = math.floor(log10_zoom)
        zoom = Decimal(str(10.0 ** (log10_zoom - whole))).scaleb(whole)
        self.span_y = (Decimal("4.0") if coordinate_mode == 1 else Decimal("2.8")) / zoom
        self.span_x = self.span_y * width / height

    def offset(self, doubled: int, size: int) -> Decimal:
        """Axis offset of a doubled position, as a fraction of the span."""
        origin = size if self.mode == 1 else size - 1
        return Decimal(doubled - origin) / Decimal(2 * size)

    def point(self, doubled_x: int, doubled_y: int) -> tuple[Decimal, Decimal]:
        return (self.x_center + self.span_x * self.offset(doubled_x, self.width),
                self.y_center - self.span_y * self.offset(doubled_y, self.height))


def local_reference_cell(width: int, height: int, log10_zoom: float, x_center: str, y_center: str,
                         cell: tuple[int, int, int, int], probe: Optional[tuple[int, int]] = None,
                         coordinate_mode: int = 0) -> tuple[int, int, int, int, str, str, float]:
    """Centre (decimal text) and zoom of a secondary reference for a cell
    (visualizer.py:7906-7963): the probe pixel's centre, or the middle of
    the cell without a probe, and the zoom at which the cell's longer side
    relative to the frame fills the frame."""
    x0, x1, y0, y1 = cell
    if not (0 <= x0 < x1 <= width and 0 <= y0 < y1 <= height):
        raise ValueError("invalid local-reference cell")
    if probe is not None and not (0 <= probe[0] < width and 0 <= probe[1] < height):
        raise ValueError("glitch probe is outside the render")
    with localcontext() as context:
        context.prec = max(96, decimal_precision(x_center, y_center, log10_zoom) + 48)
        view = _View(width, height, log10_zoom, x_center, y_center, coordinate_mode)
        doubled = (x0 + x1 - 1, y0 + y1 - 1) if probe is None else (2 * probe[0], 2 * probe[1])
        local_x, local_y = view.point(*doubled)
    fill = max(width / float(x1 - x0), height / float(y1 - y0))
    return x0, x1, y0, y1, str(local_x), str(local_y), float(log10_zoom) + math.log10(fill)


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
        view = _View(width, height, log10_zoom, x_center, y_center, coordinate_mode)
        step_x_mantissa, step_x_exponent = decimal_to_scaled(view.span_x / Decimal(width))
        step_y_mantissa, step_y_exponent = decimal_to_scaled(view.span_y / Decimal(height))
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
    """Draft-render recovery of a cell that stays unresolved: NaN pixels
    take the fallback's finite value where there is one, then grow inward
    from the finite pixels, each newly reached pixel the mean of its finite
    8-neighbours, ring by ring (the recovery of visualizer.py:8026-8087,
    as an estimate: the neighbour sums here are float64 convolutions, so
    values are not bit-identical to the reference's).  Returns the field and
    the number of pixels filled."""
    from scipy import ndimage
    result = np.asarray(local_field, dtype=np.float32).copy()
    missing = ~np.isfinite(result)
    if not missing.any():
        return result, 0
    if fallback_field is not None:
        fallback = np.asarray(fallback_field, dtype=np.float32)
        take = missing & np.isfinite(fallback)
        result[take] = fallback[take]
    ring = np.ones((3, 3))
    ring[1, 1] = 0.0
    while True:
        known = np.isfinite(result)
        if known.all():
            break
        sums = ndimage.convolve(np.where(known, result, 0.0).astype(np.float64), ring, mode="constant")
        counts = ndimage.convolve(known.astype(np.float64), ring, mode="constant")
        reached = ~known & (counts > 0.0)
        if not reached.any():
            raise RuntimeError("adaptive local reference has no spatially valid recovery")
        result[reached] = (sums[reached] / counts[reached]).astype(np.float32)
    return result, int(missing.sum())


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
