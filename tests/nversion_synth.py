"""Second, independently written restatement of the depth- and warp-driven
part of synthesize (crates/chiaro-fusion/src/synth.rs:591-725, with
band_touches_module :1688-1721, focus_consistency_weight :1469-1491,
smoothstep :1723-1726) and DenseDepthMap::sample_nearest
(crates/chiaro-fusion/src/depth.rs:152-159); test infrastructure only.

Written from the reference's Rust, not from the C oracle
(oracle/synth_oracle.c): numpy float32 one operation at a time, vectorised
over the canvas pixels, so that agreement of the two pins the oracle's
canvas mapping, band filter, nearest-node rounding, feather and focus
weights bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

from nversion import F
from nversion_resolution import Grid

BAND_ROWS = 64        # png16.rs:29


def _round_away(v: np.ndarray) -> np.ndarray:
    """f32::round on an array: halves away from zero (as float64)."""
    a = np.abs(v.astype(np.float64))
    return np.sign(v.astype(np.float64)) * np.floor(a + 0.5)


def _as_usize(v):
    """`as usize` of f32 values (NaN and negatives to 0, saturating)."""
    v = np.asarray(v, np.float64)
    v = np.where(np.isnan(v) | (v < 0), 0.0, v)
    return np.minimum(v, 2.0 ** 62).astype(np.int64)


def canvas_size(crop, scale):
    """synth.rs:611-620: None for an invalid geometry."""
    x, y, w, h = (F(c) for c in crop)
    scale = F(scale)
    if math.isnan(scale) or scale <= 0 or w < 1 or h < 1:
        return None
    width = max(float(_round_away(np.array([F(w * scale)]))[0]), 1.0)
    height = max(float(_round_away(np.array([F(h * scale)]))[0]), 1.0)
    return int(width), int(height)


class Source:
    """The fields of SynthSource (synth.rs:346-370) this stage reads."""

    def __init__(self, step, points, confidence, mosaic_width, mosaic_height, magnification,
                 confidence_weight, reference=False, focus_distance=None):
        self.grid = Grid(step, points, confidence if confidence is not None
                         else np.zeros(np.asarray(points).shape[:2], np.float32))
        self.has_confidence = confidence is not None
        self.mosaic_width = mosaic_width
        self.mosaic_height = mosaic_height
        self.magnification = F(magnification)
        self.confidence = F(confidence_weight)
        self.reference = reference
        self.focus_distance = focus_distance

    def local_confidence(self, x, y):
        """Warp::confidence (align.rs:112-131) element-wise."""
        if not self.has_confidence:
            return np.zeros(x.shape, np.float32)
        g = self.grid
        ok, c0, r0, c1, r1, tx, ty = g._cell(x, y)
        v = g._blend(g.confidence, c0, r0, c1, r1, tx, ty)
        with np.errstate(invalid="ignore"):
            v = np.where(v < 0, F(0), np.where(v > 1, F(1), v)).astype(np.float32)
        return np.where(ok, v, F(0)).astype(np.float32)


def band_touches(src: Source, start, end, crop, scale) -> bool:
    """band_touches_module (synth.rs:1688-1721)."""
    cx, cy, cw, _ = (F(c) for c in crop)
    scale = F(scale)
    step = src.grid.step
    margin = F(F(step * 2) * np.fmax(src.magnification, F(1)))
    for v in (start, (start + end) // 2, max(end - 1, 0)):
        ry = F(F(cy + F(F(F(v) + F(0.5)) / scale)) - F(0.5))
        samples = max(int(_as_usize(F(cw / F(step)))), 8) + 1
        for i in range(samples + 1):
            rx = F(cx + F(F(F(i) / F(samples)) * F(cw - F(1))))
            q = src.grid.map(rx, ry)
            if q is not None and q[0] >= -margin and q[1] >= -margin and \
                    q[0] <= F(F(src.mosaic_width) + margin) and q[1] <= F(F(src.mosaic_height) + margin):
                return True
    return False


def nearest_depth(depth_map, rx, ry):
    """sample_nearest, then node.depth.map(|d| (d, confidence)) element-wise:
    (has, depth f64, confidence f32)."""
    shape = rx.shape
    if depth_map is None:
        return np.zeros(shape, bool), np.zeros(shape), np.zeros(shape, np.float32)
    depth, conf, step = depth_map
    rows, columns = depth.shape
    with np.errstate(invalid="ignore"):
        ok = ~((rx < 0) | (ry < 0)) & (step != 0)
    s = F(step if step else 1)
    col = _as_usize(_round_away((rx / s).astype(np.float32)))
    row = _as_usize(_round_away((ry / s).astype(np.float32)))
    ok &= (col < columns) & (row < rows)
    col = np.where(ok, col, 0)
    row = np.where(ok, row, 0)
    d = depth[row, col]
    has = ok & ~np.isnan(d)
    return has, np.where(has, d, 0.0), np.where(has, conf[row, col], F(0)).astype(np.float32)


def focus_weight(src: Source, has, depth, conf):
    """focus_consistency_weight (synth.rs:1469-1491) element-wise; powi(4)
    as two squarings."""
    out = np.ones(has.shape, np.float32)
    if src.reference or src.magnification <= F(1.05) or src.focus_distance is None:
        return out
    focus = float(src.focus_distance)
    with np.errstate(all="ignore"):
        live = has & np.isfinite(depth) & (depth > 0)
        if not (math.isfinite(focus) and focus > 0):
            return out
        mismatch = ((focus / np.where(live, depth, 1.0)).astype(np.float32)
                    * F(src.magnification - F(1))).astype(np.float32)
        relative = (mismatch / F(0.75)).astype(np.float32)
        square = (relative * relative).astype(np.float32)
        supported = (F(1) / (F(1) + (square * square).astype(np.float32))).astype(np.float32)
        c = np.where(conf < 0, F(0), np.where(conf > 1, F(1), conf)).astype(np.float32)
        w = (F(1) - (c * (F(1) - supported)).astype(np.float32)).astype(np.float32)
    return np.where(live, w, F(1)).astype(np.float32)


def synthesis_geometry(depth_map, sources, crop, scale, feather_px):
    """(points [S, H, W, 2], feather, local, base, focus [S, H, W])."""
    width, height = canvas_size(crop, scale)
    cx, cy = F(crop[0]), F(crop[1])
    scale = F(scale)
    vs, us = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    rx = ((cx + ((us.astype(np.float32) + F(0.5)) / scale).astype(np.float32))
          .astype(np.float32) - F(0.5)).astype(np.float32)
    ry = ((cy + ((vs.astype(np.float32) + F(0.5)) / scale).astype(np.float32))
          .astype(np.float32) - F(0.5)).astype(np.float32)
    has, depth, dconf = nearest_depth(depth_map, rx, ry)
    bands = -(-height // BAND_ROWS)
    S = len(sources)
    points = np.full((S, height, width, 2), np.nan, np.float32)
    feather = np.zeros((S, height, width), np.float32)
    local = np.zeros_like(feather)
    base = np.zeros_like(feather)
    focus = np.zeros_like(feather)
    fscale = np.fmax(F(feather_px), F(1))
    for s, src in enumerate(sources):
        touch_rows = np.zeros(height, bool)
        for b in range(bands):
            start, end = b * BAND_ROWS, min(b * BAND_ROWS + BAND_ROWS, height)
            touch_rows[start:end] = band_touches(src, start, end, crop, scale)
        ok, qx, qy = src.grid.map_mask(rx, ry)
        wm1, hm1 = F(src.mosaic_width - 1), F(src.mosaic_height - 1)
        with np.errstate(invalid="ignore"):
            ok &= touch_rows[:, None] & (qx >= 0) & (qy >= 0) & (qx <= wm1) & (qy <= hm1)
        border = np.fmin(np.fmin(np.fmin(qx, qy), (wm1 - qx).astype(np.float32)),
                         (hm1 - qy).astype(np.float32))
        t = (border / fscale).astype(np.float32)
        with np.errstate(invalid="ignore"):
            t = np.where(t < 0, F(0), np.where(t > 1, F(1), t)).astype(np.float32)
        fe = ((t * t).astype(np.float32) * (F(3) - (F(2) * t).astype(np.float32))).astype(np.float32)
        lc = src.local_confidence(rx, ry)
        with np.errstate(invalid="ignore"):
            ok &= ~((fe <= 0) | (lc <= 0))
        m = src.magnification
        bw = ((((fe * lc).astype(np.float32) * m).astype(np.float32) * m).astype(np.float32)
              * src.confidence).astype(np.float32)
        fw = focus_weight(src, has, depth, dconf)
        points[s, ..., 0] = np.where(ok, qx, np.nan)
        points[s, ..., 1] = np.where(ok, qy, np.nan)
        feather[s] = np.where(ok, fe, F(0))
        local[s] = np.where(ok, lc, F(0))
        base[s] = np.where(ok, bw, F(0))
        focus[s] = np.where(ok, fw, F(0))
    return points, feather, local, base, focus
