"""Second, independently written restatement of refine_resolution_warp
(crates/chiaro-fusion/src/resolution.rs:232-432) and match_patch
(crates/chiaro-fusion/src/image.rs:531-640), test infrastructure only.

Written from the reference's Rust, not from the C oracle
(oracle/align_oracle.c), so that agreement of the two on seeded planes pins
the oracle's thresholds, medians, completion passes and the Gaussian
render footprint bit for bit. Arrays are numpy float32 and every f32
operation is done one at a time (numpy never fuses a multiply-add); the
window sums are running sums in the reference's row-major order
(np.cumsum is a sequential accumulate); f32::exp and f32::hypot go through
the platform libm as Rust calls them.
"""
from __future__ import annotations

import math

import numpy as np

from nversion import F, expf, total_key32

STEP = 32          # LOCAL_WARP_STEP
PATCH = 12         # LOCAL_MATCH_PATCH
RADIUS = 2         # LOCAL_MATCH_RADIUS
MAGNIFIED = 5      # MAGNIFIED_MATCH_RADIUS
MIN_SCORE = F(0.72)


def _hypot(a, b) -> np.float32:
    return np.hypot(F(a), F(b), dtype=np.float32)


def _clamp(v, lo, hi) -> np.float32:
    """f32::clamp: NaN stays NaN."""
    v = F(v)
    return F(lo) if v < lo else F(hi) if v > hi else v


def _round_away(v) -> int:
    """f32::round: halves away from zero."""
    r = math.floor(abs(float(v)) + 0.5)
    return int(r) if v >= 0 else -int(r)


def _median(values):
    """values.sort_by(f32::total_cmp); values[len / 2]."""
    return sorted(values, key=total_key32)[len(values) // 2]


def sample_mask(plane, x, y):
    """Plane::sample (image.rs:452-466) element-wise: (defined, value)."""
    h, w = plane.shape
    with np.errstate(invalid="ignore"):
        ok = (x >= 0) & (y >= 0) & (x <= F(w - 1)) & (y <= F(h - 1))
    xs = np.where(ok, x, F(0)).astype(np.float32)
    ys = np.where(ok, y, F(0)).astype(np.float32)
    x0 = np.floor(xs).astype(np.int64)
    y0 = np.floor(ys).astype(np.int64)
    x1 = np.minimum(x0 + 1, w - 1)
    y1 = np.minimum(y0 + 1, h - 1)
    tx = (xs - x0.astype(np.float32)).astype(np.float32)
    ty = (ys - y0.astype(np.float32)).astype(np.float32)
    one = F(1)
    top = plane[y0, x0] * (one - tx) + plane[y0, x1] * tx
    bottom = plane[y1, x0] * (one - tx) + plane[y1, x1] * tx
    return ok, (top * (one - ty) + bottom * ty).astype(np.float32)


class Grid:
    """Warp (align.rs:35-144) over numpy arrays: points (rows, columns, 2),
    confidence (rows, columns)."""

    def __init__(self, step, points, confidence):
        self.step = step
        self.points = np.asarray(points, np.float32)
        self.confidence = np.asarray(confidence, np.float32)
        self.rows, self.columns = self.points.shape[:2]

    def _cell(self, x, y):
        fx = (np.asarray(x, np.float32) / F(self.step)).astype(np.float32)
        fy = (np.asarray(y, np.float32) / F(self.step)).astype(np.float32)
        with np.errstate(invalid="ignore"):
            ok = ~((fx < 0) | (fy < 0))

        def index(f, n):
            fl = np.floor(np.where(ok, f, F(0)))
            fl = np.where(np.isnan(fl), 0.0, fl)       # `as usize` saturates
            return np.minimum(np.minimum(fl, 2.0 ** 62), n - 1).astype(np.int64)

        c0, r0 = index(fx, self.columns), index(fy, self.rows)
        c1, r1 = np.minimum(c0 + 1, self.columns - 1), np.minimum(r0 + 1, self.rows - 1)
        tx = (fx - c0.astype(np.float32)).astype(np.float32)
        ty = (fy - r0.astype(np.float32)).astype(np.float32)
        return ok, c0, r0, c1, r1, tx, ty

    def _blend(self, grid, c0, r0, c1, r1, tx, ty):
        one = F(1)
        top = grid[r0, c0] * (one - tx) + grid[r0, c1] * tx
        bottom = grid[r1, c0] * (one - tx) + grid[r1, c1] * tx
        return (top * (one - ty) + bottom * ty).astype(np.float32)

    def map_mask(self, x, y):
        """Warp::map (align.rs:86-111) element-wise: (defined, qx, qy)."""
        ok, c0, r0, c1, r1, tx, ty = self._cell(x, y)
        qx = self._blend(self.points[..., 0], c0, r0, c1, r1, tx, ty)
        qy = self._blend(self.points[..., 1], c0, r0, c1, r1, tx, ty)
        ok = ok & ~np.isnan(qx) & ~np.isnan(qy)
        return ok, qx, qy

    def map(self, x, y):
        ok, qx, qy = self.map_mask(np.array([x], np.float32), np.array([y], np.float32))
        return (qx[0], qy[0]) if ok[0] else None

    def confidence_at(self, x, y) -> np.float32:
        """Warp::confidence (align.rs:112-131)."""
        ok, c0, r0, c1, r1, tx, ty = self._cell(np.array([x], np.float32),
                                                np.array([y], np.float32))
        if not ok[0]:
            return F(0)
        return _clamp(self._blend(self.confidence, c0, r0, c1, r1, tx, ty)[0], 0.0, 1.0)

    def magnification(self, x, y):
        """Warp::magnification (align.rs:135-143)."""
        h = F(self.step)
        a, b, c = self.map(x, y), self.map(F(x) + h, y), self.map(x, F(y) + h)
        if a is None or b is None or c is None:
            return None
        dx = F(np.sqrt(F(F(b[0] - a[0]) * F(b[0] - a[0])) + F(F(b[1] - a[1]) * F(b[1] - a[1])))) / h
        dy = F(np.sqrt(F(F(c[0] - a[0]) * F(c[0] - a[0])) + F(F(c[1] - a[1]) * F(c[1] - a[1])))) / h
        return np.fmax(F(np.sqrt(F(dx * dy))), F(1e-6))


def match_patch(reference, target, x, y, size, radius):
    """match_patch (image.rs:531-640): (shift, score) or None."""
    rh, rw = reference.shape
    th, tw = target.shape
    if (x + size > rw or y + size > rh or x < radius or y < radius
            or x + size + radius > tw or y + size + radius > th):
        return None
    n = F(size * size)
    window = reference[y:y + size, x:x + size].ravel().astype(np.float32)
    total = np.cumsum(np.concatenate([np.array([-0.0], np.float32), window]), dtype=np.float32)[-1]
    mean = F(total / n)
    dev = (window - mean).astype(np.float32)
    sq = np.cumsum(np.concatenate([np.array([-0.0], np.float32), dev * dev]), dtype=np.float32)[-1]
    std = F(np.sqrt(F(sq / n)))
    if std < F(1e-4):
        return None
    span = 2 * radius + 1
    # every shifted target window, row-major within each
    t = np.stack([target[y + dy - radius:y + dy - radius + size,
                         x + dx - radius:x + dx - radius + size].ravel()
                  for dy in range(span) for dx in range(span)]).astype(np.float32)
    s = np.cumsum(t, axis=1, dtype=np.float32)[:, -1]
    s2 = np.cumsum(t * t, axis=1, dtype=np.float32)[:, -1]
    cross = np.cumsum(t * dev[None, :], axis=1, dtype=np.float32)[:, -1]
    m = (s / n).astype(np.float32)
    var = np.fmax((s2 / n).astype(np.float32) - (m * m).astype(np.float32), F(0))
    scores = np.full(span * span, -np.inf, np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        live = ~(var < F(1e-8))
        scores[live] = ((cross[live] / n).astype(np.float32)
                        / (np.sqrt(var[live]).astype(np.float32) * std)).astype(np.float32)
    keys = [total_key32(v) for v in scores]
    best = max(range(len(keys)), key=lambda i: (keys[i], i))   # the last maximum
    score = scores[best]
    if not np.isfinite(score):
        return None
    bx, by = best % span, best // span

    def refine(minus, centre, plus):
        den = F(F(minus - F(F(2) * centre)) + plus)
        if abs(den) < F(1e-9):
            return F(0)
        return _clamp(F(F(F(0.5) * F(minus - plus)) / den), -0.5, 0.5)

    sub_x = refine(scores[by * span + bx - 1], score, scores[by * span + bx + 1]) \
        if 0 < bx < span - 1 else F(0)
    sub_y = refine(scores[(by - 1) * span + bx], score, scores[(by + 1) * span + bx]) \
        if 0 < by < span - 1 else F(0)
    return (F(F(bx - radius) + sub_x), F(F(by - radius) + sub_y)), score


def render_at_bandwidth(target, x, y, magnification):
    """sample_at_reference_bandwidth (resolution.rs:441-465) element-wise:
    (defined, value)."""
    if magnification <= F(1.15):
        return sample_mask(target, x, y)
    sigma = np.fmax(F(F(0.42) * magnification), F(0.55))
    radius = int(_clamp(np.ceil(F(F(2) * sigma)), 1.0, 4.0))
    total = np.zeros(x.shape, np.float32)
    weight = np.zeros(x.shape, np.float32)
    for dy in range(-radius, radius + 1):
        for dx in range(-radius, radius + 1):
            d2 = F(dx * dx + dy * dy)
            kernel = expf(F(F(F(-0.5) * d2) / F(sigma * sigma)))
            ok, v = sample_mask(target, (x + F(dx)).astype(np.float32),
                                (y + F(dy)).astype(np.float32))
            total = np.where(ok, (total + (kernel * v).astype(np.float32)).astype(np.float32),
                             total)
            weight = np.where(ok, (weight + kernel).astype(np.float32), weight)
    ok = weight > F(1e-8)
    with np.errstate(invalid="ignore", divide="ignore"):
        return ok, (total / weight).astype(np.float32)


def _smoothstep(v):
    v = _clamp(v, 0.0, 1.0)
    return F(F(v * v) * F(F(3) - F(F(2) * v)))


def refine_resolution_warp(reference, target, base: Grid, width, height):
    """refine_resolution_warp (resolution.rs:232-432): (points (rows,
    columns, 2), confidence (rows, columns), report dict)."""
    rh, rw = reference.shape
    mag = base.magnification(F(width * 0.5), F(height * 0.5))
    mag = _clamp(F(1) if mag is None else mag, 0.5, 4.0)
    radius = MAGNIFIED if mag > F(1.35) else RADIUS
    ys, xs = np.meshgrid(np.arange(rh), np.arange(rw), indexing="ij")
    rx = (xs.astype(np.float32) * F(2) + F(0.5)).astype(np.float32)
    ry = (ys.astype(np.float32) * F(2) + F(0.5)).astype(np.float32)
    mapped, qx, qy = base.map_mask(rx, ry)
    u = (((qx + F(0.5)) * F(0.5)).astype(np.float32) - F(0.5)).astype(np.float32)
    v = (((qy + F(0.5)) * F(0.5)).astype(np.float32) - F(0.5)).astype(np.float32)
    u = np.where(mapped, u, F(0)).astype(np.float32)
    v = np.where(mapped, v, F(0)).astype(np.float32)
    sampled, values = render_at_bandwidth(target, u, v, mag)
    rendered = np.where(mapped & sampled, values, F(np.nan)).astype(np.float32)

    columns = -(-width // STEP) + 1
    rows = -(-height // STEP) + 1
    measured = {}
    for row in range(rows):
        for column in range(columns):
            nx, ny = F(column * STEP), F(row * STEP)
            if nx >= F(width) or ny >= F(height):
                continue
            px = _round_away(F(F(nx - F(0.5)) * F(0.5)))
            py = _round_away(F(F(ny - F(0.5)) * F(0.5)))
            x, y = px - PATCH // 2, py - PATCH // 2
            if x < radius or y < radius or x + PATCH + radius > rw or y + PATCH + radius > rh:
                continue
            if not np.isfinite(rendered[y - radius:y + PATCH + radius,
                                        x - radius:x + PATCH + radius]).all():
                continue
            found = match_patch(reference, rendered, x, y, PATCH, radius)
            if found is None or not found[1] >= MIN_SCORE:
                continue
            measured[(row, column)] = ((F(found[0][0] * F(2)), F(found[0][1] * F(2))), found[1])

    points = np.empty((rows, columns, 2), np.float32)
    confidence = np.empty((rows, columns), np.float32)
    correction = {}
    corrections = []
    confidence_sum = F(0)

    def neighbours(row, column):
        for r in range(max(row - 1, 0), min(row + 1, rows - 1) + 1):
            for c in range(max(column - 1, 0), min(column + 1, columns - 1) + 1):
                yield r, c

    for row in range(rows):
        for column in range(columns):
            nx, ny = F(column * STEP), F(row * STEP)
            refined = None
            if (row, column) in measured:
                centre, centre_score = measured[(row, column)]
                near = [measured[k][0] for k in neighbours(row, column) if k in measured]
                if len(near) >= 3:
                    med = (_median([s[0] for s in near]), _median([s[1] for s in near]))
                    if not _hypot(F(centre[0] - med[0]), F(centre[1] - med[1])) > F(1.5):
                        q = base.map(F(nx + med[0]), F(ny + med[1]))
                        if q is not None:
                            sc = _smoothstep(F(F(centre_score - MIN_SCORE) / F(0.20)))
                            refined = (q, F(sc * base.confidence_at(nx, ny)), med)
            if refined is not None:
                points[row, column] = refined[0]
                confidence[row, column] = refined[1]
                correction[(row, column)] = refined[2]
                corrections.append(_hypot(*refined[2]))
                confidence_sum = F(confidence_sum + refined[1])
            else:
                q = base.map(nx, ny)
                points[row, column] = (np.nan, np.nan) if q is None else q
                confidence[row, column] = F(0)
    for _ in range(2):
        snapshot = dict(correction)
        conf_snapshot = confidence.copy()
        for row in range(rows):
            for column in range(columns):
                if (row, column) in snapshot:
                    continue
                near = [k for k in neighbours(row, column) if k in snapshot]
                if len(near) < 3:
                    continue
                xs_ = sorted((snapshot[k][0] for k in near), key=total_key32)
                ys_ = sorted((snapshot[k][1] for k in near), key=total_key32)
                csum = F(0)
                for k in near:
                    csum = F(csum + conf_snapshot[k])
                corr = (xs_[len(near) // 2], ys_[len(near) // 2])
                # the reference zips the two independently sorted lists
                if not all(_hypot(F(a - corr[0]), F(b - corr[1])) <= F(1.5)
                           for a, b in zip(xs_, ys_)):
                    continue
                nx, ny = F(column * STEP), F(row * STEP)
                q = base.map(F(nx + corr[0]), F(ny + corr[1]))
                if q is None:
                    continue
                points[row, column] = q
                confidence[row, column] = F(F(F(0.45) * csum) / F(len(near)))
                correction[(row, column)] = corr
    corrections.sort(key=total_key32)
    verified = len(corrections)
    supported = int((confidence > 0).sum())
    nodes = F(max(columns * rows, 1))
    report = {
        "verified_fraction": F(F(verified) / nodes),
        "supported_fraction": F(F(supported) / nodes),
        "median_correction_px": corrections[verified // 2] if verified else F(0),
        "mean_confidence": F(confidence_sum / F(max(verified, 1))),
    }
    return points, confidence, report
