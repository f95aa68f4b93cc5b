"""Second, independently written restatement of depth.rs's decision logic
(test infrastructure only).

The C oracle (oracle/depth_oracle.c) is the checker of the GPU path. The
functions here restate, straight from crates/chiaro-fusion/src/depth.rs
and its helpers and not from the C oracle, the parts that no reference test
or closed form pins:

  select_depths with its confidences          depth.rs:1385-1488
  measure_direct_depths with the wide fallback depth.rs:715-876, 1062-1185
  refine_one_view and the per-target rule      depth.rs:420-547, 1634-1736
  enforce_warp_consensus, suppress_boundaries  depth.rs:582-712

together with the primitives they call (Plane::sample image.rs:452-466,
Warp::map align.rs:86-111, the pinhole camera geometry.rs:148-252 with
math.rs:11-60, map_at_depth depth.rs:1738-1759, warped_patch_zncc
depth.rs:1772-1822, aggregate depth.rs:1761-1770). f32 arithmetic is done
on numpy float32 values one operation at a time (numpy never fuses a
multiply-add), f64 on Python floats, and f32::exp through the platform
libm, as Rust calls it. Agreement of two independent restatements on
seeded scenes is what pins the oracle's threshold logic.

Only synthetic rigs are covered: canonical poses, no distortion, no
mirror flip.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

F = np.float32
_libm = C.CDLL("libm.so.6")
_libm.expf.restype = C.c_float
_libm.expf.argtypes = [C.c_float]

MISSING = None
MINIMUM_REGULARIZED_SCORE = F(0.30)
MAXIMUM_REGULARIZED_BASELINE_LOSS = F(0.04)
NEAR_DEPTH_PRIOR = F(0.10)
SGM_DIRECTIONS = F(8.0)
INFINITY_DEPTH = 1.0e8
BOUNDARY_DELTA = F(2.0)
BOUNDARY_CONTRAST = F(0.30)

UNDEFINED, GLOBAL, REFINED, OCCLUDED, BOUNDARY = range(5)


def expf(x) -> np.float32:
    return F(_libm.expf(float(F(x))))


def total_key32(v) -> int:
    """f32::total_cmp as an integer key."""
    u = int(np.array(v, np.float32).view(np.uint32))
    return (u ^ 0x80000000) if u < 0x80000000 else (~u & 0xFFFFFFFF)


def fmax32(a, b):
    """f32::max: the non-NaN operand when one is NaN."""
    return F(np.fmax(F(a), F(b)))


def clamp32(x, lo, hi):
    x = F(x)
    if x < lo:
        return F(lo)
    if x > hi:
        return F(hi)
    return x


def clamp64(x, lo, hi):
    return lo if x < lo else hi if x > hi else x


def seq_sum(values: np.ndarray) -> np.float32:
    """A f32 running sum from 0.0 in order (`let mut s = 0.0; s += v`)."""
    return np.cumsum(np.concatenate([np.zeros(1, np.float32), values.astype(np.float32)]),
                     dtype=np.float32)[-1]


# ------------------------------------------------------------ primitives

def sample(plane: np.ndarray, x, y):
    """Plane::sample (image.rs:452-466) for one point."""
    ok, v = sample_many(plane, np.array([x], np.float32), np.array([y], np.float32))
    return v[0] if ok else None


def sample_many(plane: np.ndarray, x: np.ndarray, y: np.ndarray):
    """Plane::sample at every point; (False, None) when any point is outside."""
    h, w = plane.shape
    inside = (x >= 0) & (y >= 0) & (x <= F(w - 1)) & (y <= F(h - 1))
    if not inside.all():
        return False, None
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    x1 = np.minimum(x0 + 1, w - 1)
    y1 = np.minimum(y0 + 1, h - 1)
    tx = (x - x0.astype(np.float32)).astype(np.float32)
    ty = (y - y0.astype(np.float32)).astype(np.float32)
    one = F(1.0)
    top = plane[y0, x0] * (one - tx) + plane[y0, x1] * tx
    bottom = plane[y1, x0] * (one - tx) + plane[y1, x1] * tx
    return True, (top * (one - ty) + bottom * ty).astype(np.float32)


def warp_map_many(points: np.ndarray, step: int, x: np.ndarray, y: np.ndarray):
    """Warp::map (align.rs:86-111) at every point; (False, None) when any
    point maps to None."""
    rows, columns = points.shape[:2]
    fx = (x / F(step)).astype(np.float32)
    fy = (y / F(step)).astype(np.float32)
    if (fx < 0).any() or (fy < 0).any():
        return False, None

    def index(f, n):
        fl = np.floor(f)
        fl = np.where(np.isnan(fl), 0.0, fl)       # `as usize` saturates NaN to 0
        return np.minimum(np.minimum(fl, 2.0 ** 62), n - 1).astype(np.int64)

    c0, r0 = index(fx, columns), index(fy, rows)
    c1, r1 = np.minimum(c0 + 1, columns - 1), np.minimum(r0 + 1, rows - 1)
    tx = (fx - c0.astype(np.float32)).astype(np.float32)
    ty = (fy - r0.astype(np.float32)).astype(np.float32)
    one = F(1.0)
    out = []
    for k in range(2):
        a, b = points[r0, c0, k], points[r0, c1, k]
        c, d = points[r1, c0, k], points[r1, c1, k]
        top = a * (one - tx) + b * tx
        bottom = c * (one - tx) + d * tx
        out.append((top * (one - ty) + bottom * ty).astype(np.float32))
    if np.isnan(out[0]).any() or np.isnan(out[1]).any():
        return False, None
    return True, out


def warp_map(points, step, x, y):
    ok, out = warp_map_many(points, step, np.array([F(x)]), np.array([F(y)]))
    return (out[0][0], out[1][0]) if ok else None


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _mul_vec(m, v):
    return [_dot(m[0:3], v), _dot(m[3:6], v), _dot(m[6:9], v)]


def _normalize(a):
    n = math.sqrt(_dot(a, a))
    if n == 0.0:
        return list(a)
    s = 1.0 / n
    return [a[0] * s, a[1] * s, a[2] * s]


class Camera:
    """The fields of a canonical pinhole ResolvedCamera (geometry.rs:31-62)."""

    def __init__(self, c):
        if c.has_distortion or c.flip_around_x != -1 or c.pose_kind != 0:
            raise ValueError("the restatement covers canonical pinhole cameras only")
        self.width, self.height = c.width, c.height
        self.k = list(c.k)
        self.k_inverse = list(c.k_inverse)
        self.rotation_wc = list(c.rotation_wc)
        self.rotation_cw = list(c.rotation_cw)
        self.translation_wc = list(c.translation_wc)
        self.center = list(c.center)
        self.orientation = list(c.orientation_correction)
        if self.orientation != [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]:
            raise ValueError("the restatement covers cameras without orientation correction")

    def pixel_to_ray(self, pixel):
        """geometry.rs:148-177 without distortion or flip."""
        direction = _normalize(_mul_vec(self.k_inverse, [pixel[0], pixel[1], 1.0]))
        world = _mul_vec(self.rotation_cw, direction)
        return self.center, _normalize(_mul_vec(self.orientation, _normalize(world)))

    def project(self, point):
        """geometry.rs:183-243, bounded, identity orientation."""
        r = _mul_vec(self.rotation_wc, point)
        cam = [r[0] + self.translation_wc[0], r[1] + self.translation_wc[1],
               r[2] + self.translation_wc[2]]
        if cam[2] <= 0.0:
            return None
        n = [cam[0] / cam[2], cam[1] / cam[2]]
        k = self.k
        ideal = [k[0] * n[0] + k[1] * n[1] + k[2], k[4] * n[1] + k[5]]
        mx, my = self.width * 0.5, self.height * 0.5
        if ideal[0] < -mx or ideal[1] < -my or ideal[0] > self.width + mx or \
                ideal[1] > self.height + my:
            return None
        return ideal

    def map_from(self, reference: "Camera", pixel, depth):
        """geometry.rs:249-252."""
        origin, direction = reference.pixel_to_ray(pixel)
        d = [direction[0] * depth, direction[1] * depth, direction[2] * depth]
        return self.project([origin[0] + d[0], origin[1] + d[1], origin[2] + d[2]])


class View:
    def __init__(self, module):
        self.plane = np.ascontiguousarray(module.luminance, np.float32)
        self.width, self.height = module.width, module.height
        self.camera = None if module.camera is None else Camera(module.camera)
        self.points = np.ascontiguousarray(module.warp_points, np.float32)
        self.step = module.warp_step
        self.accepted = module.accepted


def map_at_depth(reference: View, target: View, pixel, depth):
    """depth.rs:1738-1759."""
    g = warp_map(target.points, target.step, F(pixel[0]), F(pixel[1]))
    if g is None:
        return None
    corrected = reference.camera.map_from(target.camera, [float(g[0]), float(g[1])],
                                          INFINITY_DEPTH)
    if corrected is None:
        return None
    q = target.camera.map_from(reference.camera, corrected, depth)
    if q is None or not (math.isfinite(q[0]) and math.isfinite(q[1])):
        return None
    return q


def zncc(reference: View, target: View, centre, delta, radius: int):
    """warped_patch_zncc (depth.rs:1772-1822)."""
    c_ref = sample(reference.plane, F((centre[0] - 0.5) * 0.5), F((centre[1] - 0.5) * 0.5))
    if c_ref is None:
        return None
    sigma = fmax32(F(radius) * F(0.75), F(1.0))
    offsets = np.arange(-radius, radius + 1)
    dy, dx = np.meshgrid(offsets, offsets, indexing="ij")
    dx, dy = dx.reshape(-1), dy.reshape(-1)
    px = centre[0] + dx.astype(np.float64) * 2.0
    py = centre[1] + dy.astype(np.float64) * 2.0
    ok, rv = sample_many(reference.plane, ((px - 0.5) * 0.5).astype(np.float32),
                         ((py - 0.5) * 0.5).astype(np.float32))
    if not ok:
        return None
    ok, g = warp_map_many(target.points, target.step, px.astype(np.float32),
                          py.astype(np.float32))
    if not ok:
        return None
    qx = g[0].astype(np.float64) + delta[0]
    qy = g[1].astype(np.float64) + delta[1]
    ok, tv = sample_many(target.plane, ((qx - 0.5) * 0.5).astype(np.float32),
                         ((qy - 0.5) * 0.5).astype(np.float32))
    if not ok:
        return None
    den_w = F(2.0) * sigma * sigma
    weight = np.empty(rv.size, np.float32)
    for t in range(rv.size):
        distance_sq = F(int(dx[t] * dx[t] + dy[t] * dy[t]))
        spatial = expf(-distance_sq / den_w)
        rng = expf(F(-1.2) * abs(F(rv[t] - c_ref)))
        weight[t] = spatial * rng
    count = seq_sum(weight)
    wr = weight * rv
    wt = weight * tv
    s_ref, s_tgt = seq_sum(wr), seq_sum(wt)
    s_ref2, s_tgt2, s_prod = seq_sum(wr * rv), seq_sum(wt * tv), seq_sum(wr * tv)
    covariance = F(s_prod - F(F(s_ref * s_tgt) / count))
    ref_energy = F(s_ref2 - F(F(s_ref * s_ref) / count))
    tgt_energy = F(s_tgt2 - F(F(s_tgt * s_tgt) / count))
    den = F(np.sqrt(F(fmax32(ref_energy, 0.0) * fmax32(tgt_energy, 0.0))))
    if not den > F(1.0e-6):
        return None
    return clamp32(F(covariance / den), F(-1.0), F(1.0))


def score_one_view(reference: View, target: View, centre, depth, radius):
    """depth.rs:1593-1627."""
    if reference.camera is None or target.camera is None:
        return None
    g = warp_map(target.points, target.step, F(centre[0]), F(centre[1]))
    if g is None:
        return None
    if depth is None:
        mapped = [float(g[0]), float(g[1])]
    else:
        mapped = map_at_depth(reference, target, centre, depth)
        if mapped is None:
            return None
    delta = [mapped[0] - float(g[0]), mapped[1] - float(g[1])]
    if abs(delta[0]) > 128.0 or abs(delta[1]) > 128.0:
        return None
    return zncc(reference, target, centre, delta, radius)


def score_at_point(reference: View, target: View, centre, global_point, point, radius):
    """score_one_view_at_point (depth.rs:1710-1736)."""
    delta = [float(F(point[0] - global_point[0])), float(F(point[1] - global_point[1]))]
    if abs(delta[0]) > 128.0 or abs(delta[1]) > 128.0:
        return None
    return zncc(reference, target, centre, delta, radius)


def aggregate(scores, minimum_support, best_view_count):
    """depth.rs:1761-1770."""
    if len(scores) < minimum_support:
        return None
    values = sorted(scores, key=total_key32, reverse=True)[:min(best_view_count, len(scores))]
    if len(values) < minimum_support:
        return None
    total = F(-0.0)
    for v in values:
        total = F(total + v)
    return F(total / F(len(values)))


class Capture:
    """A chiaro_amd.scene.Capture in restatement form."""

    def __init__(self, cap):
        self.views = [View(m) for m in cap.modules]
        self.ref = cap.reference_index
        self.o = cap.options
        self.targets = [i for i, v in enumerate(self.views)
                        if i != self.ref and v.accepted and v.camera is not None]

    def score_views(self, centre, depth, radius):
        """score_views (depth.rs:1566-1591): targets in index order."""
        out = []
        for i in self.targets:
            s = score_one_view(self.views[self.ref], self.views[i], centre, depth, radius)
            if s is not None:
                out.append(s)
        return out

    def aggregate_at(self, centre, depth, radius):
        return aggregate(self.score_views(centre, depth, radius), self.o.minimum_support,
                         self.o.best_view_count)

    def guidance(self, pixel):
        """reference_guidance (depth.rs:1824-1832)."""
        v = sample(self.views[self.ref].plane, F((pixel[0] - 0.5) * 0.5),
                   F((pixel[1] - 0.5) * 0.5))
        return F(0.0) if v is None else v


# ------------------------------------------------------------ select_depths

def label_depths(o):
    """inverse_depth_samples reversed behind the global label
    (depth.rs:1196-1201, 1834-1843)."""
    near_inv, far_inv = 1.0 / o.near_depth, 1.0 / o.far_depth
    finite = [1.0 / (near_inv * (1.0 - i / (o.planes - 1)) + far_inv * (i / (o.planes - 1)))
              for i in range(o.planes)]
    return [None] + finite[::-1]


def sublabel_depth(labels, costs, best):
    """depth.rs:1469-1488."""
    centre_depth = labels[best]
    if best <= 1 or best + 1 >= len(labels):
        return centre_depth
    left, centre, right = F(costs[best - 1]), F(costs[best]), F(costs[best + 1])
    curvature = F(F(left - F(F(2.0) * centre)) + right)
    if not np.isfinite(curvature) or curvature <= F(1.0e-6):
        return centre_depth
    offset = float(clamp32(F(F(F(0.5) * F(left - right)) / curvature), F(-0.75), F(0.75)))
    lower_inverse = 1.0 / labels[best - 1]
    upper_inverse = 1.0 / labels[best + 1]
    inverse_step = (upper_inverse - lower_inverse) * 0.5
    return 1.0 / (1.0 / centre_depth + offset * inverse_step)


def select_depths(labels, scores, tested, regularised, o):
    """depth.rs:1385-1467. Returns ([None | (depth, confidence, improvement,
    regularized)], fillable, best_label)."""
    n, L = scores.shape
    field, fillable, best_label = [None] * n, [False] * n, [0] * n
    min_score, min_margin, min_impr = F(o.minimum_score), F(o.minimum_margin), F(o.minimum_improvement)
    for node in range(n):
        if not tested[node]:
            continue
        costs = regularised[node]
        best, best_cost = None, None
        for label in range(1, L):          # min_by: the first minimum
            c = costs[label]
            if best is None or total_key32(c) < total_key32(best_cost):
                best, best_cost = label, c
        if best is None:
            continue
        best_label[node] = best
        depth = sublabel_depth(labels, costs, best)
        score = F(scores[node, best])
        if not np.isfinite(score) or score < MINIMUM_REGULARIZED_SCORE:
            fillable[node] = True
            continue
        competing = None
        for label in range(1, L):
            if abs(label - best) > 2:
                c = costs[label]
                if competing is None or total_key32(c) < total_key32(competing):
                    competing = c
        if competing is None:
            competing = best_cost
        margin = fmax32(F(F(competing - best_cost) / SGM_DIRECTIONS), F(0.0))
        baseline = F(scores[node, 0])
        improvement = F(score - baseline) if np.isfinite(baseline) else F(0.0)
        if np.isfinite(baseline) and improvement < -MAXIMUM_REGULARIZED_BASELINE_LOSS:
            continue
        measured = (score >= min_score and improvement >= F(min_impr * F(0.5))
                    and (margin >= min_margin or improvement >= min_impr))
        score_conf = clamp32(F(F(score - min_score) / fmax32(F(F(1.0) - min_score), F(1.0e-3))),
                             0.0, 1.0)
        margin_conf = clamp32(F(margin / fmax32(min_margin, F(1.0e-3))), 0.0, 1.0)
        impr_conf = clamp32(F(improvement / fmax32(min_impr, F(1.0e-3))), 0.0, 1.0)
        if measured:
            conf = clamp32(F(F(F(0.30) + F(F(0.45) * score_conf)) + F(F(0.25) * margin_conf)),
                           0.0, 1.0)
        else:
            conf = clamp32(F(F(F(F(0.10) + F(F(0.25) * score_conf)) + F(F(0.15) * margin_conf))
                             + F(F(0.10) * impr_conf)), F(0.10), F(0.55))
        field[node] = (depth, conf, improvement, not measured)
    return field, fillable, best_label


# ------------------------------------------------- measure_direct_depths

def nearest_coarse(coarse, columns, rows, step, pixel):
    """depth.rs:1062-1076 (Rust rounds half away from zero)."""
    def rnd(v):
        return math.floor(v + 0.5) if v >= 0 else -math.floor(-v + 0.5)
    column = int(clamp64(rnd(pixel[0] / step), 0.0, float(columns - 1)))
    row = int(clamp64(rnd(pixel[1] / step), 0.0, float(rows - 1)))
    return coarse[row * columns + column]


def direct_candidates(seed, inverse_step, o):
    """depth.rs:1078-1107."""
    far_inv, near_inv = 1.0 / o.far_depth, 1.0 / o.near_depth
    if seed is None:
        return _samples_reversed(o.near_depth, o.far_depth, 32)
    depths = []
    for offset in range(-6, 7):
        inverse = clamp64(1.0 / seed + offset * inverse_step * 0.5, far_inv, near_inv)
        depth = 1.0 / inverse
        if not depths or abs(1.0 / depths[-1] - 1.0 / depth) > inverse_step * 0.1:
            depths.append(depth)
    return depths


def _samples_reversed(near, far, count):
    near_inv, far_inv = 1.0 / near, 1.0 / far
    s = [1.0 / (near_inv * (1.0 - i / (count - 1)) + far_inv * (i / (count - 1)))
         for i in range(count)]
    return s[::-1]


def _objective(depth, score, o):
    far_inv = 1.0 / o.far_depth
    inverse_range = 1.0 / o.near_depth - far_inv
    near_fraction = clamp64((1.0 / depth - far_inv) / inverse_range, 0.0, 1.0)
    return F(score - F(NEAR_DEPTH_PRIOR * F(near_fraction)))


def select_direct_depth(depths, scores, baseline, seeded, o):
    """depth.rs:1109-1165."""
    ranked = [(i, d, s, _objective(d, s, o)) for i, (d, s) in enumerate(zip(depths, scores))
              if s is not None]
    ranked.sort(key=lambda r: total_key32(r[3]), reverse=True)   # stable
    if not ranked:
        return None
    best_index, depth, score, best_objective = ranked[0]
    competing = next((r[3] for r in ranked if abs(r[0] - best_index) > 1), best_objective)
    margin = fmax32(F(best_objective - competing), F(0.0))
    improvement = F(0.0) if baseline is None else F(score - baseline)
    min_score, min_margin, min_impr = F(o.minimum_score), F(o.minimum_margin), F(o.minimum_improvement)
    floor = fmax32(F(min_score - F(0.03)), MINIMUM_REGULARIZED_SCORE)
    required = F(min_margin * F(0.6)) if seeded else min_margin
    improves = baseline is None or improvement >= F(min_impr * F(0.5))
    supported = score >= floor and improves and (margin >= required or improvement >= min_impr)
    if not supported:
        return None
    score_conf = clamp32(F(F(score - floor) / fmax32(F(F(1.0) - floor), F(1.0e-3))), 0.0, 1.0)
    margin_conf = clamp32(F(margin / fmax32(required, F(1.0e-3))), 0.0, 1.0)
    conf = clamp32(F(F(F(0.30) + F(F(0.45) * score_conf)) + F(F(0.25) * margin_conf)), 0.0, 1.0)
    return (depth, conf, improvement, False)


def best_depth_candidate(depths, scores, o):
    """depth.rs:1167-1185 (max_by: the last maximum)."""
    best = None
    for i, (d, s) in enumerate(zip(depths, scores)):
        if s is None:
            continue
        obj = _objective(d, s, o)
        if best is None or total_key32(obj) >= total_key32(best[1]):
            best = (i, obj)
    return None if best is None else best[0]


def measure_direct(cap: Capture, coarse, coarse_columns, coarse_rows, coarse_step, rows_range=None):
    """measure_direct_depths (depth.rs:715-876). Returns (field, guidance,
    tested, wide_count) over the fine grid (rows outside rows_range None)."""
    o = cap.o
    ref = cap.views[cap.ref]
    step = max(coarse_step // 2, 4)
    columns = -(-ref.width // step) + 1
    rows = -(-ref.height // step) + 1
    inverse_step = (1.0 / o.near_depth - 1.0 / o.far_depth) / (o.planes - 1)
    r_direct = max(o.patch_radius, 3)
    r_wide = max(o.patch_radius * 2, 8)
    field = [None] * (columns * rows)
    guidance = np.zeros(columns * rows, np.float32)
    tested = np.zeros(columns * rows, np.uint8)
    wide_count = 0
    for row in (rows_range or range(rows)):
        for column in range(columns):
            p = [float(column * step), float(row * step)]
            index = row * columns + column
            guidance[index] = cap.guidance(p)
            seed = nearest_coarse(coarse, coarse_columns, coarse_rows, coarse_step, p)
            candidates = direct_candidates(None if seed is None else seed[0], inverse_step, o)
            baseline = cap.aggregate_at(p, None, r_direct)
            scores = [cap.aggregate_at(p, d, r_direct) for d in candidates]
            tested[index] = baseline is not None or any(s is not None for s in scores)
            selected = select_direct_depth(candidates, scores, baseline, seed is not None, o)
            if selected is None:
                best = best_depth_candidate(candidates, scores, o)
                if best is not None and scores[best] is not None and \
                        scores[best] >= MINIMUM_REGULARIZED_SCORE:
                    wide_count += 1
                    first, last = max(best - 3, 0), min(best + 3, len(candidates) - 1)
                    wide = candidates[first:last + 1]
                    wide_baseline = cap.aggregate_at(p, None, r_wide)
                    wide_scores = [cap.aggregate_at(p, d, r_wide) for d in wide]
                    selected = select_direct_depth(wide, wide_scores, wide_baseline,
                                                   seed is not None, o)
            field[index] = selected
    return field, guidance, tested, wide_count


# ------------------------------------------------- per-target decisions

def refine_one_view(cap: Capture, target: View, centre, shared_depth, inverse_step):
    """depth.rs:1634-1708: (score, point) or None."""
    o = cap.o
    reference = cap.views[cap.ref]
    g = warp_map(target.points, target.step, F(centre[0]), F(centre[1]))
    if g is None:
        return None
    min_inv, max_inv = 1.0 / o.far_depth, 1.0 / o.near_depth
    shared_inv = 1.0 / shared_depth
    best = None
    for offset in (-1.0, -0.5, 0.0, 0.5, 1.0):
        inverse = clamp64(shared_inv + offset * inverse_step, min_inv, max_inv)
        q = map_at_depth(reference, target, centre, 1.0 / inverse)
        if q is None:
            return None
        point = (F(q[0]), F(q[1]))
        score = score_at_point(reference, target, centre, g, point, o.patch_radius)
        if score is None:
            return None
        objective = F(score - F(F(abs(offset)) * F(0.002)))
        if best is None or objective > best[0]:
            best = (objective, score, point)
    _, depth_score, depth_point = best
    best = (depth_score, depth_score, depth_point)
    for dy in (F(-1.5), F(0.0), F(1.5)):
        for dx in (F(-1.5), F(0.0), F(1.5)):
            if dx == 0 and dy == 0:
                continue
            point = (F(depth_point[0] + dx), F(depth_point[1] + dy))
            score = score_at_point(reference, target, centre, g, point, o.patch_radius)
            if score is None:
                return None
            residual = F(F(dx * dx) + F(dy * dy))
            objective = F(score - F(residual * F(0.002)))
            if objective > best[0]:
                best = (objective, score, point)
    return best[1], best[2]


def _mapping(d):
    code, glob, point = d[0], d[1], d[2]
    if code == GLOBAL:
        return glob, glob
    if code == REFINED:
        return glob, point
    return None


def _agree(kind, centre, neighbour):
    """warp_decisions_agree (depth.rs:682-712)."""
    if kind == 0 and centre[0] == REFINED and neighbour[0] == REFINED:
        cd = (F(centre[2][0] - centre[1][0]), F(centre[2][1] - centre[1][1]))
        nd = (F(neighbour[2][0] - neighbour[1][0]), F(neighbour[2][1] - neighbour[1][1]))
        dx, dy = F(cd[0] - nd[0]), F(cd[1] - nd[1])
        return F(F(dx * dx) + F(dy * dy)) <= F(9.0)
    return kind == 1 and centre[0] == OCCLUDED and neighbour[0] == OCCLUDED


def enforce_consensus(decisions, columns, rows):
    """depth.rs:582-618."""
    source = list(decisions)
    for row in range(rows):
        for column in range(columns):
            i = row * columns + column
            d = source[i]
            if d[0] == REFINED:
                kind, need = 0, 2 if d[4] else 4
            elif d[0] == OCCLUDED:
                kind, need = 1, 3 if d[4] else 5
            else:
                continue
            count = 0
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dx == 0 and dy == 0:
                        continue
                    x, y = column + dx, row + dy
                    if 0 <= x < columns and 0 <= y < rows:
                        count += _agree(kind, d, source[y * columns + x])
            if count < need:
                decisions[i] = (GLOBAL, d[1], d[1], F(1.0), False)


def suppress_boundaries(decisions, guidance, columns, rows):
    """depth.rs:620-680."""
    source = list(decisions)
    boundary = [False] * len(decisions)
    limit = F(BOUNDARY_DELTA * BOUNDARY_DELTA)
    for row in range(rows):
        for column in range(columns):
            i = row * columns + column
            m = _mapping(source[i])
            if m is None:
                continue
            delta = (F(m[1][0] - m[0][0]), F(m[1][1] - m[0][1]))
            for x, y in ((column + 1, row), (column, row + 1)):
                if x >= columns or y >= rows:
                    continue
                j = y * columns + x
                n = _mapping(source[j])
                if n is None:
                    continue
                nd = (F(n[1][0] - n[0][0]), F(n[1][1] - n[0][1]))
                diff = (F(delta[0] - nd[0]), F(delta[1] - nd[1]))
                mapping_edge = F(F(diff[0] * diff[0]) + F(diff[1] * diff[1])) > limit
                image_edge = abs(F(guidance[i] - guidance[j])) > BOUNDARY_CONTRAST
                if mapping_edge and image_edge:
                    boundary[i] = boundary[j] = True
    for i, b in enumerate(boundary):
        if b:
            m = _mapping(source[i])
            if m is not None:
                decisions[i] = (BOUNDARY, m[0], m[0], F(0.0), False)


def target_decisions(cap: Capture, field, guidance):
    """The per-target block of refine_multiview_depth (depth.rs:420-547):
    {target: (codes u8, points f32 (n, 2), confidence f32)} on the fine grid."""
    o = cap.o
    ref = cap.views[cap.ref]
    coarse_step = max(o.grid_step, 4)
    step = max(coarse_step // 2, 4)
    columns = -(-ref.width // step) + 1
    rows = -(-ref.height // step) + 1
    inverse_step = (1.0 / o.near_depth - 1.0 / o.far_depth) / (o.planes - 1)
    min_score = F(o.minimum_score)
    out = {}
    for t in cap.targets:
        target = cap.views[t]
        decisions = []
        for row in range(rows):
            for column in range(columns):
                i = row * columns + column
                p = [float(column * step), float(row * step)]
                g = warp_map(target.points, target.step, F(p[0]), F(p[1]))
                if g is None:
                    decisions.append((UNDEFINED, None, None, F(0.0), False))
                    continue
                node = field[i]
                if node is None:
                    decisions.append((GLOBAL, g, g, F(1.0), False))
                    continue
                selected = refine_one_view(cap, target, p, node[0], inverse_step)
                baseline = score_one_view(ref, target, p, None, o.patch_radius)
                supported = selected is not None and selected[0] >= min_score and \
                    (baseline is None or F(selected[0] + F(0.01)) >= baseline)
                contradicted = selected is not None and baseline is not None and \
                    F(selected[0] + F(0.12)) < baseline
                measured = not node[3]
                if contradicted:
                    decisions.append((OCCLUDED, g, g, F(0.0), measured))
                elif supported:
                    conf = F(F(node[1]) * F(0.6)) if node[3] else \
                        F(F(0.5) + F(F(0.5) * F(node[1])))
                    decisions.append((REFINED, g, selected[1], conf, measured))
                else:
                    decisions.append((GLOBAL, g, g, F(1.0), False))
        enforce_consensus(decisions, columns, rows)
        suppress_boundaries(decisions, guidance, columns, rows)
        codes = np.array([d[0] for d in decisions], np.uint8)
        points = np.full((len(decisions), 2), np.nan, np.float32)
        conf = np.zeros(len(decisions), np.float32)
        for i, d in enumerate(decisions):
            if d[0] == UNDEFINED:
                continue
            if d[0] == REFINED:
                points[i] = d[2]
                conf[i] = d[3]
            elif d[0] == GLOBAL:
                points[i] = d[1]
                conf[i] = F(1.0)
            else:                      # occluded, boundary: the global point
                points[i] = d[1]
        out[t] = (codes, points, conf)
    return out
