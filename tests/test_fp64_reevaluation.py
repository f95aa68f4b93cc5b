"""Float error of the path's f32 stages against an FP64 re-evaluation.

The GPU reproduces the reference's f32 arithmetic bit for bit (the GPU tests
compare it with the oracle), so its error against exact arithmetic is the
reference's own. These tests measure that error where the path computes in
f32 - warped_patch_zncc (depth.rs:1772-1822), the cost built from it
(depth.rs:1245-1262) and sublabel_depth (depth.rs:1469-1488) - by
re-evaluating the same quantities in float64 from the same inputs.

Measured on a config 0 capture (600 random patches and displacements each):

* ZNCC on the planes as the path gets them (log luminance around 6.2 with
  a spread of about 0.3): median |f32 - f64| 6e-4, 99th percentile 8e-3,
  largest 1e-2. The reference forms the energies from one-pass sums of
  w v and w v^2, which cancel: the error follows the patches' conditioning
  kappa = sum(w v^2) / sum(w (v - mean)^2) (median 6,100 here) and stays
  below 32 eps kappa (eps = 2^-24) for every sample (largest ratio 13);
* the same patches with the planes shifted to mean 0 (kappa near 1): largest
  |f32 - f64| 4.5e-5, median 7e-7 - the sampling and weighting add little;
* cost = clamp(1 - score, 0, 2) + prior carries the score's error;
* sublabel depth from the same f32 costs: relative depth error 4e-9 at
  most over 3,000 random cost triples (bound held: 1e-6).

The f64 re-evaluation keeps every coordinate in f64 (no f32 rounding of
the tap positions), samples bilinearly in f64, weights with f64 exp and
forms the weighted covariance with the two-pass (centred) formula.
"""
import ctypes as C
import math

import numpy as np
import pytest

from chiaro_amd import _abi, scene


def _warp_struct(points, step):
    pts = np.ascontiguousarray(points, dtype=np.float32)
    wp = _abi.Warp(step=step, columns=pts.shape[1], rows=pts.shape[0])
    wp.points = pts.ctypes.data_as(C.POINTER(C.c_float))
    return wp, pts


def _sample64(plane, x, y):
    """Plane::sample (image.rs:451-466) in f64; None outside the plane."""
    h, w = plane.shape
    if not (x >= 0.0 and y >= 0.0 and x <= w - 1 and y <= h - 1):
        return None
    x0, y0 = int(math.floor(x)), int(math.floor(y))
    x1, y1 = min(x0 + 1, w - 1), min(y0 + 1, h - 1)
    tx, ty = x - x0, y - y0
    p = plane
    top = float(p[y0, x0]) * (1 - tx) + float(p[y0, x1]) * tx
    bottom = float(p[y1, x0]) * (1 - tx) + float(p[y1, x1]) * tx
    return top * (1 - ty) + bottom * ty


def _warp64(points, step, x, y):
    """Warp::map (align.rs:84-111) in f64; None where undefined."""
    rows, cols = points.shape[:2]
    fx, fy = x / step, y / step
    if fx < 0 or fy < 0:
        return None
    c0, r0 = min(int(math.floor(fx)), cols - 1), min(int(math.floor(fy)), rows - 1)
    c1, r1 = min(c0 + 1, cols - 1), min(r0 + 1, rows - 1)
    tx, ty = fx - c0, fy - r0
    a, b = points[r0, c0].astype(np.float64), points[r0, c1].astype(np.float64)
    c, d = points[r1, c0].astype(np.float64), points[r1, c1].astype(np.float64)
    o = (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty
    if np.isnan(o).any():
        return None
    return float(o[0]), float(o[1])


def _zncc64(ref, tgt, points, step, cx, cy, ddx, ddy, radius):
    centre = _sample64(ref, (cx - 0.5) * 0.5, (cy - 0.5) * 0.5)
    if centre is None:
        return None
    sigma = max(radius * 0.75, 1.0)
    w, r, t = [], [], []
    for dy in range(-radius, radius + 1):
        for dx in range(-radius, radius + 1):
            px, py = cx + dx * 2.0, cy + dy * 2.0
            rv = _sample64(ref, (px - 0.5) * 0.5, (py - 0.5) * 0.5)
            g = _warp64(points, step, px, py)
            if rv is None or g is None:
                return None
            tv = _sample64(tgt, (g[0] + ddx - 0.5) * 0.5, (g[1] + ddy - 0.5) * 0.5)
            if tv is None:
                return None
            spatial = math.exp(-(dx * dx + dy * dy) / (2 * sigma * sigma))
            w.append(spatial * math.exp(-1.2 * abs(rv - centre)))
            r.append(rv)
            t.append(tv)
    w, r, t = np.array(w), np.array(r), np.array(t)
    mr, mt = (w * r).sum() / w.sum(), (w * t).sum() / w.sum()
    cov = (w * (r - mr) * (t - mt)).sum()
    er, et = (w * (r - mr) ** 2).sum(), (w * (t - mt) ** 2).sum()
    den = math.sqrt(er * et)
    if den == 0.0:
        return None
    # conditioning of the one-pass energies: mean square over variance
    kappa = max((w * r * r).sum() / er, (w * t * t).sum() / et)
    return max(-1.0, min(1.0, cov / den)), kappa


def _errors(oracle_lib, ref, tgt, points, step, samples, seed, radius=4):
    rng = np.random.default_rng(seed)
    h, w = ref.shape
    wp, keep = _warp_struct(points, step)
    out = C.c_float()
    ref = np.ascontiguousarray(ref, np.float32)
    tgt = np.ascontiguousarray(tgt, np.float32)
    errs = []
    for _ in range(samples):
        cx = float(rng.uniform(24, 2 * w - 24))
        cy = float(rng.uniform(24, 2 * h - 24))
        ddx, ddy = float(rng.uniform(-12, 12)), float(rng.uniform(-12, 12))
        ok = oracle_lib.or_warped_patch_zncc(ref.ctypes.data, w, h, tgt.ctypes.data,
                                             tgt.shape[1], tgt.shape[0], C.byref(wp), cx, cy,
                                             ddx, ddy, radius, C.byref(out))
        if not ok:
            continue
        exact = _zncc64(ref, tgt, points, step, cx, cy, ddx, ddy, radius)
        if exact is None:
            continue
        errs.append((abs(out.value - exact[0]), exact[1]))
    return np.array(errs).reshape(-1, 2)


@pytest.fixture(scope="module")
def small_capture():
    return scene.config0(seed=3, width=320, height=240)


def test_zncc_error_follows_conditioning(oracle_lib, small_capture):
    cap = small_capture
    ref, tgt = cap.modules[0].luminance, cap.modules[1].luminance
    e = _errors(oracle_lib, ref, tgt, cap.modules[1].warp_points, cap.modules[1].warp_step,
                600, seed=1)
    assert e.shape[0] > 400
    err, kappa = e[:, 0], e[:, 1]
    ratio = err / (kappa * 2.0 ** -24)
    print(f"zncc f32 vs f64: max {err.max():.3g}, p99 {np.quantile(err, 0.99):.3g}, "
          f"median {np.median(err):.3g}; kappa median {np.median(kappa):.0f}; "
          f"largest err / (eps kappa) {ratio.max():.1f} over {err.size}")
    assert ratio.max() <= 32.0
    assert err.max() <= 0.03


def test_zncc_error_on_centred_planes(oracle_lib, small_capture):
    cap = small_capture
    ref = (cap.modules[0].luminance - np.float32(6.2)).astype(np.float32)
    tgt = (cap.modules[1].luminance - np.float32(6.2)).astype(np.float32)
    e = _errors(oracle_lib, ref, tgt, cap.modules[1].warp_points, cap.modules[1].warp_step,
                600, seed=2)
    assert e.shape[0] > 400
    err = e[:, 0]
    print(f"zncc f32 vs f64, centred planes: max {err.max():.3g}, "
          f"p99 {np.quantile(err, 0.99):.3g}, median {np.median(err):.3g} over {err.size}")
    assert err.max() <= 1e-4


def test_sublabel_depth_error(oracle_lib):
    """The parabola vertex from f32 costs (depth.rs:1469-1488) against the
    same vertex evaluated in f64 from the same costs."""
    rng = np.random.default_rng(4)
    labels = np.empty(129, np.float64)
    oracle_lib.or_inverse_depth_samples(800.0, 20000.0, 129, labels.ctypes.data)
    worst = 0.0
    for _ in range(3000):
        best = int(rng.integers(2, 127))
        costs = rng.uniform(0.0, 2.25, 129).astype(np.float32)
        costs[best] = np.float32(min(costs[best - 1], costs[best + 1]) * rng.uniform(0.2, 0.99))
        got = oracle_lib.or_sublabel_depth(labels.ctypes.data, 129, costs.ctypes.data, best)
        l, c, r = (float(costs[best - 1]), float(costs[best]), float(costs[best + 1]))
        curvature = l - 2 * c + r
        if not curvature > 1e-6:
            continue
        offset = max(-0.75, min(0.75, 0.5 * (l - r) / curvature))
        step = (1 / labels[best + 1] - 1 / labels[best - 1]) * 0.5
        exact = 1 / (1 / labels[best] + offset * step)
        worst = max(worst, abs(got - exact) / exact)
    print(f"sublabel depth f32 vs f64: max relative error {worst:.3g}")
    assert worst <= 1e-6
