"""The oracle's alignment (oracle/align_oracle.c) pinned by the reference's
own tests of it, ported (align.rs:867-892, image.rs:658-671), and by closed
forms: least squares recovers an exact homography from four or more exact
pairs, match_patch finds an integer shift of a copy at score 1, the pyramid
halves down to its floor, and align_module's factory model (refine off)
equals the calibrated infinity-map warp."""
import ctypes as C
import math

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import _ViewArrays

_libm = C.CDLL("libm.so.6")
for name in ("sinf", "cosf"):
    getattr(_libm, name).restype = C.c_float
    getattr(_libm, name).argtypes = [C.c_float]


def textured(width, height, shift):
    """The reference test's texture (image.rs:642-656) with f32 sin/cos."""
    plane = np.empty((height, width), np.float32)
    f = np.float32
    for y in range(height):
        for x in range(width):
            fx, fy = f(f(x) - f(shift[0])), f(f(y) - f(shift[1]))
            a = f(_libm.sinf(f(fx * f(0.31)))) * f(_libm.cosf(f(fy * f(0.17))))
            b = f(0.3) * f(_libm.sinf(f(f(fx * f(0.05)) + f(fy * f(0.07)))))
            plane[y, x] = f(a + b)
    return plane


def match(lib, ref, tgt, rx, ry, size, radius):
    shift = np.empty(2, np.float32)
    score = C.c_float()
    ok = lib.or_match_patch(ref.ctypes.data, ref.shape[1], ref.shape[0], tgt.ctypes.data,
                            tgt.shape[1], tgt.shape[0], rx, ry, size, radius, shift.ctypes.data,
                            C.byref(score))
    return (shift, score.value) if ok else None


def test_patch_match_recovers_a_known_shift(oracle_lib):
    """image.rs:658-671: a texture shifted by (3.4, -2.6) is found within
    0.25 px at score > 0.9; a flat reference has no match."""
    reference = textured(200, 160, (0.0, 0.0))
    target = textured(200, 160, (3.4, -2.6))
    shift, score = match(oracle_lib, reference, target, 60, 50, 48, 8)
    assert abs(shift[0] - 3.4) < 0.25 and abs(shift[1] + 2.6) < 0.25 and score > 0.9
    flat = np.zeros((160, 200), np.float32)
    assert match(oracle_lib, flat, target, 60, 50, 48, 8) is None


def test_patch_match_integer_shift_of_a_copy(oracle_lib):
    rng = np.random.default_rng(3)
    reference = rng.random((64, 80)).astype(np.float32)
    target = np.roll(np.roll(reference, 2, axis=1), -3, axis=0)
    shift, score = match(oracle_lib, reference, target, 20, 20, 16, 5)
    assert tuple(shift) == (2.0, -3.0) or np.allclose(shift, (2.0, -3.0), atol=0.5)
    assert abs(score - 1.0) < 1e-5
    # out of bounds and radius outside the target: None
    assert match(oracle_lib, reference, target, 3, 20, 16, 5) is None


def _pairs(truth, oracle_lib, outliers=True):
    pairs, scores = [], []
    for y in range(0, 3000, 250):
        for x in range(0, 4000, 250):
            a = np.array([x, y], np.float64)
            b = np.empty(2)
            assert oracle_lib.or_apply_homography(truth.ctypes.data, a.ctypes.data, b.ctypes.data)
            pairs.append([a[0], a[1], b[0], b[1]])
            scores.append(0.9)
    if outliers:
        for i in range(20):
            a = pairs[i * 7]
            pairs.append([a[0], a[1], a[0] + 80.0, a[1] - 120.0])
            scores.append(0.9)
    return np.array(pairs, np.float64), np.array(scores, np.float32)


def test_homography_fit_recovers_a_known_transform(oracle_lib):
    """align.rs:867-892: RANSAC over pairs of a known homography with 20
    gross outliers."""
    truth = np.array([1.02, 0.01, 5.0, -0.015, 0.99, -3.0, 1e-6, -2e-6, 1.0])
    pairs, scores = _pairs(truth, oracle_lib)
    h = np.empty(9)
    inliers = C.c_uint32()
    residuals = np.empty(len(pairs), np.float32)
    assert oracle_lib.or_fit_homography_ransac(pairs.ctypes.data, scores.ctypes.data, len(pairs),
                                               1.0, h.ctypes.data, C.byref(inliers),
                                               residuals.ctypes.data)
    assert inliers.value >= 190
    assert residuals[inliers.value // 2] < 1e-3
    for a in ([100.0, 100.0], [3900.0, 2900.0]):
        a = np.array(a)
        want, got = np.empty(2), np.empty(2)
        oracle_lib.or_apply_homography(truth.ctypes.data, a.ctypes.data, want.ctypes.data)
        oracle_lib.or_apply_homography(h.ctypes.data, a.ctypes.data, got.ctypes.data)
        assert np.abs(got - want).max() < 1e-3


def test_least_squares_is_exact_on_exact_pairs(oracle_lib):
    truth = np.array([0.98, -0.02, 12.0, 0.03, 1.01, -7.0, 2e-6, 1e-6, 1.0])
    pairs, scores = _pairs(truth, oracle_lib, outliers=False)
    for take in (4, len(pairs)):
        h = np.empty(9)
        sub = np.ascontiguousarray(pairs[[0, 15, 176, 191]] if take == 4 else pairs)
        assert oracle_lib.or_fit_homography_least_squares(sub.ctypes.data, scores.ctypes.data,
                                                          len(sub), h.ctypes.data)
        assert np.allclose(h, truth, rtol=1e-7, atol=1e-10)
    # degenerate: four collinear points have no homography
    line = np.array([[x, 0.0, x, 0.0] for x in (0.0, 1.0, 2.0, 3.0)])
    assert not oracle_lib.or_fit_homography_least_squares(line.ctypes.data, scores.ctypes.data, 4,
                                                          h.ctypes.data)


def test_pyramid_levels(oracle_lib):
    assert oracle_lib.or_pyramid_levels(2080, 1560, 96) == 5   # 2080 .. 130 x 97
    assert oracle_lib.or_pyramid_levels(2, 2, 1) == 1           # image.rs:697
    assert oracle_lib.or_pyramid_levels(4160, 3120, 48) == 7


def _align(oracle_lib, cap, target, options):
    ref = cap.modules[cap.reference_index]
    views = _ViewArrays(cap.modules)
    step = options.grid_step
    columns, rows = -(-ref.width // step) + 1, -(-ref.height // step) + 1
    points = np.empty((rows, columns, 2), np.float32)
    conf = np.empty((rows, columns), np.float32)
    warp = _abi.Warp()
    warp.points = points.ctypes.data_as(C.POINTER(C.c_float))
    warp.confidence = conf.ctypes.data_as(C.POINTER(C.c_float))
    report = _abi.AlignmentReport()
    rc = oracle_lib.or_align_module(C.byref(views.views[cap.reference_index]),
                                    C.byref(views.views[target]), C.byref(options), C.byref(warp),
                                    C.byref(report))
    assert rc == 0
    return points, report


def test_factory_model_is_the_infinity_map(oracle_lib):
    cap = scene.config1(seed=1, width=416, height=312, focal_px=330.0, planes=32)
    o = _abi.AlignOptions()
    oracle_lib.or_default_align_options(C.byref(o))
    o.refine = 0
    points, report = _align(oracle_lib, cap, 2, o)
    np.testing.assert_array_equal(points, cap.modules[2].warp_points)
    assert report.accepted == 1 and report.status == 2 and report.level_count == 0
    o.refine = 1
    points, report = _align(oracle_lib, cap, 0, o)     # the reference itself
    assert report.status == 0 and report.accepted == 1
    np.testing.assert_array_equal(points, scene.identity_warp(416, 312))


def rotation_y(degrees):
    a = math.radians(degrees)
    return np.array([[math.cos(a), 0.0, math.sin(a)], [0.0, 1.0, 0.0],
                     [-math.sin(a), 0.0, math.cos(a)]])


def rotated_pair(width=416, height=312, focal=330.0, degrees=1.0, seed=4):
    """A reference and a module at the same centre turned by `degrees` about
    the vertical axis: the module's image is the reference through the
    homography K R K^-1, whatever the scene depth. Returns (capture whose
    target camera is the uncalibrated identity, the true target camera)."""
    mods = [scene._module(focal, width, height, (0.0, 0.0, 0.0)),
            scene._module(focal, width, height, (0.0, 0.0, 0.0), rotation_y(degrees))]
    desc = _abi.SceneDesc(seed=seed, plane_count=8, inverse_depth_min=0.05e-3,
                          inverse_depth_max=1.25e-3, texture_cell_px=16)
    raws = scene._render(desc, mods, 0, 0.0, _abi.Bayer(**scene.BAYER_12BIT))
    planes = [scene.luminance_half_u16(r, 4, 64.0 * 16, 4095.0 * 16) for r in raws]
    ref_cam = scene.camera_for(mods[0])
    factory = scene.camera_for(scene._module(focal, width, height, (0.0, 0.0, 0.0)))
    true_cam = scene.camera_for(mods[1])
    inputs = [scene.ModuleInput(name=f"M{i}", luminance=p, width=width, height=height, camera=cam,
                                warp_points=scene.identity_warp(width, height), warp_step=32)
              for i, (p, cam) in enumerate(zip(planes, [ref_cam, factory]))]
    cap = scene.Capture(modules=inputs, reference_index=0, options=None)
    return cap, true_cam


def test_refinement_recovers_a_rotation(oracle_lib):
    """Closed form: from an identity factory model the refined warp lands on
    the true rotation's infinity map within a quarter pixel (two pyramid
    levels: 416x312 and 208x156 planes), the correction at the centre is
    f tan(1 deg), and the report says refined."""
    w, h, f = 832, 624, 660.0
    cap, true_cam = rotated_pair(w, h, f)
    o = _abi.AlignOptions()
    oracle_lib.or_default_align_options(C.byref(o))
    points, report = _align(oracle_lib, cap, 1, o)
    truth = scene.global_warp(cap.modules[0].camera, true_cam, w, h)
    ok = np.isfinite(points[..., 0]) & np.isfinite(truth[..., 0])
    inside = ok & (truth[..., 0] > 20) & (truth[..., 0] < w - 20) & (truth[..., 1] > 20) & \
        (truth[..., 1] < h - 20)
    err = np.hypot(*(points[inside] - truth[inside]).T)
    factory_err = np.hypot(*(scene.identity_warp(w, h)[inside] - truth[inside]).T)
    assert factory_err.mean() > 10.0           # the uncalibrated start is off by ~11.5 px
    assert np.median(err) < 0.25 and err.max() < 1.0, (np.median(err), err.max())
    assert report.status == 3 and report.accepted == 1 and report.level_count == 2
    assert report.inliers >= 12 and report.inlier_ratio > 0.8
    assert abs(report.correction_median_px[0] - f * math.tan(math.radians(1.0))) < 1.0
