"""refine_resolution_warp's oracle (oracle/align_oracle.c, restating
crates/chiaro-fusion/src/resolution.rs:232-465) pinned three ways:

- the reference's own test (resolution.rs:531-559): a sub-pixel plane shift
  recovered at the centre with confidence > 0.5;
- closed forms: a target at twice the reference's scale (the Gaussian
  render footprint and the magnified search radius) recovers a known shift
  times the scale; textureless planes keep the base warp with zero
  confidence; a flat hole inside a coherent shift field is completed with
  0.45 x the neighbours' mean confidence;
- an independent numpy restatement (tests/nversion_resolution.py) agreeing
  bit for bit, including a base warp with undefined nodes and non-unit
  confidences and magnifications on every branch.
"""
import ctypes as C

import numpy as np
import pytest

from chiaro_amd import _abi

import nversion_resolution as nv

F = np.float32


def signal(x, y):
    """The reference test's texture (resolution.rs:533-537)."""
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    return (np.sin(x * F(0.27)) * np.cos(y * F(0.19)) + F(0.35) * np.sin(x * F(0.073) + y * F(0.11))
            + F(0.12) * np.cos(x * F(0.41) - y * F(0.037))).astype(np.float32)


def plane(width, height, fn):
    ys, xs = np.meshgrid(np.arange(height, dtype=np.float32), np.arange(width, dtype=np.float32),
                         indexing="ij")
    return np.ascontiguousarray(fn(xs, ys), np.float32)


def warp_from_fn(width, height, step, fn):
    """Warp::from_fn (align.rs:50-80)."""
    columns, rows = -(-width // step) + 1, -(-height // step) + 1
    points = np.full((rows, columns, 2), np.nan, np.float32)
    conf = np.zeros((rows, columns), np.float32)
    for r in range(rows):
        for c in range(columns):
            q = fn(float(c * step), float(r * step))
            if q is not None:
                points[r, c] = (np.float32(q[0]), np.float32(q[1]))
                conf[r, c] = 1.0
    return points, conf


def _warp_struct(step, points, conf):
    w = _abi.Warp()
    w.step = step
    w.rows, w.columns = points.shape[:2]
    w.points = points.ctypes.data_as(C.POINTER(C.c_float))
    w.confidence = conf.ctypes.data_as(C.POINTER(C.c_float))
    return w


def refine(fn, reference, target, base_points, base_conf, width, height, step=32):
    """Run refine_resolution_warp through a C entry (the oracle's or the
    library's, minus its context): (points, confidence, report)."""
    base_points = np.ascontiguousarray(base_points, np.float32)
    base_conf = np.ascontiguousarray(base_conf, np.float32)
    base = _warp_struct(step, base_points, base_conf)
    columns, rows = -(-width // 32) + 1, -(-height // 32) + 1
    points = np.empty((rows, columns, 2), np.float32)
    conf = np.empty((rows, columns), np.float32)
    out = _warp_struct(32, points, conf)
    report = _abi.ResolutionReport()
    rh, rw = reference.shape
    th, tw = target.shape
    rc = fn(reference.ctypes.data, rw, rh, target.ctypes.data, tw, th, C.byref(base), width,
            height, C.byref(out), C.byref(report))
    assert rc == 0
    assert (out.step, out.columns, out.rows) == (32, columns, rows)
    return points, conf, report


def oracle_refine(lib, *args, **kw):
    return refine(lib.or_refine_resolution_warp, *args, **kw)


def warp_map(points, step, x, y):
    return nv.Grid(step, points, np.zeros(points.shape[:2], np.float32)).map(F(x), F(y))


def warp_confidence(points, conf, step, x, y):
    return nv.Grid(step, points, conf).confidence_at(F(x), F(y))


def shift_case():
    shift = (0.6, -0.35)
    ref = plane(256, 192, signal)
    tgt = plane(256, 192, lambda x, y: signal(x - F(shift[0]), y - F(shift[1])))
    points, conf = warp_from_fn(512, 384, 32, lambda x, y: (x, y))
    return ref, tgt, points, conf, 512, 384, shift


def magnified_case(shift=(1.3, -0.8)):
    """Target raster at twice the reference raster: base q = 2p, target
    plane T(u, v) = signal((u - 0.25) / 2 - s) so that rendering at
    reference plane pixel x reads signal(x - s)."""
    ref = plane(256, 192, signal)
    tgt = plane(512, 384, lambda u, v: signal((u - F(0.25)) * F(0.5) - F(shift[0]),
                                              (v - F(0.25)) * F(0.5) - F(shift[1])))
    points, conf = warp_from_fn(512, 384, 32, lambda x, y: (2.0 * x, 2.0 * y))
    return ref, tgt, points, conf, 512, 384, shift


def hole_case():
    """A flat square in both planes over the window of the node at raster
    (256, 192), inside a uniform shift."""
    ref, tgt, points, conf, w, h, shift = shift_case()
    ref[88:104, 120:136] = 0.25
    tgt[88:104, 120:136] = 0.25
    return ref, tgt, points, conf, w, h, shift


def test_reference_subpixel_plane_shift(oracle_lib):
    """resolution.rs:531-559, the reference's assertions."""
    ref, tgt, bp, bc, w, h, shift = shift_case()
    points, conf, report = oracle_refine(oracle_lib, ref, tgt, bp, bc, w, h)
    q = warp_map(points, 32, 256.0, 192.0)
    assert abs(q[0] - 256.0 - shift[0] * 2.0) < 0.35
    assert abs(q[1] - 192.0 - shift[1] * 2.0) < 0.35
    assert warp_confidence(points, conf, 32, 256.0, 192.0) > 0.5
    assert report.verified_fraction > 0.5
    assert report.supported_fraction >= report.verified_fraction
    assert abs(report.median_correction_px - np.hypot(1.2, 0.7)) < 0.3


def test_magnified_target_recovers_the_scaled_shift(oracle_lib):
    """Magnification 2: the Gaussian footprint and search radius 5. The
    correction (raster px) is twice the plane shift, and the target point
    moves by the base warp's scale times that."""
    ref, tgt, bp, bc, w, h, shift = magnified_case()
    points, conf, report = oracle_refine(oracle_lib, ref, tgt, bp, bc, w, h)
    q = warp_map(points, 32, 256.0, 192.0)
    assert abs(q[0] - 512.0 - 4.0 * shift[0]) < 0.5
    assert abs(q[1] - 384.0 - 4.0 * shift[1]) < 0.5
    assert warp_confidence(points, conf, 32, 256.0, 192.0) > 0.5
    assert report.verified_fraction > 0.5
    # a 4.2 plane px shift lies beyond the unmagnified search radius (2) and
    # within the magnified one (5)
    ref2, tgt2, bp2, bc2, w2, h2, _ = magnified_case(shift=(4.2, 0.0))
    _, _, rep2 = oracle_refine(oracle_lib, ref2, tgt2, bp2, bc2, w2, h2)
    assert rep2.verified_fraction > 0.5
    assert abs(rep2.median_correction_px - 8.4) < 0.4


def test_textureless_planes_keep_the_base_warp(oracle_lib):
    ref = np.full((96, 128), 0.5, np.float32)
    tgt = np.full((96, 128), 0.5, np.float32)
    bp, bc = warp_from_fn(256, 192, 32, lambda x, y: (x + 3.0, y - 1.0))
    points, conf, report = oracle_refine(oracle_lib, ref, tgt, bp, bc, 256, 192)
    np.testing.assert_array_equal(points, bp)
    assert (conf == 0).all()
    assert (report.verified_fraction, report.supported_fraction, report.median_correction_px,
            report.mean_confidence) == (0.0, 0.0, 0.0, 0.0)


def test_flat_hole_is_completed(oracle_lib):
    ref, tgt, bp, bc, w, h, shift = hole_case()
    points, conf, report = oracle_refine(oracle_lib, ref, tgt, bp, bc, w, h)
    r, c = 6, 8          # raster (256, 192)
    centre = (32 * c, 32 * r)
    # every neighbour verified, the centre not: completion gives 0.45 x
    # their mean confidence and the median correction
    nb = [conf[rr, cc] for rr in (5, 6, 7) for cc in (7, 8, 9) if (rr, cc) != (r, c)]
    assert all(v > 0.5 for v in nb)
    expect = np.float32(0.45) * np.cumsum(np.array(nb, np.float32), dtype=np.float32)[-1] \
        / np.float32(8)
    assert conf[r, c] == pytest.approx(float(expect), rel=1e-6)
    assert abs(points[r, c, 0] - centre[0] - 2 * shift[0]) < 0.35
    assert abs(points[r, c, 1] - centre[1] - 2 * shift[1]) < 0.35
    assert report.supported_fraction > report.verified_fraction


def perturbed_case(seed, magnification):
    """A base warp with scale `magnification`, a smooth wobble, undefined
    nodes in one corner and confidences in [0, 1.2], over textured noise."""
    rng = np.random.default_rng(seed)
    rw, rh = 160, 120
    ref = plane(rw, rh, signal)
    ref = (ref + F(0.05) * rng.standard_normal(ref.shape).astype(np.float32)).astype(np.float32)
    m = magnification
    tw, th = int(np.ceil(rw * m)) + 4, int(np.ceil(rh * m)) + 4
    tgt = plane(tw, th, lambda u, v: signal((u - F(0.25)) / F(m) + F(0.4),
                                            (v - F(0.25)) / F(m) - F(0.3)))
    bp, bc = warp_from_fn(2 * rw, 2 * rh, 32,
                          lambda x, y: (m * x + 0.7 * np.sin(y / 50.0), m * y + 0.5 * np.cos(x / 60.0)))
    bp[:2, :2] = np.nan
    bc = rng.uniform(0.0, 1.2, bc.shape).astype(np.float32)
    return ref, tgt, bp, bc, 2 * rw, 2 * rh


def _nversion(ref, tgt, bp, bc, w, h):
    return nv.refine_resolution_warp(ref, tgt, nv.Grid(32, bp, bc), w, h)


def _same(label, oracle_out, nv_out):
    points, conf, report = oracle_out
    np.testing.assert_array_equal(points.view(np.uint32), nv_out[0].view(np.uint32),
                                  err_msg=f"{label} points")
    np.testing.assert_array_equal(conf.view(np.uint32), nv_out[1].view(np.uint32),
                                  err_msg=f"{label} confidence")
    for k, v in nv_out[2].items():
        assert np.float32(getattr(report, k)).view(np.uint32) == np.float32(v).view(np.uint32), \
            f"{label} {k}: {getattr(report, k)} != {v}"


@pytest.mark.parametrize("case", ["shift", "magnified", "hole", "mag0.7", "mag1.25", "mag3"])
def test_nversion_agrees_bit_for_bit(oracle_lib, case):
    if case == "shift":
        args = shift_case()[:6]
    elif case == "magnified":
        args = magnified_case()[:6]
    elif case == "hole":
        args = hole_case()[:6]
    else:
        args = perturbed_case(int(float(case[3:]) * 10), float(case[3:]))
    got = oracle_refine(oracle_lib, *args)
    want = _nversion(*args)
    _same(case, got, want)
    assert got[2].verified_fraction > 0.2, case
