"""The synthesis stage's oracle (oracle/synth_oracle.c: synthesize's depth
and warp part, synth.rs:591-725, and DenseDepthMap::sample_nearest,
depth.rs:152-159) pinned three ways:

- the reference's own test of the focus prior (synth.rs:2034-2046), run
  through the whole stage;
- closed forms: an identity warp maps canvas pixel (u, v) to (u, v) exactly,
  the feather is exactly 1 beyond feather_px from the mosaic border and
  exactly 0.5 at half of it, the base weight is feather x local confidence
  x magnification^2 x confidence, the nearest node is the brute-force
  nearest, a source beyond every band contributes nowhere, and invalid
  canvases are refused;
- an independent numpy restatement (tests/nversion_synth.py) agreeing bit
  for bit on canvases with crops, magnified scales, undefined warp nodes,
  confidences above 1, depth maps with holes and focus distances.
"""
import ctypes as C

import numpy as np
import pytest

from chiaro_amd import _abi

import nversion_synth as nv

F = np.float32


def warp_arrays(width, height, step, fn):
    columns, rows = -(-width // step) + 1, -(-height // step) + 1
    ys, xs = np.meshgrid(np.arange(rows) * float(step), np.arange(columns) * float(step),
                         indexing="ij")
    qx, qy = fn(xs, ys)
    points = np.ascontiguousarray(np.stack([qx, qy], -1).astype(np.float32))
    conf = np.where(np.isnan(points[..., 0]), 0.0, 1.0).astype(np.float32)
    return points, conf


class Stage:
    """Inputs of one synthesis-geometry call, as ctypes and as N-version
    objects."""

    def __init__(self, sources, crop, scale, feather_px, depth=None):
        self.keep = []
        self.sources = (_abi.SynthSource * len(sources))()
        self.nv_sources = []
        for i, s in enumerate(sources):
            points = np.ascontiguousarray(s["points"], np.float32)
            conf = None if s.get("confidence") is None else \
                np.ascontiguousarray(s["confidence"], np.float32)
            self.keep += [points, conf]
            d = self.sources[i]
            d.warp.step = s["step"]
            d.warp.rows, d.warp.columns = points.shape[:2]
            d.warp.points = _abi.fptr(points)
            d.warp.confidence = _abi.fptr(conf) if conf is not None else None
            d.mosaic_width, d.mosaic_height = s["mosaic"]
            d.magnification = s.get("magnification", 1.0)
            d.confidence = s.get("weight", 1.0)
            d.reference = int(s.get("reference", False))
            focus = s.get("focus")
            d.has_focus_distance = int(focus is not None)
            d.focus_distance = focus if focus is not None else 0.0
            self.nv_sources.append(nv.Source(s["step"], points, conf, s["mosaic"][0],
                                             s["mosaic"][1], d.magnification, d.confidence,
                                             bool(d.reference), focus))
        self.canvas = _abi.Canvas(crop[0], crop[1], crop[2], crop[3], scale, feather_px)
        self.crop, self.scale, self.feather_px = crop, scale, feather_px
        self.map = None
        self.nv_map = None
        if depth is not None:
            d, c, step = depth
            d = np.ascontiguousarray(d, np.float64)
            c = np.ascontiguousarray(c, np.float32)
            prov = np.zeros(d.shape, np.uint8)
            self.keep += [d, c, prov]
            m = _abi.DepthMap()
            m.rows, m.columns = d.shape
            m.step = step
            m.depth = d.ctypes.data_as(C.POINTER(C.c_double))
            m.confidence = _abi.fptr(c)
            m.provenance = prov.ctypes.data_as(C.POINTER(C.c_uint8))
            self.map = m
            self.nv_map = (d, c, step)

    def size(self, lib):
        w, h = C.c_uint32(), C.c_uint32()
        assert lib.or_canvas_size(C.byref(self.canvas), C.byref(w), C.byref(h)) == 0
        return w.value, h.value

    def run(self, fn, size):
        w, h = size
        n = len(self.nv_sources)
        points = np.empty((n, h, w, 2), np.float32)
        outs = [np.empty((n, h, w), np.float32) for _ in range(4)]
        rc = fn(C.byref(self.map) if self.map is not None else None, self.sources, n,
                C.byref(self.canvas), points.ctypes.data, *[o.ctypes.data for o in outs])
        assert rc == 0
        return (points, *outs)

    def oracle(self, lib):
        return self.run(lib.or_synthesis_geometry, self.size(lib))

    def nversion(self):
        return nv.synthesis_geometry(self.nv_map, self.nv_sources, self.crop, self.scale,
                                     self.feather_px)


def identity_source(width, height, step=32, **kw):
    points, conf = warp_arrays(width, height, step, lambda x, y: (x, y))
    return dict(step=step, points=points, confidence=conf, mosaic=(width, height), **kw)


def test_reference_focus_prior(oracle_lib):
    """synth.rs:2034-2046 through the whole stage: close content (500 mm,
    confidence 1) under a 2500 mm focus is rejected for a 2.25x source only."""
    depth = (np.full((5, 5), 500.0), np.ones((5, 5), np.float32), 8)

    def weight(magnification, depth_map, focus=2500.0):
        st = Stage([identity_source(64, 64, magnification=magnification, focus=focus)],
                   (0, 0, 32, 32), 1.0, 4.0, depth_map)
        return st.oracle(oracle_lib)[4][0, 16, 16]

    assert weight(2.25, depth) < 0.01
    assert weight(1.0, depth) == 1.0
    far = (np.full((5, 5), 10000.0), np.ones((5, 5), np.float32), 8)
    assert weight(2.25, far) > 0.97
    assert weight(2.25, None) == 1.0
    # no focus distance, or a reference source: no prior
    assert weight(2.25, depth, focus=None) == 1.0


def test_closed_forms(oracle_lib):
    W, H = 96, 64
    st = Stage([identity_source(W, H, weight=0.8)], (0, 0, W, H), 1.0, 8.0)
    points, feather, local, base, focus = st.oracle(oracle_lib)
    vs, us = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    live = ~np.isnan(points[0, ..., 0])
    # the border pixels have feather 0 and do not contribute
    assert not live[0].any() and not live[:, 0].any() and not live[-1].any()
    np.testing.assert_array_equal(points[0, live, 0], us[live].astype(np.float32))
    np.testing.assert_array_equal(points[0, live, 1], vs[live].astype(np.float32))
    border = np.minimum(np.minimum(us, vs), np.minimum(W - 1 - us, H - 1 - vs))
    assert (feather[0][border >= 8] == 1.0).all()
    assert (feather[0][border == 4] == 0.5).all()
    assert (local[0][live] == 1.0).all()
    np.testing.assert_array_equal(base[0][live], (feather[0][live] * F(0.8)).astype(np.float32))
    assert (focus[0][live] == 1.0).all()


def test_sample_nearest_is_the_nearest_node(oracle_lib):
    rng = np.random.default_rng(1)
    rows, columns, step = 7, 9, 4
    d = np.ascontiguousarray(rng.uniform(500, 5000, (rows, columns)))
    c = np.ones((rows, columns), np.float32)
    prov = np.zeros((rows, columns), np.uint8)
    m = _abi.DepthMap()
    m.rows, m.columns, m.step = rows, columns, step
    m.depth = d.ctypes.data_as(C.POINTER(C.c_double))
    m.confidence = _abi.fptr(c)
    m.provenance = prov.ctypes.data_as(C.POINTER(C.c_uint8))
    idx = C.c_size_t()
    for x, y in rng.uniform(-3, 40, (400, 2)).astype(np.float32):
        got = oracle_lib.or_sample_nearest(C.byref(m), float(x), float(y), C.byref(idx))
        if x < 0 or y < 0:
            assert not got
            continue
        cols = np.abs(np.arange(columns) * step - float(x))
        rws = np.abs(np.arange(rows) * step - float(y))
        if np.ptp(np.sort(cols)[:2]) < 1e-4 or np.ptp(np.sort(rws)[:2]) < 1e-4:
            continue   # a tie: f32::round goes away from zero
        cc, rr = int(np.argmin(cols)), int(np.argmin(rws))
        inside = float(x) < (columns - 0.5) * step and float(y) < (rows - 0.5) * step
        assert bool(got) == inside, (x, y)
        if got:
            assert idx.value == rr * columns + cc


def test_far_source_and_invalid_canvas(oracle_lib):
    far = identity_source(64, 64)
    far["points"] = (far["points"] + 1.0e5).astype(np.float32)
    st = Stage([identity_source(64, 64), far], (0, 0, 64, 64), 1.0, 4.0)
    points, feather, _, base, _ = st.oracle(oracle_lib)
    assert np.isnan(points[1]).all() and (base[1] == 0).all() and (feather[1] == 0).all()
    assert (base[0] > 0).any()
    for crop, scale in (((0, 0, 64, 64), 0.0), ((0, 0, 64, 64), float("nan")),
                        ((0, 0, 0.5, 64), 1.0)):
        bad = Stage([identity_source(64, 64)], crop, scale, 4.0)
        w, h = C.c_uint32(), C.c_uint32()
        assert oracle_lib.or_canvas_size(C.byref(bad.canvas), C.byref(w), C.byref(h)) != 0


def stage_case(name):
    rng = np.random.default_rng(sum(map(ord, name)))
    W, H = 320, 240
    ref = identity_source(W, H, reference=True)
    tele_points, tele_conf = warp_arrays(
        W, H, 16, lambda x, y: (2.2 * (x - 160) + 310 + 0.8 * np.sin(y / 40.0),
                                2.2 * (y - 120) + 230 + 0.6 * np.cos(x / 50.0)))
    tele_points[:3, :4] = np.nan
    tele_conf = rng.uniform(0.0, 1.2, tele_conf.shape).astype(np.float32)
    tele = dict(step=16, points=tele_points, confidence=tele_conf, mosaic=(640, 480),
                magnification=2.2, weight=0.7, focus=1500.0)
    side_points, side_conf = warp_arrays(W, H, 24, lambda x, y: (0.9 * x + 40.0, 0.9 * y - 12.0))
    side = dict(step=24, points=side_points, confidence=side_conf, mosaic=(W, H),
                magnification=0.9, weight=0.95, focus=None)
    nodes = (-(-H // 4) + 1, -(-W // 4) + 1)
    depth = rng.uniform(300.0, 20000.0, nodes)
    depth[rng.random(nodes) < 0.2] = np.nan
    dmap = (depth, rng.uniform(0.0, 1.1, nodes).astype(np.float32), 4)
    cases = {
        "full": ([ref, tele, side], (0, 0, W, H), 1.0, 30.0, dmap),
        "crop_magnified": ([ref, tele, side], (100.5, 70.25, 90, 60), 2.5, 12.0, dmap),
        "reduced_no_map": ([ref, tele, side], (0, 0, W, H), 0.37, 0.5, None),
        "tele_only": ([tele], (20, 10, 260, 200), 1.3, 60.0, dmap),
    }
    return cases[name]


@pytest.mark.parametrize("case", ["full", "crop_magnified", "reduced_no_map", "tele_only"])
def test_nversion_agrees_bit_for_bit(oracle_lib, case):
    st = Stage(*stage_case(case))
    got = st.oracle(oracle_lib)
    want = st.nversion()
    for label, g, w in zip(("points", "feather", "local", "base", "focus"), got, want):
        np.testing.assert_array_equal(g.view(np.uint32), w.view(np.uint32),
                                      err_msg=f"{case} {label}")
    assert (got[3] > 0).sum() > 100, case
