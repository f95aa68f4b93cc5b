"""GPU path against the CPU oracle on the same seeded captures: every
integer output (labels, flags, provenance, decision codes) and every float
output (scores, costs, aggregated SGM volume, depths, confidences, warp
points) must agree bit for bit."""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import DepthOptions, Session
from oracle import binding

from parity import assert_bitwise, assert_nodes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small_config0(seed, scale=0.25, **kw):
    w, h = int(1040 * scale) // 8 * 8, int(780 * scale) // 8 * 8
    return scene.config0(seed=seed, width=w, height=h, focal_px=1600.0 * scale, **kw)


def _full_parity(ctx, cap, label, threads=8):
    _say(f"parity case {label}: {cap.description}")
    sess = Session(ctx, cap.modules, cap.reference_index, cap.options)
    try:
        sess.run(_abi.CHD_STAGE_ALL | _abi.CHD_STAGE_KEEP_AGGREGATED)
        ctx.synchronize()
        dmap, targets = sess.results()
        scores, costs, guidance, tested = sess.cost_volume()
        aggregated = sess.aggregated()
        coarse, fillable, best = sess.coarse()
        direct = [sess.direct(k) for k in range(4)]
    finally:
        sess.close()
    omap, otargets, dbg = binding.refine(cap, threads=threads)
    assert_bitwise("scores", scores, dbg["scores"])
    assert_bitwise("costs", costs, dbg["costs"])
    assert_bitwise("guidance", guidance, dbg["guidance"])
    assert_bitwise("tested", tested, dbg["tested"])
    assert_bitwise("aggregated", aggregated, dbg["aggregated"])
    assert_nodes("coarse", coarse, dbg["coarse"])
    assert_bitwise("fillable", fillable, dbg["fillable"])
    assert_bitwise("best_label", best, dbg["best_label"])
    for k in range(4):
        assert_nodes(f"direct[{k}]", direct[k][0], dbg["direct"][k])
    assert_bitwise("fine_guidance", direct[0][1], dbg["fine_guidance"])
    assert_bitwise("fine_tested", direct[0][2], dbg["fine_tested"])
    assert_bitwise("map.depth", dmap.depth, omap["depth"])
    assert_bitwise("map.confidence", dmap.confidence, omap["confidence"])
    assert_bitwise("map.provenance", dmap.provenance, omap["provenance"])
    assert sorted(targets) == sorted(otargets)
    for t in targets:
        assert_bitwise(f"target{t}.decisions", targets[t].decisions, otargets[t]["decisions"])
        assert_bitwise(f"target{t}.points", targets[t].points, otargets[t]["points"])
        assert_bitwise(f"target{t}.confidence", targets[t].confidence, otargets[t]["confidence"])
        for key, want in otargets[t]["report"].items():
            got = targets[t].report[key]
            if key in ("median_depth", "median_score_improvement"):
                if not otargets[t]["report"]["has_" + key]:
                    assert got is None
                    continue
            assert (got == want) or (isinstance(want, float) and math.isnan(want) and math.isnan(got)), \
                f"target {t} report {key}: {got} != {want}"
    return dbg


def test_device_expf_exhaustive_on_the_used_domain(gpu_context, oracle_lib):
    """Every argument the kernels give expf is -d^2 / (2 sigma^2), -1.2 |dv|
    or -3 e: zero or negative, so the domain is +0 and every float with the
    sign bit set (2^31 + 1 values, negative NaNs and -inf included). The
    device expf equals the oracle's or_expf on all of them, and or_expf equals
    the platform expf on all 2^32 floats (test_expf_exhaustive_against_libm)."""
    import os
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    chunk = 1 << 27
    zero = np.zeros(1, np.float32)
    y0 = np.empty(1, np.float32)
    _abi.check(gpu_context.lib.chd_device_expf(gpu_context.handle, _abi.ptr(zero), _abi.ptr(y0), 1),
               "expf")
    assert oracle_lib.or_expf_compare(0, 1, _abi.ptr(y0), 1) == 0
    bits_ = np.empty(chunk, np.uint32)
    y = np.empty(chunk, np.float32)
    total = 0
    for lo in range(0x80000000, 1 << 32, chunk):
        np.add(np.arange(chunk, dtype=np.uint32), np.uint32(lo), out=bits_)
        _abi.check(gpu_context.lib.chd_device_expf(gpu_context.handle, _abi.ptr(bits_),
                                                   _abi.ptr(y), chunk), "expf")
        bad = oracle_lib.or_expf_compare(lo, chunk, _abi.ptr(y), threads)
        assert bad == 0, f"{bad} mismatches in bits {lo:#x}..{lo + chunk - 1:#x}"
        total += chunk
    _say(f"expf: {total + 1} inputs, all equal to the oracle")


def test_device_exp_matches_libm(gpu_context, oracle_lib):
    _say("exp parity on [-1.5, 0]")
    rng = np.random.default_rng(3)
    x = -1.5 * rng.random(100000)
    x = np.concatenate([x, -2.5 * rng.random(20000).astype(np.float32).astype(np.float64),
                        -np.array([0, 1, 2, 4, 5, 8], dtype=np.float64) / 6.0])
    y = np.empty_like(x)
    _abi.check(gpu_context.lib.chd_device_exp(gpu_context.handle, _abi.ptr(x), _abi.ptr(y), x.size), "exp")
    want = np.array([oracle_lib.or_libm_exp(float(v)) for v in x])
    assert_bitwise("exp", y, want)


@pytest.mark.parametrize("seed", [0, 1])
def test_config0_quarter_bit_exact(gpu_context, seed):
    cap = _small_config0(seed)
    _full_parity(gpu_context, cap, f"config0 quarter seed {seed}")


def test_config0_half_bit_exact(gpu_context):
    cap = _small_config0(7, scale=0.5)
    _full_parity(gpu_context, cap, "config0 half seed 7")


def test_config1_small_bit_exact(gpu_context):
    cap = scene.config1(seed=2, width=416, height=312, focal_px=330.0, planes=32)
    _full_parity(gpu_context, cap, "config1 416x312 32 planes")


def test_completion_and_options_bit_exact(gpu_context):
    cap = _small_config0(4)
    cap.options = DepthOptions(near_depth=800.0, far_depth=20000.0, planes=40, patch_radius=3,
                               completion_iterations=3, best_view_count=2, grid_step=12)
    _full_parity(gpu_context, cap, "config0 options variant")


def test_odd_fine_step_bit_exact(gpu_context):
    """Grid step 10: the fine grid's step is 5, so the direct and target
    kernels cannot share a 2-px tap lattice across a tile and build each
    node's taps on their own; results stay bit-exact."""
    cap = _small_config0(5)
    cap.options = DepthOptions(grid_step=10, patch_radius=3)
    _full_parity(gpu_context, cap, "config0 fine step 5")


def test_sgm_stage_on_oracle_costs(gpu_context, oracle_lib):
    _say("sgm stage on random costs")
    rng = np.random.default_rng(11)
    # wide, tall, square and single-column grids: every path order and the
    # diagonal families' two length classes
    for columns, rows, labels in [(5, 1, 4), (17, 9, 65), (31, 23, 129), (7, 40, 257),
                                  (1, 6, 4), (12, 12, 64), (3, 50, 129)]:
        costs = rng.random((rows * columns, labels), dtype=np.float32) * 2.25
        guidance = rng.random(rows * columns, dtype=np.float32) * 3
        got = np.empty_like(costs)
        _abi.check(gpu_context.lib.chd_semi_global_costs(gpu_context.handle, _abi.ptr(costs),
                                                         _abi.ptr(guidance), columns, rows, labels,
                                                         _abi.ptr(got)), "sgm")
        want = np.empty_like(costs)
        oracle_lib.or_semi_global_costs(_abi.ptr(costs), _abi.ptr(guidance), columns, rows, labels,
                                        _abi.ptr(want))
        assert_bitwise(f"sgm {columns}x{rows}x{labels}", got, want)


def _camera_variants(oracle_lib, cap):
    """Replace the target cameras by a distorted canonical camera, a mirror
    camera with a flipped axis and a camera with an orientation correction
    (geometry.rs:64-133, 143-243), with global warps from the oracle's
    align_module(refine = false)."""
    ref = cap.modules[cap.reference_index]
    w, h = ref.width, ref.height
    f = ref.camera.k[0]
    k = np.array([f, 0, (w - 1) / 2, 0, f, (h - 1) / 2, 0, 0, 1], np.float64)
    cams = []
    # distorted canonical camera, 20 mm to the right
    d = _abi.Distortion()
    d.center[:] = [(w - 1) / 2 + 3.0, (h - 1) / 2 - 2.0]
    d.normalization[:] = [f, f]
    d.coeffs[:] = [0.02, -0.05, 0.0005, -0.0004, 0.01]
    d.coeff_count = 5
    c = _abi.Camera()
    t = np.array([-20.0, 0.0, 0.0])
    assert oracle_lib.or_camera_canonical(k.ctypes.data, np.eye(3).ctypes.data, t.ctypes.data, w, h,
                                          C.byref(d), C.byref(c))
    cams.append(c)
    # mirror camera: real module looking along -x, folded by a 45 deg mirror
    m = _abi.Camera()
    m.width, m.height = w, h
    for i in range(9):
        m.k[i] = k[i]
    kinv = np.empty(9)
    oracle_lib.or_inverse3(k.ctypes.data, kinv.ctypes.data)
    for i in range(9):
        m.k_inverse[i] = kinv[i]
    m.flip_around_x = 1
    m.pose_kind = 1
    real_cw = np.array([[0, 0, -1], [0, 1, 0], [1, 0, 0]], np.float64)  # camera z -> world -x
    n = np.array([1.0, 0.0, 1.0]) / math.sqrt(2.0)
    reflect = np.eye(3) - 2.0 * np.outer(n, n)
    for i in range(9):
        m.real_cw[i] = real_cw.reshape(-1)[i]
        m.reflect[i] = reflect.reshape(-1)[i]
    m.virtual_center[:] = [17.5, 30.311, 0.0]
    for i in range(9):
        m.orientation_correction[i] = np.eye(3).reshape(-1)[i]
    m.focal_px = f
    cams.append(m)
    # canonical camera with an orientation correction of ~0.2 deg
    o = _abi.Camera()
    t = np.array([0.0, -15.0, 0.0])
    assert oracle_lib.or_camera_canonical(k.ctypes.data, np.eye(3).ctypes.data, t.ctypes.data, w, h,
                                          None, C.byref(o))
    a = math.radians(0.2)
    rot = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    for i in range(9):
        o.orientation_correction[i] = rot.reshape(-1)[i]
    cams.append(o)
    return cams


def _oracle_warp(oracle_lib, ref_cam, cam, w, h, step=32):
    columns, rows = -(-w // step) + 1, -(-h // step) + 1
    pts = np.empty((rows, columns, 2), np.float32)
    conf = np.empty((rows, columns), np.float32)
    wp = _abi.Warp()
    wp.points = pts.ctypes.data_as(C.POINTER(C.c_float))
    wp.confidence = conf.ctypes.data_as(C.POINTER(C.c_float))
    oracle_lib.or_global_warp(C.byref(ref_cam), C.byref(cam), w, h, step, C.byref(wp))
    return pts


def test_camera_model_branches_bit_exact(gpu_context, oracle_lib):
    cap = _small_config0(5)
    ref = cap.modules[0]
    cams = _camera_variants(oracle_lib, cap)
    base = cap.modules[1]
    mods = [ref]
    for i, cam in enumerate(cams):
        src = cap.modules[1 + (i % 2)]
        mods.append(scene.ModuleInput(name=f"V{i}", luminance=src.luminance, width=ref.width,
                                      height=ref.height, camera=cam,
                                      warp_points=_oracle_warp(oracle_lib, ref.camera, cam, ref.width, ref.height),
                                      warp_step=32))
    cap.modules = mods
    _full_parity(gpu_context, cap, "distortion / mirror / orientation cameras")


def test_warp_steps_bit_exact(gpu_context, oracle_lib):
    """Global warps on grids of 24 and 16 px (Warp::map divides by the step:
    the kernels multiply by the reciprocal only for power-of-two steps)."""
    for step in (24, 16):
        cap = _small_config0(7)
        ref = cap.modules[0]
        for m in cap.modules[1:]:
            m.warp_points = _oracle_warp(oracle_lib, ref.camera, m.camera, ref.width, ref.height,
                                         step)
            m.warp_step = step
        _full_parity(gpu_context, cap, f"global warps with a {step} px grid")


def test_skipped_views_and_undefined_warps_bit_exact(gpu_context):
    """A rejected alignment, an uncalibrated module and NaN (undefined) warp
    regions are skipped exactly as the reference skips them."""
    cap = scene.config1(seed=8, width=416, height=312, focal_px=330.0, planes=24)
    cap.modules[1].accepted = False
    cap.modules[2].camera = None
    w = cap.modules[3].warp_points.copy()
    w[: w.shape[0] // 3, : w.shape[1] // 2] = np.nan
    cap.modules[3].warp_points = w
    _full_parity(gpu_context, cap, "skipped views and undefined warps")


def test_none_cases(gpu_context):
    from chiaro_amd.depth import refine_multiview_depth
    cap = _small_config0(2, scale=0.125)
    for bad in [dict(planes=2), dict(far_depth=100.0), dict(best_view_count=1), dict(enabled=False)]:
        o = DepthOptions(near_depth=800.0, far_depth=20000.0, planes=16)
        for key, value in bad.items():
            setattr(o, key, value)
        dmap, targets = refine_multiview_depth(cap.modules, 0, o, gpu_context)
        assert dmap is None and targets == {}
    cap.modules[0].camera = None
    dmap, targets = refine_multiview_depth(cap.modules, 0, cap.options, gpu_context)
    assert dmap is None
    dmap, targets = refine_multiview_depth(cap.modules, 5, cap.options, gpu_context)
    assert dmap is None


def test_drop_in_matches_session(gpu_context):
    """chd_refine_multiview_depth (host buffers in and out) equals the
    session form and the oracle."""
    from chiaro_amd.depth import refine_multiview_depth
    cap = _small_config0(9)
    dmap, targets = refine_multiview_depth(cap.modules, 0, cap.options, gpu_context)
    omap, otargets, _ = binding.refine(cap)
    assert_bitwise("drop-in depth", dmap.depth, omap["depth"])
    assert_bitwise("drop-in provenance", dmap.provenance, omap["provenance"])
    for t in otargets:
        assert_bitwise(f"drop-in target {t}", targets[t].points, otargets[t]["points"])


@pytest.mark.parametrize("shape,black,white,headroom", [
    ((3120, 4160), 64.0 * 16, 4095.0 * 16, False),  # config 1 module size, 12-bit codes << 4
    ((781, 1043), 1024.0, 65520.0, False),          # odd sizes
    ((64, 96), 1024.0, 65520.0, True),              # headroom halving with wrapping
    ((16, 16), 5000.0, 100.0, False),               # white below black
])
def test_luminance_half_bit_exact(gpu_context, oracle_lib, shape, black, white, headroom):
    from chiaro_amd import luminance_half
    _say(f"luminance_half {shape} black {black} white {white} headroom {headroom}")
    rng = np.random.default_rng(11)
    samples = rng.integers(0, 65536, size=shape, dtype=np.uint16)
    samples[0, :8] = [65535, 0, 65535, 65535, 0, 0, 1, 65534]
    h, w = shape
    got = luminance_half(gpu_context, samples, black, white, headroom)
    want = np.empty((h // 2, w // 2), np.float32)
    oracle_lib.or_luminance_half_u16(samples.ctypes.data, w, h, black, white, int(headroom),
                                     want.ctypes.data)
    assert_bitwise("luminance_half", got, want)


def test_luminance_half_feeds_the_path(gpu_context):
    """The GPU planes of a config 1 crop equal the host-prepared ones the
    parity captures use."""
    from chiaro_amd import luminance_half
    cap = scene.config1(seed=2, width=416, height=312, focal_px=330.0)
    for k, raw in enumerate(cap.raws):
        got = luminance_half(gpu_context, raw << 4, 64.0 * 16, 4095.0 * 16)
        assert_bitwise(f"module {k} plane", got, cap.modules[k].luminance)
