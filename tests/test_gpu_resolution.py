"""refine_resolution_warp on the GPU (chd_refine_resolution_warp) against the
oracle (oracle/align_oracle.c) bit for bit: warp points, confidences and
every report field, for the reference's own sub-pixel shift, a target at
twice the reference's scale (Gaussian footprint, magnified search radius),
a flat hole filled by completion, perturbed base warps with undefined nodes
at magnifications 0.7, 1.25 and 3, textureless planes, and full-size
config 4 modules (2080x1560 planes: a wide, a mid and a tele module seen
through their factory warps)."""
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene

from parity import assert_bitwise
from test_resolution_oracle import (hole_case, magnified_case, oracle_refine, perturbed_case,
                                    refine, shift_case)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gpu(ctx, *args):
    def call(*a):
        rc = ctx.lib.chd_refine_resolution_warp(ctx.handle, *a)
        assert rc == 0, _abi.last_error()
        return rc
    return refine(call, *args)


def _parity(ctx, oracle_lib, label, args):
    got = _gpu(ctx, *args)
    want = oracle_refine(oracle_lib, *args)
    assert_bitwise(f"{label} points", got[0], want[0])
    assert_bitwise(f"{label} confidence", got[1], want[1])
    for name, _ in _abi.ResolutionReport._fields_:
        a, b = getattr(got[2], name), getattr(want[2], name)
        assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32), \
            f"{label} report {name}: {a} != {b}"
    r = got[2]
    _say(f"{label}: verified {r.verified_fraction:.3f}, supported {r.supported_fraction:.3f}, "
         f"median correction {r.median_correction_px:.3f} px, "
         f"mean confidence {r.mean_confidence:.3f}")
    return got


@pytest.mark.parametrize("case", ["shift", "magnified", "magnified4.2", "hole", "mag0.7",
                                  "mag1.25", "mag3", "flat"])
def test_small_cases_bit_exact(gpu_context, oracle_lib, case):
    if case == "shift":
        args = shift_case()[:6]
    elif case == "magnified":
        args = magnified_case()[:6]
    elif case == "magnified4.2":
        args = magnified_case(shift=(4.2, 0.0))[:6]
    elif case == "hole":
        args = hole_case()[:6]
    elif case == "flat":
        flat = np.full((96, 128), 0.5, np.float32)
        ys, xs = np.meshgrid(np.arange(7), np.arange(9), indexing="ij")
        bp = np.stack([xs * 32.0 + 3.0, ys * 32.0 - 1.0], -1).astype(np.float32)
        args = (flat, flat.copy(), bp, np.ones((7, 9), np.float32), 256, 192)
    else:
        args = perturbed_case(int(float(case[3:]) * 10), float(case[3:]))
    got = _parity(gpu_context, oracle_lib, case, args)
    if case != "flat":
        assert got[2].verified_fraction > 0.2


def test_full_size_config4_modules_bit_exact(gpu_context, oracle_lib):
    """The reference's call (fusion.rs:384): the reference plane, one
    module's plane and its factory warp, at the full 4160x3120 raster."""
    rig = scene.rig_config4(seed=4)
    frames = scene.Frames(gpu_context, rig)
    try:
        frames.render(4)
        cap = frames.capture()
        ref_plane = frames.read(0)[1]
        for index in (1, 5, 10):     # wide W1, mid M0, tele T0
            m = cap.modules[index]
            tgt_plane = frames.read(index)[1]
            conf = np.where(np.isnan(m.warp_points[..., 0]), 0.0, 1.0).astype(np.float32)
            args = (ref_plane, tgt_plane, m.warp_points, conf, rig.modules[0].width,
                    rig.modules[0].height)
            _say(f"config4 module {rig.names[index]}: {tgt_plane.shape[1]}x{tgt_plane.shape[0]} plane")
            _parity(gpu_context, oracle_lib, f"config4 {rig.names[index]}", args)
    finally:
        frames.close()


def test_align_then_refine_through_the_python_mirror(gpu_context, oracle_lib):
    """The reference's chain (pipeline.rs:1185): align_module, then
    refine_resolution_warp through the warp it produced, via chiaro_amd.align;
    both steps equal the oracle's."""
    from chiaro_amd import align
    from test_align_oracle import rotated_pair
    from test_gpu_align import _options, _run

    cap, _ = rotated_pair(832, 624, 660.0)
    ref, tgt = cap.modules[0], cap.modules[1]
    got = align.align_module(gpu_context, cap.modules, 0, 1)
    want = _run(oracle_lib.or_align_module, cap, 1, _options(gpu_context.lib))
    assert_bitwise("aligned warp", got.warp.points, want[0])
    assert got.report["status"] == "refined" and got.accepted
    ident = align.align_module(gpu_context, cap.modules, 0, 0)
    assert ident.report["status"] == "reference"
    rr = align.refine_resolution_warp(gpu_context, ref.luminance, tgt.luminance, got.warp,
                                      ref.width, ref.height)
    w = oracle_refine(oracle_lib, ref.luminance, tgt.luminance, want[0], want[1], ref.width,
                      ref.height)
    assert_bitwise("resolution warp points", rr.warp.points, w[0])
    assert_bitwise("resolution warp confidence", rr.warp.confidence, w[1])
    assert rr.report["verified_fraction"] == w[2].verified_fraction
    _say(f"align -> resolution: {got.report['status']}, verified "
         f"{rr.report['verified_fraction']:.3f}")


def test_device_planes_equal_host_planes(gpu_context, oracle_lib):
    """Planes already in HBM (chd_frames) give the host planes' alignment and
    resolution warp bit for bit, and both equal the oracle."""
    from chiaro_amd import align
    from test_align_oracle import rotation_y
    from test_gpu_align import _options, _run

    rig = scene.rig_config1(seed=9, width=832, height=624, focal_px=660.0, planes=32)
    frames = scene.Frames(gpu_context, rig)
    try:
        frames.render(9)
        device = frames.planes()
        host = [frames.read(i)[1] for i in range(len(rig.modules))]
        m = rig.modules[2]
        turned = scene._module(m.focal_px, m.width, m.height, tuple(m.center[:]),
                               rotation_y(0.3))
        cams = [scene.camera_for(rig.modules[0]), scene.camera_for(turned)]
        mods = {kind: [scene.ModuleInput(name=str(i), luminance=lum[k], width=m.width,
                                         height=m.height, camera=cams[i],
                                         warp_points=scene.identity_warp(m.width, m.height),
                                         warp_step=32)
                       for i, k in enumerate((0, 2))]
                for kind, lum in (("device", device), ("host", host))}
        got = align.align_module(gpu_context, mods["device"], 0, 1)
        want = align.align_module(gpu_context, mods["host"], 0, 1)
        assert_bitwise("device-plane alignment", got.warp.points, want.warp.points)
        assert got.gain == want.gain and got.report["status"] == want.report["status"]
        cap = scene.Capture(modules=mods["host"], reference_index=0, options=None)
        oracle = _run(oracle_lib.or_align_module, cap, 1, _options(gpu_context.lib))
        assert_bitwise("alignment vs oracle", got.warp.points, oracle[0])
        rd = align.refine_resolution_warp(gpu_context, device[0], device[2], got.warp,
                                          m.width, m.height)
        rh = align.refine_resolution_warp(gpu_context, host[0], host[2], want.warp,
                                          m.width, m.height)
        assert_bitwise("device-plane resolution warp", rd.warp.points, rh.warp.points)
        assert_bitwise("device-plane resolution confidence", rd.warp.confidence,
                       rh.warp.confidence)
        _say(f"device planes: {got.report['status']}, gain {got.gain:.4f}, verified "
             f"{rd.report['verified_fraction']:.3f}")
    finally:
        frames.close()
