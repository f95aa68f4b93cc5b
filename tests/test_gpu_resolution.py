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
