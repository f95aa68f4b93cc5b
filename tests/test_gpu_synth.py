"""The synthesis consumer's depth and warp stage on the GPU
(chd_synthesis_geometry) against the oracle (oracle/synth_oracle.c) bit for
bit: warp points, feathers, local confidences, base and focus weights, on
the oracle tests' synthetic cases and on the real outputs of the depth path
(its dense depth map and the refined per-target warps), from a crop of a
small capture up to a full-size config 1 capture's canvas."""
import ctypes as C
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.align import Warp
from chiaro_amd.depth import Session
from chiaro_amd.synth import CropWindow, SynthSource, synthesis_geometry

from parity import assert_bitwise
from test_synth_oracle import Stage, stage_case

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(label, got, want):
    for name, g, w in zip(("points", "feather", "local", "base", "focus"), got, want):
        assert_bitwise(f"{label} {name}", g, w)


@pytest.mark.parametrize("case", ["full", "crop_magnified", "reduced_no_map", "tele_only"])
def test_synthetic_cases_bit_exact(gpu_context, oracle_lib, case):
    st = Stage(*stage_case(case))
    size = st.size(oracle_lib)
    got = st.run(lambda *a: gpu_context.lib.chd_synthesis_geometry(gpu_context.handle, *a), size)
    _same(case, got, st.oracle(oracle_lib))
    _say(f"{case}: {size[0]}x{size[1]} canvas, {int((got[3] > 0).sum())} weighted samples")


def _sources(cap, targets, magnification=None, focus=None):
    """The capture's modules as synthesis sources: the reference through its
    identity warp, every replaced target through its refined warp."""
    out = []
    for i, m in enumerate(cap.modules):
        if i == cap.reference_index:
            pts = scene.identity_warp(m.width, m.height, 32)
            warp = Warp(32, pts, np.ones(pts.shape[:2], np.float32))
        elif i in targets:
            t = targets[i]
            warp = Warp(4, t.points, t.confidence)
        else:
            continue
        out.append(SynthSource(warp, m.width, m.height,
                               magnification=(magnification or {}).get(i, 1.0),
                               confidence=0.9, reference=i == cap.reference_index,
                               focus_distance=(focus or {}).get(i)))
    return out


def _oracle(oracle_lib, dmap, sources, crop, scale, feather_px):
    """The same call through the oracle, from the chiaro_amd.synth inputs."""
    cases = [dict(step=s.warp.step, points=s.warp.points, confidence=s.warp.confidence,
                  mosaic=(s.mosaic_width, s.mosaic_height), magnification=s.magnification,
                  weight=s.confidence, reference=s.reference, focus=s.focus_distance)
             for s in sources]
    depth = None if dmap is None else (dmap.depth, dmap.confidence, dmap.step)
    st = Stage(cases, (crop.x, crop.y, crop.width, crop.height), scale, feather_px, depth)
    return st.oracle(oracle_lib)


def _check(gpu_context, oracle_lib, label, dmap, sources, crop, scale, feather_px):
    got = synthesis_geometry(gpu_context, dmap, sources, crop, scale, feather_px)
    want = _oracle(oracle_lib, dmap, sources, crop, scale, feather_px)
    _same(label, (got.points, got.feather, got.local_confidence, got.base_weight,
                  got.focus_weight), want)
    live = got.base_weight > 0
    _say(f"{label}: {got.base_weight.shape[2]}x{got.base_weight.shape[1]} canvas, "
         f"{len(sources)} sources, weighted fraction {live.mean():.3f}, focus-suppressed "
         f"{int((got.focus_weight[live] < 0.5).sum())}")
    return got


def test_depth_path_outputs_bit_exact(gpu_context, oracle_lib):
    cap = scene.config1(seed=4, width=416, height=312, focal_px=330.0, planes=32)
    sess = Session(gpu_context, cap.modules, 0, cap.options)
    try:
        sess.run(_abi.CHD_STAGE_ALL)
        dmap, targets = sess.results()
    finally:
        sess.close()
    assert dmap is not None and targets
    sources = _sources(cap, targets, magnification={1: 1.6, 3: 2.4},
                       focus={1: 1200.0, 3: 3000.0})
    ref = cap.modules[0]
    _check(gpu_context, oracle_lib, "config1 crop full", dmap, sources,
           CropWindow.full(ref.width, ref.height), 1.0, 40.0)
    _check(gpu_context, oracle_lib, "config1 centred 0.4 at 2.5x", dmap, sources,
           CropWindow.centred(ref.width, ref.height, 0.4), 2.5, 120.0)
    _check(gpu_context, oracle_lib, "config1 without depth", None, sources,
           CropWindow.full(ref.width, ref.height), 0.5, 10.0)


def test_full_size_config1_canvas_bit_exact(gpu_context, oracle_lib):
    """A full 5 x 4160x3120 config 1 capture rendered on the GPU: its depth
    map and refined warps feed a quarter-scale canvas of the whole raster
    and a native-scale centred crop."""
    rig = scene.rig_config1(seed=2)
    frames = scene.Frames(gpu_context, rig)
    try:
        frames.render(2)
        cap = frames.capture()
        sess = Session(gpu_context, cap.modules, 0, rig.options)
        try:
            sess.run(_abi.CHD_STAGE_ALL)
            dmap, targets = sess.results()
        finally:
            sess.close()
    finally:
        frames.close()
    sources = _sources(cap, targets, magnification={2: 1.5}, focus={2: 2000.0})
    w, h = cap.modules[0].width, cap.modules[0].height
    _check(gpu_context, oracle_lib, "config1 full size at 0.25", dmap, sources,
           CropWindow.full(w, h), 0.25, 120.0)
    _check(gpu_context, oracle_lib, "config1 centred 0.2 at 1.0", dmap, sources,
           CropWindow.centred(w, h, 0.2), 1.0, 120.0)
