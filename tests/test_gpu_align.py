"""align_module's refinement on the GPU (chd_align_module, chd_match_patches)
against the oracle (oracle/align_oracle.c), bit for bit: match_patch on
random and textured windows, and whole alignments (warp points, every report
field) of a rotated module from a wrong factory model, of a baseline module
from a perturbed calibration, at the reference's own size (a 4160x3120
module: five pyramid levels), with refinement off and for the reference
itself."""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import _ViewArrays

from parity import assert_bitwise
from test_align_oracle import rotated_pair, rotation_y, textured

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _options(lib, **kw):
    o = _abi.AlignOptions()
    lib.chd_default_align_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _run(fn, cap, target, options):
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
    rc = fn(C.byref(views.views[cap.reference_index]), C.byref(views.views[target]),
            C.byref(options), C.byref(warp), C.byref(report))
    assert rc == 0, _abi.last_error()
    return points, conf, report


def _same_report(label, got, want):
    for name, _ in _abi.AlignmentReport._fields_:
        if name == "levels":
            for i in range(want.level_count):
                for f, _t in _abi.LevelReport._fields_:
                    a, b = getattr(got.levels[i], f), getattr(want.levels[i], f)
                    assert a == b or (isinstance(b, float) and math.isnan(a) and math.isnan(b)), \
                        f"{label} level {i} {f}: {a} != {b}"
        elif name == "correction_median_px":
            assert list(got.correction_median_px) == list(want.correction_median_px), label
        else:
            a, b = getattr(got, name), getattr(want, name)
            assert a == b or (isinstance(b, float) and math.isnan(a) and math.isnan(b)), \
                f"{label} report {name}: {a} != {b}"


def _parity(gpu_context, oracle_lib, cap, target, options, label):
    lib = gpu_context.lib
    got = _run(lambda *a: lib.chd_align_module(gpu_context.handle, *a), cap, target, options)
    want = _run(oracle_lib.or_align_module, cap, target, options)
    assert_bitwise(f"{label} warp points", got[0], want[0])
    assert_bitwise(f"{label} warp confidence", got[1], want[1])
    _same_report(label, got[2], want[2])
    r = got[2]
    _say(f"{label}: status {_abi.ALIGN_STATUS[r.status]}, {r.level_count} levels, "
         f"{r.inliers}/{r.patches} inliers, residual {r.residual_median_px:.3f} px, "
         f"correction {list(r.correction_median_px)}, gain {r.gain:.4f}")
    return got


def test_match_patches_bit_exact(gpu_context, oracle_lib):
    rng = np.random.default_rng(9)
    cases = [(textured(200, 160, (0.0, 0.0)), textured(200, 160, (3.4, -2.6)), 48, 8),
             (rng.random((120, 150)).astype(np.float32), None, 32, 12),
             (np.zeros((80, 80), np.float32), None, 16, 3)]
    for ref, tgt, size, radius in cases:
        if tgt is None:
            tgt = (ref + 0.01 * rng.random(ref.shape)).astype(np.float32)
            tgt = np.ascontiguousarray(np.roll(tgt, (1, -2), axis=(0, 1)))
        h, w = ref.shape
        rx = rng.integers(0, w - size, 200).astype(np.uint32)
        ry = rng.integers(0, h - size, 200).astype(np.uint32)
        shift = np.empty((200, 2), np.float32)
        score = np.empty(200, np.float32)
        valid = np.empty(200, np.uint8)
        _abi.check(gpu_context.lib.chd_match_patches(
            gpu_context.handle, ref.ctypes.data, tgt.ctypes.data, w, h, rx.ctypes.data,
            ry.ctypes.data, 200, size, radius, shift.ctypes.data, score.ctypes.data,
            valid.ctypes.data), "match patches")
        for k in range(200):
            s = np.empty(2, np.float32)
            sc = C.c_float()
            ok = oracle_lib.or_match_patch(ref.ctypes.data, w, h, tgt.ctypes.data, w, h,
                                           int(rx[k]), int(ry[k]), size, radius, s.ctypes.data,
                                           C.byref(sc))
            assert valid[k] == ok, f"window {k} validity"
            if ok:
                assert_bitwise(f"window {k} shift", shift[k], s)
                assert_bitwise(f"window {k} score", score[k:k + 1], np.array([sc.value], np.float32))


def test_rotated_module_bit_exact(gpu_context, oracle_lib):
    cap, _ = rotated_pair(832, 624, 660.0)
    points, _, report = _parity(gpu_context, oracle_lib, cap, 1, _options(gpu_context.lib),
                                "rotated module")
    assert report.status == _abi.ALIGN_STATUS.index("refined")


def test_perturbed_calibration_bit_exact(gpu_context, oracle_lib):
    """A config 1 module 30 mm from the reference whose factory rotation is
    off by 0.4 deg: refinement on a scene with parallax."""
    cap = scene.config1(seed=6, width=1040, height=780, focal_px=825.0, planes=32)
    m = cap.modules[1]
    module = scene._module(825.0, m.width, m.height, (30.0, 0.0, 0.0), rotation_y(0.4))
    cap.modules[1].camera = scene.camera_for(module)
    _parity(gpu_context, oracle_lib, cap, 1, _options(gpu_context.lib), "perturbed calibration")
    _parity(gpu_context, oracle_lib, cap, 2, _options(gpu_context.lib, patch=24, coarse_radius=8),
            "true calibration, patch 24")


def test_refine_off_and_reference(gpu_context, oracle_lib):
    cap = scene.config1(seed=1, width=416, height=312, focal_px=330.0, planes=32)
    points, _, report = _parity(gpu_context, oracle_lib, cap, 2,
                                _options(gpu_context.lib, refine=0), "factory model")
    assert_bitwise("factory model = global warp", points, cap.modules[2].warp_points)
    points, _, report = _parity(gpu_context, oracle_lib, cap, 0, _options(gpu_context.lib),
                                "reference")
    assert report.status == 0


def test_full_size_module_bit_exact(gpu_context, oracle_lib):
    """A 4160x3120 module (2080x1560 plane: five pyramid levels, about 12,000
    finest-level windows) turned by 0.3 deg against its factory model."""
    rig = scene.rig_config1(seed=3)
    mods = [rig.modules[0], scene._module(3300.0, 4160, 3120, (30.0, 0.0, 0.0), rotation_y(0.3))]
    cap = scene.config1(seed=3)
    cap.modules[1].camera = scene.camera_for(mods[1])
    got = _parity(gpu_context, oracle_lib, cap, 1, _options(gpu_context.lib), "full-size module")
    assert got[2].level_count == 5
