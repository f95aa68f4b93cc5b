"""The synthetic capture generator (no reference counterpart): determinism
per seed and independence of the thread count, a committed hash of one
render's bytes (the same bytes on every machine and build), and closed forms
of its geometry (a single fronto-parallel plane seen by a translated module
is the reference shifted by f B / Z; the ground truth is the plane's inverse
depth over the ray length)."""
import ctypes as C
import hashlib
import math

import numpy as np

from chiaro_amd import _abi, scene

# sha256 of scene.config1(seed=7, width=160, height=120, focal_px=127.0,
# planes=16): the five u16 mosaics, then the five luminance planes
RAW_SHA256 = "3759ef5cb9b601eab62b51730c01ca9a0f90e9b8aa0d6e942fe6d269d497e22f"
PLANE_SHA256 = "88b94cc7f5810111556014c3c9db832fa330a923de18e4a72b669346adedbcd9"


def small_config1():
    return scene.config1(seed=7, width=160, height=120, focal_px=127.0, planes=16)


def digest(arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_render_bytes_are_pinned():
    cap = small_config1()
    assert digest(cap.raws) == RAW_SHA256
    assert digest(m.luminance for m in cap.modules) == PLANE_SHA256


def _render(rig, threads):
    n = len(rig.modules)
    arr = (_abi.Module * n)(*rig.modules)
    images = [np.empty((m.height, m.width), np.uint16) for m in rig.modules]
    ptrs = (C.POINTER(C.c_uint16) * n)(*[a.ctypes.data_as(C.POINTER(C.c_uint16)) for a in images])
    assert _abi.scene_lib().chs_render(C.byref(rig.scene), arr, n, rig.reference, 0.0,
                                       C.byref(rig.bayer), None, ptrs, threads) == 0
    return images


def test_render_is_independent_of_threads_and_repeatable():
    rig = scene.rig_config2(seed=3, scale=0.0625)
    one = _render(rig, 1)
    many = _render(rig, 7)
    again = _render(rig, 7)
    for a, b, c in zip(one, many, again):
        assert np.array_equal(a, b) and np.array_equal(b, c)
    other = _render(rig.with_seed(4), 7)
    assert not all(np.array_equal(a, b) for a, b in zip(one, other))


def test_single_plane_parallax_closed_form():
    """One plane at Z = 1000 mm, f = 400 px, a module 20 mm to the right:
    its image is the reference image shifted by f B / Z = 8 px (noise-free
    f32 rendering), and the ground truth is (1/Z) / |K^-1 (x, y, 1)|."""
    w, h, f = 96, 64, 400.0
    mods = [scene._module(f, w, h, (0.0, 0.0, 0.0)), scene._module(f, w, h, (20.0, 0.0, 0.0))]
    desc = _abi.SceneDesc(seed=5, plane_count=1, inverse_depth_min=1e-3, inverse_depth_max=1e-3,
                          texture_cell_px=8)
    n = len(mods)
    arr = (_abi.Module * n)(*mods)
    images = [np.empty((h, w), np.float32) for _ in mods]
    ptrs = (C.POINTER(C.c_float) * n)(*[_abi.fptr(a) for a in images])
    assert _abi.scene_lib().chs_render(C.byref(desc), arr, n, 0, 0.0, None, ptrs, None, 2) == 0
    ref, target = images
    assert ref.std() > 0.02
    np.testing.assert_allclose(target[:, :w - 8], ref[:, 8:], rtol=0, atol=2e-6)
    truth = np.empty((-(-h // 8) + 1, -(-w // 8) + 1), np.float32)
    _abi.scene_lib().chs_truth_inverse_distance(C.byref(desc), C.byref(mods[0]), 8,
                                                _abi.ptr(truth))
    ys, xs = np.mgrid[0:truth.shape[0], 0:truth.shape[1]] * 8.0
    norm = np.sqrt(((xs - (w - 1) / 2) / f) ** 2 + ((ys - (h - 1) / 2) / f) ** 2 + 1.0)
    np.testing.assert_allclose(truth, (1e-3 / norm).astype(np.float32), rtol=1e-6)


def test_rig_seed_changes_only_the_scene():
    rig = scene.rig_config4(seed=1, scale=0.125)
    other = rig.with_seed(9)
    assert other.scene.seed == 9 and rig.scene.seed == 1
    assert other.scene.plane_count == rig.scene.plane_count
    assert math.isclose(other.scene.inverse_depth_min, rig.scene.inverse_depth_min)
    assert other.options.planes == rig.options.planes == 256
