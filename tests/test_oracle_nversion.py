"""Pins of the oracle's threshold logic that no reference test holds.

1. N-version: tests/nversion.py restates select_depths (with its
   confidences), measure_direct_depths (with the wide fallback) and the
   per-target refine/decision block straight from depth.rs; the C oracle
   must agree with it bit for bit on seeded scenes whose runs reach every
   branch (regularized nodes, wide fallbacks, Refined, Boundary and Occluded
   decisions).
2. Closed forms: targets that are exact shifted copies of the reference
   (a fronto-parallel plane at a known depth) come out Refined, at the
   correspondence the rig fixes and with confidence 0.5 + 0.5 c; a target
   that sees the background where the others see a near surface comes out
   Occluded.
3. Ground truth: on a full-size config 0 scene the WTA labels of interior
   nodes sit within one label of the rendered scene's true inverse depth.
"""
import numpy as np
import pytest
from numpy.lib.stride_tricks import sliding_window_view

import nversion as nv
from chiaro_amd import scene
from chiaro_amd.depth import DepthOptions
from oracle import binding

from parity import assert_bitwise

pytestmark = pytest.mark.slow


def _small(seed):
    return scene.config0(seed=seed, width=128, height=96, focal_px=200.0, planes=24)


def _nodes(arr):
    return [None if not arr["valid"][i] else
            (float(arr["depth"][i]), np.float32(arr["confidence"][i]),
             np.float32(arr["improvement"][i]), bool(arr["regularized"][i]))
            for i in range(arr.size)]


def _assert_fields(name, got, want):
    """A restated field (list of None | tuple) against an oracle node array."""
    assert len(got) == want.size
    valid = np.array([g is not None for g in got])
    assert_bitwise(name + " valid", valid.astype(np.uint8), want["valid"])
    v = np.flatnonzero(valid)
    assert_bitwise(name + " depth", np.array([got[i][0] for i in v], np.float64), want["depth"][v])
    assert_bitwise(name + " confidence", np.array([got[i][1] for i in v], np.float32),
                   want["confidence"][v])
    assert_bitwise(name + " improvement", np.array([got[i][2] for i in v], np.float32),
                   want["improvement"][v])
    assert_bitwise(name + " regularized", np.array([got[i][3] for i in v], np.uint8),
                   want["regularized"][v])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_select_depths_two_restatements_agree(seed):
    cap = _small(seed)
    _, _, dbg = binding.refine(cap, threads=4, stop_after_sgm=True)
    labels = nv.label_depths(cap.options)
    field, fillable, best = nv.select_depths(labels, dbg["scores"], dbg["tested"],
                                             dbg["aggregated"], cap.options)
    _assert_fields("select_depths", field, dbg["coarse"])
    assert_bitwise("fillable", np.array(fillable, np.uint8), dbg["fillable"])
    assert_bitwise("best label", np.array(best, np.uint16), dbg["best_label"])
    if seed == 2:
        assert sum(1 for f in field if f and f[3]) > 0, "no regularized node reached"


def test_direct_pass_two_restatements_agree():
    cap = _small(2)
    _, _, dbg = binding.refine(cap, threads=4)
    g = cap.options.grid(128, 96)
    field, guidance, tested, wide = nv.measure_direct(
        nv.Capture(cap), _nodes(dbg["coarse"]), g.coarse_columns, g.coarse_rows, g.coarse_step)
    assert wide > 50, "the wide fallback must be reached"
    _assert_fields("measure_direct_depths", field, dbg["direct"][0])
    assert_bitwise("fine guidance", guidance, dbg["fine_guidance"])
    assert_bitwise("fine tested", tested, dbg["fine_tested"])


def _assert_targets(cap, otargets, dbg):
    decisions = nv.target_decisions(nv.Capture(cap), _nodes(dbg["direct"][3]),
                                    dbg["fine_guidance"])
    assert sorted(decisions) == sorted(otargets)
    counts = np.zeros(5, int)
    for t, (codes, points, conf) in decisions.items():
        o = otargets[t]
        assert_bitwise(f"target {t} codes", codes, o["decisions"].reshape(-1))
        assert_bitwise(f"target {t} points", points, o["points"].reshape(-1, 2))
        assert_bitwise(f"target {t} confidence", conf, o["confidence"].reshape(-1))
        counts += np.bincount(codes, minlength=5)
    return counts


def test_target_decisions_two_restatements_agree():
    cap = _small(2)
    _, otargets, dbg = binding.refine(cap, threads=4)
    counts = _assert_targets(cap, otargets, dbg)
    assert counts[nv.REFINED] > 0 and counts[nv.BOUNDARY] > 0 and counts[nv.GLOBAL] > 0


# ------------------------------------------------------------ closed forms

F_PX, BASELINE_MM, SHIFT = 200.0, 20.0, 2   # plane px; 4 raster px = f B / Z at Z = 1000 mm


def _shifted_copy_capture(occluder: bool, seed: int = 0, width: int = 128, height: int = 96):
    """A reference plane of short-range texture and targets that are exact
    copies of it shifted by 2 plane px (4 raster px: the parallax of a
    fronto-parallel plane at 1000 mm for f = 200 px and a 20 mm baseline):
    one target 20 mm to the right (the copy moves left), one 20 mm down (it
    moves up). With `occluder`, a fourth view 20 mm to the left holds the
    reference unshifted: it sees the scene at infinity, as a camera looking
    past a near object at a far background would."""
    rng = np.random.default_rng(seed)
    m = 8
    h, w = height // 2, width // 2
    t = rng.normal(size=(h + 2 * m, w + 2 * m))
    k = np.array([0.25, 0.5, 0.25])
    t = np.apply_along_axis(lambda r: np.convolve(r, k, mode="same"), 1, t)
    t = np.apply_along_axis(lambda r: np.convolve(r, k, mode="same"), 0, t)
    t = (6.2 + 0.3 * t / t.std()).astype(np.float32)
    ref = t[m:m + h, m:m + w].copy()
    centres = [(0.0, 0.0, 0.0), (BASELINE_MM, 0.0, 0.0), (0.0, BASELINE_MM, 0.0)]
    planes = [ref, t[m:m + h, m + SHIFT:m + SHIFT + w].copy(),
              t[m + SHIFT:m + SHIFT + h, m:m + w].copy()]
    if occluder:
        centres.append((-BASELINE_MM, 0.0, 0.0))
        planes.append(ref.copy())
    mods = [scene._module(F_PX, width, height, c) for c in centres]
    cams = [scene.camera_for(x) for x in mods]
    inputs = []
    for i, (plane, cam) in enumerate(zip(planes, cams)):
        warp = scene.identity_warp(width, height) if i == 0 else \
            scene.global_warp(cams[0], cam, width, height)
        inputs.append(scene.ModuleInput(name=f"V{i}", luminance=plane, width=width,
                                        height=height, camera=cam, warp_points=warp,
                                        warp_step=32))
    options = DepthOptions(near_depth=800.0, far_depth=20000.0, planes=32)
    return scene.Capture(modules=inputs, reference_index=0, options=options)


def test_exact_shifted_copies_are_refined():
    cap = _shifted_copy_capture(False)
    _, otargets, dbg = binding.refine(cap, threads=4)
    field = dbg["direct"][3]
    valid = field["valid"].astype(bool)
    g = cap.options.grid(128, 96)
    ys, xs = np.mgrid[0:g.fine_rows, 0:g.fine_columns]
    x = (xs * g.fine_step).reshape(-1).astype(np.float64)
    y = (ys * g.fine_step).reshape(-1).astype(np.float64)
    assert valid.sum() > 0.5 * valid.size
    assert not field["regularized"][valid].any()
    # the fitted depth of the plane: f B / 4 px = 1000 mm, within a label
    inverse_step = (1 / 800.0 - 1 / 20000.0) / 31
    assert abs(1 / np.median(field["depth"][valid]) - 1 / 1000.0) < inverse_step * 2
    for t, (ex, ey) in ((1, (-2.0 * SHIFT, 0.0)), (2, (0.0, -2.0 * SHIFT))):
        codes = otargets[t]["decisions"].reshape(-1)
        points = otargets[t]["points"].reshape(-1, 2)
        conf = otargets[t]["confidence"].reshape(-1)
        refined = codes == nv.REFINED
        assert refined.sum() > 0.85 * valid.sum(), f"target {t}: {np.bincount(codes)}"
        assert valid[refined].all() and not (codes == nv.OCCLUDED).any()
        # measured nodes: confidence 0.5 + 0.5 c (depth.rs:473-477)
        want = (np.float32(0.5) + np.float32(0.5) * field["confidence"][refined]).astype(np.float32)
        assert_bitwise(f"target {t} refined confidence", conf[refined], want)
        # the refined point is the copy's true correspondence (x - 4, y) or
        # (x, y - 4) within the 0.5 label / 1.5 px search
        err = np.hypot(points[refined, 0] - (x[refined] + ex), points[refined, 1] - (y[refined] + ey))
        assert err.max() < 0.5, f"target {t}: refined point {err.max():.3f} px from the copy"


def test_view_of_the_background_is_occluded():
    cap = _shifted_copy_capture(True)
    _, otargets, dbg = binding.refine(cap, threads=4)
    valid = dbg["direct"][3]["valid"].astype(bool)
    occluded = otargets[3]["decisions"].reshape(-1) == nv.OCCLUDED
    assert occluded.sum() > 0.75 * valid.sum() and valid[occluded].all()
    assert not (otargets[3]["decisions"] == nv.REFINED).any()
    for t in (1, 2):
        assert not (otargets[t]["decisions"] == nv.OCCLUDED).any()
    # and the second restatement reaches the same decisions, Occluded included
    counts = _assert_targets(cap, otargets, dbg)
    assert counts[nv.OCCLUDED] == occluded.sum()


# ------------------------------------------------------------ ground truth

def true_labels(cap):
    """The label nearest to the rendered scene's inverse distance at each
    coarse node, and a mask of interior nodes (the truth varies by < 1 %
    over the 5 x 5 nodes around them: one plane)."""
    labels = nv.label_depths(cap.options)
    inverse = np.array([1.0 / d for d in labels[1:]])
    truth = cap.truth_inverse_distance.astype(np.float64)
    label = 1 + np.abs(inverse[None, :] - truth.reshape(-1, 1)).argmin(axis=1)
    win = sliding_window_view(np.pad(truth, 2, mode="edge"), (5, 5))
    interior = ((win.max(axis=(2, 3)) - win.min(axis=(2, 3))) < 0.01 * truth).reshape(-1)
    return label, interior


def test_wta_labels_follow_the_ground_truth():
    cap = scene.config0(seed=0)
    _, _, dbg = binding.refine(cap, threads=8, stop_after_sgm=True)
    label, interior = true_labels(cap)
    sel = interior & dbg["coarse"]["valid"].astype(bool)
    within = np.abs(dbg["best_label"].astype(int) - label)[sel] <= 1
    assert sel.sum() > 0.6 * sel.size
    assert within.mean() >= 0.95, f"{within.mean():.4f} of interior nodes within one label"
