"""Parity at the stated sizes of every benchmark configuration.

The oracle finishes the small configurations whole (1040x780: configs 0 and
3), so those run the full path against it. At 4160x3120 (configs 1, 2 and 4)
the GPU runs the whole capture and the oracle checks it stage by stage on
its own inputs: the cost volume on three bands of coarse rows (the first,
middle and last rows: the image borders and the interior), the whole SGM
volume recomputed from the GPU's costs, select_depths from the GPU's scores
and aggregated volume, the direct pass on three bands of fine rows from the
GPU's coarse field, and the three fine-grid filters on the GPU's direct
field. Every comparison is bit for bit.
"""
import os
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import Session, _ViewArrays

from parity import assert_bitwise, assert_nodes

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _label_depths(oracle_lib, o):
    finite = np.empty(o.planes, np.float64)
    oracle_lib.or_inverse_depth_samples(o.near_depth, o.far_depth, o.planes, finite.ctypes.data)
    return np.concatenate([[np.nan], finite[::-1]])


def _bands(rows, width):
    mid = rows // 2 - width // 2
    return [(0, width), (mid, mid + width), (rows - width, rows)]


def _staged_parity(ctx, oracle_lib, cap, label, coarse_band=4, fine_band=6):
    """The whole capture on the GPU, checked stage by stage by the oracle."""
    _say(f"full-size parity {label}: {cap.description}")
    o = cap.options
    sess = Session(ctx, cap.modules, cap.reference_index, o)
    try:
        sess.run(_abi.CHD_STAGE_ALL | _abi.CHD_STAGE_KEEP_AGGREGATED)
        ctx.synchronize()
        g = sess.grid
        scores, costs, guidance, tested = sess.cost_volume()
        aggregated = sess.aggregated()
        coarse, fillable, best = sess.coarse()
        direct = [sess.direct(k) for k in range(4)]
        dmap, targets = sess.results()
    finally:
        sess.close()
    cols, L = g.coarse_columns, g.labels
    _say(f"  grid {cols}x{g.coarse_rows} x {L} labels, fine {g.fine_columns}x{g.fine_rows}; "
         f"oracle on {THREADS} threads")
    views = _ViewArrays(cap.modules)
    n = len(cap.modules)
    copts = o.to_c()
    # cost volume on coarse row bands (depth.rs:1187-1277)
    for r0, r1 in _bands(g.coarse_rows, coarse_band):
        k = (r1 - r0) * cols
        ws = np.empty((k, L), np.float32)
        wc = np.empty((k, L), np.float32)
        wg = np.empty(k, np.float32)
        wt = np.empty(k, np.uint8)
        oracle_lib.or_cost_volume_rows(views.views, n, cap.reference_index, copts, r0, r1, THREADS,
                                       ws.ctypes.data, wc.ctypes.data, wg.ctypes.data,
                                       wt.ctypes.data)
        sl = slice(r0 * cols, r1 * cols)
        assert_bitwise(f"{label} scores rows {r0}-{r1}", scores[sl], ws)
        assert_bitwise(f"{label} costs rows {r0}-{r1}", costs[sl], wc)
        assert_bitwise(f"{label} guidance rows {r0}-{r1}", guidance[sl], wg)
        assert_bitwise(f"{label} tested rows {r0}-{r1}", tested[sl], wt)
        _say(f"  cost volume rows {r0}..{r1 - 1}: bit-exact "
             f"({int(np.isfinite(ws).sum())} finite scores)")
    # the whole SGM volume from the GPU's costs (depth.rs:1279-1383)
    want = np.empty_like(costs)
    oracle_lib.or_semi_global_costs(costs.ctypes.data, guidance.ctypes.data, cols, g.coarse_rows,
                                    L, want.ctypes.data)
    assert_bitwise(f"{label} aggregated", aggregated, want)
    _say(f"  aggregated SGM volume ({aggregated.size} values): bit-exact")
    # select_depths from the GPU's scores and aggregated volume (depth.rs:1385-1488)
    nodes = cols * g.coarse_rows
    labels = _label_depths(oracle_lib, o)
    wf = np.zeros(nodes, _abi.NODE_DTYPE)
    wfill = np.empty(nodes, np.uint8)
    wbest = np.empty(nodes, np.uint16)
    oracle_lib.or_select_depths(labels.ctypes.data, L, scores.ctypes.data, tested.ctypes.data,
                                aggregated.ctypes.data, nodes, copts, wf.ctypes.data,
                                wfill.ctypes.data, wbest.ctypes.data)
    assert_nodes(f"{label} coarse field", coarse, wf)
    assert_bitwise(f"{label} fillable", fillable, wfill)
    assert_bitwise(f"{label} best label", best, wbest)
    _say(f"  select_depths: bit-exact ({int(wf['valid'].sum())} of {nodes} nodes with a depth)")
    # direct pass on fine row bands from the GPU's coarse field (depth.rs:715-876)
    fine = g.fine_columns * g.fine_rows
    fcols = g.fine_columns
    df = np.zeros(fine, _abi.NODE_DTYPE)
    dg = np.empty(fine, np.float32)
    dt = np.empty(fine, np.uint8)
    for r0, r1 in _bands(g.fine_rows, fine_band):
        oracle_lib.or_direct_rows(views.views, n, cap.reference_index, copts, coarse.ctypes.data,
                                  r0, r1, THREADS, df.ctypes.data, dg.ctypes.data, dt.ctypes.data)
        sl = slice(r0 * fcols, r1 * fcols)
        assert_nodes(f"{label} direct rows {r0}-{r1}", direct[0][0][sl], df[sl])
        assert_bitwise(f"{label} fine guidance rows {r0}-{r1}", direct[0][1][sl], dg[sl])
        assert_bitwise(f"{label} fine tested rows {r0}-{r1}", direct[0][2][sl], dt[sl])
        _say(f"  direct pass rows {r0}..{r1 - 1}: bit-exact "
             f"({int(df['valid'][sl].sum())} measured)")
    # the filters on the GPU's direct field (depth.rs:881-1060)
    inverse_step = (1.0 / o.near_depth - 1.0 / o.far_depth) / (o.planes - 1)
    field = direct[0][0].copy()
    fg = np.ascontiguousarray(direct[0][1])
    oracle_lib.or_reject_isolated_direct_depths(field.ctypes.data, fg.ctypes.data, fcols,
                                                g.fine_rows, inverse_step * 2.5)
    assert_nodes(f"{label} isolated", direct[1][0], field)
    oracle_lib.or_reject_small_direct_components(field.ctypes.data, fg.ctypes.data, fcols,
                                                 g.fine_rows, inverse_step * 2.5, 24)
    assert_nodes(f"{label} components", direct[2][0], field)
    oracle_lib.or_fit_local_depth_planes(field.ctypes.data, fg.ctypes.data, fcols, g.fine_rows,
                                         inverse_step * 4.0)
    assert_nodes(f"{label} planes", direct[3][0], field)
    _say(f"  filters: bit-exact ({int(field['valid'].sum())} nodes kept)")
    # output assembly (depth.rs:548-579): provenance from the final field
    tested_fine = direct[0][2].astype(bool)
    valid = field["valid"].astype(bool)
    reg = field["regularized"].astype(bool)
    want_prov = np.where(valid, np.where(reg, _abi.PROVENANCE_REGULARIZED, _abi.PROVENANCE_MEASURED),
                         np.where(tested_fine, _abi.PROVENANCE_GLOBAL, _abi.PROVENANCE_UNSUPPORTED))
    assert_bitwise(f"{label} provenance", dmap.provenance.reshape(-1), want_prov.astype(np.uint8))
    assert_bitwise(f"{label} map depth", dmap.depth.reshape(-1)[valid], field["depth"][valid])
    assert np.isnan(dmap.depth.reshape(-1)[~valid]).all()
    want_conf = np.where(valid, field["confidence"], np.where(tested_fine, 1.0, 0.0))
    assert_bitwise(f"{label} map confidence", dmap.confidence.reshape(-1),
                   want_conf.astype(np.float32))
    # per-target decisions: codes, and refined points are finite
    for t, r in targets.items():
        assert r.decisions.max() <= _abi.WARP_BOUNDARY
        refined = r.decisions == _abi.WARP_REFINED
        assert np.isfinite(r.points[refined]).all()
    return g


def test_config0_full_size(gpu_context):
    """Config 0 at 1040x780 (3 views, 64 hypotheses): the whole path against
    the oracle."""
    from test_gpu_parity import _full_parity
    cap = scene.config0(seed=0)
    _full_parity(gpu_context, cap, "config0 1040x780 seed 0", threads=THREADS)


def test_config3_full_size(gpu_context):
    """Config 3 at 1040x780 with a 4 px grid and 128 hypotheses: the whole
    path against the oracle."""
    from test_gpu_parity import _full_parity
    cap = scene.config3(seed=0)
    _full_parity(gpu_context, cap, "config3 1040x780 grid 4 seed 0", threads=THREADS)


def test_config1_full_size(gpu_context, oracle_lib):
    """Config 1: five 4160x3120 modules, 128 hypotheses (521x391 nodes x 129
    labels: three SGM label chunks, a 2081-float plane pitch)."""
    cap = scene.config1(seed=0)
    g = _staged_parity(gpu_context, oracle_lib, cap, "config1")
    assert (g.coarse_columns, g.coarse_rows, g.labels) == (521, 391, 129)


def test_config2_full_size(gpu_context, oracle_lib):
    """Config 2: one capture of ten uneven modules (4160x3120, 2080x1560,
    1040x780) on a rotated rig, 128 hypotheses."""
    cap = scene.config2(seed=0)
    _staged_parity(gpu_context, oracle_lib, cap, "config2", coarse_band=3, fine_band=4)


def test_config4_full_size(gpu_context, oracle_lib):
    """Config 4: one capture of sixteen 4160x3120 modules (wide, mid, tele)
    at 256 hypotheses (257 labels: the running top-3 over 15 views)."""
    cap = scene.config4(seed=0)
    g = _staged_parity(gpu_context, oracle_lib, cap, "config4", coarse_band=3, fine_band=4)
    assert g.labels == 257


def test_config1_wta_labels_follow_the_ground_truth(gpu_context):
    """At config 1's size the GPU's WTA labels of interior nodes (one scene
    plane over the 5 x 5 nodes around them) sit within one label of the
    rendered scene's true inverse distance: a size-independent property."""
    from test_oracle_nversion import true_labels
    cap = scene.config1(seed=1)
    sess = Session(gpu_context, cap.modules, cap.reference_index, cap.options)
    try:
        sess.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM)
        field, _, best = sess.coarse()
    finally:
        sess.close()
    label, interior = true_labels(cap)
    sel = interior & field["valid"].astype(bool)
    within = (np.abs(best.astype(int) - label) <= 1)[sel]
    _say(f"config1 ground truth: {within.mean():.4f} of {sel.sum()} interior nodes within one label "
         f"(exact {(best.astype(int) == label)[sel].mean():.4f})")
    assert sel.sum() > 0.6 * sel.size
    assert within.mean() >= 0.9
