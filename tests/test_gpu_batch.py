"""Schedules that must not change results: launch shapes (tuning), the
autotuner and its cache, batched captures (two-stream pipelining, persistent
cost queue, batched SGM sweeps, hipGraph replay), and the per-view score
slices against runs of one view pair alone. Every comparison is bit for bit; the uneven-module config 2
capture is also checked against the oracle."""
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import Batch, Session

from parity import assert_bitwise, assert_nodes

pytestmark = pytest.mark.gpu

ALL_KEEP = _abi.CHD_STAGE_ALL | _abi.CHD_STAGE_KEEP_AGGREGATED


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small_config0(seed, scale=0.25):
    w, h = int(1040 * scale) // 8 * 8, int(780 * scale) // 8 * 8
    return scene.config0(seed=seed, width=w, height=h, focal_px=1600.0 * scale)


def _snapshot(sess):
    """Every output of a finished run, as host arrays."""
    scores, costs, guidance, tested = sess.cost_volume()
    out = {"scores": scores, "costs": costs, "guidance": guidance, "tested": tested,
           "aggregated": sess.aggregated()}
    field, fillable, best = sess.coarse()
    out.update(coarse=field, fillable=fillable, best=best)
    for k in range(4):
        out[f"direct{k}"] = sess.direct(k)[0]
    dmap, targets = sess.results()
    out.update(depth=dmap.depth, confidence=dmap.confidence, provenance=dmap.provenance)
    for t, r in targets.items():
        out[f"t{t}.decisions"] = r.decisions
        out[f"t{t}.points"] = r.points
        out[f"t{t}.confidence"] = r.confidence
    return out


def _same(label, got, want):
    assert sorted(got) == sorted(want), f"{label}: outputs differ in keys"
    for key in want:
        if want[key].dtype == _abi.NODE_DTYPE:
            assert_nodes(f"{label} {key}", got[key], want[key])
        else:
            assert_bitwise(f"{label} {key}", got[key], want[key])


def _eager(ctx, cap):
    sess = Session(ctx, cap.modules, cap.reference_index, cap.options)
    try:
        sess.run(ALL_KEEP)
        return _snapshot(sess)
    finally:
        sess.close()


def test_tuning_shapes_bit_exact(gpu_context):
    cap = _small_config0(3)
    _say("tuning shapes on config0 quarter seed 3")
    want = _eager(gpu_context, cap)
    sess = Session(gpu_context, cap.modules, cap.reference_index, cap.options)
    try:
        for nb, slab, wpb, order in [(1, 0, 1, 0), (4, 0, 2, 1), (3, 17, 4, 0), (12, 33, 1, 1),
                                     (16, 65, 2, 0), (7, 0, 4, 1)]:
            _say(f"  nodes/block {nb} slab {slab} paths/block {wpb} xcd order {order}")
            sess.set_tuning(nb, slab, wpb, order)
            assert sess.tuning() == {"cost_nodes_per_block": nb, "cost_label_slab": slab,
                                     "sgm_paths_per_block": wpb, "cost_tile_order": order}
            sess.run(ALL_KEEP)
            _same(f"tuning {nb}/{slab}/{wpb}/{order}", _snapshot(sess), want)
    finally:
        sess.close()


def test_tuning_rejects_bad_shapes(gpu_context):
    cap = _small_config0(3)
    sess = Session(gpu_context, cap.modules, cap.reference_index, cap.options)
    try:
        with pytest.raises(RuntimeError):
            sess.set_tuning(0, 0, 3)
        with pytest.raises(RuntimeError):
            sess.set_tuning(65, 0, 0)
        with pytest.raises(RuntimeError):
            sess.set_tuning(0, 0, 0, 3)
    finally:
        sess.close()


def test_autotune_and_cache(gpu_context, tmp_path):
    cap = _small_config0(4)
    _say("autotune on config0 quarter seed 4")
    want = _eager(gpu_context, cap)
    cache = tmp_path / "tuning.txt"
    sess = Session(gpu_context, cap.modules, cap.reference_index, cap.options)
    try:
        tuned = sess.autotune(cache)
        assert tuned["cost_nodes_per_block"] > 0 and tuned["sgm_paths_per_block"] in (1, 2, 4)
        lines = cache.read_text().splitlines()
        assert len(lines) == 1 and lines[0].startswith("chd-tune-2 views=2 taps=81 labels=65")
        assert sess.autotune(cache) == tuned       # read back, not timed again
        assert len(cache.read_text().splitlines()) == 1
        sess.run(ALL_KEEP)
        _same("autotuned", _snapshot(sess), want)
    finally:
        sess.close()


@pytest.mark.parametrize("mode", ["pipelined", "one_sweep", "persistent"])
def test_batch_matches_single_sessions_on_replays(gpu_context, mode):
    """Three captures (two sizes) batched in each schedule (two streams, one
    SGM sweep launch, persistent cost queue); the captured graph is replayed
    three times with eager work of another session in between, and every
    replay equals the single-session results."""
    caps = [_small_config0(5), _small_config0(6, scale=0.3), _small_config0(7)]
    _say(f"batch of 3 config0 captures, {mode}")
    wants = [_eager(gpu_context, c) for c in caps]
    sessions = [Session(gpu_context, c.modules, c.reference_index, c.options) for c in caps]
    sessions[0].set_tuning(0, 0, 0, 1)  # the queue in XCD order
    sessions[1].set_tuning(6, 33, 2)    # a different shape inside the same batch
    other_cap = _small_config0(8)
    other = Session(gpu_context, other_cap.modules, 0, other_cap.options)
    batch = Batch(gpu_context, sessions, persistent=mode == "persistent",
                  one_sweep=mode == "one_sweep")
    try:
        for replay in range(3):
            _say(f"  replay {replay}")
            batch.run(ALL_KEEP)
            other.run(_abi.CHD_STAGE_ALL)        # eager work between replays
            gpu_context.synchronize()
            for k, (s, want) in enumerate(zip(sessions, wants)):
                _same(f"replay {replay} capture {k}", _snapshot(s), want)
    finally:
        batch.close()
        for s in sessions:
            s.close()
        other.close()


def test_batch_stage_order_is_checked(gpu_context):
    cap = _small_config0(5)
    sess = Session(gpu_context, cap.modules, 0, cap.options)
    batch = Batch(gpu_context, [sess])
    try:
        with pytest.raises(RuntimeError):
            batch.run(_abi.CHD_STAGE_DIRECT)
    finally:
        batch.close()
        sess.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_view_scores_equal_single_pair_runs(gpu_context, persistent):
    """Each target's ZNCC slice in a batched run equals that of a session
    holding only the reference and that target."""
    cap = _small_config0(10)
    _say(f"view scores vs single pairs, config0 quarter seed 10, persistent {persistent}")
    sess = Session(gpu_context, cap.modules, 0, cap.options)
    batch = Batch(gpu_context, [sess], persistent=persistent)
    try:
        batch.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_KEEP_VIEW_SCORES)
        slices = sess.view_scores()
        assert slices.shape[0] == 2
        for k, t in enumerate([1, 2]):
            pair = Session(gpu_context, [cap.modules[0], cap.modules[t]], 0, cap.options)
            try:
                pair.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_KEEP_VIEW_SCORES)
                assert_bitwise(f"view {t} slice", slices[k], pair.view_scores()[0])
            finally:
                pair.close()
        # the aggregate of the slices is the cost volume's score
        scores = sess.cost_volume()[0]
        assert np.isfinite(slices).any() and np.array_equal(np.isnan(scores),
                                                            np.isnan(slices).any(axis=0))
    finally:
        batch.close()
        sess.close()


def test_config2_batch_bit_exact_with_oracle(gpu_context):
    """Two small config 2 captures (ten modules of three sizes, rotated rig)
    batched: each equals its single-session run, and the first equals the
    oracle."""
    from test_gpu_parity import _full_parity
    caps = [scene.config2(seed=s, scale=0.125) for s in (0, 1)]
    _say("config2 small: single-session parity with the oracle")
    _full_parity(gpu_context, caps[0], "config2 scale 0.125 seed 0")
    wants = [_eager(gpu_context, c) for c in caps]
    sessions = [Session(gpu_context, c.modules, 0, c.options) for c in caps]
    batch = Batch(gpu_context, sessions, persistent=True)
    try:
        batch.run(ALL_KEEP)
        gpu_context.synchronize()
        for k, (s, want) in enumerate(zip(sessions, wants)):
            _same(f"config2 capture {k}", _snapshot(s), want)
    finally:
        batch.close()
        for s in sessions:
            s.close()


def test_config4_batch_bit_exact_with_oracle(gpu_context):
    """Two small config 4 captures (sixteen modules: wide, mid and tele
    rings with nested fields of view, 256 hypotheses) batched: each equals
    its single-session run, and the first equals the oracle."""
    from test_gpu_parity import _full_parity
    caps = [scene.config4(seed=s, scale=0.125) for s in (0, 1)]
    _say("config4 small: single-session parity with the oracle")
    _full_parity(gpu_context, caps[0], "config4 scale 0.125 seed 0")
    wants = [_eager(gpu_context, c) for c in caps]
    sessions = [Session(gpu_context, c.modules, 0, c.options) for c in caps]
    batch = Batch(gpu_context, sessions)
    try:
        batch.run(ALL_KEEP)
        gpu_context.synchronize()
        for k, (s, want) in enumerate(zip(sessions, wants)):
            _same(f"config4 capture {k}", _snapshot(s), want)
    finally:
        batch.close()
        for s in sessions:
            s.close()


def _banded_session(ctx, cap, seed, transform_from=None):
    """A session with band windows from a seeded random proposal (about half
    of the tiles banded); the input transform from `transform_from`'s planes
    (default: cap's)."""
    from chiaro_amd.proposal import Proposal, input_transform, layer_specs
    sess = Session(ctx, cap.modules, cap.reference_index, cap.options)
    factor = max(cap.options.grid_step, 4) // 2
    h, w = (n // factor for n in cap.modules[cap.reference_index].luminance.shape)
    net = Proposal(ctx, w, h, max_batch=1, seed=seed)
    net.set_label_count(sess.grid.labels)
    net.set_input_transform(*input_transform([transform_from or cap]))
    rng = np.random.default_rng(seed)
    parts = []
    for (cin, cout, *_r) in layer_specs(w, h):
        cp_in, cp_out = (cin + 15) // 16 * 16, (cout + 15) // 16 * 16
        wt = np.zeros((9 * cp_in, cp_out), np.float32)
        for t in range(9):
            wt[t * cp_in:t * cp_in + cin, :cout] = rng.normal(size=(cin, cout)) * 0.3
        parts += [wt.reshape(-1), np.zeros(cp_out, np.float32)]
    net.set_parameters(np.concatenate(parts))
    net.apply(sess, threshold=0.5, half_width=8, tile=4, max_half_width=24)
    return sess, net


@pytest.mark.parametrize("persistent", [False, True])
def test_band_batches_match_single_sessions(gpu_context, persistent):
    """Band sweeps (windows from a proposal) batched equal the same sessions
    run alone: two captures of two views (band items with all views kept),
    two of nine views (band items with a running top-K) and the mixed pair
    (with the persistent queue: its per-kind and mixed kernels)."""
    small3 = [scene.config3(seed=s, width=520, height=392, focal_px=800.0) for s in (40, 41)]
    small2 = []
    for s in (42, 43):
        c = scene.config2(seed=s, scale=0.125)
        c.options.grid_step = 4
        small2.append(c)
    for label, caps in (("two views", small3), ("nine views", small2),
                        ("mixed", [small3[0], small2[0]])):
        _say(f"band batch ({label}), persistent {persistent}")
        wants = []
        for k, cap in enumerate(caps):
            sess, net = _banded_session(gpu_context, cap, 7 + k)
            try:
                sess.run(ALL_KEEP)
                wants.append(_snapshot(sess))
            finally:
                net.close()
                sess.close()
        pairs = [_banded_session(gpu_context, cap, 7 + k) for k, cap in enumerate(caps)]
        batch = Batch(gpu_context, [s for s, _ in pairs], persistent=persistent)
        try:
            batch.run(ALL_KEEP)
            gpu_context.synchronize()
            for k, ((s, _), want) in enumerate(zip(pairs, wants)):
                _same(f"band batch {label} capture {k}", _snapshot(s), want)
        finally:
            batch.close()
            for s, net in pairs:
                net.close()
                s.close()


def test_attached_proposal_runs_inside_the_batch(gpu_context):
    """A proposal attached to the sessions runs as the batch's PROPOSE stage:
    after new planes are set, a batched PROPOSE run sets each capture's
    windows from its new planes (inside the graph) and every output equals a
    fresh session of those planes given the same windows by apply(); with
    the proposal destroyed, a PROPOSE run is refused."""
    import copy
    caps = []
    for s in (42, 43):
        c = scene.config2(seed=s, scale=0.125)
        c.options.grid_step = 4
        caps.append(c)
    # the same rigs seen with other planes: each capture's modules with the
    # luminance of the other capture (same sizes)
    swapped = []
    for k, cap in enumerate(caps):
        other = caps[1 - k]
        c = copy.copy(cap)
        c.modules = [copy.copy(m) for m in cap.modules]
        for m, o in zip(c.modules, other.modules):
            assert m.luminance.shape == o.luminance.shape
            m.luminance = o.luminance
        swapped.append(c)
    wants = []
    for k, cap in enumerate(swapped):
        # the network as attached below (input transform of the first planes)
        sess, net = _banded_session(gpu_context, cap, 7 + k, transform_from=caps[k])
        try:
            sess.run(ALL_KEEP)
            wants.append(_snapshot(sess))
        finally:
            net.close()
            sess.close()
    pairs = [_banded_session(gpu_context, cap, 7 + k) for k, cap in enumerate(caps)]
    batch = Batch(gpu_context, [s for s, _ in pairs])
    try:
        for s, net in pairs:
            net.attach(s, threshold=0.5, half_width=8, tile=4, max_half_width=24)
        batch.run(ALL_KEEP | _abi.CHD_STAGE_PROPOSE)
        for k, (s, _) in enumerate(pairs):
            s.set_planes([m.luminance for m in swapped[k].modules])
        for _ in range(2):  # the captured graph, replayed
            batch.run(ALL_KEEP | _abi.CHD_STAGE_PROPOSE)
            gpu_context.synchronize()
            for k, ((s, _), want) in enumerate(zip(pairs, wants)):
                _same(f"attached proposal capture {k}", _snapshot(s), want)
        pairs[0][1].close()
        with pytest.raises(RuntimeError):
            batch.run(ALL_KEEP | _abi.CHD_STAGE_PROPOSE)
        with pytest.raises(RuntimeError):
            pairs[0][0].run(ALL_KEEP | _abi.CHD_STAGE_PROPOSE)
    finally:
        batch.close()
        for s, net in pairs:
            net.close()
            s.close()
