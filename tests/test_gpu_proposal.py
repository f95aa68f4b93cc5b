"""The learned band proposal on the GPU against its FP64 restatement
(forward, gradients, Adam), and the band sweep it drives: with every tile
falling back it equals the plain full sweep bit for bit; with bands it runs
and is scored against the full sweep."""
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import Session
from chiaro_amd.proposal import (Proposal, ReferenceNet, adam_reference, agreement,
                                 build_dataset, capture_inputs, clear, constant_proposal,
                                 from_bf16, input_transform, layer_specs, to_bf16)

from parity import assert_bitwise

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _params(width, height, rng, scale=0.2):
    out = []
    for (cin, cout, *_r) in layer_specs(width, height):
        cp_in, cp_out = (cin + 15) // 16 * 16, (cout + 15) // 16 * 16
        w = np.zeros((9 * cp_in, cp_out), np.float32)
        for t in range(9):
            w[t * cp_in:t * cp_in + cin, :cout] = rng.normal(size=(cin, cout)) * scale
        b = np.zeros(cp_out, np.float32)
        b[:cout] = rng.normal(size=cout) * 0.1
        out += [w.reshape(-1), b]
    return np.concatenate(out).astype(np.float32)


def assert_band_equals_full(name, session, scores, costs, want_scores, want_costs):
    """Inside each node's window (label 0 and lo..hi) the band sweep's scores
    and costs equal the full sweep's bit for bit; outside they are NaN and
    +inf (depth.rs:1224-1247 evaluated on the window only)."""
    lo, hi = session.band()
    labels = np.arange(scores.shape[1])[None, :]
    inside = (labels == 0) | ((labels >= lo[:, None]) & (labels <= hi[:, None]))
    assert_bitwise(name + " scores in window", scores[inside], want_scores[inside])
    assert_bitwise(name + " costs in window", costs[inside], want_costs[inside])
    assert np.isnan(scores[~inside]).all(), name + ": a score outside the window"
    assert np.all(costs[~inside] == np.inf), name + ": a cost outside the window"
    return int(inside.sum())


def _inputs(rng, n, h, w):
    x = np.zeros((n, h, w, 16), np.float32)
    x[..., :3] = rng.uniform(0, 1, size=(n, h, w, 3))
    return to_bf16(x)


def test_forward_matches_fp64_restatement(gpu_context):
    rng = np.random.default_rng(0)
    w, h = 33, 25
    net = Proposal(gpu_context, w, h, max_batch=2, seed=1)
    try:
        params = _params(w, h, rng)
        net.set_parameters(params)
        assert np.array_equal(net.parameters(), params)
        x = _inputs(rng, 2, h, w)
        got = net.predict(x)
        ref = ReferenceNet(w, h, params, bf16=True).forward(from_bf16(x).astype(np.float64))[-1]
        centre = 64.0 + 16.0 * ref[..., 0]
        scale = np.abs(ref[..., :3]).max()
        _say(f"forward: max |o| {scale:.3f}")
        assert np.abs((got[..., 0] - 64.0) / 16.0 - ref[..., 0]).max() <= 0.02 * scale + 1e-3
        assert np.abs(got[..., 1] - ref[..., 1]).max() <= 0.02 * scale + 1e-3
        assert np.abs(got[..., 2] - 1 / (1 + np.exp(-ref[..., 2]))).max() <= 0.01 + 0.005 * scale
        assert np.abs(got[..., 0] - centre).max() <= 16 * (0.02 * scale + 1e-3)
    finally:
        net.close()


def test_gradients_match_fp64_restatement(gpu_context):
    rng = np.random.default_rng(2)
    w, h = 17, 13
    net = Proposal(gpu_context, w, h, max_batch=2, seed=3)
    try:
        params = _params(w, h, rng)
        net.set_parameters(params)
        x = _inputs(rng, 2, h, w)
        labels = rng.integers(1, 129, size=(2, h, w)).astype(np.uint16)
        labels[1, :2] = 0
        g, loss = net.gradients(x, labels)
        ref = ReferenceNet(w, h, params, bf16=True)
        acts = ref.forward(from_bf16(x).astype(np.float64))
        want_loss, dout = ref.loss_grad(acts[-1], labels)
        want = ref.backward(acts, dout)
        cos = float(np.dot(g, want) / (np.linalg.norm(g) * np.linalg.norm(want)))
        rel = float(np.linalg.norm(g - want) / np.linalg.norm(want))
        _say(f"gradients: cosine {cos:.5f}, relative error {rel:.4f}, loss {loss:.3f} vs {want_loss:.3f}")
        assert abs(loss - want_loss) <= 1e-2 * abs(want_loss)
        assert cos > 0.995 and rel < 0.06
    finally:
        net.close()


def test_adam_step_matches_reference(gpu_context):
    rng = np.random.default_rng(4)
    w, h = 16, 12
    net = Proposal(gpu_context, w, h, max_batch=1, seed=5)
    try:
        p0 = net.parameters()
        m = np.zeros_like(p0)
        v = np.zeros_like(p0)
        p = p0
        for step in (1, 2, 3):
            g = rng.normal(size=p0.shape).astype(np.float32)
            net.adam_step(g, 1e-3)
            p, m, v = adam_reference(p, g, m, v, step, 1e-3)
            got = net.parameters()
            assert np.allclose(got, p, rtol=2e-6, atol=1e-7), f"step {step}"
    finally:
        net.close()


def test_chunked_training_equals_one_run(gpu_context):
    """Training in calls of a few epochs (chd_proposal_train_epochs, with
    progress between them) is the same run as one call: the same shuffles
    and step sizes. The backward pass sums weight gradients with float
    atomics, so two runs of either kind agree to rounding only; the losses
    and parameters are compared at that level, against a run whose schedule
    restarts at every call (three separate chd_proposal_train calls), which
    must stray further."""
    rng = np.random.default_rng(6)
    w, h, count = 24, 16, 12
    inputs = rng.integers(0x3c00, 0x4000, (count, h, w, 16), dtype=np.uint16)
    labels = rng.integers(1, 60, (count, h, w), dtype=np.uint16)
    nets = [Proposal(gpu_context, w, h, max_batch=4, seed=9) for _ in range(3)]
    try:
        for net in nets:
            net.set_label_count(65)
            net.set_dataset(inputs, labels)
        seen = []
        one = np.array(nets[0].train(epochs=7, batch=4, lr=3e-3, seed=3))
        chunked = np.array(nets[1].train(epochs=7, batch=4, lr=3e-3, seed=3, chunk=3,
                                         log=lambda done, l: seen.append(done)))
        restarted = np.concatenate([nets[2].train(epochs=n, batch=4, lr=3e-3, seed=3)
                                    for n in (3, 3, 1)])
        assert seen == [3, 6, 7]
        p0, p1, p2 = (n.parameters() for n in nets)
        chunk_gap = np.abs(p1 - p0).max()
        restart_gap = np.abs(p2 - p0).max()
        _say(f"chunked vs one run: losses {np.abs(chunked - one).max():.2e}, parameters "
             f"{chunk_gap:.2e}; restarted schedule: parameters {restart_gap:.2e}")
        np.testing.assert_allclose(chunked, one, rtol=1e-3)
        assert chunk_gap < 1e-3 and restart_gap > 5 * chunk_gap
    finally:
        for net in nets:
            net.close()


def _cap(seed):
    return scene.config3(seed=seed, width=520, height=392, focal_px=800.0)


def test_training_and_band_sweep(gpu_context, oracle_lib):
    """Train on a few small config 3 scenes: the loss falls; every tile in
    fallback gives the plain full sweep bit for bit; with bands the cost
    volume holds +inf outside each node's window and the windowed SGM equals
    the oracle's full SGM over that volume bit for bit."""
    _say("building 12 small config3 scenes")
    offset, scale = input_transform([_cap(s) for s in range(4)])
    x, y = build_dataset(gpu_context, range(12), _cap, offset, scale)
    h, w = x.shape[1:3]
    net = Proposal(gpu_context, w, h, max_batch=4, seed=7)
    try:
        net.set_input_transform(offset, scale)
        net.set_dataset(x, y)
        losses = net.train(epochs=8, batch=4, lr=1e-3, seed=1)
        _say(f"losses {['%.3f' % l for l in losses]}")
        assert np.isfinite(losses).all() and losses[-1] < losses[0]
        cap = _cap(100)
        full = Session(gpu_context, cap.modules, 0, cap.options)
        full.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM | _abi.CHD_STAGE_KEEP_AGGREGATED)
        want_scores, want_costs, _, want_tested = full.cost_volume()
        want_best = full.coarse()[2]
        # every tile in fallback: the plain sweep
        fb, tiles = net.apply(full, threshold=2.0)
        assert fb == tiles
        full.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM)
        s, c, _, t = full.cost_volume()
        assert_bitwise("fallback scores", s, want_scores)
        assert_bitwise("fallback costs", c, want_costs)
        assert_bitwise("fallback tested", t, want_tested)
        assert_bitwise("fallback labels", full.coarse()[2], want_best)
        # bands everywhere: windows of +-8 around the centre, then widths
        # from the predicted scale (+-8 .. +-24)
        for max_half, most in ((0, 18), (24, 50)):
            fb, tiles = net.apply(full, threshold=0.0, max_half_width=max_half)
            assert fb == 0
            full.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM | _abi.CHD_STAGE_KEEP_AGGREGATED)
            band_scores, costs, guidance, _ = full.cost_volume()
            swept = assert_band_equals_full(f"band (max half {max_half})", full, band_scores,
                                            costs, want_scores, want_costs)
            _say(f"max half {max_half}: {swept} (node, label) pairs equal the full sweep")
            swept = np.isfinite(costs).sum(axis=1)
            assert swept.max() <= most and swept.min() >= 1
            if max_half == 0:
                # the device builds the network input from the session's planes
                # exactly as capture_inputs does on the host: the windows equal
                # the ones the host-side prediction gives
                g = full.grid
                pred = net.predict(capture_inputs(cap, offset, scale)[None])[0]
                r = np.minimum(np.arange(g.coarse_rows), pred.shape[0] - 1)
                c = np.minimum(np.arange(g.coarse_columns), pred.shape[1] - 1)
                centre = np.clip(np.rint(pred[np.ix_(r, c)][..., 0]), 1, g.labels - 1)
                want_lo = np.maximum(1, centre - 8).reshape(-1)
                want_hi = np.minimum(g.labels - 1, centre + 8).reshape(-1)
                fin = np.isfinite(costs[:, 1:])
                got_lo = 1 + np.argmax(fin, axis=1)
                got_hi = g.labels - 1 - np.argmax(fin[:, ::-1], axis=1)
                assert np.array_equal(got_lo, want_lo) and np.array_equal(got_hi, want_hi)
            g = full.grid
            want = np.empty_like(costs)
            oracle_lib.or_semi_global_costs(_abi.ptr(costs), _abi.ptr(guidance), g.coarse_columns,
                                            g.coarse_rows, g.labels, _abi.ptr(want))
            assert_bitwise(f"band SGM (max half {max_half})", full.aggregated(), want)
        clear(full)
        full.close()
        stats = agreement(gpu_context, _cap(101), net, threshold=0.5)
        _say(f"band vs full: {stats}")
        assert 0.0 <= stats["agreement"] <= 1.0 and stats["band_ms"] > 0
    finally:
        net.close()


def test_ablation_network_beats_constant_band(gpu_context):
    """Ablation at an equal hypothesis budget (+-8 labels at every node, no
    fallback): the band centred by the trained network against one centred
    on the training set's median label. The network must cover the full
    sweep's label, and reproduce it within +-1, far more often."""
    _say("building 1000 small config3 scenes")
    offset, scale = input_transform([_cap(s) for s in range(1000, 1008)])
    x, y = build_dataset(gpu_context, range(1000, 2000), _cap, offset, scale)
    h, w = x.shape[1:3]
    net = Proposal(gpu_context, w, h, max_batch=32, seed=11)
    const = constant_proposal(gpu_context, w, h, float(np.median(y[y > 0])))
    try:
        net.set_input_transform(offset, scale)
        net.set_dataset(x, y)
        losses = net.train(epochs=60, batch=32, lr=1e-3, seed=11)
        _say(f"losses {losses[0]:.3f} -> {losses[-1]:.3f}")
        with_net, without = [], []
        for seed in (3000, 3001, 3002):
            cap = _cap(seed)
            with_net.append(agreement(gpu_context, cap, net, threshold=0.0, steps=1))
            without.append(agreement(gpu_context, cap, const, threshold=0.0, steps=1))
        mean = lambda rows, k: float(np.mean([r[k] for r in rows]))
        a_net, a_const = mean(with_net, "agreement"), mean(without, "agreement")
        c_net, c_const = mean(with_net, "window_covers_label"), mean(without, "window_covers_label")
        _say(f"network: agreement {a_net:.3f} coverage {c_net:.3f}; "
             f"constant centre: agreement {a_const:.3f} coverage {c_const:.3f}")
        assert all(r["fallback_fraction"] == 0.0 for r in with_net + without)
        assert c_net > 2.0 * c_const and a_net > 1.5 * a_const
    finally:
        net.close()
        const.close()


def test_band_on_an_8px_grid_with_257_labels(gpu_context, oracle_lib):
    """A network on the coarse grid of a config 4 capture (8 px step: 4x4
    plane cells per network pixel, 257 labels): the device-built input gives
    the host's windows exactly and the band SGM equals the oracle's."""
    cap = scene.config4(seed=3, scale=0.125)
    sess = Session(gpu_context, cap.modules, 0, cap.options)
    g = sess.grid
    sess.run(_abi.CHD_STAGE_COST)
    want_scores, want_costs, _, _ = sess.cost_volume()
    # four work-groups per CU for every ZNCC kernel (the tap loops' register
    # budget assumes it; a block a few bytes over its LDS quarter runs three)
    assert sess.occupancy() == {"cost_volume": 4, "direct": 4, "targets": 4}
    ref = cap.modules[0].luminance
    h, w = ref.shape[0] // 4, ref.shape[1] // 4
    net = Proposal(gpu_context, w, h, max_batch=1, seed=5)
    try:
        net.set_label_count(g.labels)
        offset, scale = input_transform([cap])
        net.set_input_transform(offset, scale)
        params = _params(w, h, np.random.default_rng(6), scale=0.3)
        net.set_parameters(params)
        fb, tiles = net.apply(sess, threshold=0.0)
        assert fb == 0 and tiles > 0
        assert sess.occupancy()["cost_volume"] == 4
        sess.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM | _abi.CHD_STAGE_KEEP_AGGREGATED)
        band_scores, costs, guidance, _ = sess.cost_volume()
        assert_band_equals_full("band at 257 labels", sess, band_scores, costs, want_scores,
                                want_costs)
        pred = net.predict(capture_inputs(cap, offset, scale)[None])[0]
        r = np.minimum(np.arange(g.coarse_rows), h - 1)
        c = np.minimum(np.arange(g.coarse_columns), w - 1)
        centre = np.clip(np.rint(pred[np.ix_(r, c)][..., 0]), 1, g.labels - 1)
        fin = np.isfinite(costs[:, 1:])
        assert np.array_equal(1 + np.argmax(fin, axis=1), np.maximum(1, centre - 8).reshape(-1))
        assert np.array_equal(g.labels - 1 - np.argmax(fin[:, ::-1], axis=1),
                              np.minimum(g.labels - 1, centre + 8).reshape(-1))
        want = np.empty_like(costs)
        oracle_lib.or_semi_global_costs(_abi.ptr(costs), _abi.ptr(guidance), g.coarse_columns,
                                        g.coarse_rows, g.labels, _abi.ptr(want))
        assert_bitwise("band SGM at 257 labels", sess.aggregated(), want)
        # some tiles in fallback beside banded ones: the path sweep moves
        # every label chunk at the fallback nodes and the window's chunks at
        # the others; the SGM still equals the oracle's over the volume
        mixed = None
        for threshold in (0.5, 0.3, 0.7, 0.4, 0.6):
            fb, tiles = net.apply(sess, threshold=threshold)
            if 0 < fb < tiles:
                mixed = threshold
                break
        assert mixed is not None, "no threshold gives a mix of banded and fallback tiles"
        sess.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM | _abi.CHD_STAGE_KEEP_AGGREGATED)
        band_scores, costs, guidance, _ = sess.cost_volume()
        assert_band_equals_full(f"band with {fb} of {tiles} tiles in fallback", sess, band_scores,
                                costs, want_scores, want_costs)
        oracle_lib.or_semi_global_costs(_abi.ptr(costs), _abi.ptr(guidance), g.coarse_columns,
                                        g.coarse_rows, g.labels, _abi.ptr(want))
        assert_bitwise("band SGM with fallback tiles", sess.aggregated(), want)
        # every tile in fallback with a +-20 fallback window: centre +- 20
        fb, tiles = net.apply(sess, threshold=2.0, fallback_half_width=20)
        assert fb == tiles
        assert sess.occupancy()["cost_volume"] == 4
        sess.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM | _abi.CHD_STAGE_KEEP_AGGREGATED)
        band_scores, costs, guidance, _ = sess.cost_volume()
        assert_band_equals_full("fallback windows", sess, band_scores, costs, want_scores,
                                want_costs)
        lo, hi = sess.band()
        assert np.array_equal(lo, np.maximum(1, centre - 20).reshape(-1))
        assert np.array_equal(hi, np.minimum(g.labels - 1, centre + 20).reshape(-1))
        oracle_lib.or_semi_global_costs(_abi.ptr(costs), _abi.ptr(guidance), g.coarse_columns,
                                        g.coarse_rows, g.labels, _abi.ptr(want))
        assert_bitwise("band SGM with fallback windows", sess.aggregated(), want)
        clear(sess)
    finally:
        net.close()
        sess.close()
