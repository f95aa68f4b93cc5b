"""The generator on the GPU (chd_frames) against the host generator: the
same Bayer codes and luminance planes byte for byte, for a translated ring
(config 1), a rotated rig of uneven modules (config 2) and nested wide, mid
and tele fields (config 4); sessions made from the device planes, or fed
them through chd_session_set_planes, give the host-plane session's results
bit for bit; the network input built from device planes equals the host's."""
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import Session

from parity import assert_bitwise, assert_nodes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


RIGS = {
    "config1": lambda: scene.rig_config1(seed=7, width=416, height=312, focal_px=330.0, planes=32),
    "config2": lambda: scene.rig_config2(seed=3, scale=0.125),
    "config4": lambda: scene.rig_config4(seed=5, scale=0.125),
}


@pytest.mark.parametrize("name", sorted(RIGS))
def test_device_render_equals_host_render(gpu_context, name):
    rig = RIGS[name]()
    host = scene.render_host(rig)
    frames = scene.Frames(gpu_context, rig)
    try:
        frames.render(rig.scene.seed)
        for i, raw in enumerate(host.raws):
            mosaic, plane = frames.read(i)
            assert_bitwise(f"{name} module {i} codes", mosaic, (raw << 4).astype(np.uint16))
            assert_bitwise(f"{name} module {i} plane", plane, host.modules[i].luminance)
        _say(f"{name}: {len(host.raws)} modules equal the host render")
    finally:
        frames.close()


def test_full_size_config4_render_equals_host(gpu_context):
    """One full 16-module config 4 capture: every module's codes."""
    rig = scene.rig_config4(seed=11)
    host = scene.render_host(rig)
    frames = scene.Frames(gpu_context, rig)
    try:
        frames.render(11)
        for i, raw in enumerate(host.raws):
            mosaic, _ = frames.read(i)
            assert_bitwise(f"config4 module {i} codes", mosaic, (raw << 4).astype(np.uint16))
    finally:
        frames.close()


def _outputs(sess):
    sess.run(_abi.CHD_STAGE_ALL)
    scores, costs, _, tested = sess.cost_volume()
    field, _, best = sess.coarse()
    dmap, targets = sess.results()
    return scores, costs, best, field, dmap, targets


def test_sessions_on_device_planes(gpu_context):
    rig = RIGS["config1"]()
    host = scene.render_host(rig)
    want = Session(gpu_context, host.modules, 0, host.options)
    w = _outputs(want)
    want.close()
    frames = scene.Frames(gpu_context, rig)
    sess = None
    try:
        frames.render(rig.scene.seed)
        sess = Session(gpu_context, frames.capture().modules, 0, rig.options)
        got = _outputs(sess)
        for label, g, x in (("scores", got[0], w[0]), ("costs", got[1], w[1]),
                            ("best", got[2], w[2]), ("depth", got[4].depth, w[4].depth)):
            assert_bitwise(f"device planes {label}", g, x)
        assert_nodes("device planes coarse", got[3], w[3])
        # a second scene through set_planes: equals a fresh host session
        host2 = scene.render_host(rig.with_seed(8))
        frames.render(8)
        sess.set_planes(frames.planes())
        got2 = _outputs(sess)
        ref2 = Session(gpu_context, host2.modules, 0, host2.options)
        w2 = _outputs(ref2)
        ref2.close()
        assert_bitwise("set_planes best", got2[2], w2[2])
        assert_bitwise("set_planes costs", got2[1], w2[1])
        for t in w2[5]:
            assert_bitwise(f"set_planes target {t}", got2[5][t].decisions, w2[5][t].decisions)
        # host planes through set_planes as well
        sess.set_planes([m.luminance for m in host.modules])
        got3 = _outputs(sess)
        assert_bitwise("host set_planes best", got3[2], w[2])
    finally:
        if sess is not None:
            sess.close()
        frames.close()


def test_network_input_from_device_planes(gpu_context):
    from chiaro_amd.proposal import Proposal, capture_inputs, input_transform
    rig = scene.rig_config4(seed=2, scale=0.25, wide_only=True)
    rig.options.grid_step = 4
    host = scene.render_host(rig)
    offset, scale = input_transform([host])
    frames = scene.Frames(gpu_context, rig)
    try:
        frames.render(2)
        sess = Session(gpu_context, frames.capture().modules, 0, rig.options)
        h, w = host.modules[0].luminance.shape
        net = Proposal(gpu_context, w // 2, h // 2, max_batch=1)
        net.set_label_count(sess.grid.labels)
        net.set_input_transform(offset, scale)
        got = net.session_input(sess)
        want = capture_inputs(host, offset, scale)
        assert_bitwise("network input", got, want)
        net.close()
        sess.close()
    finally:
        frames.close()
