"""Object lifetimes over the C ABI: contexts, sessions, batches and
proposals destroyed in every order (the context outliving nothing), a batch
whose session was destroyed refusing to run, no new object on a destroyed
context; autotuning a session held by a persistent batch; and proposal
checkpoints (parameters plus Adam state) that restore the optimiser
exactly."""
import ctypes as C
import itertools
import sys

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import Batch, Context, Session, _ViewArrays
from chiaro_amd.proposal import Proposal

from parity import assert_bitwise

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cap(seed=5):
    return scene.config0(seed=seed, width=128, height=96, focal_px=200.0, planes=16)


def _raw_objects(lib, cap):
    """A context with one session, one batch over it and one proposal, made
    through the C ABI directly (no Python-side bookkeeping)."""
    ctx = C.c_void_p()
    _abi.check(lib.chd_context_create(0, C.byref(ctx)), "context")
    views = _ViewArrays(cap.modules)
    sess = C.c_void_p()
    _abi.check(lib.chd_session_create(ctx, views.views, len(cap.modules), cap.reference_index,
                                      C.byref(cap.options.to_c()), C.byref(sess)), "session")
    arr = (C.c_void_p * 1)(sess.value)
    batch = C.c_void_p()
    _abi.check(lib.chd_batch_create(ctx, arr, 1, 0, C.byref(batch)), "batch")
    prop = C.c_void_p()
    _abi.check(lib.chd_proposal_create(ctx, 16, 12, 1, 3, C.byref(prop)), "proposal")
    return {"context": ctx, "session": sess, "batch": batch, "proposal": prop}, views


DESTROY = {"context": "chd_context_destroy", "session": "chd_session_destroy",
           "batch": "chd_batch_destroy", "proposal": "chd_proposal_destroy"}


@pytest.mark.parametrize("order", list(itertools.permutations(DESTROY)))
def test_any_destroy_order(order):
    lib = _abi.depth_lib()
    _say(f"destroy order {order}")
    objs, views = _raw_objects(lib, _cap())
    assert lib.chd_batch_run(objs["batch"], _abi.CHD_STAGE_ALL) == 0
    gone = set()
    for name in order:
        getattr(lib, DESTROY[name])(objs[name])
        gone.add(name)
        if "session" in gone and "batch" not in gone:
            # the batch outlives its session: it refuses to run
            assert lib.chd_batch_run(objs["batch"], _abi.CHD_STAGE_COST) == _abi.CHD_ERR_STATE
            assert "destroyed" in _abi.last_error()
        if "context" in gone and "session" not in gone:
            # a released context keeps serving the objects made on it
            assert lib.chd_session_run(objs["session"], _abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM) == 0
    del views


def test_no_new_objects_on_a_destroyed_context():
    lib = _abi.depth_lib()
    cap = _cap()
    objs, views = _raw_objects(lib, cap)
    lib.chd_context_destroy(objs["context"])
    sess = C.c_void_p()
    rc = lib.chd_session_create(objs["context"], views.views, len(cap.modules),
                                cap.reference_index, C.byref(cap.options.to_c()), C.byref(sess))
    assert rc == _abi.CHD_ERR_STATE and not sess.value
    prop = C.c_void_p()
    assert lib.chd_proposal_create(objs["context"], 16, 12, 1, 3, C.byref(prop)) == _abi.CHD_ERR_STATE
    for name in ("batch", "session", "proposal"):
        getattr(lib, DESTROY[name])(objs[name])


def test_python_objects_closed_in_any_order():
    cap = _cap(6)
    ctx = Context(0)
    sess = Session(ctx, cap.modules, cap.reference_index, cap.options)
    batch = Batch(ctx, [sess])
    net = Proposal(ctx, 16, 12, max_batch=1)
    batch.run()
    sess.close()
    with pytest.raises(RuntimeError):
        batch.run()
    ctx.close()          # closes the batch and the proposal too
    assert batch.handle is None and net.handle is None
    net.close()
    batch.close()


def test_autotune_under_a_persistent_batch(gpu_context, tmp_path):
    """A persistent batch holds a session; autotuning the session changes its
    launch shape, and the next batch run with per-view slices must rebuild its
    job tables: every output equals a fresh single session's."""
    cap = scene.config0(seed=9, width=256, height=192, focal_px=400.0)
    want = Session(gpu_context, cap.modules, cap.reference_index, cap.options)
    keep = _abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM | _abi.CHD_STAGE_KEEP_VIEW_SCORES
    want.run(keep | _abi.CHD_STAGE_KEEP_AGGREGATED)
    want_views = want.view_scores()
    want_agg = want.aggregated()
    want.close()
    sess = Session(gpu_context, cap.modules, cap.reference_index, cap.options)
    batch = Batch(gpu_context, [sess], persistent=True)
    try:
        sess.set_tuning(2, 0, 1, 0)
        batch.run(keep)
        tuned = sess.autotune(tmp_path / "tune.txt")
        _say(f"autotuned under the batch: {tuned}")
        assert sess.tuning() == tuned
        batch.run(keep | _abi.CHD_STAGE_KEEP_AGGREGATED)
        assert_bitwise("view scores after autotune", sess.view_scores(), want_views)
        assert_bitwise("aggregated after autotune", sess.aggregated(), want_agg)
    finally:
        batch.close()
        sess.close()


def test_checkpoint_restores_the_optimiser(gpu_context, tmp_path):
    rng = np.random.default_rng(8)
    a = Proposal(gpu_context, 16, 12, max_batch=1, seed=2)
    b = Proposal(gpu_context, 16, 12, max_batch=1, seed=99)
    try:
        a.set_input_transform(6.0, 3.5)
        for _ in range(3):
            a.adam_step(rng.normal(size=a.parameter_count).astype(np.float32), 1e-3)
        path = tmp_path / "net.npz"
        a.save(path)
        b.load(path)
        moments_a, step_a = a.optimizer_state()
        moments_b, step_b = b.optimizer_state()
        assert step_a == step_b == 3
        assert_bitwise("moments", moments_b, moments_a)
        assert (b.input_offset, b.input_scale) == (a.input_offset, a.input_scale)
        g = rng.normal(size=a.parameter_count).astype(np.float32)
        a.adam_step(g, 1e-3)
        b.adam_step(g, 1e-3)
        assert_bitwise("parameters after the next step", b.parameters(), a.parameters())
    finally:
        a.close()
        b.close()
