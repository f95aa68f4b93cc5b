"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle): two tiny captures with every
intermediate and output. The oracle must still reproduce them bit for bit
(CPU), and so must the GPU path, read straight from the fixture without
running the oracle."""
import re
from pathlib import Path

import numpy as np
import pytest

from chiaro_amd import _abi, scene
from chiaro_amd.depth import DepthOptions, Session
from oracle import binding

from parity import assert_bitwise, assert_nodes

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        data = {k: f[k] for k in f.files}
    raw = _abi.Options.from_buffer_copy(data["options"].tobytes())
    options = DepthOptions(**{k: (bool(getattr(raw, k)) if k == "enabled" else getattr(raw, k))
                              for k, _ in _abi.Options._fields_})
    modules = []
    for i in range(int(data["modules"])):
        w, h = (int(v) for v in data[f"m{i}.raster"])
        modules.append(scene.ModuleInput(
            name=f"M{i}", luminance=data[f"m{i}.luminance"], width=w, height=h,
            camera=_abi.Camera.from_buffer_copy(data[f"m{i}.camera"].tobytes()),
            warp_points=data[f"m{i}.warp"], warp_step=int(data[f"m{i}.warp_step"])))
    cap = scene.Capture(modules=modules, reference_index=int(data["reference_index"]),
                        options=options)
    return cap, data


def test_fixtures_exist():
    assert len(CASES) >= 2


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_the_golden_outputs(name):
    cap, data = load(name)
    omap, targets, dbg = binding.refine(cap, threads=4)
    for k in ("scores", "costs", "guidance", "tested", "aggregated", "fillable", "best_label",
              "fine_guidance", "fine_tested"):
        assert_bitwise(f"{name} {k}", dbg[k], data["out." + k])
    assert_nodes(f"{name} coarse", dbg["coarse"], data["out.coarse"])
    assert_nodes(f"{name} field", dbg["direct"][3], data["out.field"])
    for k in ("depth", "confidence", "provenance"):
        assert_bitwise(f"{name} map {k}", omap[k], data["out.map." + k])
    want_targets = sorted({int(m.group(1)) for m in map(re.compile(r"out\.t(\d+)\.").match, data)
                           if m})
    assert sorted(targets) == want_targets
    for t in targets:
        for k in ("decisions", "points", "confidence"):
            assert_bitwise(f"{name} target {t} {k}", targets[t][k], data[f"out.t{t}.{k}"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_reproduces_the_golden_outputs(gpu_context, name):
    cap, data = load(name)
    sess = Session(gpu_context, cap.modules, cap.reference_index, cap.options)
    try:
        sess.run(_abi.CHD_STAGE_ALL | _abi.CHD_STAGE_KEEP_AGGREGATED)
        scores, costs, guidance, tested = sess.cost_volume()
        coarse, fillable, best = sess.coarse()
        field = sess.direct(3)[0]
        aggregated = sess.aggregated()
        dmap, targets = sess.results()
    finally:
        sess.close()
    assert_bitwise(f"{name} scores", scores, data["out.scores"])
    assert_bitwise(f"{name} costs", costs, data["out.costs"])
    assert_bitwise(f"{name} aggregated", aggregated, data["out.aggregated"])
    assert_bitwise(f"{name} best label", best, data["out.best_label"])
    assert_nodes(f"{name} coarse", coarse, data["out.coarse"])
    assert_nodes(f"{name} field", field, data["out.field"])
    assert_bitwise(f"{name} map depth", dmap.depth, data["out.map.depth"])
    assert_bitwise(f"{name} map provenance", dmap.provenance, data["out.map.provenance"])
    for t in targets:
        assert_bitwise(f"{name} target {t} decisions", targets[t].decisions,
                       data[f"out.t{t}.decisions"])
        assert_bitwise(f"{name} target {t} points", targets[t].points, data[f"out.t{t}.points"])
