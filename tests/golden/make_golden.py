"""Writes the golden fixtures of the depth path: two tiny seeded captures
(inputs: planes, camera structs, global warps, options) with the oracle's
outputs (cost volume, aggregated SGM volume, coarse field and WTA labels,
the final fine field, the depth map and every target's decisions, points
and confidences). The fixtures are data only (.npz, no pickled objects);
tests/test_golden.py replays them against the oracle and the GPU.

    python tests/golden/make_golden.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from chiaro_amd import scene  # noqa: E402
from oracle import binding  # noqa: E402

CASES = {
    "config0_tiny_seed3": lambda: scene.config0(seed=3, width=64, height=48, focal_px=100.0,
                                                planes=16),
    "config1_tiny_seed5": lambda: scene.config1(seed=5, width=96, height=64, focal_px=76.0,
                                                planes=20),
}


def capture_arrays(cap) -> dict:
    out = {"reference_index": np.array(cap.reference_index, np.uint32)}
    o = cap.options
    out["options"] = np.frombuffer(bytes(o.to_c()), np.uint8).copy()
    for i, m in enumerate(cap.modules):
        out[f"m{i}.luminance"] = np.ascontiguousarray(m.luminance, np.float32)
        out[f"m{i}.raster"] = np.array([m.width, m.height], np.uint32)
        out[f"m{i}.camera"] = np.frombuffer(bytes(m.camera), np.uint8).copy()
        out[f"m{i}.warp"] = np.ascontiguousarray(m.warp_points, np.float32)
        out[f"m{i}.warp_step"] = np.array(m.warp_step, np.uint32)
    out["modules"] = np.array(len(cap.modules), np.uint32)
    return out


def main():
    for name, make in CASES.items():
        cap = make()
        omap, targets, dbg = binding.refine(cap, threads=4)
        data = capture_arrays(cap)
        for k in ("scores", "costs", "guidance", "tested", "aggregated", "fillable",
                  "best_label", "fine_guidance", "fine_tested"):
            data["out." + k] = dbg[k]
        data["out.coarse"] = dbg["coarse"]
        data["out.field"] = dbg["direct"][3]
        for k in ("depth", "confidence", "provenance"):
            data["out.map." + k] = omap[k]
        for t, r in targets.items():
            for k in ("decisions", "points", "confidence"):
                data[f"out.t{t}.{k}"] = r[k]
        np.savez_compressed(HERE / f"{name}.npz", **data)
        print(name, sum(v.nbytes for v in data.values()), "bytes")


if __name__ == "__main__":
    main()
