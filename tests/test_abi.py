"""The C ABI: the product library loads without a GPU, exports every
function include/chiaro_depth.h declares, and the ctypes structs match the C
compiler's layout."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from chiaro_amd import _abi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include/chiaro_depth.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    if not _abi.DEPTH_LIB.exists():
        from chiaro_amd import build
        build.build_depth()
    lib = C.CDLL(str(_abi.DEPTH_LIB))
    names = declared_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(_abi.DEPTH_SIGNATURES) == names


def test_cpu_only_calls():
    lib = _abi.depth_lib()
    assert lib.chd_abi_version() == 1
    o = _abi.Options()
    lib.chd_default_options(C.byref(o))
    assert (o.grid_step, o.planes, o.patch_radius, o.best_view_count) == (8, 96, 4, 3)
    assert lib.chd_options_valid(C.byref(o)) == 1
    o.planes = 2
    assert lib.chd_options_valid(C.byref(o)) == 0
    o.planes = 128
    g = _abi.GridSizes()
    assert lib.chd_grid_sizes_for(C.byref(o), 4160, 3120, C.byref(g)) == 0
    assert (g.coarse_columns, g.coarse_rows, g.labels) == (521, 391, 129)
    assert (g.fine_columns, g.fine_rows, g.fine_step) == (1041, 781, 4)


LAYOUT = {
    "chd_options": (_abi.Options, ["near_depth", "planes", "minimum_score", "completion_iterations"]),
    "chd_distortion": (_abi.Distortion, ["normalization", "coeffs", "coeff_count"]),
    "chd_camera": (_abi.Camera, ["k_inverse", "distortion", "flip_around_x", "rotation_wc", "center",
                                 "orientation_correction", "focal_px"]),
    "chd_warp": (_abi.Warp, ["points", "confidence"]),
    "chd_view": (_abi.View, ["camera", "global_warp", "accepted"]),
    "chd_depth_report": (_abi.DepthReport, ["reconstructed_fraction", "median_depth",
                                            "median_score_improvement"]),
    "chd_target_result": (_abi.TargetResult, ["decisions", "report"]),
    "chd_depth_map": (_abi.DepthMap, ["near_depth", "depth", "provenance"]),
    "chd_node_depth": (_abi.NodeDepth, ["confidence", "valid", "regularized"]),
    "chd_grid_sizes": (_abi.GridSizes, ["labels", "fine_rows"]),
    "chd_node_warp": (_abi.NodeWarp, ["point", "measured"]),
    "chd_kernel_time": (_abi.KernelTime, ["launches", "total_ms"]),
    "chd_work": (_abi.Work, ["direct_zncc", "wide_nodes"]),
    "chd_align_options": (_abi.AlignOptions, ["min_score", "min_inlier_ratio", "refine"]),
    "chd_level_report": (_abi.LevelReport, ["inliers", "median_residual_px"]),
    "chd_alignment_report": (_abi.AlignmentReport, ["coverage", "levels", "inliers",
                                                    "residual_p90_px", "accepted", "gain"]),
    "chd_synth_source": (_abi.SynthSource, ["mosaic_width", "magnification", "reference",
                                            "focus_distance"]),
    "chd_canvas": (_abi.Canvas, ["crop_height", "feather_px"]),
    "chd_resolution_report": (_abi.ResolutionReport, ["supported_fraction", "mean_confidence"]),
    "chd_synthetic_scene": (_abi.SceneDesc, ["plane_count", "inverse_depth_max",
                                             "texture_cell_px"]),
    "chd_synthetic_module": (_abi.Module, ["width", "center", "rotation_wc"]),
    "chd_bayer": (_abi.Bayer, ["exposure", "read_noise_dn"]),
}


def test_struct_layouts_match_the_c_compiler(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, (_, members) in LAYOUT.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for m in members:
            lines.append(f'printf("{cname}.{m} %zu\\n", offsetof({cname}, {m}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.split("\n") if l)
    for cname, (cls, members) in LAYOUT.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for m in members:
            field = "global_" if m == "global" else m
            assert int(out[f"{cname}.{m}"]) == getattr(cls, field).offset, f"{cname}.{m}"


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits non-zero with a message,
    before touching a device."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    proc = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"],
                          capture_output=True, text=True, timeout=120, cwd=root)
    assert proc.returncode != 0
    assert "one GPU" in proc.stdout + proc.stderr
