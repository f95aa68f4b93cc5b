"""ctypes mirror of include/chiaro_depth.h and loaders of the built libraries.

The structs below follow the header field by field; tests check their sizes
against the C compiler's. Libraries are loaded from chiaro_amd/lib/ (built
in-tree by chiaro_amd.build); a missing library raises instead of falling
back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PACKAGE = Path(__file__).resolve().parent
LIB_DIR = PACKAGE / "lib"
DEPTH_LIB = LIB_DIR / "libchiaro_depth.so"
SCENE_LIB = LIB_DIR / "libchiaro_scene.so"

CHD_OK = 0
CHD_NONE = 1
CHD_ERR_ARGUMENT = -1
CHD_ERR_HIP = -2
CHD_ERR_NO_DEVICE = -3
CHD_ERR_LIMIT = -4
CHD_ERR_STATE = -5
CHD_MAX_VIEWS = 32
CHD_NOALLOC_MAX_NODES = 2097152

CHD_STAGE_COST = 1
CHD_STAGE_SGM = 2
CHD_STAGE_DIRECT = 4
CHD_STAGE_TARGETS = 8
CHD_STAGE_ALL = 15
CHD_STAGE_PROPOSE = 16
CHD_STAGE_KEEP_AGGREGATED = 256
CHD_STAGE_KEEP_VIEW_SCORES = 512
CHD_BATCH_PERSISTENT = 1
CHD_BATCH_ONE_SWEEP = 2

WARP_UNDEFINED, WARP_GLOBAL, WARP_REFINED, WARP_OCCLUDED, WARP_BOUNDARY = range(5)
PROVENANCE_UNSUPPORTED, PROVENANCE_GLOBAL, PROVENANCE_MEASURED, PROVENANCE_REGULARIZED = range(4)


class Options(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("grid_step", C.c_uint32),
        ("near_depth", C.c_double),
        ("far_depth", C.c_double),
        ("planes", C.c_uint32),
        ("patch_radius", C.c_uint32),
        ("minimum_support", C.c_uint32),
        ("best_view_count", C.c_uint32),
        ("minimum_score", C.c_float),
        ("minimum_margin", C.c_float),
        ("minimum_improvement", C.c_float),
        ("minimum_neighbour_support", C.c_uint32),
        ("completion_iterations", C.c_uint32),
    ]


class Distortion(C.Structure):
    _fields_ = [
        ("center", C.c_double * 2),
        ("normalization", C.c_double * 2),
        ("coeffs", C.c_double * 5),
        ("coeff_count", C.c_uint32),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("k", C.c_double * 9),
        ("k_inverse", C.c_double * 9),
        ("has_distortion", C.c_int32),
        ("distortion", Distortion),
        ("flip_around_x", C.c_int32),
        ("pose_kind", C.c_int32),
        ("rotation_wc", C.c_double * 9),
        ("rotation_cw", C.c_double * 9),
        ("translation_wc", C.c_double * 3),
        ("center", C.c_double * 3),
        ("real_cw", C.c_double * 9),
        ("reflect", C.c_double * 9),
        ("virtual_center", C.c_double * 3),
        ("orientation_correction", C.c_double * 9),
        ("focal_px", C.c_double),
    ]


class Warp(C.Structure):
    _fields_ = [
        ("step", C.c_uint32),
        ("columns", C.c_uint32),
        ("rows", C.c_uint32),
        ("points", C.POINTER(C.c_float)),
        ("confidence", C.POINTER(C.c_float)),
    ]


class View(C.Structure):
    _fields_ = [
        ("luminance", C.POINTER(C.c_float)),
        ("plane_width", C.c_uint32),
        ("plane_height", C.c_uint32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("camera", C.POINTER(Camera)),
        ("global_warp", Warp),
        ("accepted", C.c_int32),
    ]


class DepthReport(C.Structure):
    _fields_ = [
        ("tested_nodes", C.c_uint64),
        ("measured_nodes", C.c_uint64),
        ("regularized_nodes", C.c_uint64),
        ("fallback_nodes", C.c_uint64),
        ("refined_nodes", C.c_uint64),
        ("occluded_nodes", C.c_uint64),
        ("boundary_nodes", C.c_uint64),
        ("reconstructed_fraction", C.c_float),
        ("refined_fraction", C.c_float),
        ("occluded_fraction", C.c_float),
        ("has_median_depth", C.c_int32),
        ("median_depth", C.c_double),
        ("has_median_score_improvement", C.c_int32),
        ("median_score_improvement", C.c_float),
    ]


class TargetResult(C.Structure):
    _fields_ = [
        ("replaced", C.c_int32),
        ("points", C.POINTER(C.c_float)),
        ("confidence", C.POINTER(C.c_float)),
        ("decisions", C.POINTER(C.c_uint8)),
        ("report", DepthReport),
    ]


class DepthMap(C.Structure):
    _fields_ = [
        ("columns", C.c_uint32),
        ("rows", C.c_uint32),
        ("step", C.c_uint32),
        ("near_depth", C.c_double),
        ("far_depth", C.c_double),
        ("depth", C.POINTER(C.c_double)),
        ("confidence", C.POINTER(C.c_float)),
        ("provenance", C.POINTER(C.c_uint8)),
    ]


class NodeDepth(C.Structure):
    _fields_ = [
        ("depth", C.c_double),
        ("confidence", C.c_float),
        ("improvement", C.c_float),
        ("valid", C.c_uint8),
        ("regularized", C.c_uint8),
        ("pad", C.c_uint8 * 6),
    ]


NODE_DTYPE = np.dtype(
    [("depth", "<f8"), ("confidence", "<f4"), ("improvement", "<f4"), ("valid", "u1"),
     ("regularized", "u1"), ("pad", "u1", (6,))]
)
assert NODE_DTYPE.itemsize == C.sizeof(NodeDepth) == 24


class GridSizes(C.Structure):
    _fields_ = [
        ("coarse_step", C.c_uint32),
        ("coarse_columns", C.c_uint32),
        ("coarse_rows", C.c_uint32),
        ("labels", C.c_uint32),
        ("fine_step", C.c_uint32),
        ("fine_columns", C.c_uint32),
        ("fine_rows", C.c_uint32),
    ]


class NodeWarp(C.Structure):
    _fields_ = [
        ("code", C.c_uint32),
        ("global_", C.c_float * 2),
        ("point", C.c_float * 2),
        ("confidence", C.c_float),
        ("measured", C.c_uint32),
    ]


WARP_DTYPE = np.dtype(
    [("code", "<u4"), ("global", "<f4", (2,)), ("point", "<f4", (2,)), ("confidence", "<f4"),
     ("measured", "<u4")]
)
assert WARP_DTYPE.itemsize == C.sizeof(NodeWarp) == 28


class Work(C.Structure):
    _fields_ = [("cost_zncc", C.c_uint64), ("direct_zncc", C.c_uint64), ("wide_zncc", C.c_uint64),
                ("target_zncc", C.c_uint64), ("wide_nodes", C.c_uint64)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint32), ("total_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("cost_nodes_per_block", C.c_uint32), ("cost_label_slab", C.c_uint32),
                ("sgm_paths_per_block", C.c_uint32), ("cost_tile_order", C.c_uint32)]


# ------------------------------------------------------------- loaders

_depth = None
_scene = None


def depth_lib() -> C.CDLL:
    """The HIP product library. Raises if it has not been built."""
    global _depth
    if _depth is None:
        if not DEPTH_LIB.exists():
            raise RuntimeError(
                f"{DEPTH_LIB} is missing: build it with `python -m chiaro_amd.build` "
                "(the depth path has no other backend)"
            )
        lib = C.CDLL(str(DEPTH_LIB))
        _declare_depth(lib)
        _depth = lib
    return _depth


def scene_lib() -> C.CDLL:
    global _scene
    if _scene is None:
        if not SCENE_LIB.exists():
            raise RuntimeError(f"{SCENE_LIB} is missing: build it with `python -m chiaro_amd.build`")
        lib = C.CDLL(str(SCENE_LIB))
        _declare_scene(lib)
        _scene = lib
    return _scene


P = C.POINTER
VP = C.c_void_p


def _fn(lib, name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


class AlignOptions(C.Structure):
    _fields_ = [
        ("grid_step", C.c_uint32),
        ("patch", C.c_uint32),
        ("coarse_radius", C.c_uint32),
        ("min_score", C.c_float),
        ("inlier_px", C.c_float),
        ("min_inlier_ratio", C.c_float),
        ("refine", C.c_int32),
    ]


class LevelReport(C.Structure):
    _fields_ = [
        ("scale", C.c_uint32),
        ("patches", C.c_uint32),
        ("inliers", C.c_uint32),
        ("median_residual_px", C.c_float),
    ]


CHD_MAX_ALIGN_LEVELS = 16
ALIGN_STATUS = ["reference", "rejected", "factory model only", "refined", "weak refinement",
                "no correlation support; factory model kept"]


class AlignmentReport(C.Structure):
    _fields_ = [
        ("initialised_from", C.c_uint32),
        ("coverage", C.c_float),
        ("correction_median_px", C.c_float * 2),
        ("level_count", C.c_uint32),
        ("levels", LevelReport * CHD_MAX_ALIGN_LEVELS),
        ("inliers", C.c_uint32),
        ("patches", C.c_uint32),
        ("residual_median_px", C.c_float),
        ("residual_p90_px", C.c_float),
        ("inlier_ratio", C.c_float),
        ("accepted", C.c_int32),
        ("status", C.c_uint32),
        ("gain", C.c_float),
    ]


class SynthSource(C.Structure):
    _fields_ = [
        ("warp", Warp),
        ("mosaic_width", C.c_uint32),
        ("mosaic_height", C.c_uint32),
        ("magnification", C.c_float),
        ("confidence", C.c_float),
        ("reference", C.c_int32),
        ("has_focus_distance", C.c_int32),
        ("focus_distance", C.c_double),
    ]


class Canvas(C.Structure):
    _fields_ = [
        ("crop_x", C.c_float),
        ("crop_y", C.c_float),
        ("crop_width", C.c_float),
        ("crop_height", C.c_float),
        ("scale", C.c_float),
        ("feather_px", C.c_float),
    ]


class ResolutionReport(C.Structure):
    _fields_ = [
        ("verified_fraction", C.c_float),
        ("supported_fraction", C.c_float),
        ("median_correction_px", C.c_float),
        ("mean_confidence", C.c_float),
    ]


# Every function declared in include/chiaro_depth.h, with its signature.
DEPTH_SIGNATURES = {
    "chd_last_error": (C.c_char_p,),
    "chd_abi_version": (C.c_int,),
    "chd_default_options": (None, P(Options)),
    "chd_options_valid": (C.c_int, P(Options)),
    "chd_grid_sizes_for": (C.c_int, P(Options), C.c_uint32, C.c_uint32, P(GridSizes)),
    "chd_camera_canonical": (C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), C.c_uint32,
                             C.c_uint32, P(Distortion), P(Camera)),
    "chd_context_create": (C.c_int, C.c_int, P(VP)),
    "chd_context_destroy": (None, VP),
    "chd_context_synchronize": (C.c_int, VP),
    "chd_refine_multiview_depth": (C.c_int, VP, P(View), C.c_uint32, C.c_uint32, P(Options),
                                   P(DepthMap), P(TargetResult)),
    "chd_session_create": (C.c_int, VP, P(View), C.c_uint32, C.c_uint32, P(Options), P(VP)),
    "chd_session_destroy": (None, VP),
    "chd_session_run": (C.c_int, VP, C.c_uint32),
    "chd_session_grid": (C.c_int, VP, P(GridSizes)),
    "chd_session_read_results": (C.c_int, VP, P(DepthMap), P(TargetResult)),
    "chd_session_read_cost_volume": (C.c_int, VP, VP, VP, VP, VP),
    "chd_session_read_aggregated": (C.c_int, VP, VP),
    "chd_session_read_coarse": (C.c_int, VP, VP, VP, VP),
    "chd_session_read_direct": (C.c_int, VP, C.c_uint32, VP, VP, VP),
    "chd_session_work": (C.c_int, VP, P(Work)),
    "chd_session_enable_timing": (C.c_int, VP, C.c_int),
    "chd_session_kernel_times": (C.c_int, VP, P(KernelTime), C.c_uint32, P(C.c_uint32)),
    "chd_session_reset_kernel_times": (C.c_int, VP),
    "chd_session_read_view_scores": (C.c_int, VP, VP),
    "chd_session_set_tuning": (C.c_int, VP, P(Tuning)),
    "chd_session_get_tuning": (C.c_int, VP, P(Tuning)),
    "chd_session_occupancy": (C.c_int, VP, P(C.c_uint32)),
    "chd_autotune": (C.c_int, VP, C.c_char_p, P(Tuning)),
    "chd_batch_create": (C.c_int, VP, P(VP), C.c_uint32, C.c_uint32, P(VP)),
    "chd_batch_run": (C.c_int, VP, C.c_uint32),
    "chd_batch_destroy": (None, VP),
    "chd_proposal_create": (C.c_int, VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, P(VP)),
    "chd_proposal_destroy": (None, VP),
    "chd_proposal_parameters": (C.c_int, VP, VP, VP, P(C.c_size_t)),
    "chd_proposal_optimizer_state": (C.c_int, VP, VP, VP, P(C.c_size_t), P(C.c_uint64)),
    "chd_proposal_set_dataset": (C.c_int, VP, VP, VP, C.c_uint32),
    "chd_proposal_train": (C.c_int, VP, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, VP),
    "chd_proposal_train_epochs": (C.c_int, VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_float, C.c_uint64, VP),
    "chd_proposal_predict": (C.c_int, VP, VP, C.c_uint32, VP),
    "chd_proposal_gradients": (C.c_int, VP, VP, VP, C.c_uint32, VP, P(C.c_float)),
    "chd_proposal_adam_step": (C.c_int, VP, VP, C.c_float),
    "chd_proposal_set_input_transform": (C.c_int, VP, C.c_float, C.c_float),
    "chd_proposal_set_label_count": (C.c_int, VP, C.c_uint32),
    "chd_proposal_get_input_transform": (C.c_int, VP, P(C.c_float), P(C.c_float)),
    "chd_session_set_proposal": (C.c_int, VP, VP, C.c_float, C.c_uint32, C.c_uint32,
                                 C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_uint32)),
    "chd_session_attach_proposal": (C.c_int, VP, VP, C.c_float, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_uint32)),
    "chd_session_clear_proposal": (C.c_int, VP),
    "chd_session_read_band": (C.c_int, VP, VP, VP),
    "chd_semi_global_costs": (C.c_int, VP, VP, VP, C.c_uint32, C.c_uint32, C.c_uint32, VP),
    "chd_select_depths": (C.c_int, VP, VP, C.c_uint32, VP, VP, VP, C.c_uint32, P(Options), VP, VP,
                          VP),
    "chd_reject_isolated_direct_depths": (C.c_int, VP, VP, VP, C.c_uint32, C.c_uint32,
                                          C.c_double),
    "chd_reject_small_direct_components": (C.c_int, VP, VP, VP, C.c_uint32, C.c_uint32,
                                           C.c_double, C.c_uint32),
    "chd_reject_small_direct_components_noalloc": (C.c_int, VP, VP, VP, C.c_uint32, C.c_uint32,
                                                   C.c_double, C.c_uint32, VP),
    "chd_fit_local_depth_planes": (C.c_int, VP, VP, VP, C.c_uint32, C.c_uint32, C.c_double),
    "chd_enforce_warp_consensus": (C.c_int, VP, VP, C.c_uint32, C.c_uint32),
    "chd_suppress_warp_boundaries": (C.c_int, VP, VP, VP, C.c_uint32, C.c_uint32),
    "chd_luminance_half": (C.c_int, VP, VP, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                           C.c_int, VP),
    "chd_luminance_half_device": (C.c_int, VP, VP, C.c_uint32, C.c_uint32, C.c_float,
                                  C.c_float, C.c_int, VP),
    "chd_default_align_options": (None, P(AlignOptions)),
    "chd_align_module": (C.c_int, VP, P(View), P(View), P(AlignOptions), P(Warp),
                         P(AlignmentReport)),
    "chd_match_patches": (C.c_int, VP, VP, VP, C.c_uint32, C.c_uint32, VP, VP, C.c_uint32,
                          C.c_uint32, C.c_uint32, VP, VP, VP),
    "chd_refine_resolution_warp": (C.c_int, VP, VP, C.c_uint32, C.c_uint32, VP, C.c_uint32,
                                   C.c_uint32, P(Warp), C.c_uint32, C.c_uint32, P(Warp),
                                   P(ResolutionReport)),
    "chd_canvas_size": (C.c_int, P(Canvas), P(C.c_uint32), P(C.c_uint32)),
    "chd_synthesis_geometry": (C.c_int, VP, P(DepthMap), P(SynthSource), C.c_uint32, P(Canvas),
                               VP, VP, VP, VP, VP),
    "chd_frames_create": (C.c_int, VP, VP, C.c_uint32, C.c_uint32, VP, C.c_uint32, C.c_float,
                          C.c_float, P(VP)),
    "chd_frames_render": (C.c_int, VP, VP),
    "chd_frames_plane": (C.c_int, VP, C.c_uint32, P(VP), P(C.c_uint32), P(C.c_uint32)),
    "chd_frames_read": (C.c_int, VP, C.c_uint32, VP, VP),
    "chd_frames_destroy": (None, VP),
    "chd_session_create_device": (C.c_int, VP, P(View), C.c_uint32, C.c_uint32, P(Options),
                                  P(VP)),
    "chd_session_set_planes": (C.c_int, VP, P(VP), C.c_uint32, C.c_int),
    "chd_session_network_input": (C.c_int, VP, VP, VP),
    "chd_device_expf": (C.c_int, VP, VP, VP, C.c_size_t),
    "chd_device_exp": (C.c_int, VP, VP, VP, C.c_size_t),
}


def _declare_depth(lib):
    for name, sig in DEPTH_SIGNATURES.items():
        _fn(lib, name, sig[0], *sig[1:])


class Module(C.Structure):
    _fields_ = [
        ("focal_px", C.c_double),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("center", C.c_double * 3),
        ("rotation_wc", C.c_double * 9),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("plane_count", C.c_uint32),
        ("inverse_depth_min", C.c_double),
        ("inverse_depth_max", C.c_double),
        ("texture_cell_px", C.c_uint32),
    ]


class Bayer(C.Structure):
    _fields_ = [
        ("black_level", C.c_float),
        ("white_level", C.c_float),
        ("exposure", C.c_float),
        ("electrons_per_dn", C.c_float),
        ("read_noise_dn", C.c_float),
    ]


def _declare_scene(lib):
    _fn(lib, "chs_render", C.c_int, P(SceneDesc), P(Module), C.c_uint32, C.c_uint32, C.c_float,
        P(Bayer), P(P(C.c_float)), P(P(C.c_uint16)), C.c_uint32)
    _fn(lib, "chs_truth_inverse_distance", C.c_int, P(SceneDesc), P(Module), C.c_uint32, VP)
    _fn(lib, "chs_luminance_half_f32", C.c_int, VP, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
        VP)
    _fn(lib, "chs_luminance_half_u16", C.c_int, VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
        C.c_float, VP)
    _fn(lib, "chs_camera", C.c_int, P(Module), P(Camera))
    _fn(lib, "chs_global_warp", C.c_int, P(Camera), P(Camera), C.c_uint32, C.c_uint32, C.c_uint32,
        VP, VP)


def ptr(a: np.ndarray) -> int:
    """Address of a C-contiguous numpy array (kept alive by the caller)."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P(C.c_float))


def last_error() -> str:
    return depth_lib().chd_last_error().decode(errors="replace")


def check(code: int, what: str) -> int:
    if code < 0:
        raise RuntimeError(f"{what} failed ({code}): {last_error()}")
    return code
