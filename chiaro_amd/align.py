"""Alignment on the GPU in the reference's shape: align_module
(crates/chiaro-fusion/src/align.rs:270-519) and refine_resolution_warp
(crates/chiaro-fusion/src/resolution.rs:232-432), through
chd_align_module and chd_refine_resolution_warp."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .depth import Context, DevicePlane, _ViewArrays


@dataclass
class Warp:
    """Warp (align.rs:35-45): points (rows, columns, 2) f32 with NaN where
    undefined, one confidence per node."""
    step: int
    points: np.ndarray
    confidence: np.ndarray

    @property
    def columns(self) -> int:
        return self.points.shape[1]

    @property
    def rows(self) -> int:
        return self.points.shape[0]

    @classmethod
    def empty(cls, width: int, height: int, step: int) -> "Warp":
        columns, rows = -(-width // step) + 1, -(-height // step) + 1
        return cls(step, np.empty((rows, columns, 2), np.float32),
                   np.empty((rows, columns), np.float32))

    def ffi(self) -> _abi.Warp:
        for a in (self.points, self.confidence):
            if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
                raise ValueError("warp arrays must be C-contiguous float32")
        w = _abi.Warp()
        w.step, w.columns, w.rows = self.step, self.columns, self.rows
        w.points = _abi.fptr(self.points)
        w.confidence = _abi.fptr(self.confidence)
        return w


@dataclass
class AlignOptions:
    """AlignOptions (align.rs:210-244) without the depth sub-options."""
    grid_step: int = 32
    patch: int = 32
    coarse_radius: int = 12
    min_score: float = 0.5
    inlier_px: float = 3.0
    min_inlier_ratio: float = 0.45
    refine: bool = True

    def ffi(self) -> _abi.AlignOptions:
        o = _abi.AlignOptions()
        o.grid_step, o.patch, o.coarse_radius = self.grid_step, self.patch, self.coarse_radius
        o.min_score, o.inlier_px = self.min_score, self.inlier_px
        o.min_inlier_ratio, o.refine = self.min_inlier_ratio, int(self.refine)
        return o


@dataclass
class ModuleAlignment:
    """ModuleAlignment (align.rs:147-157): the warp, the photometric gain and
    the report's numeric fields with its status string."""
    warp: Warp
    gain: float
    report: dict = field(default_factory=dict)

    @property
    def accepted(self) -> bool:
        return bool(self.report.get("accepted"))


def _status_text(r: _abi.AlignmentReport) -> str:
    """AlignmentReport::status as align.rs:488-501 words it."""
    if r.status == _abi.ALIGN_STATUS.index("rejected"):
        return f"rejected: {r.inlier_ratio * 100.0:.0f}% correspondence consensus"
    return _abi.ALIGN_STATUS[r.status]


def align_module(context: Context, modules, reference_index: int, target_index: int,
                 options: AlignOptions | None = None) -> ModuleAlignment:
    """align_module for modules[target_index] against
    modules[reference_index] (chiaro_amd.scene.ModuleInput-like objects with
    luminance, width, height and camera). Both need cameras; the reference
    aligned to itself gives its identity warp."""
    options = options or AlignOptions()
    if options.grid_step <= 0:
        raise ValueError("grid_step must be positive")
    ref = modules[reference_index]
    views = _ViewArrays([ref, modules[target_index]])
    warp = Warp.empty(ref.width, ref.height, options.grid_step)
    report = _abi.AlignmentReport()
    target_view = views.views[0] if target_index == reference_index else views.views[1]
    _abi.check(context.lib.chd_align_module(context.handle, C.byref(views.views[0]),
                                            C.byref(target_view), C.byref(options.ffi()),
                                            C.byref(warp.ffi()), C.byref(report)),
               "align_module")
    out = {name: getattr(report, name) for name, _ in _abi.AlignmentReport._fields_
           if name not in ("levels", "correction_median_px")}
    out["correction_median_px"] = list(report.correction_median_px)
    out["levels"] = [{f: getattr(report.levels[i], f) for f, _ in _abi.LevelReport._fields_}
                     for i in range(report.level_count)]
    out["initialised_from"] = "reference" if report.initialised_from == 0 else "calibration"
    out["status"] = _status_text(report)
    out["accepted"] = bool(report.accepted)
    return ModuleAlignment(warp, report.gain, out)


@dataclass
class ResolutionWarp:
    """ResolutionWarp (resolution.rs:101-105)."""
    warp: Warp
    report: dict


def refine_resolution_warp(context: Context, reference: np.ndarray, target: np.ndarray,
                           base: Warp, width: int, height: int) -> ResolutionWarp:
    """refine_resolution_warp: local 12x12 NCC on a 32 px grid of the
    width x height reference raster between the reference plane and the
    target plane rendered through `base`. Planes are numpy arrays or
    DevicePlanes (already in HBM)."""
    def plane(p):
        if isinstance(p, DevicePlane):
            return p.address, p.shape
        a = np.ascontiguousarray(p, np.float32)
        keep.append(a)
        return a.ctypes.data, a.shape

    keep = []
    ref_ptr, (rh, rw) = plane(reference)
    tgt_ptr, (th, tw) = plane(target)
    base = Warp(base.step, np.ascontiguousarray(base.points, np.float32),
                np.ascontiguousarray(base.confidence, np.float32))
    out = Warp.empty(width, height, 32)
    report = _abi.ResolutionReport()
    _abi.check(context.lib.chd_refine_resolution_warp(
        context.handle, ref_ptr, rw, rh, tgt_ptr, tw, th,
        C.byref(base.ffi()), width, height, C.byref(out.ffi()), C.byref(report)),
        "refine_resolution_warp")
    return ResolutionWarp(out, {name: getattr(report, name)
                                for name, _ in _abi.ResolutionReport._fields_})
