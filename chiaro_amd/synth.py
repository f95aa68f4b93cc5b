"""The synthesis consumer's depth and warp stage on the GPU, in the
reference's shape (crates/chiaro-fusion/src/synth.rs): CropWindow,
canvas_scale and the per-pixel, per-source geometry that synthesize derives
from the depth map and the warps (chd_synthesis_geometry)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from .align import Warp
from .depth import Context, DenseDepthMap

FEATHER_PX = 120.0  # SynthOptions::default().feather_px (synth.rs:92)


@dataclass
class CropWindow:
    """CropWindow (synth.rs:106-134), reference raster pixels."""
    x: float
    y: float
    width: float
    height: float

    @classmethod
    def full(cls, width: int, height: int) -> "CropWindow":
        return cls(0.0, 0.0, float(width), float(height))

    @classmethod
    def centred(cls, width: int, height: int, fraction: float) -> "CropWindow":
        f = np.float32(min(max(fraction, 0.05), 1.0))
        w, h = np.float32(width) * f, np.float32(height) * f
        return cls(float((np.float32(width) - w) / np.float32(2)),
                   float((np.float32(height) - h) / np.float32(2)), float(w), float(h))


def canvas_scale(crop: CropWindow, reference_width: int, mode: str = "native",
                 finest_magnification: float = 1.0, max_megapixels: float = 0.0,
                 scale: float = 1.0) -> float:
    """canvas_scale (synth.rs:568-589) for CanvasMode native, maximum
    (with max_megapixels) or scale, in f32."""
    F = np.float32
    if mode == "native":
        return float(F(reference_width) / F(crop.width))
    if mode == "maximum":
        wanted = np.fmax(F(finest_magnification), F(reference_width) / F(crop.width))
        pixels = F(F(F(F(crop.width) * F(crop.height)) * wanted) * wanted)
        cap = np.fmax(F(F(max_megapixels) * F(1.0e6)), F(1.0))
        return float(F(wanted * np.sqrt(F(cap / pixels)))) if pixels > cap else float(wanted)
    if mode == "scale":
        return float(np.fmax(F(scale), F(0.05)))
    raise ValueError(f"unknown canvas mode {mode!r}")


@dataclass
class SynthSource:
    """The fields of SynthSource (synth.rs:346-370) the stage reads."""
    warp: Warp
    mosaic_width: int
    mosaic_height: int
    magnification: float = 1.0
    confidence: float = 1.0
    reference: bool = False
    focus_distance: float | None = None


@dataclass
class SynthesisGeometry:
    """Per source and canvas pixel ([source][row][column]): the warp point
    (NaN where the source does not contribute), border feather, local warp
    confidence, base weight and focus weight (0 where it does not)."""
    points: np.ndarray
    feather: np.ndarray
    local_confidence: np.ndarray
    base_weight: np.ndarray
    focus_weight: np.ndarray


def _canvas(crop: CropWindow, scale: float, feather_px: float) -> _abi.Canvas:
    return _abi.Canvas(crop.x, crop.y, crop.width, crop.height, scale, feather_px)


def canvas_size(context: Context, crop: CropWindow, scale: float) -> tuple[int, int]:
    w, h = C.c_uint32(), C.c_uint32()
    _abi.check(context.lib.chd_canvas_size(C.byref(_canvas(crop, scale, 1.0)), C.byref(w),
                                           C.byref(h)), "canvas size")
    return w.value, h.value


def synthesis_geometry(context: Context, depth_map: DenseDepthMap | None, sources,
                       crop: CropWindow, scale: float,
                       feather_px: float = FEATHER_PX) -> SynthesisGeometry:
    """synthesize's depth and warp stage (synth.rs:633-725) over the canvas
    of `crop` at `scale`."""
    keep = []
    arr = (_abi.SynthSource * len(sources))()
    for i, s in enumerate(sources):
        w = Warp(s.warp.step, np.ascontiguousarray(s.warp.points, np.float32),
                 np.ascontiguousarray(s.warp.confidence, np.float32))
        keep.append(w)
        d = arr[i]
        d.warp = w.ffi()
        d.mosaic_width, d.mosaic_height = s.mosaic_width, s.mosaic_height
        d.magnification, d.confidence = s.magnification, s.confidence
        d.reference = int(s.reference)
        d.has_focus_distance = int(s.focus_distance is not None)
        d.focus_distance = s.focus_distance if s.focus_distance is not None else 0.0
    m = None
    if depth_map is not None:
        depth = np.ascontiguousarray(depth_map.depth, np.float64)
        conf = np.ascontiguousarray(depth_map.confidence, np.float32)
        prov = np.ascontiguousarray(depth_map.provenance, np.uint8)
        keep += [depth, conf, prov]
        m = _abi.DepthMap()
        m.columns, m.rows, m.step = depth_map.columns, depth_map.rows, depth_map.step
        m.near_depth, m.far_depth = depth_map.near_depth, depth_map.far_depth
        m.depth = depth.ctypes.data_as(C.POINTER(C.c_double))
        m.confidence = _abi.fptr(conf)
        m.provenance = prov.ctypes.data_as(C.POINTER(C.c_uint8))
    width, height = canvas_size(context, crop, scale)
    n = len(sources)
    points = np.empty((n, height, width, 2), np.float32)
    outs = [np.empty((n, height, width), np.float32) for _ in range(4)]
    _abi.check(context.lib.chd_synthesis_geometry(
        context.handle, C.byref(m) if m is not None else None, arr, n,
        C.byref(_canvas(crop, scale, feather_px)), points.ctypes.data,
        *[o.ctypes.data for o in outs]), "synthesis geometry")
    return SynthesisGeometry(points, *outs)
