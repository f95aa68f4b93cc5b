"""Seeded synthetic captures for the depth path (the reference has no scene
generator; see chiaro_amd/csrc/scene.c for the model).

A Capture holds, per module, exactly what refine_multiview_depth reads
(crates/chiaro-fusion/src/depth.rs:312-317): the half-resolution log
luminance plane, raster size, calibrated camera, the global warp of
align_module(refine = false) and the accepted flag.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .depth import DepthOptions


@dataclass
class ModuleInput:
    name: str
    luminance: np.ndarray          # (plane_height, plane_width) float32
    width: int                     # raster size
    height: int
    camera: _abi.Camera | None
    warp_points: np.ndarray        # (rows, columns, 2) float32, NaN = undefined
    warp_step: int
    accepted: bool = True


@dataclass
class Capture:
    modules: list[ModuleInput]
    reference_index: int
    options: DepthOptions
    description: dict = field(default_factory=dict)
    truth_step: int = 8
    truth_inverse_distance: np.ndarray | None = None  # 1/mm on the coarse grid
    raws: list | None = None  # sensor codes (u16 Bayer) when the capture has them


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _module(focal_px, width, height, center_mm, rotation=None) -> _abi.Module:
    m = _abi.Module()
    m.focal_px = focal_px
    m.width = width
    m.height = height
    for i in range(3):
        m.center[i] = center_mm[i]
    r = np.eye(3) if rotation is None else np.asarray(rotation, dtype=np.float64)
    for i in range(9):
        m.rotation_wc[i] = r.reshape(-1)[i]
    return m


def camera_for(module: _abi.Module) -> _abi.Camera:
    cam = _abi.Camera()
    if _abi.scene_lib().chs_camera(C.byref(module), C.byref(cam)) != 0:
        raise RuntimeError("singular intrinsics")
    return cam


def identity_warp(width: int, height: int, step: int = 32) -> np.ndarray:
    """Warp::from_fn(width, height, step, Some) (align.rs:299-300)."""
    columns = -(-width // step) + 1
    rows = -(-height // step) + 1
    ys, xs = np.meshgrid(np.arange(rows) * step, np.arange(columns) * step, indexing="ij")
    return np.stack([xs, ys], axis=-1).astype(np.float32)


def global_warp(reference: _abi.Camera, target: _abi.Camera, width: int, height: int,
                step: int = 32) -> np.ndarray:
    columns = -(-width // step) + 1
    rows = -(-height // step) + 1
    points = np.empty((rows, columns, 2), dtype=np.float32)
    confidence = np.empty((rows, columns), dtype=np.float32)
    _abi.scene_lib().chs_global_warp(C.byref(reference), C.byref(target), width, height, step,
                                     _abi.ptr(points), _abi.ptr(confidence))
    return points


def _render(desc: _abi.SceneDesc, modules: list[_abi.Module], reference: int, noise: float,
            bayer: _abi.Bayer | None):
    n = len(modules)
    arr = (_abi.Module * n)(*modules)
    if bayer is None:
        images = [np.empty((m.height, m.width), dtype=np.float32) for m in modules]
        ptrs = (C.POINTER(C.c_float) * n)(*[_abi.fptr(a) for a in images])
        rc = _abi.scene_lib().chs_render(C.byref(desc), arr, n, reference, noise, None, ptrs, None,
                                         _threads())
    else:
        images = [np.empty((m.height, m.width), dtype=np.uint16) for m in modules]
        ptrs = (C.POINTER(C.c_uint16) * n)(*[a.ctypes.data_as(C.POINTER(C.c_uint16)) for a in images])
        rc = _abi.scene_lib().chs_render(C.byref(desc), arr, n, reference, 0.0, C.byref(bayer),
                                         None, ptrs, _threads())
    if rc != 0:
        raise RuntimeError("scene rendering failed")
    return images


def luminance_half_f32(image: np.ndarray, black: float = 0.0, white: float = 1.0) -> np.ndarray:
    h, w = image.shape
    out = np.empty((h // 2, w // 2), dtype=np.float32)
    _abi.scene_lib().chs_luminance_half_f32(_abi.ptr(np.ascontiguousarray(image)), w, h, black,
                                            white, _abi.ptr(out))
    return out


def luminance_half_u16(raw: np.ndarray, shift: int, black: float, white: float) -> np.ndarray:
    h, w = raw.shape
    out = np.empty((h // 2, w // 2), dtype=np.float32)
    _abi.scene_lib().chs_luminance_half_u16(_abi.ptr(np.ascontiguousarray(raw)), w, h, shift,
                                            black, white, _abi.ptr(out))
    return out


def _assemble(modules, planes, reference, options, description, desc, names=None) -> Capture:
    cams = [camera_for(m) for m in modules]
    ref = modules[reference]
    inputs = []
    for i, (m, plane, cam) in enumerate(zip(modules, planes, cams)):
        if i == reference:
            warp = identity_warp(ref.width, ref.height)
        else:
            warp = global_warp(cams[reference], cam, ref.width, ref.height)
        inputs.append(ModuleInput(name=(names[i] if names else f"M{i}"), luminance=plane,
                                  width=m.width, height=m.height, camera=cam, warp_points=warp,
                                  warp_step=32))
    step = max(options.grid_step, 4)
    columns = -(-ref.width // step) + 1
    rows = -(-ref.height // step) + 1
    truth = np.empty((rows, columns), dtype=np.float32)
    _abi.scene_lib().chs_truth_inverse_distance(C.byref(desc), C.byref(ref), step, _abi.ptr(truth))
    return Capture(modules=inputs, reference_index=reference, options=options,
                   description=description, truth_step=step, truth_inverse_distance=truth)


def config0(seed: int = 0, width: int = 1040, height: int = 780, focal_px: float = 1600.0,
            planes: int = 64, plane_count: int = 8) -> Capture:
    """Three views: reference plus neighbours at 20 mm / 0 deg and 35 mm / 60
    deg; pinhole, principal point at the image centre; single-channel f32
    images in [0, 1] with Gaussian noise 0.01; Voronoi scene of fronto-parallel
    planes with inverse depth U[0.05, 1.25] 1/m; 64 hypotheses, near 800 mm,
    far 20000 mm. The image is taken as the raster with black 0 and white 1."""
    modules = [
        _module(focal_px, width, height, (0.0, 0.0, 0.0)),
        _module(focal_px, width, height, (20.0, 0.0, 0.0)),
        _module(focal_px, width, height, (35.0 * math.cos(math.radians(60)),
                                          35.0 * math.sin(math.radians(60)), 0.0)),
    ]
    desc = _abi.SceneDesc(seed=seed, plane_count=plane_count, inverse_depth_min=0.05e-3,
                          inverse_depth_max=1.25e-3, texture_cell_px=32)
    images = _render(desc, modules, 0, 0.01, None)
    planes_ = [luminance_half_f32(img) for img in images]
    options = DepthOptions(near_depth=800.0, far_depth=20000.0, planes=planes)
    return _assemble(modules, planes_, 0, options,
                     {"workload": "config0", "seed": seed, "width": width, "height": height,
                      "views": 3, "planes": planes}, desc)


BAYER_12BIT = dict(black_level=64.0, white_level=4095.0, exposure=0.85, electrons_per_dn=0.5,
                   read_noise_dn=2.0)


def rig_config1(seed: int = 0, width: int = 4160, height: int = 3120, focal_px: float = 3300.0,
                planes: int = 128, plane_count: int = 16, ring_mm: float = 30.0) -> "Rig":
    """Five modules: the reference at the centre and four on a 30 mm ring at
    0/90/180/270 deg; f = 3300 px; 12-bit RGGB Bayer codes (black 64, Poisson
    shot noise at 0.5 e-/DN, read noise 2 DN) converted by luminance_half with
    codes shifted left by 4 (black 64*16, white 4095*16); 16 planes; 128
    hypotheses, near 800 mm, far 20000 mm."""
    modules = [_module(focal_px, width, height, (0.0, 0.0, 0.0))]
    for k in range(4):
        a = math.radians(90.0 * k)
        modules.append(_module(focal_px, width, height, (ring_mm * math.cos(a), ring_mm * math.sin(a), 0.0)))
    desc = _abi.SceneDesc(seed=seed, plane_count=plane_count, inverse_depth_min=0.05e-3,
                          inverse_depth_max=1.25e-3, texture_cell_px=32)
    options = DepthOptions(near_depth=800.0, far_depth=20000.0, planes=planes)
    return Rig(modules, 0, desc, _abi.Bayer(**BAYER_12BIT), options,
               {"workload": "config1", "seed": seed, "width": width, "height": height,
                "views": 5, "planes": planes})


def config1(seed: int = 0, **kw) -> Capture:
    """Config 1 rendered on the host (see rig_config1)."""
    return render_host(rig_config1(seed, **kw))


def _small_rotation(rng: np.random.Generator, degrees: float) -> np.ndarray:
    """Rotation about a uniformly random axis by an angle U[-degrees, degrees]."""
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    a = math.radians(rng.uniform(-degrees, degrees))
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(a) * k + (1.0 - math.cos(a)) * (k @ k)


def rig_config2(seed: int = 0, scale: float = 1.0, planes: int = 128,
                plane_count: int = 16) -> "Rig":
    """One capture of the batched workload: ten modules of uneven size, the
    reference (4160x3120, f = 3300 px) at the origin plus four more of that
    kind, three 2080x1560 at f = 4125 px and two 1040x780 at f = 4425 px;
    each non-reference module at a baseline U[10, 60] mm in a random
    direction of the sensor plane with a random rotation of up to 1 degree.
    Rig parameters come from numpy's PCG64 seeded with (seed, 2). 12-bit
    Bayer codes as in config1, 16 planes, 128 hypotheses; `scale` shrinks
    every module's raster and focal length together (for small test
    captures)."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, 2]))
    specs = [(3300.0, 4160, 3120)] * 5 + [(4125.0, 2080, 1560)] * 3 + [(4425.0, 1040, 780)] * 2

    def size(n):
        return max(16, int(round(n * scale)) // 8 * 8)

    modules = []
    for k, (f, w, h) in enumerate(specs):
        if k == 0:
            modules.append(_module(f * scale, size(w), size(h), (0.0, 0.0, 0.0)))
            continue
        baseline = rng.uniform(10.0, 60.0)
        angle = rng.uniform(0.0, 2.0 * math.pi)
        rotation = _small_rotation(rng, 1.0)
        modules.append(_module(f * scale, size(w), size(h),
                               (baseline * math.cos(angle), baseline * math.sin(angle), 0.0),
                               rotation))
    desc = _abi.SceneDesc(seed=seed, plane_count=plane_count, inverse_depth_min=0.05e-3,
                          inverse_depth_max=1.25e-3, texture_cell_px=32)
    options = DepthOptions(near_depth=800.0, far_depth=20000.0, planes=planes)
    return Rig(modules, 0, desc, _abi.Bayer(**BAYER_12BIT), options,
               {"workload": "config2", "seed": seed, "scale": scale, "views": len(specs),
                "planes": planes})


def config2(seed: int = 0, **kw) -> Capture:
    """Config 2 rendered on the host (see rig_config2)."""
    return render_host(rig_config2(seed, **kw))


def config3(seed: int = 0, width: int = 1040, height: int = 780, focal_px: float = 1600.0) -> Capture:
    """The learned-proposal workload: config0's three-view scene at 1040x780
    swept with 128 hypotheses on a grid of step 4, so that the coarse grid
    (261 x 196 nodes) matches the network's 260 x 195 pixels (the reference
    plane at half size)."""
    cap = config0(seed=seed, width=width, height=height, focal_px=focal_px, planes=128)
    cap.options.grid_step = 4
    cap.description["workload"] = "config3"
    return cap


def rig_config4(seed: int = 0, scale: float = 1.0, planes: int = 256,
                plane_count: int = 24, wide_only: bool = False) -> "Rig":
    """One capture of the full-capture workload: sixteen 4160x3120 modules
    viewing nested fields of one 24-plane scene - five wide (f = 3300 px: the
    reference at the origin and four on a 30 mm ring at 0/90/180/270 deg),
    five mid (f = 8250 px, on a 20 mm ring at 36 + 72k deg) and six tele
    (f = 17700 px, on a 45 mm ring at 30 + 60k deg), all with the reference's
    orientation. Inverse depths U[0.02, 1.25] 1/m, 12-bit Bayer codes as in
    config1, 256 hypotheses, near 800 mm, far 50000 mm. `scale` shrinks every
    raster, focal length and the texture's cell size together (small test
    captures; at 0.5 with a 4 px grid the coarse grid matches the full
    capture's at 8 px). `wide_only` keeps the five wide modules (the same
    scene and reference, a third of the rendering)."""
    def size(n):
        return max(16, int(round(n * scale)) // 8 * 8)

    w, h = size(4160), size(3120)
    modules = [_module(3300.0 * scale, w, h, (0.0, 0.0, 0.0))]
    rings = [(3300.0, 30.0, 0.0, 4), (8250.0, 20.0, 36.0, 5), (17700.0, 45.0, 30.0, 6)]
    if wide_only:
        rings = rings[:1]
    for f, radius, first, count in rings:
        step = 90.0 if count == 4 else 360.0 / count
        for k in range(count):
            a = math.radians(first + step * k)
            modules.append(_module(f * scale, w, h, (radius * math.cos(a), radius * math.sin(a), 0.0)))
    desc = _abi.SceneDesc(seed=seed, plane_count=plane_count, inverse_depth_min=0.02e-3,
                          inverse_depth_max=1.25e-3,
                          texture_cell_px=max(4, int(round(32 * scale))))
    options = DepthOptions(near_depth=800.0, far_depth=50000.0, planes=planes)
    names = ([f"W{k}" for k in range(5)] + [f"M{k}" for k in range(5)]
             + [f"T{k}" for k in range(6)])[:len(modules)]
    return Rig(modules, 0, desc, _abi.Bayer(**BAYER_12BIT), options,
               {"workload": "config4", "seed": seed, "scale": scale, "views": len(modules),
                "planes": planes}, names)


def config4(seed: int = 0, **kw) -> Capture:
    """Config 4 rendered on the host (see rig_config4)."""
    return render_host(rig_config4(seed, **kw))


# ------------------------------------------------------------ rigs

@dataclass
class Rig:
    """What a Bayer capture is made of: modules, reference, the scene
    description (its seed is the capture's), the sensor model and the
    depth options."""

    modules: list
    reference: int
    scene: _abi.SceneDesc
    bayer: _abi.Bayer
    options: DepthOptions
    description: dict
    names: list | None = None

    def with_seed(self, seed: int) -> "Rig":
        d = _abi.SceneDesc()
        C.pointer(d)[0] = self.scene
        d.seed = seed
        desc = dict(self.description, seed=seed)
        return Rig(self.modules, self.reference, d, self.bayer, DepthOptions(**vars(self.options)),
                   desc, self.names)


def render_host(rig: Rig) -> Capture:
    """The capture rendered by the host generator (scene.c): Bayer codes,
    luminance_half of the codes shifted left by 4 (black 64*16, white
    4095*16)."""
    raws = _render(rig.scene, rig.modules, rig.reference, 0.0, rig.bayer)
    planes_ = [luminance_half_u16(r, 4, 64.0 * 16, 4095.0 * 16) for r in raws]
    cap = _assemble(rig.modules, planes_, rig.reference, DepthOptions(**vars(rig.options)),
                    dict(rig.description), rig.scene, rig.names)
    cap.raws = raws
    return cap


class Frames:
    """A rig's frames rendered on the GPU (chd_frames): the same codes as the
    host generator, shifted left by 4 and converted by luminance_half on the
    device, planes kept in HBM for sessions (chiaro_amd.depth.DevicePlane)."""

    SHIFT, BLACK, WHITE = 4, 64.0 * 16, 4095.0 * 16

    def __init__(self, context, rig: Rig):
        from .depth import DevicePlane  # noqa: F401  (the planes' type)
        self.context = context
        self.lib = context.lib
        self.rig = rig
        n = len(rig.modules)
        arr = (_abi.Module * n)(*rig.modules)
        handle = C.c_void_p()
        _abi.check(self.lib.chd_frames_create(context.handle, arr, n, rig.reference,
                                              C.byref(rig.bayer), self.SHIFT, self.BLACK,
                                              self.WHITE, C.byref(handle)), "frames create")
        self.handle = handle
        context._frames.add(self)
        self.seed = None
        self._cameras = [camera_for(m) for m in rig.modules]
        self._warps = None

    def close(self):
        if self.handle:
            self.lib.chd_frames_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, seed: int):
        """Draws the rig's scene with this seed (asynchronous on the
        context's stream)."""
        d = _abi.SceneDesc()
        C.pointer(d)[0] = self.rig.scene
        d.seed = seed
        _abi.check(self.lib.chd_frames_render(self.handle, C.byref(d)), "frames render")
        self.seed = seed

    def planes(self):
        from .depth import DevicePlane
        out = []
        for i in range(len(self.rig.modules)):
            p, w, h = C.c_void_p(), C.c_uint32(), C.c_uint32()
            _abi.check(self.lib.chd_frames_plane(self.handle, i, C.byref(p), C.byref(w),
                                                 C.byref(h)), "frames plane")
            out.append(DevicePlane(p.value, (h.value, w.value), self))
        return out

    def read(self, index: int):
        """(mosaic u16 of shifted codes, plane f32) of one module."""
        m = self.rig.modules[index]
        mosaic = np.empty((m.height, m.width), np.uint16)
        plane = np.empty((m.height // 2, m.width // 2), np.float32)
        _abi.check(self.lib.chd_frames_read(self.handle, index, mosaic.ctypes.data,
                                            plane.ctypes.data), "frames read")
        return mosaic, plane

    def capture(self) -> Capture:
        """The last render as a Capture whose planes stay on the device."""
        rig = self.rig.with_seed(self.seed)
        ref = rig.modules[rig.reference]
        if self._warps is None:
            self._warps = [identity_warp(ref.width, ref.height) if i == rig.reference
                           else global_warp(self._cameras[rig.reference], cam, ref.width, ref.height)
                           for i, cam in enumerate(self._cameras)]
        inputs = [ModuleInput(name=(rig.names[i] if rig.names else f"M{i}"), luminance=plane,
                              width=m.width, height=m.height, camera=self._cameras[i],
                              warp_points=self._warps[i], warp_step=32)
                  for i, (m, plane) in enumerate(zip(rig.modules, self.planes()))]
        return Capture(modules=inputs, reference_index=rig.reference, options=rig.options,
                       description=rig.description)
