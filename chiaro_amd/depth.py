"""Python face of the depth hot path: the reference's refine_multiview_depth
(crates/chiaro-fusion/src/depth.rs:312-580) over the C ABI of
libchiaro_depth.so, plus a device-resident Session for repeated runs.

The arithmetic runs in the HIP kernels; this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, fields

import numpy as np

from . import _abi


@dataclass
class DepthOptions:
    """DepthOptions and its Default (depth.rs:40-91)."""

    enabled: bool = True
    grid_step: int = 8
    near_depth: float = 500.0
    far_depth: float = 10_000_000.0
    planes: int = 96
    patch_radius: int = 4
    minimum_support: int = 2
    best_view_count: int = 3
    minimum_score: float = 0.45
    minimum_margin: float = 0.01
    minimum_improvement: float = 0.01
    minimum_neighbour_support: int = 2
    completion_iterations: int = 0

    def to_c(self) -> _abi.Options:
        o = _abi.Options()
        for f in fields(self):
            setattr(o, f.name, int(getattr(self, f.name)) if f.name == "enabled" else getattr(self, f.name))
        return o

    def grid(self, width: int, height: int) -> _abi.GridSizes:
        g = _abi.GridSizes()
        _abi.check(_abi.depth_lib().chd_grid_sizes_for(C.byref(self.to_c()), width, height,
                                                        C.byref(g)), "grid sizes")
        return g


@dataclass
class DenseDepthMap:
    """DenseDepthMap (depth.rs:133-159): NaN depth is None."""

    columns: int
    rows: int
    step: int
    near_depth: float
    far_depth: float
    depth: np.ndarray        # (rows, columns) float64
    confidence: np.ndarray   # (rows, columns) float32
    provenance: np.ndarray   # (rows, columns) uint8

    def sample_nearest(self, x: float, y: float):
        if x < 0 or y < 0 or self.step == 0:
            return None
        column = int(np.round(np.float32(x) / np.float32(self.step)))
        row = int(np.round(np.float32(y) / np.float32(self.step)))
        if column >= self.columns or row >= self.rows:
            return None
        return (self.depth[row, column], self.confidence[row, column], self.provenance[row, column])


@dataclass
class TargetResult:
    points: np.ndarray       # (rows, columns, 2) float32, NaN = undefined
    confidence: np.ndarray   # (rows, columns) float32
    decisions: np.ndarray    # (rows, columns) uint8 NodeWarp codes
    report: dict


class Context:
    """A HIP stream on one device (one per thread of use)."""

    def __init__(self, device: int = 0):
        self.lib = _abi.depth_lib()
        handle = C.c_void_p()
        _abi.check(self.lib.chd_context_create(device, C.byref(handle)), "context create")
        self.handle = handle
        # objects living on this context, closed with it (batches first); the
        # library also accepts any destroy order
        self._batches = weakref.WeakSet()
        self._sessions = weakref.WeakSet()
        self._proposals = weakref.WeakSet()
        self._frames = weakref.WeakSet()

    def synchronize(self):
        _abi.check(self.lib.chd_context_synchronize(self.handle), "synchronize")

    def close(self):
        if self.handle:
            for b in list(self._batches):
                b.close()
            for s in list(self._sessions):
                s.close()
            for p in list(self._proposals):
                p.close()
            for f in list(self._frames):
                f.close()
            self.lib.chd_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class DevicePlane:
    """A luminance plane in device memory (row pitch = width), owned by
    whoever made it (e.g. chiaro_amd.scene.Frames)."""

    address: int
    shape: tuple  # (height, width)
    owner: object = None  # kept alive with the plane


def device_planes(modules) -> bool:
    """True when every module's luminance is a DevicePlane (False when none
    is; a mix is an error)."""
    kinds = {isinstance(m.luminance, DevicePlane) for m in modules}
    if len(kinds) > 1:
        raise ValueError("modules mix host and device planes")
    return kinds == {True}


class _ViewArrays:
    """Keeps the numpy buffers behind a chd_view array alive."""

    def __init__(self, modules):
        self.keep = []
        self.cams = []
        n = len(modules)
        self.views = (_abi.View * n)()
        for i, m in enumerate(modules):
            warp = np.ascontiguousarray(m.warp_points, dtype=np.float32)
            conf = np.where(np.isnan(warp[..., 0]), 0.0, 1.0).astype(np.float32)
            self.keep += [warp, conf]
            v = self.views[i]
            if isinstance(m.luminance, DevicePlane):
                v.luminance = C.cast(C.c_void_p(m.luminance.address), C.POINTER(C.c_float))
                v.plane_height, v.plane_width = m.luminance.shape
                self.keep.append(m.luminance)
            else:
                plane = np.ascontiguousarray(m.luminance, dtype=np.float32)
                self.keep.append(plane)
                v.luminance = _abi.fptr(plane)
                v.plane_height, v.plane_width = plane.shape
            v.width, v.height = m.width, m.height
            if m.camera is not None:
                cam = _abi.Camera()
                C.pointer(cam)[0] = m.camera
                self.cams.append(cam)
                v.camera = C.pointer(cam)
            v.global_warp.step = m.warp_step
            v.global_warp.rows, v.global_warp.columns = warp.shape[:2]
            v.global_warp.points = _abi.fptr(warp)
            v.global_warp.confidence = _abi.fptr(conf)
            v.accepted = 1 if m.accepted else 0


def _report_dict(r: _abi.DepthReport) -> dict:
    out = {name: getattr(r, name) for name, _ in _abi.DepthReport._fields_}
    out["median_depth"] = r.median_depth if r.has_median_depth else None
    out["median_score_improvement"] = (
        r.median_score_improvement if r.has_median_score_improvement else None)
    return out


class Session:
    """Device-resident inputs and intermediates of one call (chd_session)."""

    def __init__(self, context: Context, modules, reference_index: int, options: DepthOptions):
        self.context = context
        self.lib = context.lib
        self.modules = modules
        self.options = options
        self._views = _ViewArrays(modules)
        handle = C.c_void_p()
        create = (self.lib.chd_session_create_device if device_planes(modules)
                  else self.lib.chd_session_create)
        _abi.check(create(context.handle, self._views.views, len(modules), reference_index,
                          C.byref(options.to_c()), C.byref(handle)), "session create")
        self.handle = handle
        context._sessions.add(self)
        self._reference = reference_index
        g = _abi.GridSizes()
        self.none = self.lib.chd_session_grid(handle, C.byref(g)) == _abi.CHD_NONE
        self.grid = g

    def close(self):
        """Frees the session; a Batch that holds it can no longer run."""
        if self.handle:
            self.lib.chd_session_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_planes(self, planes):
        """New luminance for every view (a new capture of the same rig):
        numpy arrays or DevicePlanes, one per module."""
        device = isinstance(planes[0], DevicePlane)
        keep = []
        arr = (C.c_void_p * len(planes))()
        for i, p in enumerate(planes):
            if isinstance(p, DevicePlane) != device:
                raise ValueError("planes mix host and device memory")
            if device:
                arr[i] = p.address
            else:
                a = np.ascontiguousarray(p, np.float32)
                keep.append(a)
                arr[i] = a.ctypes.data
        _abi.check(self.lib.chd_session_set_planes(self.handle, arr, len(planes), int(device)),
                   "set planes")
        self._planes = planes

    def run(self, stages: int = _abi.CHD_STAGE_ALL) -> int:
        return _abi.check(self.lib.chd_session_run(self.handle, stages), "session run")

    def enable_timing(self, on: bool = True):
        _abi.check(self.lib.chd_session_enable_timing(self.handle, int(on)), "timing")

    def kernel_times(self) -> dict:
        cap = 64
        arr = (_abi.KernelTime * cap)()
        count = C.c_uint32()
        _abi.check(self.lib.chd_session_kernel_times(self.handle, arr, cap, C.byref(count)), "times")
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(count.value)}

    def work(self) -> dict:
        w = _abi.Work()
        _abi.check(self.lib.chd_session_work(self.handle, C.byref(w)), "work")
        return {name: getattr(w, name) for name, _ in _abi.Work._fields_}

    def reset_kernel_times(self):
        _abi.check(self.lib.chd_session_reset_kernel_times(self.handle), "reset times")

    def set_tuning(self, cost_nodes_per_block: int = 0, cost_label_slab: int = 0,
                   sgm_paths_per_block: int = 0, cost_tile_order: int = 0):
        t = _abi.Tuning(cost_nodes_per_block, cost_label_slab, sgm_paths_per_block,
                        cost_tile_order)
        _abi.check(self.lib.chd_session_set_tuning(self.handle, C.byref(t)), "set tuning")

    def tuning(self) -> dict:
        t = _abi.Tuning()
        _abi.check(self.lib.chd_session_get_tuning(self.handle, C.byref(t)), "get tuning")
        return {name: getattr(t, name) for name, _ in _abi.Tuning._fields_}

    def occupancy(self) -> dict:
        """Work-groups per CU of the ZNCC kernels at this session's launch
        shapes (the band sweep while windows are set)."""
        out = (C.c_uint32 * 3)()
        _abi.check(self.lib.chd_session_occupancy(self.handle, out), "occupancy")
        return {"cost_volume": out[0], "direct": out[1], "targets": out[2]}

    def autotune(self, cache_path: str | None = None) -> dict:
        t = _abi.Tuning()
        path = None if cache_path is None else str(cache_path).encode()
        _abi.check(self.lib.chd_autotune(self.handle, path, C.byref(t)), "autotune")
        return {name: getattr(t, name) for name, _ in _abi.Tuning._fields_}

    # --- intermediate results (parity checks)
    def view_scores(self):
        """Per-view ZNCC slices of the last run with CHD_STAGE_KEEP_VIEW_SCORES:
        (views, nodes, labels) f32, views in index order over the targets."""
        g = self.grid
        views = sum(1 for i, m in enumerate(self.modules)
                    if i != self._reference and m.accepted and m.camera is not None)
        out = np.empty((views, g.coarse_columns * g.coarse_rows, g.labels), np.float32)
        _abi.check(self.lib.chd_session_read_view_scores(self.handle, _abi.ptr(out)),
                   "read view scores")
        return out

    def band(self):
        """The label window (lo, hi) of every coarse node set by the last
        proposal (labels 0 and lo..hi are swept)."""
        g = self.grid
        n = g.coarse_columns * g.coarse_rows
        lo = np.empty(n, np.uint16)
        hi = np.empty(n, np.uint16)
        _abi.check(self.lib.chd_session_read_band(self.handle, _abi.ptr(lo), _abi.ptr(hi)),
                   "read band")
        return lo, hi

    def cost_volume(self):
        g = self.grid
        nodes = g.coarse_columns * g.coarse_rows
        scores = np.empty((nodes, g.labels), np.float32)
        costs = np.empty((nodes, g.labels), np.float32)
        guidance = np.empty(nodes, np.float32)
        tested = np.empty(nodes, np.uint8)
        _abi.check(self.lib.chd_session_read_cost_volume(
            self.handle, _abi.ptr(scores), _abi.ptr(costs), _abi.ptr(guidance), _abi.ptr(tested)),
            "read cost volume")
        return scores, costs, guidance, tested

    def tested(self):
        """The cost volume's tested flags alone (the volumes stay on the device)."""
        g = self.grid
        tested = np.empty(g.coarse_columns * g.coarse_rows, np.uint8)
        _abi.check(self.lib.chd_session_read_cost_volume(self.handle, None, None, None,
                                                         _abi.ptr(tested)), "read tested")
        return tested

    def aggregated(self):
        g = self.grid
        out = np.empty((g.coarse_columns * g.coarse_rows, g.labels), np.float32)
        _abi.check(self.lib.chd_session_read_aggregated(self.handle, _abi.ptr(out)), "read aggregated")
        return out

    def coarse(self):
        g = self.grid
        nodes = g.coarse_columns * g.coarse_rows
        field = np.zeros(nodes, _abi.NODE_DTYPE)
        fillable = np.empty(nodes, np.uint8)
        best = np.empty(nodes, np.uint16)
        _abi.check(self.lib.chd_session_read_coarse(self.handle, _abi.ptr(field), _abi.ptr(fillable),
                                                    _abi.ptr(best)), "read coarse")
        return field, fillable, best

    def direct(self, stage: int):
        g = self.grid
        nodes = g.fine_columns * g.fine_rows
        field = np.zeros(nodes, _abi.NODE_DTYPE)
        guidance = np.empty(nodes, np.float32)
        tested = np.empty(nodes, np.uint8)
        _abi.check(self.lib.chd_session_read_direct(self.handle, stage, _abi.ptr(field),
                                                    _abi.ptr(guidance), _abi.ptr(tested)),
                   "read direct")
        return field, guidance, tested

    def results(self):
        """(DenseDepthMap, {view index: TargetResult}) or (None, {})."""
        if self.none:
            return None, {}
        g = self.grid
        shape = (g.fine_rows, g.fine_columns)
        depth = np.empty(shape, np.float64)
        conf = np.empty(shape, np.float32)
        prov = np.empty(shape, np.uint8)
        m = _abi.DepthMap()
        m.depth = depth.ctypes.data_as(C.POINTER(C.c_double))
        m.confidence = _abi.fptr(conf)
        m.provenance = prov.ctypes.data_as(C.POINTER(C.c_uint8))
        n = len(self.modules)
        targets = (_abi.TargetResult * n)()
        keep = []
        for t in targets:
            pts = np.empty(shape + (2,), np.float32)
            cf = np.empty(shape, np.float32)
            dec = np.empty(shape, np.uint8)
            keep.append((pts, cf, dec))
            t.points = _abi.fptr(pts)
            t.confidence = _abi.fptr(cf)
            t.decisions = dec.ctypes.data_as(C.POINTER(C.c_uint8))
        rc = _abi.check(self.lib.chd_session_read_results(self.handle, C.byref(m), targets),
                        "read results")
        if rc == _abi.CHD_NONE:
            return None, {}
        dmap = DenseDepthMap(m.columns, m.rows, m.step, m.near_depth, m.far_depth, depth, conf, prov)
        out = {}
        for i, t in enumerate(targets):
            if t.replaced:
                pts, cf, dec = keep[i]
                out[i] = TargetResult(pts, cf, dec, _report_dict(t.report))
        return dmap, out


class Batch:
    """Several sessions run together (chd_batch), captured once per stage
    mask into a hipGraph: by default pipelined over two streams (capture k's
    SGM and later stages beside capture k + 1's cost volume); one_sweep=True
    runs every SGM path sweep in one launch after the cost volumes;
    persistent=True runs the cost volumes as one persistent work-queue
    launch."""

    def __init__(self, context: Context, sessions, persistent: bool = False,
                 one_sweep: bool = False):
        self.context = context
        self.lib = context.lib
        self.sessions = list(sessions)
        arr = (C.c_void_p * len(self.sessions))(*[s.handle.value for s in self.sessions])
        handle = C.c_void_p()
        flags = (_abi.CHD_BATCH_PERSISTENT if persistent else 0) | \
            (_abi.CHD_BATCH_ONE_SWEEP if one_sweep else 0)
        _abi.check(self.lib.chd_batch_create(context.handle, arr, len(self.sessions), flags,
                                             C.byref(handle)), "batch create")
        self.handle = handle
        context._batches.add(self)

    def set_planes(self, planes):
        """New luminance for every view (a new capture of the same rig):
        numpy arrays or DevicePlanes, one per module."""
        device = isinstance(planes[0], DevicePlane)
        keep = []
        arr = (C.c_void_p * len(planes))()
        for i, p in enumerate(planes):
            if isinstance(p, DevicePlane) != device:
                raise ValueError("planes mix host and device memory")
            if device:
                arr[i] = p.address
            else:
                a = np.ascontiguousarray(p, np.float32)
                keep.append(a)
                arr[i] = a.ctypes.data
        _abi.check(self.lib.chd_session_set_planes(self.handle, arr, len(planes), int(device)),
                   "set planes")
        self._planes = planes

    def run(self, stages: int = _abi.CHD_STAGE_ALL) -> int:
        return _abi.check(self.lib.chd_batch_run(self.handle, stages), "batch run")

    def close(self):
        if self.handle:
            self.lib.chd_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def refine_multiview_depth(modules, reference_index: int, options: DepthOptions,
                           context: Context | None = None):
    """refine_multiview_depth (depth.rs:312-317) on the GPU.

    Returns (DenseDepthMap | None, {target index: TargetResult}); a target
    result replaces that module's warp and report as the reference does
    (depth.rs:526-546)."""
    own = context is None
    ctx = Context() if own else context
    try:
        session = Session(ctx, modules, reference_index, options)
        try:
            if session.none:
                return None, {}
            session.run(_abi.CHD_STAGE_ALL)
            return session.results()
        finally:
            session.close()
    finally:
        if own:
            ctx.close()


def luminance_half(context: Context, samples: np.ndarray, black_q6: float, white_q6: float,
                   reserve_headroom: bool = False) -> np.ndarray:
    """Mosaic::luminance_half (image.rs:401-426) of a u16 mosaic on the GPU,
    with a unit flat field; reserve_headroom first applies
    reserve_highlight_headroom (image.rs:69-75). Returns the (h/2, w/2) f32
    plane."""
    raw = np.ascontiguousarray(samples, dtype=np.uint16)
    if raw.ndim != 2:
        raise ValueError("samples must be a 2-D array")
    height, width = raw.shape
    plane = np.empty((height // 2, width // 2), np.float32)
    _abi.check(context.lib.chd_luminance_half(
        context.handle, raw.ctypes.data, width, height, float(black_q6), float(white_q6),
        int(bool(reserve_headroom)), plane.ctypes.data), "luminance_half")
    return plane
