"""The allocation-free entry point (chd_reject_small_direct_components_noalloc)
against the allocating one (chd_reject_small_direct_components) through the
C ABI: bit for bit, on two calls in a row with the second on a smaller grid
(the shared scratch keeps the first call's contents), and the size limit.
Device buffers for the test come from the HIP runtime directly."""
import ctypes as C

import numpy as np
import pytest

from chiaro_amd import _abi

from parity import assert_nodes


def _hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipFree.argtypes = [C.c_void_p]
    return lib


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, hip, host: np.ndarray):
        self.hip = hip
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _field(rng, columns, rows):
    """A fine-grid field of depth plateaus with holes, small islands and
    guidance edges, so that components of every size occur."""
    n = columns * rows
    field = np.zeros(n, _abi.NODE_DTYPE)
    y, x = np.mgrid[0:rows, 0:columns]
    plateau = ((x // 37) + 3 * (y // 23)) % 7
    depth = 900.0 + 400.0 * plateau + rng.normal(0, 0.05, size=(rows, columns))
    valid = rng.random((rows, columns)) > 0.25
    guidance = (plateau * 0.9 + rng.normal(0, 0.02, size=(rows, columns))).astype(np.float32)
    field["depth"] = depth.reshape(-1)
    field["confidence"] = rng.random(n).astype(np.float32)
    field["improvement"] = rng.random(n).astype(np.float32)
    field["valid"] = valid.reshape(-1)
    field["regularized"] = (rng.random(n) > 0.8)
    return field, guidance.reshape(-1)


def test_noalloc_size_limit_needs_no_device():
    lib = _abi.depth_lib()
    rc = lib.chd_reject_small_direct_components_noalloc(None, None, None, 2049, 1024, 1e-6, 24, None)
    assert rc == _abi.CHD_ERR_LIMIT
    assert "CHD_NOALLOC_MAX_NODES" in _abi.last_error()


@pytest.mark.gpu
def test_noalloc_components_equal_the_allocating_entry(gpu_context):
    lib = gpu_context.lib
    hip = _hip()
    rng = np.random.default_rng(21)
    tol = 2.5 * ((1.0 / 800.0 - 1.0 / 20000.0) / 127.0)
    for columns, rows in ((1041, 781), (300, 201)):  # the second call is smaller
        print(f"no-alloc components {columns}x{rows}", flush=True)
        field, guidance = _field(rng, columns, rows)
        want = field.copy()
        _abi.check(lib.chd_reject_small_direct_components(gpu_context.handle, _abi.ptr(want),
                                                          _abi.ptr(guidance), columns, rows, tol,
                                                          24), "components")
        d_field, d_guid = DeviceArray(hip, field), DeviceArray(hip, guidance)
        d_out = DeviceArray(hip, np.zeros_like(field))
        try:
            _abi.check(lib.chd_reject_small_direct_components_noalloc(
                gpu_context.handle, d_field.ptr, d_guid.ptr, columns, rows, tol, 24, d_out.ptr),
                "components no-alloc")
            gpu_context.synchronize()
            got = d_out.get(field)
        finally:
            for d in (d_field, d_guid, d_out):
                d.free()
        removed = int(field["valid"].sum() - want["valid"].sum())
        assert removed > 0, "the case must remove small components"
        assert_nodes(f"no-alloc {columns}x{rows}", got, want)
