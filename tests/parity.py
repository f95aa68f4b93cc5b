"""Comparison helpers shared by the parity tests."""
import numpy as np


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    return a.view({2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def assert_bitwise(name: str, got: np.ndarray, want: np.ndarray):
    """Bit equality, with every NaN counted equal to every other NaN."""
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    if got.dtype.kind == "f":
        gn, wn = np.isnan(got), np.isnan(want)
        same = (bits(got) == bits(want)) | (gn & wn)
    else:
        same = got == want
    bad = np.flatnonzero(~same.reshape(-1))
    if bad.size:
        i = bad[0]
        raise AssertionError(
            f"{name}: {bad.size} of {same.size} differ; first at {i}: got {got.reshape(-1)[i]!r} "
            f"want {want.reshape(-1)[i]!r}")


def assert_nodes(name: str, got: np.ndarray, want: np.ndarray):
    """chd_node_depth arrays: same validity, bit-equal payload where valid."""
    assert_bitwise(name + ".valid", got["valid"], want["valid"])
    v = want["valid"].astype(bool)
    assert_bitwise(name + ".regularized", got["regularized"][v], want["regularized"][v])
    assert_bitwise(name + ".depth", got["depth"][v], want["depth"][v])
    assert_bitwise(name + ".confidence", got["confidence"][v], want["confidence"][v])
    assert_bitwise(name + ".improvement", got["improvement"][v], want["improvement"][v])
