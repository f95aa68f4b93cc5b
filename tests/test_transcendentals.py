"""The path's transcendentals against the platform libm the reference calls.

f32::exp and f64::exp lower to the platform libm (glibc on
x86_64-linux-gnu). The oracle and the device kernels carry a restatement of
glibc's algorithm (ARM optimized-routines expf / exp); these tests pin it:
exhaustively for expf, on the path's argument domain for exp, and by
recomputing the tables from their definition."""
import re
import struct
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_expf_exhaustive_against_libm(oracle_lib):
    """Every f32 input (NaN payloads compared as NaN)."""
    assert oracle_lib.or_expf_mismatches(0, 0xFFFFFFFF, 1, 8) == 0


def test_exp_against_libm_on_the_path_domain(oracle_lib):
    rng = np.random.default_rng(0)
    # plane-fit weights: exp(-2.5 |dg|) with |dg| an f32 <= 0.6, and exp(-d^2/6)
    dg = (rng.random(400_000, dtype=np.float32) * np.float32(0.6)).astype(np.float64)
    xs = np.concatenate([-2.5 * dg, -np.array([0, 1, 2, 4, 5, 8], float) / 6.0,
                         -1.5 * rng.random(400_000), [0.0, -0.0, -1e-300]])
    for x in xs:
        a, b = oracle_lib.or_exp(float(x)), oracle_lib.or_libm_exp(float(x))
        assert struct.pack("<d", a) == struct.pack("<d", b), x


def _exp2f_table():
    getcontext().prec = 60
    ln2 = Decimal(2).ln()
    out = []
    for i in range(32):
        bits = struct.unpack("<Q", struct.pack("<d", float((Decimal(i) / 32 * ln2).exp())))[0]
        out.append((bits - (i << 47)) & 0xFFFFFFFFFFFFFFFF)
    return out


def _hex_block(text, start):
    block = text[text.index(start):]
    block = block[:block.index("}")]
    return [int(h, 16) for h in re.findall(r"0x([0-9a-f]{16})ULL", block)]


def test_expf_table_is_its_definition():
    want = _exp2f_table()
    oracle = (ROOT / "oracle/depth_oracle.c").read_text()
    device = (ROOT / "chiaro_amd/csrc/depth_device.hpp").read_text()
    assert _hex_block(oracle, "EXP2F_TAB[32]") == want
    assert _hex_block(device, "kExp2fTable[32]") == want


def test_exp_table_is_its_definition():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", ROOT / "chiaro_amd/csrc/gen_exp_table.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    want = [v for row in gen.table() for v in row]
    inc = (ROOT / "chiaro_amd/csrc/exp_table.inc").read_text()
    assert [int(h, 16) for h in re.findall(r"0x([0-9a-f]{16})ULL", inc)] == want
    oracle = (ROOT / "oracle/depth_oracle.c").read_text()
    assert _hex_block(oracle, "EXP_TAB[256]") == want
