"""Writes exp_table.inc: the 128-entry 2^(k/128) table of the f64 exp used by
the plane-fit weights, H[k] (1 + T[k]) with H[k] the correctly rounded double
of 2^(k/128) and T[k] the double nearest (2^(k/128) - H[k]) / H[k], computed
with 80-digit decimal arithmetic."""
import struct
import sys
from decimal import Decimal, getcontext
from pathlib import Path


def table():
    getcontext().prec = 80
    ln2 = Decimal(2).ln()
    rows = []
    for k in range(128):
        exact = (Decimal(k) / 128 * ln2).exp()
        head = float(exact)
        tail = float((exact - Decimal(head)) / Decimal(head))
        hb = struct.unpack("<Q", struct.pack("<d", head))[0]
        tb = struct.unpack("<Q", struct.pack("<d", tail))[0]
        rows.append((tb, (hb - (k << 45)) & 0xFFFFFFFFFFFFFFFF))
    return rows


def main(out: Path):
    lines = ["    0x%016xULL, 0x%016xULL," % row for row in table()]
    out.write_text(
        "// 2^(k/128) = H[k] (1 + T[k]), k = 0..127, as (bits of T[k], bits of H[k] - (k << 45)),\n"
        "// H[k] the correctly rounded double of 2^(k/128), T[k] the double nearest\n"
        "// (2^(k/128) - H[k]) / H[k]; generated by gen_exp_table.py (80-digit decimal).\n"
        + "\n".join(lines) + "\n")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path(__file__).with_name("exp_table.inc"))
