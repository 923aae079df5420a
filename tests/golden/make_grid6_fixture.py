"""Extract the reference's 6x6 weight-grid sampling tables as a data fixture.

ASTC_Encode.hlsl:268-304 (niepp/astc_encoder, MIT) holds two constant arrays
that decide how a 6x6 block's 36 texels are sampled onto its 4x4 weight grid:
idx_grids[16] (four texel indices per grid point) and wt_grids[16] (the four
tap weights).  This script reads those numbers out of the shader text and
writes them, as numbers, to grid6_taps.json.  It runs once, in a container
that has the reference mounted; the tests only read the JSON.

Usage: python tests/golden/make_grid6_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _array_rows(text, name, ctor):
    """The argument lists of `ctor(...)` inside `static const ... name[16] = { ... };`."""
    m = re.search(r"\b" + name + r"\s*\[\s*16\s*\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise SystemExit(f"{name} not found")
    rows = re.findall(ctor + r"\s*\(([^)]*)\)", m.group(1))
    if len(rows) != 16:
        raise SystemExit(f"{name}: expected 16 rows, found {len(rows)}")
    return [[s.strip() for s in r.split(",")] for r in rows]


def main(ref_root):
    path = os.path.join(ref_root, "ASTC_Encode.hlsl")
    with open(path, encoding="utf-8", errors="replace") as f:
        text = f.read()
    idx = [[int(v) for v in r] for r in _array_rows(text, "idx_grids", "uint4")]
    wt = [[float(v) for v in r] for r in _array_rows(text, "wt_grids", "float4")]
    assert all(len(r) == 4 for r in idx + wt)
    out = {
        "source": "niepp/astc_encoder ASTC_Encode.hlsl:268-304 idx_grids / wt_grids (MIT)",
        "grid": "4x4 weight grid of a 6x6 block, grid point g = gy * 4 + gx, texel k = y * 6 + x",
        "idx": idx,
        "wt": wt,
    }
    with open(os.path.join(HERE, "grid6_taps.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote grid6_taps.json")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
