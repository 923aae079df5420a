"""Pins of the 6x6 weight-grid sampling and of the decoder's weight infill.

A 6x6 block carries a 4x4 weight grid.  The encoder samples the block at each
grid point with four taps (idx_grids / wt_grids, ASTC_Encode.hlsl:268-304,
used by sample_texel :307-314 and calculate_normal_weights :335-351); the
decoder spreads the grid back over the texels by the specification's bilinear
infill (ASTC specification C.2.18).  Both are checked here against what does
not come from the code under test:

  * the reference's own tap tables, extracted as numbers into
    tests/golden/grid6_taps.json by tests/golden/make_grid6_fixture.py, against
    the oracle's taps and the kernels' tap functions (astc_format.h);
  * closed forms: a block that is a linear ramp along x (or y) is sampled at
    effective positions near + (far - near) / 3 (taps 4/9, 2/9, 2/9, 1/9 of
    0.999), i.e. at 1/3, 5/3, 10/3, 14/3, so its quantised weights follow from
    exact fractions; a linear weight grid infills to the ramp's values at each
    texel's grid coordinate s * (G - 1) / (B - 1), rounded to the nearest
    sixteenth of a grid step, and a constant grid to a constant.
Swapping a near and a far tap, transposing the grid or moving a tap index
fails at least one of these.
"""
import json
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from tests import oracle_ref as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RANGE = {True: 5, False: 11}
QUANT_5, QUANT_256 = 3, 20


@pytest.fixture(scope="module")
def ref_taps(golden_dir):
    with open(os.path.join(golden_dir, "grid6_taps.json")) as f:
        d = json.load(f)
    return np.array(d["idx"], np.int32), np.array(d["wt"], np.float64)


def test_oracle_taps_equal_the_reference_tables(ref_taps):
    idx, wt = orc.grid6_taps()
    np.testing.assert_array_equal(idx, ref_taps[0])
    # the shader's literals as binary32 constants
    np.testing.assert_array_equal(wt, ref_taps[1].astype(np.float32))


def test_kernel_tap_functions_equal_the_reference_tables(ref_taps, tmp_path):
    """astc_format.h's grid6_tap_index / grid6_tap_far (what the kernels unroll)
    give the reference's indices and, through {0.444, 0.222, 0.111}[far], its weights."""
    src = tmp_path / "taps.cpp"
    src.write_text(
        '#include <cstdio>\n#include "astc_format.h"\n'
        "int main() { for (int g = 0; g < 16; ++g) for (int t = 0; t < 4; ++t)\n"
        '  std::printf("%d %d\\n", astc::grid6_tap_index(g, t), astc::grid6_tap_far(g, t)); }\n')
    exe = tmp_path / "taps"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "astc_encoder_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    out = np.array([list(map(int, ln.split())) for ln in subprocess.run(
        [str(exe)], check=True, capture_output=True, text=True).stdout.split("\n") if ln], np.int64)
    np.testing.assert_array_equal(out[:, 0].reshape(16, 4), ref_taps[0])
    np.testing.assert_array_equal(np.array([0.444, 0.222, 0.111])[out[:, 1]].reshape(16, 4), ref_taps[1])


def test_reference_taps_are_bilinear_at_one_third():
    """The fixture's structure: each grid point's taps are a 2x2 texel square with
    weights 0.999 * (2/3, 1/3) x (2/3, 1/3), the 2/3 on the texel nearer to 5g/3."""
    with open(os.path.join(ROOT, "tests", "golden", "grid6_taps.json")) as f:
        d = json.load(f)
    for g, (ix, w) in enumerate(zip(d["idx"], d["wt"])):
        gy, gx = divmod(g, 4)
        ys, xs = [i // 6 for i in ix], [i % 6 for i in ix]
        assert xs[1] == xs[0] + 1 and ys[2] == ys[0] + 1 and xs[0] == xs[2] and ys[0] == ys[1]
        near_x = round(5 * gx / 3)
        near_y = round(5 * gy / 3)
        for k in range(4):
            fx = Fraction(2, 3) if xs[k] == near_x else Fraction(1, 3)
            fy = Fraction(2, 3) if ys[k] == near_y else Fraction(1, 3)
            assert abs(w[k] - 0.999 * float(fx * fy)) < 1e-12


def _expected_ramp_weights(rng_levels):
    """Quantised grid weights of a linear ramp sampled at near + (far - near)/3."""
    pos = []
    for g in range(4):
        near = round(Fraction(5 * g, 3))
        c0 = [0, 1, 3, 4][g]
        far = c0 + 1 if near == c0 else c0
        pos.append(near + Fraction(far - near, 3))
    lo, hi = pos[0], pos[3]
    exact = [(p - lo) / (hi - lo) * rng_levels for p in pos]
    # none of the values sits near a rounding tie
    assert all(abs((e - int(e)) - Fraction(1, 2)) > Fraction(1, 100) for e in exact)
    return [int(e + Fraction(1, 2)) for e in exact]


@pytest.mark.parametrize("alpha", [True, False])
@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("sched", [orc.SCHED_CANONICAL, orc.SCHED_SOURCE])
@pytest.mark.parametrize("f64", [False, True])
def test_6x6_linear_ramp_weights(alpha, axis, sched, f64):
    """R = 20 + 40 * coordinate (x or y), other channels constant: the axis is red,
    the endpoints the ramp's ends (20, 220), and grid weight g along the ramp is
    round(range * (p_g - p_0) / (p_3 - p_0)) at p = 1/3, 5/3, 10/3, 14/3."""
    ys, xs = np.divmod(np.arange(36), 6)
    coord = xs if axis == "x" else ys
    tex = np.stack([20.0 + 40.0 * coord, np.full(36, 100.0), np.full(36, 50.0), np.full(36, 255.0)], 1)
    _, tr = orc.encode_block(tex, 6, alpha, sched, f64=f64)
    assert list(tr.ep_q[:6]) == [20, 220, 100, 100, 50, 50]
    line = _expected_ramp_weights(RANGE[alpha])
    assert line == ([0, 3, 8, 11] if not alpha else [0, 2, 3, 5])
    grid = np.array(tr.w_q).reshape(4, 4)
    want = np.tile(line, (4, 1)) if axis == "x" else np.tile(line, (4, 1)).T
    np.testing.assert_array_equal(grid, want)


def _block_4x4_grid(weight_syms, quant, ep=(0, 255, 0, 255, 0, 255)):
    """A single-plane CEM 8 block with a 4x4 weight grid of the given ISE symbols."""
    ep_ise, _ = orc.bise(np.array(ep, np.uint32), QUANT_256)
    wt_ise, _ = orc.bise(np.array(weight_syms, np.uint32).reshape(-1), quant)
    out = np.zeros(16, np.uint8)
    orc.lib().oracle_assemble_block(orc.lib().oracle_blockmode(quant), 8, orc._ptr(ep_ise), orc._ptr(wt_ise),
                                    orc._ptr(out))
    return out


@pytest.mark.parametrize("syms", [2, 0, 4])
def test_infill_of_a_constant_grid_is_constant(syms):
    blk = _block_4x4_grid(np.full(16, syms), QUANT_5)
    ok, _, w = orc.decode_block_weights(blk, 6)
    assert ok == 1
    assert np.all(w == 16 * syms)  # QUANT_5 symbols unquantise to 0, 16, 32, 48, 64


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("dim", [6, 4, 5, 8])
def test_infill_of_a_linear_grid_is_the_ramp_at_each_texel(axis, dim):
    """Grid values 0, 16, 32, 48 along one axis: the bilinear infill of a linear
    function is linear, so texel s gets 48 * s / (dim - 1) (its grid coordinate
    s * 3 / (dim - 1) in sixteenths), rounded to the nearest integer."""
    g = np.arange(4)
    syms = np.tile(g, (4, 1)) if axis == "x" else np.tile(g, (4, 1)).T
    blk = _block_4x4_grid(syms, QUANT_5)
    ok, _, w = orc.decode_block_weights(blk, dim)
    assert ok == 1
    exact = [Fraction(48 * s, dim - 1) for s in range(dim)]
    assert all(abs((e - int(e)) - Fraction(1, 2)) > Fraction(1, 50) for e in exact)
    line = np.array([int(e + Fraction(1, 2)) for e in exact])
    want = np.tile(line, (dim, 1)) if axis == "x" else np.tile(line, (dim, 1)).T
    np.testing.assert_array_equal(w, want)
    if dim == 6:
        np.testing.assert_array_equal(line, [0, 10, 19, 29, 38, 48])


def test_infill_on_a_matching_grid_is_the_identity():
    syms = np.arange(16) % 5
    blk = _block_4x4_grid(syms, QUANT_5)
    ok, _, w = orc.decode_block_weights(blk, 4)
    assert ok == 1
    np.testing.assert_array_equal(w, 16 * syms.reshape(4, 4))
