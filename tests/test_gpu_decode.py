"""The GPU ASTC decoder against the oracle's specification decoder.

Byte-identical decodes for: the encoder's output in every mode; the
reference's own leaf blocks (which also decode to ~38 dB against the leaf
image, pinning the interpretation of endpoints and weights); and random blocks
of several block sizes with random block modes / endpoint modes (the error
colour decisions included).
"""
import numpy as np
import pytest

import astc_encoder_amd as astc
from astc_encoder_amd import synth
from tests import oracle_ref as orc
from tests.conftest import tiles_to_image

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def enc():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = astc.Encoder(0)
    yield e
    e.close()


@pytest.mark.parametrize("mode", list(synth.MODES))
def test_decode_encoder_output(enc, mode):
    gen, dim, flags, _ = synth.MODES[mode]
    img = gen(5, 203, 117)
    blocks = enc.encode(torch.from_numpy(img).cuda(), block_dim=dim, flags=flags)
    got = enc.decode(blocks, 203, 117, dim).cpu().numpy()
    ok, want = orc.decode_image(blocks.cpu().numpy(), 203, 117, dim)
    np.testing.assert_array_equal(got, want)
    # the decode is close to the source texels in the encoded channels
    assert orc.psnr(img, flags, got) > 25.0


def test_decode_reference_leaf_blocks(enc, leaf_sample):
    blocks = torch.from_numpy(leaf_sample["blocks"]).cuda()
    img = tiles_to_image(leaf_sample["tiles"], 128)
    h, w, _ = img.shape
    # arrange reference blocks in the same 128-blocks-per-row layout as the tiles
    got = enc.decode(blocks, w, h, 4).cpu().numpy()
    ok, want = orc.decode_image(leaf_sample["blocks"], w, h, 4)
    np.testing.assert_array_equal(got, want)
    assert 37.0 < orc.psnr(img, astc.FLAG_ALPHA, got) < 40.0


def valid_modes(bw, bh, cem):
    """Block modes the specification decoder accepts for this block size and CEM."""
    out = []
    for m in range(2048):
        blk = np.zeros(16, np.uint8)
        v = m | (cem << 13)
        blk[0], blk[1], blk[2] = v & 255, (v >> 8) & 255, (v >> 16) & 255
        ok, _, info = orc.decode_block(blk, bw, bh)
        if ok and not info.void_extent:
            out.append(m)
    return np.array(out, np.uint64)


@pytest.mark.parametrize("bw,bh", [(4, 4), (6, 6), (8, 5), (10, 8), (12, 12)])
def test_decode_random_blocks(enc, bw, bh):
    rng = np.random.default_rng(bw * 100 + bh)
    n = 4096
    raw = rng.integers(0, 256, size=(n, 16), dtype=np.uint8)
    v = raw.view(np.uint64).copy()
    lo = v[:, 0]
    cems = rng.choice(np.array([0, 4, 8, 12], np.uint64), n)
    one_part = rng.random(n) < 0.9
    lo = np.where(one_part, lo & ~np.uint64(0x3 << 11), lo)          # single partition
    lo = (lo & ~np.uint64(0xF << 13)) | (cems << np.uint64(13))      # LDR direct endpoint modes
    modes = {int(c): valid_modes(bw, bh, int(c)) for c in (0, 4, 8, 12)}
    use_valid = rng.random(n) < 0.8
    for i in np.nonzero(use_valid)[0]:
        ms = modes[int(cems[i])]
        if len(ms):
            lo[i] = (lo[i] & ~np.uint64(0x7FF)) | ms[rng.integers(0, len(ms))]
    v[:, 0] = lo
    blocks = v.view(np.uint8).reshape(n, 16)
    cols = 64
    w, h = cols * bw, (n // cols) * bh
    got = enc.decode(torch.from_numpy(blocks).cuda(), w, h, bw, bh).cpu().numpy()
    valid = 0
    for i in range(n):
        ok, tile, info = orc.decode_block(blocks[i], bw, bh)
        valid += ok
        bx, by = i % cols, i // cols
        np.testing.assert_array_equal(got[by * bh:(by + 1) * bh, bx * bw:(bx + 1) * bw], tile, err_msg=f"block {i}")
    assert valid > n // 2
