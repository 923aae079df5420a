"""The three-channel path for four-channel blocks with a flat alpha.

With the plain tuning flat_alpha = 1, a wave whose blocks all have one alpha
byte c in every texel, with UNORM value exactly c, is encoded by the
three-channel encoder with alpha endpoints c; that must give the four-channel
encoder's bytes, which the oracle computes.  Cases:
opaque images (every wave flat), per-block constant alphas drawn from the bytes
whose UNORM value is an integer and from those whose value is not (which must
take the four-channel path), waves that mix both, ragged sizes whose edge blocks
read zeros outside the image, both block sizes, RGB and RGBA modes, linear and
sRGB, plain launches of every tuning and batches.
"""
import numpy as np
import pytest

import astc_encoder_amd as astc
from tests import oracle_ref as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def enc():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = astc.Encoder(0)
    yield e
    e.close()


def integer_unorm_bytes():
    u = orc.unorm_table()
    c = np.arange(256)
    return c[u == c.astype(np.float32)], c[u != c.astype(np.float32)]


def flat_alpha_image(rng, w, h, dim, kind):
    img = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    good, bad = integer_unorm_bytes()
    if kind == "opaque":
        img[..., 3] = 255
    else:
        bx, by = (w + dim - 1) // dim, (h + dim - 1) // dim
        if kind == "integer":
            vals = rng.choice(good, size=(by, bx))
        else:  # rows of integer-valued blocks, then rows mixing in the other bytes
            vals = rng.choice(good, size=(by, bx))
            mixed = rng.random((by, bx)) < 0.3
            vals[by // 2:][mixed[by // 2:]] = rng.choice(bad, size=int(mixed[by // 2:].sum()))
        a = np.repeat(np.repeat(vals, dim, axis=0), dim, axis=1)[:h, :w]
        img[..., 3] = a
    return img


@pytest.mark.parametrize("dim", [4, 6])
@pytest.mark.parametrize("flags", [0, 1, 4, 5])
@pytest.mark.parametrize("kind", ["opaque", "integer", "mixed"])
@pytest.mark.parametrize("size", [(384, 192), (203, 131)])
def test_flat_alpha_blocks_match_the_oracle(enc, dim, flags, kind, size):
    rng = np.random.default_rng(dim * 1000 + flags * 100 + len(kind) * 10 + size[1])
    img = flat_alpha_image(rng, size[0], size[1], dim, kind)
    want = orc.oracle_encode(img, dim, flags)
    tex = torch.from_numpy(img).cuda()
    tunings = [None, (1, 0, 0, 1), (1, 1, 1, 1), (1, 8, 0, 1)] + ([(2, 0, 0)] if dim == 6 else [])
    for t in tunings:
        got = enc.encode(tex, block_dim=dim, flags=flags, tuning=t)
        torch.cuda.synchronize()
        got = got.cpu().numpy()
        if not np.array_equal(got, want):
            bad = np.nonzero(~np.all(got == want, axis=1))[0]
            raise AssertionError(f"{kind} {size} dim {dim} flags {flags} tuning {t}: {len(bad)} blocks differ, "
                                 f"first {bad[:5].tolist()}")


@pytest.mark.parametrize("split", [0, 1])
def test_four_channel_images_in_batches(enc, split):
    rng = np.random.default_rng(77 + split)
    images, wants = [], []
    for k, (dim, flags, kind) in enumerate([(6, 5, "opaque"), (6, 1, "mixed"), (4, 1, "integer"), (4, 0, "opaque"),
                                            (6, 4, "integer"), (4, 5, "mixed")]):
        w, h = int(rng.integers(40, 300)), int(rng.integers(40, 300))
        img = flat_alpha_image(rng, w, h, dim, kind)
        images.append((torch.from_numpy(img).cuda(), dim, flags))
        wants.append(orc.oracle_encode(img, dim, flags))
    b = astc.Batch(enc, images)
    b.set_tuning(4, 2, 0, split)
    for o in b.outputs:
        o.fill_(0xA5)
    b.launch()
    torch.cuda.synchronize()
    for i, (o, w) in enumerate(zip(b.outputs, wants)):
        np.testing.assert_array_equal(o.cpu().numpy(), w, err_msg=f"image {i}")
    b.close()
