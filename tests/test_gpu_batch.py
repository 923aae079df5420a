"""Batched persistent encoding and device mip chains.

The batched kernel must give every image exactly the blocks of a plain launch
(bit-exact, any tuning, every graph replay), and the device box filter must
reproduce the host generator's mip levels byte for byte.
"""
import numpy as np
import pytest

import astc_encoder_amd as astc
from astc_encoder_amd import synth
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


@pytest.mark.parametrize("channels,w,h", [(4, 64, 64), (3, 256, 128), (2, 96, 40), (4, 7, 33), (3, 5, 4)])
def test_device_mip_chain_matches_host(enc, channels, w, h):
    rng = np.random.default_rng(w * h + channels)
    img = rng.integers(0, 256, size=(h, w, channels), dtype=np.uint8)
    host = synth.mip_chain(img)
    dev = enc.mip_chain(torch.from_numpy(img).cuda())
    torch.cuda.synchronize()
    assert len(dev) == len(host)
    for a, b in zip(dev, host):
        np.testing.assert_array_equal(a.cpu().numpy(), b)
    info = astc.mip_chain_info(astc.make_desc(w, h, {4: 0, 3: 1, 2: 2}[channels], 4, 0))
    assert [(d.width, d.height) for d in info] == [(l.shape[1], l.shape[0]) for l in host]


def mixed_images(enc, seed, n):
    """n textures of the mixed distribution scaled down, each with its device mip chain."""
    images = []
    for mode, w, h, s in synth.draw_texture_specs(seed, n):
        gen, dim, flags, fmt = synth.MODES[mode]
        w, h = max(4, w // 16), max(4, h // 16)
        base = torch.from_numpy(gen(s, w, h)).cuda()
        for lvl in enc.mip_chain(base):
            images.append((lvl, dim, flags))
    return images


def poison(b):
    """Fill the batch's outputs with a sentinel, so that a block the launch does
    not write cannot pass for the expected one (freed plain outputs of the same
    size are often handed back to the batch by the allocator)."""
    for o in b.outputs:
        o.fill_(0xA5)
    torch.cuda.synchronize()


def plain_outputs(enc, images):
    outs = [enc.encode(t, block_dim=d, flags=f) for t, d, f in images]
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


def test_batch_matches_plain_and_oracle(enc):
    images = mixed_images(enc, 1, 40)
    # odd shapes and an unaligned pitch among them
    raw = torch.from_numpy(np.random.default_rng(0).integers(0, 256, (37, 23, 3), dtype=np.uint8)).cuda()
    images += [(raw, 6, astc.FLAG_ALPHA), (raw[:, :21], 4, 0), (raw[:1, :1], 4, astc.FLAG_NORMAL)]
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    info = b.info()
    assert info["blocks"] == sum(len(w) for w in want)
    poison(b)
    b.launch()
    torch.cuda.synchronize()
    for i, (o, w) in enumerate(zip(b.outputs, want)):
        np.testing.assert_array_equal(o.cpu().numpy(), w, err_msg=f"image {i}")
    # and against the oracle on a few of them
    for i in (0, 1, len(images) - 3, len(images) - 1):
        t, d, f = images[i]
        np.testing.assert_array_equal(b.outputs[i].cpu().numpy(), orc.oracle_encode(t.cpu().numpy(), d, f))
    b.close()


def test_batch_with_largest_dimensions(enc):
    """A row and a column of 2^24 - 1 texels (the .astc header's limit) beside
    small images: thousands of work items from one image, every block padded."""
    rng = np.random.default_rng(24)
    wide = torch.from_numpy(rng.integers(0, 256, (1, 16777215, 4), dtype=np.uint8)).cuda()
    tall = torch.from_numpy(rng.integers(0, 256, (16777215, 1, 3), dtype=np.uint8)).cuda()
    images = [(wide, 4, astc.FLAG_ALPHA), (tall, 6, astc.FLAG_SRGB)] + mixed_images(enc, 6, 4)
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    poison(b)
    b.launch()
    torch.cuda.synchronize()
    for i, (o, w) in enumerate(zip(b.outputs, want)):
        np.testing.assert_array_equal(o.cpu().numpy(), w, err_msg=f"image {i}")
    b.close()


@pytest.mark.parametrize("bpl,wpc,pf,split,lanes", [(1, 1, 0, 0, 1), (2, 2, 1, 0, 1), (4, 2, 0, 0, 1), (8, 1, 1, 0, 1),
                                                   (8, 4, 1, 0, 1), (1, 1, 0, 1, 1), (2, 2, 1, 1, 1), (4, 2, 1, 1, 1),
                                                   (8, 4, 0, 1, 1), (1, 3, 0, 0, 2), (2, 2, 1, 0, 2), (4, 3, 0, 1, 2),
                                                   (8, 1, 1, 1, 2), (8, 2, 0, 0, 2)])
def test_every_tuning_gives_the_same_blocks(enc, bpl, wpc, pf, split, lanes):
    images = mixed_images(enc, 2, 24)
    # four-channel 4x4 images (which the split form runs in the 6x6 kernel)
    raw = torch.from_numpy(np.random.default_rng(bpl * 10 + wpc).integers(0, 256, (75, 61, 4), dtype=np.uint8)).cuda()
    images += [(raw, 4, astc.FLAG_ALPHA), (raw[:, :, :3].contiguous(), 4, 0), (raw[:9, :7], 4, astc.FLAG_SRGB)]
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    b.set_tuning(bpl, wpc, pf, split, lanes)
    assert b.tuning() == {"blocks_per_lane": bpl, "workgroups_per_cu": wpc, "prefetch": pf, "split_modes": split,
                          "lanes_per_block": lanes}
    poison(b)
    b.launch()
    torch.cuda.synchronize()
    for o, w in zip(b.outputs, want):
        np.testing.assert_array_equal(o.cpu().numpy(), w)
    b.close()


@pytest.mark.parametrize("seed", range(6))
def test_random_batches(enc, seed):
    """Seeded random batches: sizes, formats, block sizes and flags mixed freely
    (some images as narrowed views, so with padded rows), under a random tuning."""
    rng = np.random.default_rng(500 + seed)
    images = []
    for _ in range(int(rng.integers(5, 40))):
        w, h = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        c = int(rng.choice([4, 3, 2]))
        t = torch.from_numpy(rng.integers(0, 256, size=(h, w + int(rng.integers(0, 3)), c), dtype=np.uint8)).cuda()
        images.append((t[:, :w], int(rng.choice([4, 6])), int(rng.integers(0, 8))))
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    tuning = (int(rng.choice([1, 2, 4, 8])), int(rng.integers(1, 3)), int(rng.integers(0, 2)), int(rng.integers(0, 2)),
              int(rng.integers(1, 3)))
    print(f"case seed={seed} images={len(images)} tuning={tuning}", flush=True)
    b.set_tuning(*tuning)
    poison(b)
    b.launch()
    torch.cuda.synchronize()
    for i, (o, w) in enumerate(zip(b.outputs, want)):
        np.testing.assert_array_equal(o.cpu().numpy(), w, err_msg=f"image {i}")
    b.close()


@pytest.mark.parametrize("split,lanes", [(0, 1), (1, 1), (0, 2), (1, 2)])
def test_graph_replays_with_eager_work_between(enc, split, lanes):
    images = mixed_images(enc, 3, 16)
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    b.set_tuning(4, 2, 1, split, lanes)
    for rep in range(4):
        for o in b.outputs:
            o.zero_()
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        for o, w in zip(b.outputs, want):
            np.testing.assert_array_equal(o.cpu().numpy(), w, err_msg=f"replay {rep}")
    b.close()


@pytest.mark.parametrize("dim", [4, 6])
def test_split_batch_with_one_block_size(enc, dim):
    """Split launches where one of the two kernels finds no work at all."""
    images = [(t, dim, f) for t, _, f in mixed_images(enc, 6, 12)]
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    b.set_tuning(2, 2, 1, 1)
    for rep in range(2):
        for o in b.outputs:
            o.zero_()
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        for o, w in zip(b.outputs, want):
            np.testing.assert_array_equal(o.cpu().numpy(), w, err_msg=f"replay {rep}")
    b.close()


def test_plain_form_over_a_batch(enc):
    images = mixed_images(enc, 4, 12)
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    for o in b.outputs:
        o.zero_()
    b.launch_plain()
    torch.cuda.synchronize()
    for o, w in zip(b.outputs, want):
        np.testing.assert_array_equal(o.cpu().numpy(), w)
    b.close()


def test_autotune_picks_a_valid_tuning(enc):
    images = mixed_images(enc, 5, 16)
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    t = b.autotune(reps=2)
    assert t["blocks_per_lane"] in (1, 2, 4, 8) and t["workgroups_per_cu"] in (1, 2)
    assert t["split_modes"] in (0, 1) and t["lanes_per_block"] in (1, 2)
    for o in b.outputs:
        o.zero_()
    b.launch()
    torch.cuda.synchronize()
    for o, w in zip(b.outputs, want):
        np.testing.assert_array_equal(o.cpu().numpy(), w)
    b.close()


@pytest.mark.parametrize("split", [0, 1])
def test_max_accumulation_images_in_a_batch(enc, split):
    """Images with the max-accumulation search (plain launches inside the batch's
    graph) among work-list images: every image equals its plain encode, over graph
    replays, and the search's blocks equal the oracle's."""
    rng = np.random.default_rng(40 + split)
    images = mixed_images(enc, 8, 6)
    for k, (c, dim, flags) in enumerate([(4, 4, astc.FLAG_ALPHA), (3, 6, astc.FLAG_SRGB), (2, 4, astc.FLAG_NORMAL),
                                         (4, 6, astc.FLAG_ALPHA | astc.FLAG_SRGB), (3, 4, 0)]):
        h, w = int(rng.integers(1, 90)), int(rng.integers(1, 90))
        t = torch.from_numpy(rng.integers(0, 256, size=(h, w, c), dtype=np.uint8)).cuda()
        images.insert(2 * k, (t, dim, flags | astc.FLAG_MAX_ACCUMULATION))
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    b.set_tuning(4, 2, 1, split)
    assert b.info()["blocks"] == sum(len(w) for w in want)
    for rep in range(3):
        poison(b)
        b.launch()
        torch.cuda.synchronize()
        for i, (o, w) in enumerate(zip(b.outputs, want)):
            np.testing.assert_array_equal(o.cpu().numpy(), w, err_msg=f"replay {rep} image {i}")
    for i, (t, d, f) in enumerate(images):
        if f & astc.FLAG_MAX_ACCUMULATION and t.shape[0] * t.shape[1] < 4000:
            np.testing.assert_array_equal(b.outputs[i].cpu().numpy(), orc.oracle_encode(t.cpu().numpy(), d, f))
    b.close()


def test_batch_of_max_accumulation_images_only(enc):
    rng = np.random.default_rng(41)
    images = [(torch.from_numpy(rng.integers(0, 256, size=(33, 47, 4), dtype=np.uint8)).cuda(), d,
               astc.FLAG_ALPHA | astc.FLAG_MAX_ACCUMULATION) for d in (4, 6)]
    want = plain_outputs(enc, images)
    b = astc.Batch(enc, images)
    assert b.info()["items"] == 0
    poison(b)
    b.launch()
    torch.cuda.synchronize()
    for o, w in zip(b.outputs, want):
        np.testing.assert_array_equal(o.cpu().numpy(), w)
    b.close()


def test_batch_rejects_bad_images(enc):
    t = torch.zeros((8, 8, 4), dtype=torch.uint8, device="cuda")
    with pytest.raises(astc.AstcError):
        astc.Batch(enc, [(t, 5, 0)])
    # output blocks must start on a 16-byte boundary
    raw = torch.zeros(4 * 16 + 1, dtype=torch.uint8, device="cuda")
    with pytest.raises(astc.AstcError, match="bad src/dst pointer"):
        astc.Batch(enc, [(t, 4, 0)], outputs=[raw[1:].view(4, 16)])


@pytest.mark.parametrize("tuning", [(3, 2, 1, 0), (4, 0, 1, 0), (4, 5, 1, 0), (4, 2, 2, 0), (4, 2, 1, 2), (4, 2, 1, 0, 3)])
def test_tuning_out_of_range_is_rejected_and_kept(enc, tuning):
    t = torch.zeros((8, 8, 4), dtype=torch.uint8, device="cuda")
    b = astc.Batch(enc, [(t, 4, 0)])
    before = b.tuning()
    with pytest.raises(astc.AstcError, match="tuning out of range"):
        b.set_tuning(*tuning)
    assert b.tuning() == before
    b.launch()
    torch.cuda.synchronize()
    b.close()
