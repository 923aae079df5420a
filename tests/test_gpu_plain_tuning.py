"""Plain-kernel tunings (astc_encode_tuned, astc_plain_autotune) against the oracle.

Every tuning must give the oracle's blocks bit for bit: one or two lanes per
6x6 block (the two-lane encoder sums, projects and samples half a block per
lane and combines the halves by lane exchange), blocks per lane, next-block
prefetch, and packed texels (the fetched bytes converted again in every pass
of the encode), over every format, flag and channel class, ragged and tiled
sizes, unaligned sources with padded rows, the largest dimensions, seeded
random cases and the full-size 4096x4096 RGBA8 sRGB texture.
"""
import numpy as np
import pytest

import astc_encoder_amd as astc
from astc_encoder_amd import synth
from tests import oracle_ref as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TUNINGS = [(2, 0, 0), (2, 0, 1), (2, 1, 1), (2, 8, 0), (1, 0, 0), (1, 0, 1), (1, 2, 0),
           (1, 0, 0, 0, 1), (1, 1, 0, 0, 1), (1, 8, 0, 0, 1)]
PACKED = [(1, 0, 0, 0, 1), (1, 2, 0, 0, 1)]


@pytest.fixture(scope="module")
def enc():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = astc.Encoder(0)
    yield e
    e.close()


def tuned(enc, tex, dim, flags, tuning):
    out = enc.encode(tex, block_dim=dim, flags=flags, tuning=tuning)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def assert_same(got, want, what):
    if not np.array_equal(got, want):
        bad = np.nonzero(~np.all(got == want, axis=1))[0]
        raise AssertionError(f"{what}: {len(bad)} of {len(want)} blocks differ, first {bad[:5].tolist()}")


@pytest.mark.parametrize("channels", [4, 3, 2])
@pytest.mark.parametrize("flags", range(8))
@pytest.mark.parametrize("shape", [(45, 38), (36, 48)])
def test_six_by_six_every_format_flag_and_class(enc, channels, flags, shape):
    rng = np.random.default_rng(channels * 100 + flags * 10 + shape[0])
    img = rng.integers(0, 256, size=shape + (channels,), dtype=np.uint8)
    # two-level tiles as well as noise: flat and near-flat blocks take the short-axis paths
    img[: shape[0] // 2, : shape[1] // 3] = img[0, 0]
    want = orc.oracle_encode(img, 6, flags)
    tex = torch.from_numpy(img).cuda()
    for t in TUNINGS:
        assert_same(tuned(enc, tex, 6, flags, t), want, f"{channels}ch flags {flags} {shape} tuning {t}")


@pytest.mark.parametrize("channels", [4, 3, 2])
def test_unaligned_source_and_padded_pitch(enc, channels):
    rng = np.random.default_rng(90 + channels)
    raw = torch.from_numpy(rng.integers(0, 256, size=(53, 47 + 3, channels), dtype=np.uint8)).cuda()
    flat = raw.view(-1)[1:1 + 53 * 49 * channels].view(53, 49, channels)  # source 1 byte past a word
    for tex in (raw[:, :47], flat[:, :46]):
        want = orc.oracle_encode(tex.cpu().numpy(), 6, astc.FLAG_ALPHA)
        for t in TUNINGS:
            assert_same(tuned(enc, tex, 6, astc.FLAG_ALPHA, t), want, f"{channels}ch tuning {t}")


@pytest.mark.parametrize("shape", [(1, 16777215), (16777215, 1)])
def test_largest_dimensions_two_lanes_and_packed(enc, shape):
    rng = np.random.default_rng(shape[0] % 97)
    img = rng.integers(0, 256, size=shape + (4,), dtype=np.uint8)
    want = orc.oracle_encode(img, 6, astc.FLAG_ALPHA | astc.FLAG_SRGB, threads=16)
    tex = torch.from_numpy(img).cuda()
    for t in [(2, 0, 1), (1, 0, 0, 0, 1)]:
        assert_same(tuned(enc, tex, 6, astc.FLAG_ALPHA | astc.FLAG_SRGB, t), want, f"{shape} tuning {t}")


@pytest.mark.parametrize("channels", [4, 3, 2])
@pytest.mark.parametrize("flags", range(8))
@pytest.mark.parametrize("shape", [(44, 36), (37, 29)])
def test_four_by_four_packed_texels(enc, channels, flags, shape):
    rng = np.random.default_rng(5000 + channels * 100 + flags * 10 + shape[0])
    img = rng.integers(0, 256, size=shape + (channels,), dtype=np.uint8)
    img[: shape[0] // 2, : shape[1] // 3] = img[0, 0]
    want = orc.oracle_encode(img, 4, flags)
    tex = torch.from_numpy(img).cuda()
    for t in PACKED:
        assert_same(tuned(enc, tex, 4, flags, t), want, f"4x4 {channels}ch flags {flags} {shape} tuning {t}")


@pytest.mark.parametrize("seed", range(24))
def test_random_cases_random_tunings(enc, seed):
    rng = np.random.default_rng(7000 + seed)
    w, h = int(rng.integers(1, 400)), int(rng.integers(1, 400))
    c = int(rng.choice([4, 3, 2]))
    flags = int(rng.integers(0, 8))
    kind = int(rng.integers(0, 3))
    img = rng.integers(0, 256, size=(h, w, c), dtype=np.uint8)
    if kind == 1:
        img[:] = img[0, 0]
    elif kind == 2:
        img = np.where(rng.random((h, w, 1)) < 0.5, img[0, 0], img[-1, -1]).astype(np.uint8)
    t = (int(rng.choice([1, 2])), int(rng.choice([0, 1, 2, 4, 8])), int(rng.integers(0, 2)))
    if seed % 3 == 2:
        t = (1, t[1], 0, 0, 1)  # packed texels
    dim = 4 if seed % 4 == 3 else 6
    print(f"case seed={seed} {w}x{h}x{c} dim={dim} flags={flags} kind={kind} tuning={t}", flush=True)
    assert_same(tuned(enc, torch.from_numpy(img).cuda(), dim, flags, t), orc.oracle_encode(img, dim, flags),
                f"seed {seed}")


def test_full_size_texture_two_lanes(enc):
    img = synth.noise_rgba8(0, 4096, 4096)
    flags = astc.FLAG_ALPHA | astc.FLAG_SRGB
    want = orc.oracle_encode(img, 6, flags, threads=16)
    tex = torch.from_numpy(img).cuda()
    for t in [(2, 0, 0), (2, 0, 1), (1, 0, 0), (1, 0, 0, 0, 1)]:
        assert_same(tuned(enc, tex, 6, flags, t), want, f"4096x4096 tuning {t}")


def test_autotune_picks_a_valid_tuning(enc):
    img = synth.noise_rgba8(3, 512, 512)
    flags = astc.FLAG_ALPHA | astc.FLAG_SRGB
    tex = torch.from_numpy(img).cuda()
    best = enc.autotune_plain(tex, 6, flags, reps=3)
    assert best["lanes_per_block"] in (1, 2) and best["blocks_per_lane"] in (1, 2, 4, 8)
    assert best["prefetch"] in (0, 1) and best["flat_alpha"] in (0, 1) and best["packed_texels"] in (0, 1)
    got = tuned(enc, tex, 6, flags, (best["lanes_per_block"], best["blocks_per_lane"], best["prefetch"],
                                     best["flat_alpha"], best["packed_texels"]))
    assert_same(got, orc.oracle_encode(img, 6, flags), "autotuned")
    # 4x4 images and the max-accumulation search have one lane per block only
    best4 = enc.autotune_plain(tex, 4, flags, reps=1)
    assert best4["lanes_per_block"] == 1
    bestm = enc.autotune_plain(tex, 6, flags | astc.FLAG_MAX_ACCUMULATION, reps=1)
    assert bestm["lanes_per_block"] == 1 and bestm["packed_texels"] == 0


@pytest.mark.parametrize("tuning", [(3, 0, 0), (0, 0, 0), (2, 3, 0), (2, 16, 0), (1, 0, 2), (2, -1, 0), (1, 0, 0, 2),
                                    (1, 0, 0, 0, 2), (2, 0, 0, 0, 1), (1, 0, 1, 0, 1), (1, 0, 0, 1, 1)])
def test_out_of_range_tunings_are_rejected(enc, tuning):
    tex = torch.zeros((12, 12, 4), dtype=torch.uint8, device="cuda")
    with pytest.raises(astc.AstcError, match="plain tuning out of range"):
        enc.encode(tex, block_dim=6, flags=0, tuning=tuning)


@pytest.mark.parametrize("dim", [4, 6])
def test_packed_tuning_with_max_accumulation_runs_the_default_kernel(enc, dim):
    rng = np.random.default_rng(61 + dim)
    img = rng.integers(0, 256, size=(50, 43, 4), dtype=np.uint8)
    flags = astc.FLAG_ALPHA | astc.FLAG_MAX_ACCUMULATION
    want = orc.oracle_encode(img, dim, flags)
    assert_same(tuned(enc, torch.from_numpy(img).cuda(), dim, flags, (1, 0, 0, 0, 1)), want, f"max-acc {dim}")


@pytest.mark.parametrize("channels,dim,flags,shape,want", [
    (4, 6, astc.FLAG_ALPHA | astc.FLAG_SRGB, (4096, 4096), dict(prefetch=1, flat_alpha=1, blocks_per_lane=4)),
    (4, 4, astc.FLAG_ALPHA, (1024, 1024), dict(prefetch=1, flat_alpha=1, blocks_per_lane=1)),
    (3, 6, 0, (96, 96), dict(prefetch=0, flat_alpha=0)),          # tiled RGB8
    (3, 6, 0, (97, 96), dict(prefetch=1, flat_alpha=1)),          # ragged RGB8: four-channel path
    (2, 4, astc.FLAG_NORMAL, (256, 256), dict(prefetch=1, flat_alpha=0)),
    (4, 6, astc.FLAG_ALPHA | astc.FLAG_MAX_ACCUMULATION, (128, 128), dict(prefetch=1, flat_alpha=0)),
])
def test_default_tuning_is_reported_and_matches_astc_encode(enc, channels, dim, flags, shape, want):
    rng = np.random.default_rng(channels * 10 + dim)
    img = rng.integers(0, 256, size=(shape[1], shape[0], channels), dtype=np.uint8)
    tex = torch.from_numpy(img).cuda()
    t = enc.default_plain_tuning(tex, dim, flags)
    assert t["lanes_per_block"] == 1 and t["packed_texels"] == 0
    for k, v in want.items():
        assert t[k] == v, (k, t)
    default = enc.encode(tex, block_dim=dim, flags=flags)
    tuned = enc.encode(tex, block_dim=dim, flags=flags,
                       tuning=(t["lanes_per_block"], t["blocks_per_lane"], t["prefetch"], t["flat_alpha"], 0))
    torch.cuda.synchronize()
    assert torch.equal(default, tuned)
