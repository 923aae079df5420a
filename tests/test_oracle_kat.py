"""Known-answer tests of the oracle's floating-point stage.

Each case has an answer fixed by the mathematics of ASTC_Encode.hlsl:93-168,
316-382 rather than by re-running the same formula:
  constant block      -> e0 = e1 = the colour, every weight 0
  two-colour block    -> the axis is the colour difference, endpoints are the
                         two colours, weights are 0 / max by membership
  linear ramp         -> endpoints at the ramp's ends, weights round(range*i/15)
  alpha-only variance -> the start vector has no alpha component, so M v = 0,
                         the iteration returns the zero axis, e0 = e1 = mean
  darker first        -> the swap orders endpoints by rounded RGB sum
  RGB mode            -> endpoint alpha forced to 255, 6 endpoint bytes
  out-of-bounds       -> padded texels fetch as 0, normal maps force b = a = 255
plus FP64-twin and schedule-sensitivity statements on random blocks.
"""
import numpy as np
import pytest

from tests import oracle_ref as orc

RANGE = {True: 5, False: 11}


def decode_fields(blk, dim):
    ok, rgba, info = orc.decode_block(blk, dim)
    assert ok == 1
    return rgba, info


@pytest.mark.parametrize("dim", [4, 6])
@pytest.mark.parametrize("alpha", [True, False])
@pytest.mark.parametrize("sched", [orc.SCHED_CANONICAL, orc.SCHED_SOURCE])
def test_constant_block(dim, alpha, sched):
    col = np.array([37.0, 141.0, 200.0, 99.0 if alpha else 255.0], np.float32)
    tex = np.tile(col, (dim * dim, 1))
    blk, tr = orc.encode_block(tex, dim, alpha, sched)
    assert list(tr.w_q) == [0] * 16
    assert list(tr.ep_q[:6]) == [37, 37, 141, 141, 200, 200]
    rgba, info = decode_fields(blk, dim)
    assert np.all(rgba[..., :3] == col[:3].astype(np.uint8))
    assert info.cem == (12 if alpha else 8)


@pytest.mark.parametrize("dim", [4, 6])
@pytest.mark.parametrize("alpha", [True, False])
def test_two_colour_block(dim, alpha):
    a = np.array([10.0, 20.0, 30.0, 255.0], np.float32)
    b = np.array([210.0, 120.0, 60.0, 255.0], np.float32)
    n = dim * dim
    member = (np.arange(n) % 3 == 0)
    tex = np.where(member[:, None], b, a).astype(np.float32)
    blk, tr = orc.encode_block(tex, dim, alpha)
    ax = np.array(tr.axis[:3])
    d = (b - a)[:3]
    cosang = abs(ax @ d) / (np.linalg.norm(ax) * np.linalg.norm(d))
    assert cosang > 1 - 1e-6
    assert list(tr.ep_q[:6]) == [10, 210, 20, 120, 30, 60]
    if dim == 4:
        expect = np.where(member, RANGE[alpha], 0)
        assert list(tr.w_q) == list(expect)


@pytest.mark.parametrize("alpha", [True, False])
def test_linear_ramp_4x4(alpha):
    i = np.arange(16, dtype=np.float32)
    tex = np.stack([i * 10 + 20, i * 5 + 40, 200 - i * 8, np.full(16, 255.0, np.float32)], 1).astype(np.float32)
    blk, tr = orc.encode_block(tex, 4, alpha)
    assert list(tr.ep_q[:6]) == [20, 170, 40, 115, 200, 80]
    r = RANGE[alpha]
    # exact ties at x.5 round to even (HLSL round)
    assert list(tr.w_q) == [int(np.rint(r * k / 15.0)) for k in range(16)]


@pytest.mark.parametrize("dim", [4, 6])
def test_alpha_only_variance_collapses_to_mean(dim):
    n = dim * dim
    tex = np.zeros((n, 4), np.float32)
    tex[:, :3] = [90.0, 60.0, 30.0]
    tex[:, 3] = np.where(np.arange(n) % 2 == 0, 0.0, 255.0)
    blk, tr = orc.encode_block(tex, dim, True)
    assert np.allclose(tr.axis, 0.0)
    mean_a = float(np.float32(tex[:, 3].sum()) * np.float32(1.0 / n))
    assert list(tr.ep_q) == [90, 90, 60, 60, 30, 30, int(np.rint(mean_a)), int(np.rint(mean_a))]
    assert list(tr.w_q) == [0] * 16


def test_darker_endpoint_first():
    i = np.arange(16, dtype=np.float32)
    tex = np.stack([250 - i * 10, 250 - i * 10, 250 - i * 10, np.full(16, 255.0, np.float32)], 1)
    _, tr = orc.encode_block(tex.astype(np.float32), 4, False)
    e0, e1 = tr.ep_q[0::2], tr.ep_q[1::2]
    assert sum(e0[:3]) <= sum(e1[:3])
    assert list(tr.ep_q[:2]) == [100, 250]


def test_rgb_mode_forces_endpoint_alpha_and_ignores_it_in_weights():
    rng = np.random.default_rng(3)
    tex = rng.integers(0, 256, size=(16, 4)).astype(np.float32)
    blk, tr = orc.encode_block(tex, 4, False)
    assert tr.e0[3] == 255.0 and tr.e1[3] == 255.0
    _, info = decode_fields(blk, 4)
    assert info.cem == 8
    rgba, _ = decode_fields(blk, 4)
    assert np.all(rgba[..., 3] == 255)


def test_fetch_out_of_bounds_and_normal_override():
    img = np.full((5, 7, 4), 77, np.uint8)
    t = orc.fetch_block(img, 4, 0, 1, 1)          # covers x 4..7, y 4..7 of a 7x5 image
    inside = np.array([(x < 7 and y < 5) for y in range(4, 8) for x in range(4, 8)])
    assert np.all(t[inside] == 77.0) and np.all(t[~inside] == 0.0)
    tn = orc.fetch_block(img, 4, orc_flags("normal"), 1, 1)
    assert np.all(tn[:, 2:] == 255.0) and np.all(tn[~inside, :2] == 0.0)
    # narrower formats: RGB8 -> alpha 255, RG8 -> blue 0 alpha 255
    t3 = orc.fetch_block(np.full((4, 4, 3), 9, np.uint8), 4, 0, 0, 0)
    assert np.all(t3 == [9, 9, 9, 255])
    t2 = orc.fetch_block(np.full((4, 4, 2), 9, np.uint8), 4, 0, 0, 0)
    assert np.all(t2 == [9, 9, 0, 255])


def orc_flags(name):
    return {"alpha": 1, "normal": 2, "srgb": 4}[name]


def test_srgb_table_endpoints_and_monotone():
    lut = orc.srgb_table()
    assert lut[0] == 0.0 and lut[255] == np.float32(255.0)
    assert np.all(np.diff(lut) > 0)
    c = np.arange(256) / 255.0
    lin = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4) * 255.0
    assert np.max(np.abs(lut - lin)) < 1e-4
    # a correctly rounded c / 255 would give every byte back exactly after x255
    u = (np.arange(256, dtype=np.float32) / np.float32(255.0)).astype(np.float32) * np.float32(255.0)
    assert np.array_equal(u, np.arange(256, dtype=np.float32))


def test_unorm_fetch_models_the_hardware_conversion():
    """UNORM8 texels convert as binary32 c * fl(1/255), then the shader's x255:
    within one ulp of c, exact for 0 and 255 (so constant-255 channels stay exact)."""
    f = np.float32
    c = np.arange(256, dtype=np.float32)
    want = ((c * (f(1) / f(255))).astype(f) * f(255)).astype(f)
    got = orc.unorm_table()
    np.testing.assert_array_equal(got, want)
    assert got[0] == 0 and got[255] == 255
    off = got != c
    assert off.sum() == 126
    assert np.all(np.abs(got[off] - c[off]) <= np.spacing(c[off]))
    # the fetch uses it for UNORM channels, the sRGB table for sRGB colour
    img = np.arange(64, dtype=np.uint8).reshape(4, 4, 4) * 3
    t = orc.fetch_block(img, 4, 0, 0, 0)
    np.testing.assert_array_equal(t, got[img.reshape(16, 4)])
    ts = orc.fetch_block(img, 4, 4, 0, 0)  # sRGB: colour by table, alpha UNORM
    np.testing.assert_array_equal(ts[:, :3], orc.srgb_table()[img.reshape(16, 4)[:, :3]])
    np.testing.assert_array_equal(ts[:, 3], got[img.reshape(16, 4)[:, 3]])


@pytest.mark.parametrize("dim,alpha", [(4, False), (4, True), (6, True), (6, False)])
def test_fp64_twin_bounds_fp32_error(dim, alpha):
    """The binary32 run stays close to the binary64 run of the same method."""
    rng = np.random.default_rng(11)
    n = dim * dim
    worst_axis = worst_ep = 0.0
    same = 0
    trials = 300
    for _ in range(trials):
        base = rng.uniform(0, 255, 4)
        tex = np.clip(base + rng.normal(0, 25, (n, 4)), 0, 255).round()
        b32, t32 = orc.encode_block(tex.astype(np.float32), dim, alpha)
        b64, t64 = orc.encode_block(tex.astype(np.float64), dim, alpha, f64=True)
        a32, a64 = np.array(t32.axis), np.array(t64.axis)
        if np.linalg.norm(a64) > 0:
            worst_axis = max(worst_axis, min(np.linalg.norm(a32 - a64), np.linalg.norm(a32 + a64)))
        worst_ep = max(worst_ep, np.max(np.abs(np.array(t32.e0) - np.array(t64.e0))),
                       np.max(np.abs(np.array(t32.e1) - np.array(t64.e1))))
        same += np.array_equal(b32, b64)
    assert worst_axis < 3e-4  # power-iteration axis; ill-conditioned when two eigenvalues are close
    assert worst_ep < 2e-2  # in 8-bit units: far below the 0.5 that flips a rounded endpoint
    assert same >= 0.95 * trials


@pytest.mark.parametrize("dim,alpha", [(4, False), (6, True)])
def test_canonical_and_source_schedules_agree_on_most_blocks(dim, alpha):
    rng = np.random.default_rng(5)
    n = dim * dim
    same = 0
    for _ in range(500):
        tex = np.clip(rng.uniform(0, 255, 4) + rng.normal(0, 20, (n, 4)), 0, 255).round().astype(np.float32)
        a, _ = orc.encode_block(tex, dim, alpha, orc.SCHED_CANONICAL)
        b, _ = orc.encode_block(tex, dim, alpha, orc.SCHED_SOURCE)
        same += np.array_equal(a, b)
    assert same >= 0.95 * 500


def _blocks(rng, dim, count):
    """Random, two-colour, constant and alpha-only blocks (texel values 0..255)."""
    n = dim * dim
    out = []
    for i in range(count):
        kind = i % 4
        if kind == 0:
            tex = np.clip(rng.uniform(0, 255, 4) + rng.normal(0, 30, (n, 4)), 0, 255).round()
        elif kind == 1:
            a, b = rng.integers(0, 256, 4), rng.integers(0, 256, 4)
            tex = np.where(rng.random(n)[:, None] < 0.5, a, b).astype(float)
        elif kind == 2:
            tex = np.tile(rng.integers(0, 256, 4).astype(float), (n, 1))
        else:
            tex = np.tile(rng.integers(0, 256, 4).astype(float), (n, 1))
            tex[:, 3] = rng.integers(0, 256, n)
        out.append(tex)
    return out


@pytest.mark.parametrize("dim,alpha", [(4, False), (4, True), (6, True), (6, False)])
def test_canonical_rearrangements_equal_the_shader_formulas_in_binary64(dim, alpha):
    """The canonical schedule computes the principal axis from S = M M with one
    length test, min/max as dot(texel, v) - dot(mean, v), and weights from texel
    projections.  In binary64 these must give the shader's literal formulas'
    axis, endpoints and weights (up to rounding): a dropped term or a wrong
    index in one of the rearranged formulas shows up here as a large difference.
    Both schedules read the same 6x6 tap table, so a wrong tap cannot show up
    here; test_oracle_grid6.py pins the table against the reference's."""
    rng = np.random.default_rng(21)
    worst_axis = worst_ep = 0.0
    diff_random = 0
    blocks = _blocks(rng, dim, 400)
    for i, tex in enumerate(blocks):
        bc, tc = orc.encode_block(tex, dim, alpha, orc.SCHED_CANONICAL, f64=True)
        bs, ts = orc.encode_block(tex, dim, alpha, orc.SCHED_SOURCE, f64=True)
        ac, as_ = np.array(tc.axis), np.array(ts.axis)
        worst_axis = max(worst_axis, np.max(np.abs(ac - as_)))
        worst_ep = max(worst_ep, np.max(np.abs(np.array(tc.e0) - np.array(ts.e0))),
                       np.max(np.abs(np.array(tc.e1) - np.array(ts.e1))))
        if not np.array_equal(bc, bs):
            diff_random += i % 4 == 0
            # only a weight on a rounding tie may differ, by one step; two-colour
            # blocks put 6x6 grid weights exactly on .5 steps (tap sums such as
            # 0.444 / 0.888), where the last bit decides
            assert list(tc.ep_q) == list(ts.ep_q)
            assert np.max(np.abs(np.array(tc.w_q, int) - np.array(ts.w_q, int))) <= 1
    assert worst_axis < 1e-9
    assert worst_ep < 1e-9
    assert diff_random <= 0.02 * len(blocks) / 4


def test_power_iteration_length_test_only_fires_on_the_first_step():
    """|M v|^2 = v.(M M)v is the Rayleigh quotient of the PSD matrix M M at the
    unit iterate v; along v <- normalize(M M v) it never decreases."""
    rng = np.random.default_rng(3)
    v0 = np.array([0.26726, 0.80178, 0.53452, 0.0])
    for _ in range(2000):
        x = rng.normal(0, rng.uniform(0.01, 50), (16, 4)) * (rng.random(4) < 0.8)
        d = x - x.mean(0)
        m = d.T @ d / 15
        v = v0 / np.linalg.norm(v0)
        q = [v @ m @ m @ v]
        for _ in range(8):
            w = m @ (m @ v)
            if np.linalg.norm(w) == 0:
                break
            v = w / np.linalg.norm(w)
            q.append(v @ m @ m @ v)
        q = np.array(q)
        assert np.all(np.diff(q) >= -1e-9 * max(1.0, q.max()))


def test_weight_clamp_never_acts():
    """(w - min) * fl(1 / max(1e-5, max - min)) is at most 1 + 2^-22 in binary32,
    so round(w * range) never exceeds range (the kernel omits the clamp)."""
    rng = np.random.default_rng(8)
    f = np.float32
    for rng_scale in (1e-6, 1e-3, 1.0, 300.0):
        p = (rng.random((20000, 16)) * rng_scale).astype(f)
        p[:, 0] = p.min(1)  # duplicate extremes
        lo, hi = p.min(1, keepdims=True), p.max(1, keepdims=True)
        inv = (f(1) / np.maximum(f(1e-5), hi - lo)).astype(f)
        wn = ((p - lo).astype(f) * inv).astype(f)
        for r in (5, 11):
            assert np.rint((wn * f(r)).astype(f)).max() <= r
