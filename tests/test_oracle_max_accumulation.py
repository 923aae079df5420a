"""The optional max-accumulation endpoint search (ASTC_Encode.hlsl:170-227).

The reference leaves the call commented out (:514), so no reference output
exists for it; these checks pin the oracle's restatement against the formula
recomputed independently in numpy (binary64) and against closed forms.
"""
import numpy as np
import pytest

from tests import oracle_ref as orc

MA = orc.METHOD_MAX_ACCUMULATION


def numpy_axis(tex, has_alpha):
    """max_accumulation_pixel_direction's vec_k, from its definition."""
    mean = tex.mean(0)
    dt = tex - mean
    sums = [dt[dt[:, c] > 0].sum(0) if np.any(dt[:, c] > 0) else np.zeros(4) for c in range(4)]
    dots = [s @ s for s in sums]
    best = 0
    for c in range(1, 4 if has_alpha else 3):
        if dots[c] > dots[best]:
            best = c
    v = sums[best]
    n = np.linalg.norm(v)
    return v if n < 1e-5 else v / n


@pytest.mark.parametrize("dim", [4, 6])
@pytest.mark.parametrize("alpha", [True, False])
def test_axis_matches_the_formula(dim, alpha):
    rng = np.random.default_rng(dim * 10 + alpha)
    for _ in range(200):
        tex = np.clip(rng.uniform(0, 255, 4) + rng.normal(0, 30, (dim * dim, 4)) * (rng.random(4) < 0.8), 0, 255).round()
        _, tr = orc.encode_block(tex, dim, alpha, orc.SCHED_SOURCE, f64=True, method=MA)
        np.testing.assert_allclose(np.array(tr.axis), numpy_axis(tex, alpha), atol=1e-12)
        _, tc = orc.encode_block(tex, dim, alpha, orc.SCHED_CANONICAL, f64=True, method=MA)
        np.testing.assert_allclose(np.array(tc.axis), numpy_axis(tex, alpha), atol=1e-12)


@pytest.mark.parametrize("dim", [4, 6])
def test_two_colour_block_gives_the_two_colours(dim):
    a, b = np.array([20.0, 200.0, 90.0, 255.0]), np.array([180.0, 40.0, 130.0, 255.0])
    n = dim * dim
    tex = np.where((np.arange(n) % 3 == 0)[:, None], b, a)
    blk, tr = orc.encode_block(tex.astype(np.float32), dim, False, method=MA)
    ends = sorted([tuple(np.round(tr.e0[:3])), tuple(np.round(tr.e1[:3]))])
    assert ends == sorted([tuple(a[:3]), tuple(b[:3])])
    ok, dec, _ = orc.decode_block(blk, dim)
    assert ok
    if dim == 4:  # one weight per texel: the block reproduces the two colours
        np.testing.assert_array_equal(dec.reshape(-1, 4)[:, :3], tex[:, :3].astype(np.uint8))


def test_constant_block_has_zero_axis_and_weights():
    tex = np.tile([12.0, 34.0, 56.0, 78.0], (16, 1))
    _, tr = orc.encode_block(tex.astype(np.float32), 4, True, method=MA)
    assert np.all(np.array(tr.axis) == 0)
    assert list(tr.ep_q) == [12, 12, 34, 34, 56, 56, 78, 78]
    assert list(tr.w_q) == [0] * 16


def test_alpha_competes_only_in_alpha_mode():
    rng = np.random.default_rng(4)
    tex = np.tile([100.0, 100.0, 100.0, 0.0], (16, 1))
    tex[:, 0] += rng.integers(-3, 4, 16)          # small red variation
    tex[:, 3] = np.where(np.arange(16) < 8, 0.0, 255.0)  # large alpha variation
    _, ta = orc.encode_block(tex, 4, True, orc.SCHED_SOURCE, f64=True, method=MA)
    _, tn = orc.encode_block(tex, 4, False, orc.SCHED_SOURCE, f64=True, method=MA)
    dt = tex - tex.mean(0)
    sum_a, sum_r = dt[dt[:, 3] > 0].sum(0), dt[dt[:, 0] > 0].sum(0)
    assert sum_a @ sum_a > sum_r @ sum_r
    np.testing.assert_allclose(np.array(ta.axis), sum_a / np.linalg.norm(sum_a), atol=1e-12)  # alpha's sum wins
    np.testing.assert_allclose(np.array(tn.axis), sum_r / np.linalg.norm(sum_r), atol=1e-12)  # RGB mode: red's


def test_image_flag_selects_the_search():
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (24, 20, 4), dtype=np.uint8)
    pca = orc.oracle_encode(img, 4, 1)
    ma = orc.oracle_encode(img, 4, 1 | orc.FLAG_MAX_ACCUMULATION)
    assert not np.array_equal(pca, ma)
    for i, (by, bx) in enumerate([(0, 0), (2, 3), (5, 4)]):
        tex = orc.fetch_block(img, 4, 1, bx, by)
        blk, _ = orc.encode_block(tex, 4, True, method=MA)
        np.testing.assert_array_equal(ma[by * 5 + bx], blk)
