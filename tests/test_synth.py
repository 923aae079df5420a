"""The seeded workload generators: deterministic bytes and the stated statistics."""
import hashlib

import numpy as np

from astc_encoder_amd import synth


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def test_deterministic_bytes():
    # identical on every x86-64 host: integer PRNG + IEEE double arithmetic only
    assert digest(synth.sinusoid_rgb8(0, 256, 128)) == digest(synth.sinusoid_rgb8(0, 256, 128))
    assert digest(synth.sinusoid_rgb8(0, 1024, 1024)) == "2b3b9abf4ec45239"
    assert digest(synth.normal_rg8(0, 2048, 2048)) == "0ad822ba26ae106a"
    assert digest(synth.noise_rgba8(0, 512, 512)) != digest(synth.noise_rgba8(1, 512, 512))


def test_noise_rgba8_alpha_statistics():
    img = synth.noise_rgba8(0, 1024, 1024)
    a = img[..., 3]
    assert abs((a == 0).mean() - 0.30) < 0.002
    assert abs((a == 255).mean() - 0.20) < 0.002


def test_normal_rg8_is_unit_normal_encoding():
    img = synth.normal_rg8(0, 256, 256).astype(np.float64)
    n = img / 255.0 * 2.0 - 1.0
    assert np.all(n[..., 0] ** 2 + n[..., 1] ** 2 <= 1.0 + 0.02)


def test_mip_chain_dims_and_box_filter():
    img = synth.sinusoid_rgb8(3, 64, 32)
    chain = synth.mip_chain(img)
    dims = [c.shape[:2] for c in chain]
    assert dims == [(32, 64), (16, 32), (8, 16), (4, 8), (4, 4)]
    a = chain[0].astype(int)
    expect = (a[0::2, 0::2] + a[0::2, 1::2] + a[1::2, 0::2] + a[1::2, 1::2] + 2) >> 2
    np.testing.assert_array_equal(chain[1], expect)
    # 8x4 -> 4x4 halves only the x axis
    b = chain[3].astype(int)
    np.testing.assert_array_equal(chain[4], (b[:, 0::2] + b[:, 1::2] + 1) >> 1)


def test_texture_set_distribution():
    specs = synth.draw_texture_specs(0, 512)
    modes = [s[0] for s in specs]
    assert 0.5 < modes.count("rgb4x4") / 512 < 0.7
    assert all(64 <= max(s[1], s[2]) <= 4096 for s in specs)
    assert specs == synth.draw_texture_specs(0, 512)
