"""The batched and serving workloads at their full stated sizes.

Mixed set: 512 textures drawn with seed 0 (side 2^k, k uniform in 6..12,
square or 2:1; 60% sinusoid RGB8 4x4, 25% noise RGBA8 sRGB 6x6, 15% normal
RG8 4x4), each with its full mip chain built on the device: 4,124 images,
1.67 G texels, 94.1 M blocks in one persistent-kernel graph.  Every image's
blocks must equal its plain launch (astc_encode_each) bit for bit, under the
autotuned tuning and under the other kernel form, and the largest image of each
mode plus a seeded sample of the others must equal the CPU oracle.

Serving: a Poisson stream at 50,000 requests/s into a 64 GB ring, each request
one full-size mip-chained texture of the same distribution (a pool of 64 drawn
with seed 1).  Every request's blocks must equal its plain encodes, served alone
and in the mix.
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


def full_set(enc, seed, n):
    """[(levels, mode, dim, flags)] for n textures of the mixed distribution."""
    out = []
    for mode, w, h, s in synth.draw_texture_specs(seed, n):
        gen, dim, flags, _ = synth.MODES[mode]
        levels = enc.mip_chain(torch.from_numpy(gen(s, w, h)).cuda())
        out.append((levels, mode, dim, flags))
    torch.cuda.synchronize()
    return out


def test_mixed_set_full_size_batched_equals_plain_and_oracle(enc):
    textures = full_set(enc, 0, 512)
    images = [(lvl, dim, flags) for levels, _, dim, flags in textures for lvl in levels]
    modes = [mode for levels, mode, _, _ in textures for _ in levels]
    b = astc.Batch(enc, images)
    info = b.info()
    print(f"mixed set: {len(images)} images, {info['texels']} texels, {info['blocks']} blocks", flush=True)
    assert len(images) == 4124 and info["texels"] == 1672631488 and info["blocks"] == 94105304
    tuned = b.autotune(reps=2)
    print(f"autotuned: {tuned}", flush=True)
    # the plain form: one launch per image into its own outputs
    plain = [torch.empty_like(o) for o in b.outputs]
    b2 = astc.Batch(enc, images, outputs=plain)
    for o in plain:
        o.fill_(0x5A)
    b2.launch_plain()
    b2.close()
    other = dict(tuned)
    other["split_modes"] = 1 - tuned["split_modes"]
    other["lanes_per_block"] = 3 - tuned["lanes_per_block"]
    for tuning in (tuned, other):
        b.set_tuning(tuning["blocks_per_lane"], tuning["workgroups_per_cu"], tuning["prefetch"],
                     tuning["split_modes"], tuning["lanes_per_block"])
        for o in b.outputs:
            o.fill_(0xA5)
        b.launch()
        torch.cuda.synchronize()
        bad = [i for i, (o, p) in enumerate(zip(b.outputs, plain)) if not torch.equal(o, p)]
        assert not bad, f"tuning {tuning}: images {bad[:10]} differ from their plain launches"
    # the oracle: the largest image of each mode, and a seeded sample of the rest
    largest = {}
    for i, (t, _, _) in enumerate(images):
        if modes[i] not in largest or t.numel() > images[largest[modes[i]]][0].numel():
            largest[modes[i]] = i
    assert sorted(largest) == sorted(synth.MODES)
    sample = set(largest.values()) | set(np.random.default_rng(3).choice(len(images), 40, replace=False).tolist())
    for i in sorted(sample):
        t, dim, flags = images[i]
        want = orc.oracle_encode(t.cpu().numpy(), dim, flags, threads=16)
        np.testing.assert_array_equal(b.outputs[i].cpu().numpy(), want, err_msg=f"image {i} ({modes[i]})")
    b.close()


def test_serving_full_size_at_50k_requests_per_second(enc):
    pool = []
    for levels, mode, dim, flags in full_set(enc, 1, 64):
        packed = astc.PackedImages(levels, dim, flags)
        want = torch.cat([enc.encode(l, block_dim=dim, flags=flags).view(-1) for l in levels])
        pool.append((packed, want))
        del levels
    torch.cuda.synchronize()
    assert all(p.out_bytes == w.numel() for p, w in pool)
    srv = astc.Server(enc, 64 << 30, max_inflight=3, max_batch_requests=256)
    # alone: every pool texture as a one-request trace
    for k, (p, want) in enumerate(pool):
        dst = torch.full((p.out_bytes,), 0xA5, dtype=torch.uint8, device="cuda")
        lat, st = srv.run_trace([(p, 0.0, dst)])
        torch.cuda.synchronize()
        assert st["batches"] == 1 and torch.equal(dst, want), f"pool texture {k} alone"
    # in the mix: Poisson arrivals at 50k/s
    n = 2500
    rng = np.random.default_rng(50)
    t = np.cumsum(rng.exponential(1.0 / 50000, n))
    picks = rng.integers(0, len(pool), n)
    reqs = [(pool[k][0], float(a), torch.full((pool[k][0].out_bytes,), 0xA5, dtype=torch.uint8, device="cuda"))
            for k, a in zip(picks, t)]
    print(f"serving: {n} requests, {sum(pool[k][0].texels for k in picks)} texels", flush=True)
    lat, st = srv.run_trace(reqs)
    torch.cuda.synchronize()
    print(f"serving stats: {st}", flush=True)
    assert st["requests"] == n and np.all(lat > 0)
    assert st["mean_batch_requests"] > 1.0 and st["peak_inflight"] >= 2
    assert st["peak_ring_bytes"] <= 64 << 30
    bad = [i for i, (k, (_, _, d)) in enumerate(zip(picks, reqs)) if not torch.equal(d, pool[k][1])]
    assert not bad, f"requests {bad[:10]} differ from their plain encodes"
    srv.close()
