"""Benchmark of the MI355X ASTC encoder; prints one JSON line.

Default workload (one step = one plain per-texture launch over a resident
texture): 4096x4096 RGBA8 sRGB, ASTC 6x6 with alpha, seeded value-noise content
with rectangles and a cut-out alpha (466,489 blocks).  The line carries:
  value           Mtexels/s over real texels, inputs already in HBM
  roofline_hbm    algorithmic HBM bytes per launch / average kernel duration
                  (HIP events on the launch stream) against 8 TB/s; `traffic`
                  is the measured HBM bytes per launch from rocprofv3 counters
                  (profiles/hbm_traffic_encode_plain_6x6.txt) when present
  roofline_valu   the same launch priced in canonical FP32 operations against
                  MI355X's 157.3 TFLOP/s vector peak
  roofline        whichever of the two the kernel comes closest to
  fp64_twin       binary32 vs binary64 runs of the same method on the three
                  plain workloads (axis and endpoint error)
  cpu_baseline    the CPU oracle (scalar restatement of the reference encoder,
                  OpenMP over blocks) on this host's cores, on the same texture
  extras          the other plain workloads (and the batched / serving forms
                  when built)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mtexels/s 4x4/6x6, batched vs plain at ≤0.05 dB PSNR Δ; p50/p99 req latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

# Canonical FP32 operation counts per block (FMA = 2; add/sub/mul/min/max/round/
# sqrt/div = 1; loads, conversions and integer packing = 0), per mode:
# mean C(N+1), covariance N(C + C(C+1)) + C(C+1)/2, power iteration
# 8(2C(2C-1) + 2(2C-1) + 2 + C), min/max N(3C+1) + 8C + 11, endpoints 2C,
# axis 6C, projections N(3C+1) (4x4) or G(10C+1) (6x6), renormalise 3 + 2G,
# quantise 2G, with C active channels, N texels and G = 16 grid points.
# server tuning for the saturation runs: (blocks_per_lane, workgroups_per_cu, prefetch, split_modes)
SATURATION_TUNING = (8, 1, 0, 1)
FLOPS_PER_BLOCK = {"rgb4x4": 1103, "rgba6x6": 2896, "normal4x4": 675, "rgba4x4": 1628, "rgb6x6": 1999}


def workloads():
    import astc_encoder_amd as astc
    from astc_encoder_amd import synth
    return {
        "noise_rgba8_srgb_6x6": dict(name="noise RGBA8 sRGB 4096x4096 ASTC 6x6", gen=synth.noise_rgba8, w=4096,
                                     h=4096, dim=6, flags=astc.FLAG_ALPHA | astc.FLAG_SRGB, bpp=4,
                                     flops_key="rgba6x6"),
        "sinusoid_rgb8_4x4": dict(name="sinusoid RGB8 1024x1024 ASTC 4x4", gen=synth.sinusoid_rgb8, w=1024, h=1024,
                                  dim=4, flags=0, bpp=3, flops_key="rgb4x4"),
        "normal_rg8_4x4": dict(name="normal RG8 2048x2048 ASTC 4x4", gen=synth.normal_rg8, w=2048, h=2048, dim=4,
                               flags=astc.FLAG_NORMAL, bpp=2, flops_key="normal4x4"),
    }


def read_traffic(path):
    """HBM bytes per launch measured by rocprofv3 counters (see the file)."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
    return float(m.group(1)) if m else None


def time_plain(enc, wl, steps, warmup, tuning=None):
    import ctypes

    import torch
    import astc_encoder_amd as astc
    img = wl["gen"](0, wl["w"], wl["h"])
    src = torch.from_numpy(img).cuda()
    out = enc.encode(src, block_dim=wl["dim"], flags=wl["flags"])
    stream = torch.cuda.current_stream()
    # one launch per step through the C ABI with everything precomputed
    desc = astc.make_desc(wl["w"], wl["h"], {4: astc.FMT_RGBA8, 3: astc.FMT_RGB8, 2: astc.FMT_RG8}[img.shape[2]],
                          wl["dim"], wl["flags"])
    ctx, dref = enc.handle, ctypes.byref(desc)
    sp, dp, st = ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream.cuda_stream)
    from astc_encoder_amd._capi import PlainTuning
    if tuning is None:
        fn, args = astc.lib().astc_encode, (ctx, dref, sp, dp, st)
        used = PlainTuning()
        if astc.lib().astc_plain_default_tuning(ctx, dref, ctypes.byref(used)) != 0:
            raise astc.AstcError(-2, astc.lib().astc_last_error().decode())
        tuning_used = (used.lanes_per_block, used.blocks_per_lane, used.prefetch, used.flat_alpha, used.packed_texels)
    else:
        tuning_used = tuple(tuning) + (0,) * (5 - len(tuning))
        t = PlainTuning(*tuning_used)
        fn, args = astc.lib().astc_encode_tuned, (ctx, dref, sp, dp, ctypes.byref(t), st)

    def launch():
        if fn(*args) != 0:
            raise astc.AstcError(-2, astc.lib().astc_last_error().decode())

    for _ in range(warmup):
        launch()
    # timed region: K launches back to back, as a caller encoding textures one
    # after another issues them; an event pair on the launch stream brackets it
    b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b0.record(stream)
    for _ in range(steps):
        launch()
    b1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_event_ms = b0.elapsed_time(b1) / steps
    # kernel duration: the same launches, each between its own event pair (what a
    # kernel trace reports per dispatch; the pair adds the dispatch's start-up and
    # completion, so this is longer than a back-to-back step)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    for i in range(steps):
        starts[i].record(stream)
        launch()
        ends[i].record(stream)
    torch.cuda.synchronize()
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    texels = wl["w"] * wl["h"]
    nblocks = out.shape[0]
    return dict(img=img, blocks=out.cpu().numpy(), wall_s=wall, kernel_ms=kernel_ms, step_event_ms=step_event_ms,
                texels=texels, tuning=tuning_used,
                nblocks=nblocks, mtexels_s=texels * steps / wall / 1e6, psnr_db=gpu_psnr(enc, wl, src, out))


def opaque_autotuned(enc, steps, warmup):
    """The headline's colours with an opaque alpha (every byte 255, what the
    command line's RGBA8 load of an opaque image gives), RGBA mode: plain
    launches at the default tuning and at the plain autotuner's choice."""
    import torch
    import astc_encoder_amd as astc
    from astc_encoder_amd import synth

    def opaque(seed, w, h):
        img = synth.noise_rgba8(seed, w, h)
        img[..., 3] = 255
        return img

    wl = dict(name="opaque RGBA8 sRGB 4096x4096 ASTC 6x6", gen=opaque, w=4096, h=4096, dim=6,
              flags=astc.FLAG_ALPHA | astc.FLAG_SRGB)
    best = enc.autotune_plain(torch.from_numpy(opaque(0, 4096, 4096)).cuda(), 6, wl["flags"], reps=20)
    default = time_plain(enc, wl, steps, warmup)
    tuned = time_plain(enc, wl, steps, warmup,
                       (best["lanes_per_block"], best["blocks_per_lane"], best["prefetch"], best["flat_alpha"]))
    same = bool(np.array_equal(default["blocks"], tuned["blocks"]))
    return {"workload": wl["name"], "default_mtexels_s": round(default["mtexels_s"], 1),
            "default_kernel_ms": round(default["kernel_ms"], 4), "autotuned": best,
            "autotuned_mtexels_s": round(tuned["mtexels_s"], 1), "autotuned_kernel_ms": round(tuned["kernel_ms"], 4),
            "same_blocks": same}


def time_host_resident(enc, wl, steps):
    """The headline texture starting in (pinned) host memory: host-to-device copy,
    encode, blocks back to (pinned) host, per step, on one stream."""
    import torch
    img = wl["gen"](0, wl["w"], wl["h"])
    host_src = torch.from_numpy(img).pin_memory()
    dev_src = torch.empty(host_src.shape, dtype=torch.uint8, device="cuda")
    out = enc.encode(dev_src.copy_(host_src), block_dim=wl["dim"], flags=wl["flags"])
    host_out = torch.empty(out.shape, dtype=torch.uint8).pin_memory()
    stream = torch.cuda.current_stream()

    def step():
        dev_src.copy_(host_src, non_blocking=True)
        enc.encode(dev_src, block_dim=wl["dim"], flags=wl["flags"], out=out)
        host_out.copy_(out, non_blocking=True)

    for _ in range(3):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return {"workload": wl["name"] + ", source in pinned host memory, blocks back to pinned host",
            "ms_per_step": round(ms, 4), "mtexels_s": round(wl["w"] * wl["h"] / (ms * 1e-3) / 1e6, 1),
            "h2d_bytes": int(img.nbytes), "d2h_bytes": int(out.numel())}


def gpu_psnr(enc, wl, src, blocks):
    """PSNR of the encoded texture, decoded by the library's GPU decoder, against the
    encoder's view of the source (sRGB linearised; the mode's channels: RG for
    normal maps, RGB, or RGBA)."""
    import torch
    import astc_encoder_amd as astc
    dec = enc.decode(blocks, wl["w"], wl["h"], wl["dim"])
    nch = 2 if wl["flags"] & astc.FLAG_NORMAL else 4 if wl["flags"] & astc.FLAG_ALPHA else 3
    ref = src[..., :nch].long()
    if wl["flags"] & astc.FLAG_SRGB:
        lut = torch.from_numpy(astc.srgb_table()).to(src.device).double()
        t = lut[ref]
        if nch == 4:
            t[..., 3] = ref[..., 3].double()
    else:
        t = ref.double()
    e = t - dec[..., :nch].double()
    sse = float((e * e).sum())
    return round(10.0 * np.log10(255.0 * 255.0 * e.numel() / sse), 4) if sse > 0 else float("inf")


def batched_vs_plain(enc, steps, warmup, n_textures=512, tuning=None):
    """512 seeded textures (side 2^k, k in 6..12, square or 2:1; 60% sinusoid RGB8
    4x4 / 25% noise RGBA8 sRGB 6x6 / 15% normal RG8 4x4), each with a full mip
    chain generated on the device: one autotuned persistent-kernel graph replay
    per step vs one plain launch per level (issued from C++)."""
    import torch
    import astc_encoder_amd as astc
    from astc_encoder_amd import synth
    t0 = time.perf_counter()
    images = []
    for mode, w, h, seed in synth.draw_texture_specs(0, n_textures):
        gen, dim, flags, _ = synth.MODES[mode]
        base = torch.from_numpy(gen(seed, w, h)).cuda()
        for lvl in enc.mip_chain(base):
            images.append((lvl, dim, flags))
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    b = astc.Batch(enc, images)
    if tuning is None:
        tuning = b.autotune(reps=10)
    else:
        b.set_tuning(*tuning)
        tuning = b.tuning()
    info = b.info()
    stream = torch.cuda.current_stream()

    def timed(fn, n):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n, e0.elapsed_time(e1) / n / 1e3

    batched_wall, batched_gpu = timed(b.launch, steps)
    snapshot = [o.clone() for o in b.outputs]
    for o in b.outputs:
        o.zero_()
    plain_wall, plain_gpu = timed(b.launch_plain, max(3, steps // 4))
    torch.cuda.synchronize()
    same = all(torch.equal(a, o) for a, o in zip(snapshot, b.outputs))
    b.close()
    texels = info["texels"]
    return {"workload": f"{n_textures} mixed textures with mip chains (seed 0)", "images": len(images),
            "texels": texels, "blocks": info["blocks"], "work_items": info["items"], "tuning": tuning,
            "batched_mtexels_s": round(texels / batched_wall / 1e6, 1),
            "plain_mtexels_s": round(texels / plain_wall / 1e6, 1),
            "batched_ms": round(batched_wall * 1e3, 3), "plain_ms": round(plain_wall * 1e3, 3),
            "batched_vs_plain": round(plain_wall / batched_wall, 2), "bit_exact_batched_vs_plain": bool(same),
            "generation_s": round(gen_s, 1)}


def serving_sweep(enc, rates=(1000, 2000, 5000, 10000, 20000, 50000), seconds=0.5, pool_size=64):
    """Poisson request streams at each rate; every request is one mip-chained
    texture drawn from the mixed distribution (a pool of device-resident sources,
    copied into the server's ring on admission).  64 GB HBM budget."""
    import torch
    import astc_encoder_amd as astc
    from astc_encoder_amd import synth
    pool = []
    for mode, w, h, seed in synth.draw_texture_specs(1, pool_size):
        gen, dim, flags, _ = synth.MODES[mode]
        levels = enc.mip_chain(torch.from_numpy(gen(seed, w, h)).cuda())
        pool.append(astc.PackedImages(levels, dim, flags))
        del levels
    torch.cuda.synchronize()
    srv = astc.Server(enc, 64 << 30, max_inflight=3, max_batch_requests=256)
    rng = np.random.default_rng(0)
    warm = [(pool[i % pool_size], 0.0, None) for i in range(2 * pool_size)]
    srv.run_trace(warm)
    # and a short Poisson stream at the top rate, so that every slot has served
    # batches of every size before anything is timed
    warm_t = np.cumsum(np.random.default_rng(99).exponential(1.0 / rates[-1], 5000))
    srv.run_trace([(pool[i % pool_size], float(a), None) for i, a in enumerate(warm_t)])
    results = []
    for lam in rates:
        n = max(200, int(lam * seconds))
        t = np.cumsum(rng.exponential(1.0 / lam, n))
        picks = rng.integers(0, pool_size, n)
        lat, st = srv.run_trace([(pool[k], float(a), None) for k, a in zip(picks, t)])
        keep = lat[n // 10:]
        results.append({"rate_per_s": lam, "requests": n, "p50_ms": round(float(np.percentile(keep, 50)) * 1e3, 3),
                        "p99_ms": round(float(np.percentile(keep, 99)) * 1e3, 3),
                        "mtexels_s": round(st["texels"] / st["wall_s"] / 1e6, 1),
                        "mean_texels_per_request": int(st["texels"] / n),
                        "batches": st["batches"], "mean_batch_requests": round(st["mean_batch_requests"], 2),
                        "peak_inflight": st["peak_inflight"], "max_loop_gap_ms": round(st["max_loop_gap_s"] * 1e3, 3)})
    srv.close()
    # saturation with the batched form's throughput tuning (8 blocks per lane,
    # one 6x6 workgroup per CU, split kernels), which the low-rate sweep's
    # latency-oriented default trades away
    srv = astc.Server(enc, 64 << 30, max_inflight=3, max_batch_requests=256, tuning=SATURATION_TUNING)
    srv.run_trace([(pool[i % pool_size], 0.0, None) for i in range(4 * pool_size)])
    sat = saturation(srv, pool, rng)
    sat["tuning"] = dict(zip(("blocks_per_lane", "workgroups_per_cu", "prefetch", "split_modes", "lanes_per_block"),
                             SATURATION_TUNING + (1,)))
    srv.close()
    return {"workload": f"Poisson arrivals, {pool_size} mip-chained textures of the mixed distribution (seed 1)",
            "hbm_budget_gb": 64, "max_inflight_batches": 3, "sweep": results, "saturation": sat}


def saturation(srv, pool, rng, n=1500):
    """Sustained throughput with n requests all waiting at once, two ways:
    copy-in (astc_server_run_trace: each request's sources are copied into the
    ring on admission) and zero-copy (online calls: the sources were written
    into reserved ring space beforehand, as a producer writing its texture there
    would, and every request submitted; the timed part is the server draining
    them: batching and encoding, from the first poll to the last completion)."""
    import torch
    picks = rng.integers(0, len(pool), n)
    texels = int(sum(pool[k].texels for k in picks))
    lat, st = srv.run_trace([(pool[k], 0.0, None) for k in picks])
    copy_in = {"requests": n, "mtexels_s": round(texels / st["wall_s"] / 1e6, 1), "wall_ms": round(st["wall_s"] * 1e3, 2),
               "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3), "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
               "mean_batch_requests": round(st["mean_batch_requests"], 1)}
    tickets = []
    for k in picks:
        ticket, ring = srv.reserve(pool[k])
        ring.copy_(pool[k].buffer)
        tickets.append(ticket)
    torch.cuda.synchronize()
    for t in tickets:
        srv.submit(t, None)
    t0 = time.perf_counter()
    done, finish = 0, {}
    while done < n:
        for t in srv.poll():
            finish[t] = time.perf_counter() - t0
            done += 1
    wall = max(finish.values())
    lat = np.array([finish[t] for t in tickets])
    for t in tickets:
        srv.release(t)
    zero_copy = {"requests": n, "mtexels_s": round(texels / wall / 1e6, 1), "wall_ms": round(wall * 1e3, 2),
                 "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                 "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}
    return {"copy_in": copy_in, "zero_copy": zero_copy, "texels": texels}


def _time_runs(fn, min_s, min_runs=5, max_runs=2000):
    """Run fn (after one warm-up) until min_s seconds and min_runs runs have passed;
    return the per-run times."""
    fn()
    times = []
    start = time.perf_counter()
    while len(times) < max_runs and (len(times) < min_runs or time.perf_counter() - start < min_s):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    return times


def cpu_baseline(wl, img, gpu_blocks, min_s=10.0):
    """The oracle (OpenMP over blocks, all cores of this host) on the same texture:
    one warm-up, then whole-texture runs for at least min_s seconds; the median run."""
    from tests import oracle_ref as orc
    cores = len(os.sched_getaffinity(0))
    ref = orc.oracle_encode(img, wl["dim"], wl["flags"], threads=cores)
    times = _time_runs(lambda: orc.oracle_encode(img, wl["dim"], wl["flags"], threads=cores), min_s)
    med = float(np.median(times))
    match = float(np.mean(np.all(ref == gpu_blocks, axis=1)))
    # the oracle's own blocks through the oracle's spec decoder: the reference PSNR
    _, dec = orc.decode_image(ref, wl["w"], wl["h"], wl["dim"])
    return dict(value=round(wl["w"] * wl["h"] / med / 1e6, 2), unit="Mtexels/s", cores=cores, kind="port",
                host_cpu=cpu_model(),
                sample=f"whole texture ({wl['name']}), median of {len(times)} runs over {sum(times):.1f} s "
                       f"after 1 warm-up",
                gpu_block_match=match, oracle_psnr_db=round(orc.psnr(img, wl["flags"], dec), 4))


def fp64_twin(wls, threads):
    """The oracle's binary32 run (the schedule the GPU runs) against a binary64 run
    of the same method, over every block of each plain workload: principal-axis
    error (unit axes, up to sign) and unrounded endpoint error in 8-bit units."""
    from tests import oracle_ref as orc
    out = {}
    for key, wl in wls.items():
        img = wl["gen"](0, wl["w"], wl["h"])
        ax, ep, diff = orc.fp64_twin(img, wl["dim"], wl["flags"], threads=threads)
        a = ax[~np.isnan(ax)]
        out[key] = {"blocks": int(len(ep)), "axis_error_max": float(a.max()) if len(a) else 0.0,
                    "axis_error_p999": float(np.percentile(a, 99.9)) if len(a) else 0.0,
                    "axis_error_mean": float(a.mean()) if len(a) else 0.0,
                    "blocks_axis_error_over_1e-3": int((a > 1e-3).sum()),
                    "endpoint_error_max": float(ep.max()), "endpoint_error_p999": float(np.percentile(ep, 99.9)),
                    "endpoint_error_mean": float(ep.mean()), "blocks_differing": round(float(diff.mean()), 6)}
    return out


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_small_case(wl, min_s=3.0):
    """The oracle on the 1024x1024 RGB8 4x4 texture (all host cores; median run over
    at least min_s seconds)."""
    from tests import oracle_ref as orc
    cores = len(os.sched_getaffinity(0))
    img = wl["gen"](0, wl["w"], wl["h"])
    times = _time_runs(lambda: orc.oracle_encode(img, wl["dim"], wl["flags"], threads=cores), min_s)
    return {"workload": wl["name"], "mtexels_s": round(wl["w"] * wl["h"] / float(np.median(times)) / 1e6, 1),
            "cores": cores, "runs": len(times)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="only the headline workload")
    ap.add_argument("--no-batched", action="store_true", help="skip the mixed-texture batched vs plain run")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sweep")
    ap.add_argument("--plain-tuning", default=None,
                    help="lanes_per_block,blocks_per_lane,prefetch[,flat_alpha[,packed_texels]] for the headline's plain launches "
                         "(default: the library's choice)")
    ap.add_argument("--batched-tuning", default=None,
                    help="blocks_per_lane,workgroups_per_cu,prefetch,split_modes,lanes_per_block for the mixed set "
                         "(default: the autotuner's choice)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 3 device copies of 256 MiB (a known-size kernel for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (textures are independent and "
              "would be sharded one process per GPU with no collective)", file=sys.stderr)
        sys.exit(2)

    import torch
    import astc_encoder_amd as astc

    enc = astc.Encoder(0)
    if args.copy_calibration:
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda").fill_(1)
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        del a, c
    wls = workloads()
    head = wls["noise_rgba8_srgb_6x6"]
    # The other GPU legs run first, so that the headline is timed with the GPU at
    # its steady clock (a few hundred microseconds of launches after idle run
    # slower while the clock rises).
    extras = {}
    if not args.no_extras:
        for key in ("sinusoid_rgb8_4x4", "normal_rg8_4x4"):
            wl = wls[key]
            e = time_plain(enc, wl, max(10, args.steps), args.warmup)
            extras[key] = {"mtexels_s": round(e["mtexels_s"], 1), "kernel_ms": round(e["kernel_ms"], 4),
                           "psnr_db": e["psnr_db"],
                           "valu_tflops": round(FLOPS_PER_BLOCK[wl["flops_key"]] * e["nblocks"] /
                                                (e["kernel_ms"] * 1e-3) / 1e12, 2)}
        extras["opaque_rgba8_6x6_autotuned"] = opaque_autotuned(enc, max(10, args.steps), args.warmup)
        extras["host_resident_headline"] = time_host_resident(enc, head, 20)
        if not args.no_batched:
            fixed = tuple(int(v) for v in args.batched_tuning.split(",")) if args.batched_tuning else None
            extras["mixed_textures_batched_vs_plain"] = batched_vs_plain(enc, max(5, args.steps // 5), 2, tuning=fixed)
        if not args.no_serving:
            extras["serving"] = serving_sweep(enc)
    head_tuning = tuple(int(v) for v in args.plain_tuning.split(",")) if args.plain_tuning else None
    r = time_plain(enc, head, args.steps, args.warmup, head_tuning)
    algo_bytes = r["texels"] * head["bpp"] + r["nblocks"] * 16
    achieved_gbs = algo_bytes / (r["kernel_ms"] * 1e-3) / 1e9
    flops = FLOPS_PER_BLOCK[head["flops_key"]] * r["nblocks"]
    achieved_tf = flops / (r["kernel_ms"] * 1e-3) / 1e12
    traffic = read_traffic(os.path.join(ROOT, "profiles", "hbm_traffic_encode_plain_6x6.txt"))
    hbm = {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic}
    valu = {"bound": "valu", "achieved": round(achieved_tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved_tf / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
            "flops_per_block": FLOPS_PER_BLOCK[head["flops_key"]]}
    # the roofline that binds is the one the kernel comes closest to
    bound = dict(hbm if hbm["frac"] >= valu["frac"] else valu)
    # the kernel the launches ran: from the tuning they used (the library's default
    # for the image when none is given, as astc_plain_default_tuning reports it)
    used = r["tuning"]
    if used[0] == 2:
        kernel = "encode_plain_pair_kernel<true, true, 4>"
    elif used[4]:
        kernel = "encode_plain_packed_kernel<6, true, true, 4>"
    else:
        kernel = f"encode_plain_kernel<6, true, true, 4, 0, {'true' if used[3] else 'false'}>"
    tuning_fields = ("lanes_per_block", "blocks_per_lane", "prefetch", "flat_alpha", "packed_texels")
    bound.update({"kernel": kernel,
                  "plain_tuning": dict(zip(tuning_fields, used), source="given" if head_tuning else "library default"),
                  "kernel_ms": round(r["kernel_ms"], 4),
                  "step_event_ms": round(r["step_event_ms"], 4), "algorithmic_bytes": algo_bytes,
                  "algorithmic_flops": flops})

    line = {
        "metric": METRIC,
        "value": round(r["mtexels_s"], 1),
        "unit": "Mtexels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall_s"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seed 0)",
        "config": {"workload": head["name"] + ", plain per-texture launch", "width": head["w"],
                   "height": head["h"], "block": "6x6", "flags": "alpha|srgb", "blocks": r["nblocks"],
                   "parallelism": "single GPU"},
        "roofline": bound,
        "roofline_hbm": hbm,
        "roofline_valu": valu,
    }
    line["psnr_db"] = r["psnr_db"]
    if not args.no_extras:
        line["extras"] = extras
    # the CPU leg last: its 16 busy host threads stay away from the serving loop's timing
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(head, r["img"], r["blocks"])
        small = wls["sinusoid_rgb8_4x4"]  # the reference's CPU-sized case: 1024^2 RGB8 4x4
        line["cpu_baseline"]["small_case"] = cpu_small_case(small)
        line["psnr_delta_vs_oracle_db"] = round(r["psnr_db"] - line["cpu_baseline"].pop("oracle_psnr_db"), 6)
        line["fp64_twin"] = fp64_twin(wls, line["cpu_baseline"]["cores"])
    enc.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
