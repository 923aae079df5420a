"""Seeded synthetic workloads (libastc_synth.so): identical bytes on every host.

  sinusoid_rgb8(seed, w, h)  -> (h, w, 3) uint8   RGB, linear, 4x4
  noise_rgba8(seed, w, h)    -> (h, w, 4) uint8   RGBA, sRGB, 6x6
  normal_rg8(seed, w, h)     -> (h, w, 2) uint8   normal map, 4x4
  mip_chain(img)             -> [level0, level1, ...] (2x2 box filter; power-of-two sides end at 4x4)
  texture_set(seed, n)       -> a list of mixed textures with their modes
"""
import ctypes

import numpy as np

from ._capi import FLAG_ALPHA, FLAG_NORMAL, FLAG_SRGB, FMT_RG8, FMT_RGB8, FMT_RGBA8, synth_lib


def _gen(fn, seed, w, h, c):
    out = np.empty((h, w, c), np.uint8)
    fn(ctypes.c_uint64(seed), int(w), int(h), out.ctypes.data_as(ctypes.c_void_p))
    return out


def sinusoid_rgb8(seed, w, h):
    return _gen(synth_lib().astc_synth_sinusoid_rgb8, seed, w, h, 3)


def noise_rgba8(seed, w, h):
    return _gen(synth_lib().astc_synth_noise_rgba8, seed, w, h, 4)


def normal_rg8(seed, w, h):
    return _gen(synth_lib().astc_synth_normal_rg8, seed, w, h, 2)


def downsample(img):
    h, w, c = img.shape
    ow = w >> 1 if w >= 8 else w
    oh = h >> 1 if h >= 8 else h
    out = np.empty((oh, ow, c), np.uint8)
    nw, nh = ctypes.c_int(), ctypes.c_int()
    src = np.ascontiguousarray(img)
    synth_lib().astc_synth_downsample(src.ctypes.data_as(ctypes.c_void_p), w, h, c,
                                      out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nw), ctypes.byref(nh))
    return out


def mip_chain(img):
    """Levels while a side can halve (a side halves while it is >= 8)."""
    levels = [img]
    while levels[-1].shape[0] >= 8 or levels[-1].shape[1] >= 8:
        levels.append(downsample(levels[-1]))
    return levels


# (name, generator, block_dim, flags, src_format) of the three texture modes
MODES = {
    "rgb4x4": (sinusoid_rgb8, 4, 0, FMT_RGB8),
    "rgba6x6srgb": (noise_rgba8, 6, FLAG_ALPHA | FLAG_SRGB, FMT_RGBA8),
    "normal4x4": (normal_rg8, 4, FLAG_NORMAL, FMT_RG8),
}


def draw_texture_specs(seed, n):
    """The mixed-texture distribution: side 2^k with k uniform in 6..12, square or
    2:1 (50/50), mode 60% rgb4x4 / 25% rgba6x6srgb / 15% normal4x4.
    Returns [(mode, width, height, content_seed)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    specs = []
    for i in range(n):
        k = int(rng.integers(6, 13))
        square = bool(rng.integers(0, 2))
        u = float(rng.random())
        mode = "rgb4x4" if u < 0.60 else ("rgba6x6srgb" if u < 0.85 else "normal4x4")
        w = 1 << k
        h = w if square else w >> 1
        specs.append((mode, w, h, int(rng.integers(0, 2**62))))
    return specs
