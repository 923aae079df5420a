"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) for tests.

Test infrastructure: the oracle is what the HIP path is compared against, never part
of the product path.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

SCHED_CANONICAL, SCHED_SOURCE = 0, 1
METHOD_PCA, METHOD_MAX_ACCUMULATION = 0, 1
FLAG_MAX_ACCUMULATION = 8
FMT = {4: 0, 3: 1, 2: 2}

P = ctypes.c_void_p
I = ctypes.c_int


class Trace(ctypes.Structure):
    _fields_ = [("mean", ctypes.c_double * 4), ("axis", ctypes.c_double * 4), ("e0", ctypes.c_double * 4),
                ("e1", ctypes.c_double * 4), ("ep_q", ctypes.c_uint32 * 8), ("w_q", ctypes.c_uint32 * 16)]


class BlockInfo(ctypes.Structure):
    _fields_ = [("valid", I), ("void_extent", I), ("grid_w", I), ("grid_h", I), ("dual_plane", I),
                ("partitions", I), ("cem", I), ("weight_quant", I), ("endpoint_quant", I),
                ("ep_q", ctypes.c_uint32 * 8), ("w_sym", ctypes.c_uint32 * 64), ("w_unq", ctypes.c_uint32 * 64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise ImportError(f"{ORACLE_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(ORACLE_PATH)
        sig = {
            "oracle_encode_block_f32": (I, [P, I, I, I, P, P]),
            "oracle_encode_block_f64": (I, [P, I, I, I, P, P]),
            "oracle_encode_block_method_f32": (I, [P, I, I, I, I, P, P]),
            "oracle_encode_block_method_f64": (I, [P, I, I, I, I, P, P]),
            "oracle_fetch_block_f32": (None, [P, I, I, I, I, I, I, I, I, P]),
            "oracle_fetch_block_f64": (None, [P, I, I, I, I, I, I, I, I, P]),
            "oracle_encode_image": (ctypes.c_int64, [P, I, I, I, I, I, I, I, I, P]),
            "oracle_fp64_twin": (None, [P, I, I, I, I, I, I, I, P, P, P]),
            "oracle_srgb_table": (None, [P]),
            "oracle_unorm_table": (None, [P]),
            "oracle_tables": (None, [P, P, P, P]),
            "oracle_bise": (I, [P, I, I, P]),
            "oracle_blockmode": (ctypes.c_uint32, [I]),
            "oracle_assemble_block": (None, [ctypes.c_uint32, ctypes.c_uint32, P, P, P]),
            "oracle_pack_block": (None, [P, P, I, P]),
            "oracle_decode_block": (I, [P, I, I, P, P]),
            "oracle_decode_block_weights": (I, [P, I, I, P, P]),
            "oracle_grid6_taps": (None, [P, P]),
            "oracle_decode_image": (I, [P, I, I, I, P]),
            "oracle_ise_decode": (I, [P, I, I, I, P]),
            "oracle_psnr": (ctypes.c_double, [P, I, I, I, I, I, P]),
            "oracle_sq_error": (None, [P, I, I, I, I, I, P, P, P]),
            "oracle_length_threshold": (ctypes.c_float, []),
            "oracle_rsqrt_nr": (ctypes.c_float, [ctypes.c_float]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(P)


def oracle_encode(img, dim, flags, sched=SCHED_CANONICAL, threads=0):
    """Encode an (H, W, C) uint8 image; returns (nblocks, 16) uint8."""
    img = np.ascontiguousarray(img)
    h, w, c = img.shape
    n = ((w + dim - 1) // dim) * ((h + dim - 1) // dim)
    out = np.zeros((n, 16), np.uint8)
    rc = lib().oracle_encode_image(_ptr(img), w, h, w * c, FMT[c], dim, flags, sched, threads, _ptr(out))
    assert rc == n
    return out


def fp64_twin(img, dim, flags, threads=0):
    """Binary32 vs binary64 runs of the same method over every block of an image:
    per-block axis error (NaN where the binary64 axis is zero), endpoint error
    (8-bit units) and byte difference."""
    img = np.ascontiguousarray(img)
    h, w, c = img.shape
    n = ((w + dim - 1) // dim) * ((h + dim - 1) // dim)
    ax, ep, diff = np.zeros(n), np.zeros(n), np.zeros(n, np.uint8)
    lib().oracle_fp64_twin(_ptr(img), w, h, w * c, FMT[c], dim, flags, threads, _ptr(ax), _ptr(ep), _ptr(diff))
    ax[ax < 0] = np.nan
    return ax, ep, diff.astype(bool)


def encode_block(texels, dim, has_alpha, sched=SCHED_CANONICAL, f64=False, method=METHOD_PCA):
    """Encode one block from fetched texels (N, 4); returns (block bytes, Trace)."""
    dt = np.float64 if f64 else np.float32
    t = np.ascontiguousarray(texels, dtype=dt).reshape(-1)
    out = np.zeros(16, np.uint8)
    tr = Trace()
    fn = lib().oracle_encode_block_method_f64 if f64 else lib().oracle_encode_block_method_f32
    assert fn(_ptr(t), dim, int(has_alpha), sched, method, _ptr(out), ctypes.byref(tr)) == 0
    return out, tr


def fetch_block(img, dim, flags, bx, by, f64=False):
    img = np.ascontiguousarray(img)
    h, w, c = img.shape
    dt = np.float64 if f64 else np.float32
    out = np.zeros(dim * dim * 4, dt)
    fn = lib().oracle_fetch_block_f64 if f64 else lib().oracle_fetch_block_f32
    fn(_ptr(img), w, h, w * c, FMT[c], dim, flags, bx, by, _ptr(out))
    return out.reshape(dim * dim, 4)


def decode_block(blk, dim_x, dim_y=None):
    dim_y = dim_x if dim_y is None else dim_y
    b = np.ascontiguousarray(blk, dtype=np.uint8)
    rgba = np.zeros(dim_x * dim_y * 4, np.uint8)
    info = BlockInfo()
    ok = lib().oracle_decode_block(_ptr(b), dim_x, dim_y, _ptr(rgba), ctypes.byref(info))
    return ok, rgba.reshape(dim_y, dim_x, 4), info


def decode_block_weights(blk, dim_x, dim_y=None):
    """Decode one block; returns (ok, rgba, per-texel infilled weights of plane 0 as (dim_y, dim_x))."""
    dim_y = dim_x if dim_y is None else dim_y
    b = np.ascontiguousarray(blk, dtype=np.uint8)
    rgba = np.zeros(dim_x * dim_y * 4, np.uint8)
    w = np.full(2 * dim_x * dim_y, -1, np.int32)
    ok = lib().oracle_decode_block_weights(_ptr(b), dim_x, dim_y, _ptr(rgba), _ptr(w))
    return ok, rgba.reshape(dim_y, dim_x, 4), w[:dim_x * dim_y].reshape(dim_y, dim_x)


def grid6_taps():
    """The oracle's 6x6 sampling taps: (16, 4) texel indices and (16, 4) float32 weights."""
    idx = np.zeros(64, np.int32)
    wt = np.zeros(64, np.float32)
    lib().oracle_grid6_taps(_ptr(idx), _ptr(wt))
    return idx.reshape(16, 4), wt.reshape(16, 4)


def decode_image(blocks, w, h, dim):
    b = np.ascontiguousarray(blocks, dtype=np.uint8)
    rgba = np.zeros((h, w, 4), np.uint8)
    ok = lib().oracle_decode_image(_ptr(b), w, h, dim, _ptr(rgba))
    return ok, rgba


def psnr(img, flags, decoded):
    img = np.ascontiguousarray(img)
    h, w, c = img.shape
    return lib().oracle_psnr(_ptr(img), w, h, w * c, FMT[c], flags, _ptr(np.ascontiguousarray(decoded)))


def sq_error(img, flags, decoded):
    img = np.ascontiguousarray(img)
    h, w, c = img.shape
    sse, cnt = ctypes.c_double(), ctypes.c_double()
    lib().oracle_sq_error(_ptr(img), w, h, w * c, FMT[c], flags, _ptr(np.ascontiguousarray(decoded)),
                          ctypes.byref(sse), ctypes.byref(cnt))
    return sse.value, cnt.value


def tables():
    tr = np.zeros(243, np.int32)
    qu = np.zeros(125, np.int32)
    sc = np.zeros(12 * 32, np.int32)
    btq = np.zeros(21 * 3, np.int32)
    lib().oracle_tables(_ptr(tr), _ptr(qu), _ptr(sc), _ptr(btq))
    return tr, qu, sc.reshape(12, 32), btq.reshape(21, 3)


def bise(numbers, quant):
    nums = np.ascontiguousarray(numbers, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    nbits = lib().oracle_bise(_ptr(nums), len(nums), quant, _ptr(out))
    return out, nbits


def ise_decode(words, bit_offset, count, quant):
    b = np.ascontiguousarray(words, dtype=np.uint32).view(np.uint8)
    out = np.zeros(count, np.uint32)
    nbits = lib().oracle_ise_decode(_ptr(b), bit_offset, count, quant, _ptr(out))
    return out, nbits


def pack_block(ep_q, w_q, has_alpha):
    e = np.ascontiguousarray(ep_q, dtype=np.uint32)
    w = np.ascontiguousarray(w_q, dtype=np.uint32)
    out = np.zeros(16, np.uint8)
    lib().oracle_pack_block(_ptr(e), _ptr(w), int(has_alpha), _ptr(out))
    return out


def srgb_table():
    out = np.zeros(256, np.float32)
    lib().oracle_srgb_table(_ptr(out))
    return out


def unorm_table():
    out = np.zeros(256, np.float32)
    lib().oracle_unorm_table(_ptr(out))
    return out
