"""ctypes binding of libastc_mi355x.so (include/astc_mi355x.h).

The HIP library is the only implementation: if it is not built, importing the
encoder raises instead of falling back to anything else.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libastc_mi355x.so")
SYNTH_PATH = os.path.join(LIB_DIR, "libastc_synth.so")
CLI_PATH = os.path.join(LIB_DIR, "astc_cs_enc")

ASTC_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "out of memory", -4: "unsupported",
          -5: "I/O error", -6: "over budget"}

FMT_RGBA8, FMT_RGB8, FMT_RG8 = 0, 1, 2
BYTES_PER_TEXEL = {FMT_RGBA8: 4, FMT_RGB8: 3, FMT_RG8: 2}
FLAG_ALPHA, FLAG_NORMAL, FLAG_SRGB, FLAG_MAX_ACCUMULATION = 1, 2, 4, 8


class AstcError(RuntimeError):
    """An ASTC_ERR_* return code with the library's message."""

    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class ImageDesc(ctypes.Structure):
    """astc_image_desc"""
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("row_pitch", ctypes.c_int32),
                ("src_format", ctypes.c_int32), ("block_dim", ctypes.c_int32), ("flags", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    """astc_tuning"""
    _fields_ = [("blocks_per_lane", ctypes.c_int32), ("workgroups_per_cu", ctypes.c_int32),
                ("prefetch", ctypes.c_int32), ("split_modes", ctypes.c_int32), ("lanes_per_block", ctypes.c_int32)]

    def as_dict(self):
        return {"blocks_per_lane": self.blocks_per_lane, "workgroups_per_cu": self.workgroups_per_cu,
                "prefetch": self.prefetch, "split_modes": self.split_modes,
                "lanes_per_block": max(1, self.lanes_per_block)}


class PlainTuning(ctypes.Structure):
    """astc_plain_tuning"""
    _fields_ = [("lanes_per_block", ctypes.c_int32), ("blocks_per_lane", ctypes.c_int32),
                ("prefetch", ctypes.c_int32), ("flat_alpha", ctypes.c_int32), ("packed_texels", ctypes.c_int32)]

    def as_dict(self):
        return {"lanes_per_block": self.lanes_per_block, "blocks_per_lane": self.blocks_per_lane,
                "prefetch": self.prefetch, "flat_alpha": self.flat_alpha, "packed_texels": self.packed_texels}


class ServerConfig(ctypes.Structure):
    """astc_server_config"""
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_inflight", ctypes.c_int32),
                ("max_batch_requests", ctypes.c_int32), ("tuning", Tuning)]


class Request(ctypes.Structure):
    """astc_request"""
    _fields_ = [("images", ctypes.POINTER(ImageDesc)), ("nimages", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("src_dev", ctypes.c_void_p), ("dst_dev", ctypes.c_void_p), ("arrival_s", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    """astc_server_stats"""
    _fields_ = [("requests", ctypes.c_int64), ("images", ctypes.c_int64), ("texels", ctypes.c_int64),
                ("blocks", ctypes.c_int64), ("batches", ctypes.c_int64), ("wall_s", ctypes.c_double),
                ("mean_batch_requests", ctypes.c_double), ("peak_ring_bytes", ctypes.c_int64),
                ("peak_inflight", ctypes.c_int32), ("reserved", ctypes.c_int32), ("max_loop_gap_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


P = ctypes.c_void_p
I32 = ctypes.c_int
I64 = ctypes.c_int64
PDESC = ctypes.POINTER(ImageDesc)

# name -> (restype, argtypes); every function declared in include/astc_mi355x.h
SIGNATURES = {
    "astc_last_error": (ctypes.c_char_p, []),
    "astc_ctx_create": (I32, [I32, ctypes.POINTER(P)]),
    "astc_ctx_destroy": (I32, [P]),
    "astc_output_size": (I64, [ctypes.POINTER(ImageDesc)]),
    "astc_encode": (I32, [P, ctypes.POINTER(ImageDesc), P, P, P]),
    "astc_encode_tuned": (I32, [P, ctypes.POINTER(ImageDesc), P, P, ctypes.POINTER(PlainTuning), P]),
    "astc_plain_autotune": (I32, [P, ctypes.POINTER(ImageDesc), P, P, P, I32, ctypes.POINTER(PlainTuning)]),
    "astc_plain_default_tuning": (I32, [P, ctypes.POINTER(ImageDesc), ctypes.POINTER(PlainTuning)]),
    "astc_decode": (I32, [P, I32, I32, I32, I32, P, P, P]),
    "astc_encode_each": (I32, [P, PDESC, ctypes.POINTER(P), ctypes.POINTER(P), I32, P]),
    "astc_write_file": (I32, [ctypes.c_char_p, I32, I32, I32, I32, P, I64]),
    "astc_read_file_header": (I32, [ctypes.c_char_p, ctypes.POINTER(I32), ctypes.POINTER(I32),
                                    ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(I64)]),
    "astc_srgb_table": (I32, [P]),
    "astc_load_image": (I32, [ctypes.c_char_p, I32, ctypes.POINTER(I32), ctypes.POINTER(I32),
                              ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))]),
    "astc_free_image": (None, [ctypes.POINTER(ctypes.c_uint8)]),
    "astc_mip_chain_info": (I32, [PDESC, I32, PDESC, ctypes.POINTER(I32)]),
    "astc_downsample": (I32, [P, PDESC, P, P, P]),
    "astc_batch_create": (I32, [P, PDESC, ctypes.POINTER(P), ctypes.POINTER(P), I32, ctypes.POINTER(P)]),
    "astc_batch_set_tuning": (I32, [P, ctypes.POINTER(Tuning)]),
    "astc_batch_get_tuning": (I32, [P, ctypes.POINTER(Tuning)]),
    "astc_batch_autotune": (I32, [P, P, I32, ctypes.POINTER(Tuning)]),
    "astc_batch_launch": (I32, [P, P]),
    "astc_batch_info": (I32, [P, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I32)]),
    "astc_batch_destroy": (I32, [P]),
    "astc_encode_batch_noalloc": (I32, [P, PDESC, ctypes.POINTER(P), ctypes.POINTER(P), I32, ctypes.POINTER(Tuning),
                                        P]),
    "astc_request_bytes": (I64, [PDESC, I32]),
    "astc_server_create": (I32, [P, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]),
    "astc_server_run_trace": (I32, [P, ctypes.POINTER(Request), I32, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ServerStats)]),
    "astc_server_destroy": (I32, [P]),
    "astc_server_reserve": (I32, [P, PDESC, I32, ctypes.POINTER(P), ctypes.POINTER(I64)]),
    "astc_server_submit": (I32, [P, I64, P]),
    "astc_server_poll": (I32, [P, ctypes.POINTER(I64), I32, ctypes.POINTER(I32)]),
    "astc_server_blocks": (I32, [P, I64, ctypes.POINTER(P), ctypes.POINTER(I64)]),
    "astc_server_release": (I32, [P, I64]),
}

_lib = None


def lib():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build the HIP library first "
                              "(python -c 'import __graft_entry__ as g; g.build()')")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    """Raise AstcError for a negative return code, else return rc."""
    if rc < 0:
        msg = lib().astc_last_error()
        raise AstcError(int(rc), msg.decode() if msg else "")
    return rc


_synth = None


def synth_lib():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise ImportError(f"{SYNTH_PATH} is missing: build the libraries first")
        s = ctypes.CDLL(SYNTH_PATH)
        U64 = ctypes.c_uint64
        for name in ("astc_synth_sinusoid_rgb8", "astc_synth_noise_rgba8", "astc_synth_normal_rg8"):
            getattr(s, name).argtypes = [U64, I32, I32, P]
            getattr(s, name).restype = None
        s.astc_synth_downsample.argtypes = [P, I32, I32, I32, P, ctypes.POINTER(I32), ctypes.POINTER(I32)]
        s.astc_synth_downsample.restype = None
        _synth = s
    return _synth
