"""MI355X-native ASTC block encoder (drop-in for niepp/astc_encoder's GPU path).

The encoder is libastc_mi355x.so (hand-written gfx950 HIP kernels behind a C
ABI, include/astc_mi355x.h); this package is its Python binding and the
reference-shaped host interface (EncodeOption, encode_astc, save_astc).
"""
from ._capi import (FLAG_ALPHA, FLAG_MAX_ACCUMULATION, FLAG_NORMAL, FLAG_SRGB, FMT_RG8, FMT_RGB8, FMT_RGBA8, AstcError, ImageDesc,
                    LIB_PATH, lib)
from .encoder import (Batch, PackedImages, Server, load_tex, EncodeOption, Encoder, encode_astc, make_desc, mip_chain_info, output_size,
                      read_astc_header, save_astc, srgb_table)

__all__ = [
    "FLAG_ALPHA", "FLAG_MAX_ACCUMULATION", "FLAG_NORMAL", "FLAG_SRGB", "FMT_RG8", "FMT_RGB8", "FMT_RGBA8", "AstcError", "ImageDesc",
    "LIB_PATH", "lib", "Batch", "PackedImages", "Server", "load_tex", "mip_chain_info", "EncodeOption", "Encoder", "encode_astc", "make_desc", "output_size", "read_astc_header",
    "save_astc", "srgb_table",
]
