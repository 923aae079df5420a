"""Extract the committed leaf fixtures from the reference's only golden pair.

The reference ships textures/leaf.png and textures/leaf.astc (MIT, niedap), the
latter produced by `astc_cs_enc leaf.png -alpha -4x4` on a D3D11 GPU: 1024x1024,
4x4 blocks, CEM 12 / QUANT_6 weights, image flipped vertically on load
(main.cpp:24), linear (no sRGB) fetch.  This script is run once, in a container
that has the reference mounted, and writes small data files:

  leaf_sample.npz   8192 seeded-random blocks: the 4x4 RGBA8 tile each block was
                    encoded from (flipped orientation, as the encoder saw it),
                    its block coordinates and the reference's 16-byte block.
  leaf_crop.png     a 128x128 crop of leaf.png (file orientation), written with
                    zlib by this script, for end-to-end runs of the CLI.
  leaf_crop.astc    the reference blocks covering that crop, with the .astc
                    header, i.e. what `astc_cs_enc leaf_crop.png -alpha -4x4`
                    must reproduce.
  leaf_uncoupled_alpha.npz
                    every block of the image whose tile has exactly zero
                    covariance between alpha and each colour channel (integer
                    arithmetic) and an alpha variance above every colour
                    eigenvalue: its tile and the reference's block.  In exact
                    arithmetic the shader's power iteration, which starts with
                    no alpha component, cannot reach the alpha axis there.
  leaf_stats.json   whole-image statistics of the reference output.

Usage: python tests/golden/make_leaf_fixture.py [/root/reference]
"""
import json
import os
import struct
import sys
import zlib

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))


def write_png_rgba(path, rgba):
    """Minimal PNG writer (RGBA8, filter 0 per row)."""
    h, w, _ = rgba.shape
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 9)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def main(ref_root):
    png = np.array(Image.open(os.path.join(ref_root, "textures/leaf.png")).convert("RGBA"))
    raw = np.fromfile(os.path.join(ref_root, "textures/leaf.astc"), dtype=np.uint8)
    header, blocks = raw[:16], raw[16:].reshape(-1, 16)
    h, w, _ = png.shape
    assert (w, h) == (1024, 1024) and blocks.shape[0] == (w // 4) * (h // 4)
    flipped = png[::-1]                       # what the encoder sees (stbi flip on load)
    xb = w // 4

    rng = np.random.default_rng(0)
    ids = np.sort(rng.choice(blocks.shape[0], size=8192, replace=False))
    bx, by = ids % xb, ids // xb
    tiles = np.stack([flipped[y * 4:y * 4 + 4, x * 4:x * 4 + 4] for x, y in zip(bx, by)])
    np.savez_compressed(os.path.join(HERE, "leaf_sample.npz"), tiles=tiles, bx=bx.astype(np.int32),
                        by=by.astype(np.int32), blocks=blocks[ids], header=header)

    # 128x128 crop at file rows [r0, r0+128), columns [c0, c0+128).  After the
    # flip its rows land at 1024-r0-128 .. 1024-r0, a multiple of 4.
    r0, c0, n = 448, 512, 128
    crop = png[r0:r0 + n, c0:c0 + n]
    write_png_rgba(os.path.join(HERE, "leaf_crop.png"), np.ascontiguousarray(crop))
    fy0 = h - r0 - n
    sel = [(fy0 // 4 + j) * xb + (c0 // 4 + i) for j in range(n // 4) for i in range(n // 4)]
    hdr = bytes([0x13, 0xAB, 0xA1, 0x5C, 4, 4, 1, n & 255, n >> 8, 0, n & 255, n >> 8, 0, 1, 0, 0])
    with open(os.path.join(HERE, "leaf_crop.astc"), "wb") as f:
        f.write(hdr + blocks[sel].tobytes())

    # tiles with no colour-alpha coupling and a dominant alpha variance
    all_tiles = flipped.reshape(h // 4, 4, xb, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4).astype(np.int64)
    s1 = all_tiles.sum(1)                                         # per channel sums
    cross = 16 * np.einsum("nkc,nk->nc", all_tiles[:, :, :3], all_tiles[:, :, 3]) - s1[:, :3] * s1[:, 3:4]
    uncoupled = np.all(cross == 0, axis=1)
    d = all_tiles - all_tiles.mean(1, keepdims=True)
    cov = np.einsum("nki,nkj->nij", d, d) / 15.0
    top_colour = np.linalg.eigvalsh(cov[:, :3, :3])[:, -1]
    pick = np.nonzero(uncoupled & (cov[:, 3, 3] > top_colour))[0]
    np.savez_compressed(os.path.join(HERE, "leaf_uncoupled_alpha.npz"), ids=pick.astype(np.int32),
                        tiles=all_tiles[pick].reshape(-1, 4, 4, 4).astype(np.uint8), blocks=blocks[pick])

    modes = blocks[:, 0].astype(np.uint32) | (blocks[:, 1].astype(np.uint32) << 8)
    stats = {
        "source": "niepp/astc_encoder textures/leaf.png + textures/leaf.astc (MIT)",
        "options": "-alpha -4x4, linear, vertical flip on load",
        "width": w, "height": h, "blocks": int(blocks.shape[0]),
        "header_hex": header.tobytes().hex(),
        "blockmode_0x043_fraction": float(np.mean((modes & 0x7FF) == 0x043)),
        "sample_block_ids_seed": 0,
        "uncoupled_alpha_blocks": int(pick.size),
    }
    with open(os.path.join(HERE, "leaf_stats.json"), "w") as f:
        json.dump(stats, f, indent=1)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
