"""Host-side mirror of the reference's encoder interface on top of the C ABI.

Reference names kept (niepp/astc_encoder):
  encode_option   astc_encode.h:14-28    -> EncodeOption
  encode_astc     astc_encode.h:87-194   -> encode_astc(texture, option)
  save_astc       astc_save.h:52-76      -> save_astc(path, xdim, ydim, xsize, ysize, buffer)

Device memory and streams come from PyTorch (plumbing only); the work is done
by libastc_mi355x.so.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._capi import (FLAG_ALPHA, FLAG_MAX_ACCUMULATION, FLAG_NORMAL, FLAG_SRGB, FMT_RG8, FMT_RGB8, FMT_RGBA8, ImageDesc,
                    BYTES_PER_TEXEL, PlainTuning, Request, ServerConfig, ServerStats, Tuning, check, lib)

_CHANNELS_TO_FMT = {4: FMT_RGBA8, 3: FMT_RGB8, 2: FMT_RG8}


@dataclass
class EncodeOption:
    """encode_option (astc_encode.h:14-28).

    The reference initialises is4x4 = true and never clears it, so its -6x6
    flag has no effect (main.cpp:156-160, astc_encode.h:59,124).  Here is6x6
    selects 6x6 blocks, which is what the flag documents (README.md:29).
    """
    is4x4: bool = True
    is6x6: bool = False
    is_normal_map: bool = False
    has_alpha: bool = False
    srgb: bool = False
    # endpoint search of ASTC_Encode.hlsl:170-227, commented out in the reference
    # (:514); not part of its options
    max_accumulation: bool = False

    @property
    def block_dim(self):
        return 6 if self.is6x6 else 4

    def flags(self):
        """Runtime flags for a texture loaded the way main.cpp:214 loads it:
        the texture is sRGB only when srgb is set and the map is not a normal map."""
        f = 0
        if self.has_alpha:
            f |= FLAG_ALPHA
        if self.is_normal_map:
            f |= FLAG_NORMAL
        if self.srgb and not self.is_normal_map:
            f |= FLAG_SRGB
        if self.max_accumulation:
            f |= FLAG_MAX_ACCUMULATION
        return f


def make_desc(width, height, src_format, block_dim, flags, row_pitch=None):
    if row_pitch is None:
        row_pitch = width * BYTES_PER_TEXEL[src_format]
    return ImageDesc(int(width), int(height), int(row_pitch), int(src_format), int(block_dim), int(flags))


def output_size(desc):
    return check(lib().astc_output_size(ctypes.byref(desc)))


def _torch_stream(stream):
    import torch
    return torch.cuda.current_stream() if stream is None else stream


def _stream_handle(stream):
    return ctypes.c_void_p(_torch_stream(stream).cuda_stream)


class Encoder:
    """An astc_ctx on one HIP device."""

    def __init__(self, device=0):
        import weakref
        self._ctx = ctypes.c_void_p()
        check(lib().astc_ctx_create(int(device), ctypes.byref(self._ctx)))
        self.device = device
        # batches and servers made on this context: closed before it is destroyed
        self._children = weakref.WeakSet()

    @property
    def handle(self):
        return self._ctx

    def close(self):
        if self._ctx:
            for child in list(self._children):
                child.close()
            lib().astc_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def encode_raw(self, desc, src_ptr, dst_ptr, stream=None):
        """astc_encode on raw device pointers (asynchronous on `stream`)."""
        check(lib().astc_encode(self._ctx, ctypes.byref(desc), ctypes.c_void_p(src_ptr),
                                ctypes.c_void_p(dst_ptr), _stream_handle(stream)))

    def _texture_desc(self, texture, block_dim, flags):
        import torch
        if texture.dtype != torch.uint8 or texture.dim() != 3 or texture.shape[2] not in _CHANNELS_TO_FMT:
            raise ValueError("texture must be an (H, W, 4|3|2) uint8 tensor")
        if not texture.is_cuda:
            raise ValueError("texture must be in device memory")
        h, w, c = texture.shape
        return make_desc(w, h, _CHANNELS_TO_FMT[c], block_dim, flags, row_pitch=texture.stride(0))

    @staticmethod
    def _dense_rows(texture, stream):
        """The texture itself if its rows hold contiguous texels, else a dense copy
        kept alive for the launch stream (the caller's reference ends before the
        asynchronous kernel does)."""
        c = texture.shape[2]
        if texture.stride(2) == 1 and texture.stride(1) == c:
            return texture
        dense = texture.contiguous()
        dense.record_stream(_torch_stream(stream))
        return dense

    def _output_for(self, desc, texture, out):
        import torch
        n = output_size(desc)
        if out is None:
            return torch.empty((n // 16, 16), dtype=torch.uint8, device=texture.device)
        if out.dtype != torch.uint8 or out.numel() < n or not out.is_contiguous() or out.device != texture.device:
            raise ValueError("out must be a contiguous uint8 tensor of at least output_size bytes on the texture's device")
        return out

    def encode(self, texture, block_dim=4, flags=0, out=None, stream=None, tuning=None):
        """Encode a device texture (H, W, C) uint8 tensor, C in {4, 3, 2}, whose
        rows hold contiguous texels (any row pitch).  tuning: None (the library's
        default) or (lanes_per_block, blocks_per_lane, prefetch[, flat_alpha[, packed_texels]]).

        Returns a (nblocks, 16) uint8 device tensor in row-major block order."""
        self._texture_desc(texture, block_dim, flags)
        texture = self._dense_rows(texture, stream)
        desc = self._texture_desc(texture, block_dim, flags)
        out = self._output_for(desc, texture, out)
        if tuning is None:
            self.encode_raw(desc, texture.data_ptr(), out.data_ptr(), stream)
        else:
            t = PlainTuning(*(int(v) for v in tuple(tuning) + (0,) * (5 - len(tuning))))
            check(lib().astc_encode_tuned(self._ctx, ctypes.byref(desc), ctypes.c_void_p(texture.data_ptr()),
                                          ctypes.c_void_p(out.data_ptr()), ctypes.byref(t), _stream_handle(stream)))
        return out

    def autotune_plain(self, texture, block_dim=4, flags=0, reps=10, stream=None):
        """astc_plain_autotune on this texture: the fastest plain tuning as a dict."""
        self._texture_desc(texture, block_dim, flags)
        texture = self._dense_rows(texture, stream)
        desc = self._texture_desc(texture, block_dim, flags)
        out = self._output_for(desc, texture, None)
        t = PlainTuning()
        check(lib().astc_plain_autotune(self._ctx, ctypes.byref(desc), ctypes.c_void_p(texture.data_ptr()),
                                        ctypes.c_void_p(out.data_ptr()), _stream_handle(stream), int(reps),
                                        ctypes.byref(t)))
        return t.as_dict()


    def default_plain_tuning(self, texture, block_dim=4, flags=0):
        """astc_plain_default_tuning: the tuning encode() uses for this texture, as a dict."""
        desc = self._texture_desc(texture, block_dim, flags)
        t = PlainTuning()
        check(lib().astc_plain_default_tuning(self._ctx, ctypes.byref(desc), ctypes.byref(t)))
        return t.as_dict()

    def decode(self, blocks, width, height, block_dim_x=4, block_dim_y=None, out=None, stream=None):
        """Decode (nblocks, 16) device blocks to an (H, W, 4) uint8 device tensor."""
        import torch
        block_dim_y = block_dim_x if block_dim_y is None else block_dim_y
        nb = ((width + block_dim_x - 1) // block_dim_x) * ((height + block_dim_y - 1) // block_dim_y)
        if blocks.dtype != torch.uint8 or not blocks.is_cuda or blocks.numel() < nb * 16:
            raise ValueError(f"blocks must be a uint8 device tensor of at least {nb * 16} bytes")
        if not blocks.is_contiguous():
            blocks = blocks.contiguous()
            blocks.record_stream(_torch_stream(stream))
        if out is None:
            out = torch.empty((height, width, 4), dtype=torch.uint8, device=blocks.device)
        elif out.dtype != torch.uint8 or not out.is_contiguous() or out.numel() < height * width * 4:
            raise ValueError("out must be a contiguous uint8 tensor of at least height * width * 4 bytes")
        check(lib().astc_decode(self._ctx, int(width), int(height), int(block_dim_x), int(block_dim_y),
                                ctypes.c_void_p(blocks.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                _stream_handle(stream)))
        return out

    def downsample(self, level, stream=None):
        """Next mip level of a device (H, W, C) uint8 tensor (2x2 box filter)."""
        import torch
        if level.dtype != torch.uint8 or level.dim() != 3 or level.shape[2] not in _CHANNELS_TO_FMT or not level.is_cuda:
            raise ValueError("level must be an (H, W, 4|3|2) uint8 device tensor")
        h, w, c = level.shape
        desc = make_desc(w, h, _CHANNELS_TO_FMT[c], 4, 0)
        ow = w >> 1 if w >= 8 else w
        oh = h >> 1 if h >= 8 else h
        if not level.is_contiguous():
            level = level.contiguous()
            level.record_stream(_torch_stream(stream))
        out = torch.empty((oh, ow, c), dtype=torch.uint8, device=level.device)
        check(lib().astc_downsample(self._ctx, ctypes.byref(desc), ctypes.c_void_p(level.data_ptr()),
                                    ctypes.c_void_p(out.data_ptr()), _stream_handle(stream)))
        return out

    def encode_batch_noalloc(self, images, outputs, tuning=None, stream=None):
        """astc_encode_batch_noalloc: encode images [(device tensor (H, W, C)
        uint8, block_dim, flags), ...] into outputs [(nblocks, 16) uint8 device
        tensors] with the persistent kernel, allocating nothing (the work list
        lives in the library's fixed device array).  tuning: None or
        (blocks_per_lane, workgroups_per_cu, prefetch[, lanes_per_block])."""
        n = len(images)
        if len(outputs) != n:
            raise ValueError("one output per image")
        descs = (ImageDesc * n)()
        srcs = (ctypes.c_void_p * n)()
        dsts = (ctypes.c_void_p * n)()
        for i, ((tex, dim, flags), out) in enumerate(zip(images, outputs)):
            h, w, c = tex.shape
            if tex.stride(2) != 1 or tex.stride(1) != c:
                raise ValueError("image rows must have contiguous texels")
            descs[i] = make_desc(w, h, _CHANNELS_TO_FMT[c], dim, flags, row_pitch=tex.stride(0))
            if out.numel() < output_size(descs[i]) or not out.is_contiguous():
                raise ValueError(f"output {i} is too small or not contiguous")
            srcs[i] = tex.data_ptr()
            dsts[i] = out.data_ptr()
        t = None if tuning is None else ctypes.byref(
            Tuning(int(tuning[0]), int(tuning[1]), int(tuning[2]), 0, int(tuning[3]) if len(tuning) > 3 else 1))
        check(lib().astc_encode_batch_noalloc(self._ctx, descs, srcs, dsts, n, t, _stream_handle(stream)))

    def mip_chain(self, base, stream=None):
        """[base, level1, ...] while a side can halve, generated on the device."""
        levels = [base]
        while levels[-1].shape[0] >= 8 or levels[-1].shape[1] >= 8:
            levels.append(self.downsample(levels[-1], stream))
        return levels


def mip_chain_info(desc):
    n = ctypes.c_int()
    check(lib().astc_mip_chain_info(ctypes.byref(desc), 0, None, ctypes.byref(n)))
    arr = (ImageDesc * n.value)()
    check(lib().astc_mip_chain_info(ctypes.byref(desc), n.value, arr, ctypes.byref(n)))
    return list(arr)


class Batch:
    """astc_batch: many images encoded by one persistent kernel replayed as a graph.

    images: list of (device tensor (H, W, C) uint8, block_dim, flags).  Output
    tensors (one per image, (nblocks, 16) uint8) are allocated here and kept."""

    def __init__(self, encoder, images, outputs=None):
        import torch
        n = len(images)
        self._keep = []
        descs = (ImageDesc * n)()
        srcs = (ctypes.c_void_p * n)()
        dsts = (ctypes.c_void_p * n)()
        self.outputs = []
        for i, (tex, dim, flags) in enumerate(images):
            h, w, c = tex.shape
            if tex.stride(2) != 1 or tex.stride(1) != c:
                tex = tex.contiguous()
            self._keep.append(tex)
            descs[i] = make_desc(w, h, _CHANNELS_TO_FMT[c], dim, flags, row_pitch=tex.stride(0))
            nb = output_size(descs[i]) // 16
            if outputs is not None:
                out = outputs[i]
                if out.dtype != torch.uint8 or not out.is_contiguous() or out.numel() < nb * 16:
                    raise ValueError(f"output {i} must be a contiguous uint8 tensor of at least {nb * 16} bytes")
            else:
                out = torch.empty((nb, 16), dtype=torch.uint8, device=tex.device)
            self.outputs.append(out)
            srcs[i] = tex.data_ptr()
            dsts[i] = out.data_ptr()
        self._h = ctypes.c_void_p()
        check(lib().astc_batch_create(encoder.handle, descs, srcs, dsts, n, ctypes.byref(self._h)))
        self.encoder = encoder
        encoder._children.add(self)
        self._descs, self._srcs, self._dsts, self._n = descs, srcs, dsts, n

    def launch(self, stream=None):
        """One replay of the batch's graph (the persistent kernel over every image)."""
        check(lib().astc_batch_launch(self._h, _stream_handle(stream)))

    def launch_plain(self, stream=None):
        """The same images encoded by one plain launch each (astc_encode_each)."""
        check(lib().astc_encode_each(self.encoder.handle, self._descs, self._srcs, self._dsts, self._n,
                                     _stream_handle(stream)))

    def autotune(self, reps=3, stream=None):
        t = Tuning()
        check(lib().astc_batch_autotune(self._h, _stream_handle(stream), int(reps), ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, blocks_per_lane, workgroups_per_cu, prefetch, split_modes=0, lanes_per_block=1):
        t = Tuning(blocks_per_lane, workgroups_per_cu, prefetch, split_modes, lanes_per_block)
        check(lib().astc_batch_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self):
        t = Tuning()
        check(lib().astc_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def info(self):
        b, t, i = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        check(lib().astc_batch_info(self._h, ctypes.byref(b), ctypes.byref(t), ctypes.byref(i)))
        return {"blocks": b.value, "texels": t.value, "items": i.value}

    def close(self):
        if self._h:
            lib().astc_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _align256(n):
    return (n + 255) // 256 * 256


class PackedImages:
    """A request's images (e.g. one mip chain) packed back to back in one device
    buffer, image i at offset sum_{j<i} align256(bytes_j), the buffer ending with
    the last image's bytes: the layout astc_server_run_trace expects.  Also
    sizes the matching block layout (image after image, no gaps)."""

    def __init__(self, tensors, block_dim, flags):
        import torch
        n = len(tensors)
        self.descs = (ImageDesc * n)()
        offs, total, out_offs, out_total, end = [], 0, [], 0, 0
        for i, t in enumerate(tensors):
            h, w, c = t.shape
            self.descs[i] = make_desc(w, h, _CHANNELS_TO_FMT[c], block_dim, flags)
            offs.append(total)
            end = total + h * w * c
            total += _align256(h * w * c)
            out_offs.append(out_total)
            out_total += output_size(self.descs[i])
        self.buffer = torch.empty(end, dtype=torch.uint8, device=tensors[0].device)
        for t, o in zip(tensors, offs):
            self.buffer[o:o + t.numel()].copy_(t.contiguous().view(-1))
        self.nimages = n
        self.src_offsets = offs
        self.out_offsets = out_offs
        self.out_bytes = out_total
        self.texels = sum(int(t.shape[0]) * int(t.shape[1]) for t in tensors)
        self.block_counts = [output_size(d) // 16 for d in self.descs]

    def blocks_of(self, out, i):
        """Image i's (nblocks, 16) blocks from a destination buffer of out_bytes."""
        o = self.out_offsets[i]
        return out[o:o + self.block_counts[i] * 16].view(-1, 16)


class Server:
    """astc_server: continuous batching of requests under an HBM budget."""

    def __init__(self, encoder, hbm_budget_bytes, max_inflight=3, max_batch_requests=64,
                 tuning=(4, 2, 1, 0)):
        t = tuple(tuning) + (0, 1)[len(tuning) - 3:] if len(tuning) < 5 else tuple(tuning)
        cfg = ServerConfig(int(hbm_budget_bytes), int(max_inflight), int(max_batch_requests),
                           Tuning(t[0], t[1], t[2], t[3], t[4]))
        self._h = ctypes.c_void_p()
        check(lib().astc_server_create(encoder.handle, ctypes.byref(cfg), ctypes.byref(self._h)))
        encoder._children.add(self)

    def run_trace(self, requests):
        """requests: list of (PackedImages, arrival_s, dst tensor or None), arrival
        times non-decreasing.  Runs in real time on this thread; returns (latency
        seconds per request as a numpy array, stats dict)."""
        n = len(requests)
        arr = (Request * n)()
        for i, (p, t, dst) in enumerate(requests):
            arr[i] = Request(p.descs, p.nimages, 0, ctypes.c_void_p(p.buffer.data_ptr()),
                             ctypes.c_void_p(dst.data_ptr() if dst is not None else 0), float(t))
        lat = np.zeros(n, np.float64)
        st = ServerStats()
        check(lib().astc_server_run_trace(self._h, arr, n, lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          ctypes.byref(st)))
        return lat, st.as_dict()

    # ---- online serving with zero-copy admission ----
    def reserve(self, images):
        """astc_server_reserve for a PackedImages layout: returns (ticket, ring),
        where ring is a uint8 device tensor over the ring space for the sources
        (write them there, then submit), or None when the ring has no room now."""
        src = ctypes.c_void_p()
        ticket = ctypes.c_int64()
        rc = lib().astc_server_reserve(self._h, images.descs, images.nimages, ctypes.byref(src), ctypes.byref(ticket))
        if rc == -6 and b"no ring space now" in lib().astc_last_error():
            return None
        check(rc)
        return ticket.value, _device_view(src.value, images.buffer.numel(), images.buffer.device)

    def submit(self, ticket, stream=None):
        """Queue a reserved request; its sources are in place once the work
        queued on `stream` (default: the current stream) so far has run."""
        check(lib().astc_server_submit(self._h, int(ticket), _stream_handle(stream)))

    def poll(self, max_done=1024):
        """Complete finished batches and launch new ones; returns the tickets
        completed since the last call (at most max_done)."""
        done = (ctypes.c_int64 * max_done)()
        n = ctypes.c_int()
        check(lib().astc_server_poll(self._h, done, int(max_done), ctypes.byref(n)))
        return list(done[:n.value])

    def blocks(self, ticket, device=None):
        """A completed request's blocks in the ring (uint8 device tensor view,
        image after image), valid until release(ticket)."""
        import torch
        ptr = ctypes.c_void_p()
        nbytes = ctypes.c_int64()
        check(lib().astc_server_blocks(self._h, int(ticket), ctypes.byref(ptr), ctypes.byref(nbytes)))
        return _device_view(ptr.value, nbytes.value, device or torch.device("cuda", torch.cuda.current_device()))

    def release(self, ticket):
        check(lib().astc_server_release(self._h, int(ticket)))

    def close(self):
        if self._h:
            lib().astc_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DeviceBytes:
    """A device address range described by __cuda_array_interface__ (a view,
    owned elsewhere), for torch.as_tensor."""

    def __init__(self, ptr, nbytes):
        self.__cuda_array_interface__ = {"shape": (int(nbytes),), "typestr": "|u1", "data": (int(ptr), False),
                                         "version": 2, "strides": None}


def _device_view(ptr, nbytes, device):
    import torch
    return torch.as_tensor(_DeviceBytes(ptr, nbytes), device=device)


_default = {}


def _encoder_for(device_index):
    enc = _default.get(device_index)
    if enc is None:
        enc = _default[device_index] = Encoder(device_index)
    return enc


def encode_astc(texture, option):
    """encode_astc (astc_encode.h:87-194): encode an RGBA8 device texture with an
    EncodeOption.  The texture's rows are in memory order (row 0 first), as the
    reference's D3D11 texture after the loader's flip.  Returns the blocks as a
    device tensor; the caller owns it (the reference returns a new ID3D11Buffer)."""
    dev = texture.device.index if texture.device.index is not None else 0
    return _encoder_for(dev).encode(texture, option.block_dim, option.flags())


def save_astc(astc_path, xdim, ydim, xsize, ysize, buffer):
    """save_astc (astc_save.h:52-76): 16-byte header + blocks.  Raises on I/O errors."""
    try:
        import torch
        if isinstance(buffer, torch.Tensor):
            buffer = buffer.cpu().numpy()
    except ImportError:
        pass
    data = np.ascontiguousarray(np.asarray(buffer, dtype=np.uint8))
    check(lib().astc_write_file(str(astc_path).encode(), int(xdim), int(ydim), int(xsize), int(ysize),
                                data.ctypes.data_as(ctypes.c_void_p), data.nbytes))


def read_astc_header(path):
    bx, by, w, h = (ctypes.c_int() for _ in range(4))
    n = ctypes.c_int64()
    check(lib().astc_read_file_header(str(path).encode(), ctypes.byref(bx), ctypes.byref(by),
                                      ctypes.byref(w), ctypes.byref(h), ctypes.byref(n)))
    return bx.value, by.value, w.value, h.value, n.value


def load_tex(path, flip=True):
    """load_tex's image step (main.cpp:19-56): a PNG as (H, W, 4) uint8 RGBA,
    flipped vertically by default as the reference's loader does."""
    w, h = ctypes.c_int(), ctypes.c_int()
    buf = ctypes.POINTER(ctypes.c_uint8)()
    check(lib().astc_load_image(str(path).encode(), int(bool(flip)), ctypes.byref(w), ctypes.byref(h),
                                ctypes.byref(buf)))
    try:
        return np.ctypeslib.as_array(buf, shape=(h.value, w.value, 4)).copy()
    finally:
        lib().astc_free_image(buf)


def srgb_table():
    out = np.zeros(256, np.float32)
    check(lib().astc_srgb_table(out.ctypes.data_as(ctypes.c_void_p)))
    return out
