// astc_capi.cpp -- extern "C" entry points of libastc_mi355x.so (see include/astc_mi355x.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "astc_internal.h"

namespace astc {

namespace {
// A fixed buffer per thread: recording an error never allocates, so the error
// path of an entry point cannot itself fail (for instance after a host
// allocation has failed).  Longer messages are truncated.
thread_local char g_last_error[512];
}

const char* last_error_cstr() { return g_last_error; }

void set_error(const char* msg) { std::snprintf(g_last_error, sizeof(g_last_error), "%s", msg ? msg : ""); }

void set_error(const std::string& msg) { set_error(msg.c_str()); }

int fail(int code, const char* msg) {
    set_error(msg);
    return code;
}

int fail(int code, const std::string& msg) { return fail(code, msg.c_str()); }

int hip_fail(hipError_t e, const char* what) {
    std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
    return ASTC_ERR_HIP;
}

int check_desc(const astc_image_desc* d, ImageGeom* g) {
    if (!d) return fail(ASTC_ERR_INVALID_ARG, "image description is NULL");
    const int kMaxSize = (1 << 24) - 1;  // 24-bit sizes in the .astc header (astc_save.h:5-14)
    if (d->width < 1 || d->height < 1 || d->width > kMaxSize || d->height > kMaxSize)
        return fail(ASTC_ERR_INVALID_ARG, "width/height must be in 1..16777215");
    if (d->block_dim != 4 && d->block_dim != 6) return fail(ASTC_ERR_INVALID_ARG, "block_dim must be 4 or 6");
    int bpp = 0;
    switch (d->src_format) {
    case ASTC_FMT_RGBA8: bpp = 4; break;
    case ASTC_FMT_RGB8: bpp = 3; break;
    case ASTC_FMT_RG8: bpp = 2; break;
    default: return fail(ASTC_ERR_INVALID_ARG, "unknown src_format");
    }
    if ((int64_t)d->row_pitch < (int64_t)d->width * bpp)
        return fail(ASTC_ERR_INVALID_ARG, "row_pitch is smaller than width * bytes per texel");
    if (d->flags & ~(ASTC_FLAG_ALPHA | ASTC_FLAG_NORMAL | ASTC_FLAG_SRGB | ASTC_FLAG_MAX_ACCUMULATION))
        return fail(ASTC_ERR_INVALID_ARG, "unknown flag bits");
    const int xb = (d->width + d->block_dim - 1) / d->block_dim;
    const int yb = (d->height + d->block_dim - 1) / d->block_dim;
    const int64_t n = (int64_t)xb * yb;
    if (n > (int64_t)0x7FFFFF00) return fail(ASTC_ERR_UNSUPPORTED, "more than 2^31-256 blocks in one image");
    if (g) {
        g->xblocks = xb;
        g->yblocks = yb;
        g->nblocks = n;
        g->bpp = bpp;
    }
    return ASTC_OK;
}

void build_tables(DeviceTables* t) {
    for (int c = 0; c < 256; ++c) {
        const double u = c / 255.0;
        const double lin = u <= 0.04045 ? u / 12.92 : std::pow((u + 0.055) / 1.055, 2.4);
        t->srgb[c] = (float)lin * 255.0f;  // the texture unit's float, then the shader's x255 (:579)
    }
    std::memcpy(t->trit, kTritEncode.code, 256);
}

}  // namespace astc

using namespace astc;

extern "C" {

const char* astc_last_error(void) { return last_error_cstr(); }

int astc_ctx_create(int device, astc_ctx** out) {
    if (!out) return fail(ASTC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= count) return fail(ASTC_ERR_INVALID_ARG, "device index out of range");
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    astc_ctx* ctx = new (std::nothrow) astc_ctx();
    if (!ctx) return fail(ASTC_ERR_NO_MEMORY, "host allocation failed");
    ctx->device = device;
    if ((e = hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) {
        delete ctx;
        return hip_fail(e, "hipDeviceGetAttribute(multiprocessor count)");
    }
    build_tables(&ctx->h_tables);
    if ((e = hipMalloc(&ctx->d_tables, sizeof(DeviceTables))) != hipSuccess) {
        delete ctx;
        return hip_fail(e, "hipMalloc(tables)");
    }
    if ((e = hipMemcpy(ctx->d_tables, &ctx->h_tables, sizeof(DeviceTables), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMalloc(&ctx->d_dec_tables, sizeof(TritQuintDecodeTable))) != hipSuccess ||
        (e = hipMemcpy(ctx->d_dec_tables, &kDecodeTables, sizeof(TritQuintDecodeTable), hipMemcpyHostToDevice)) !=
            hipSuccess) {
        astc_ctx_destroy(ctx);
        return hip_fail(e, "context tables");
    }
    if (int rc = noalloc_setup(ctx); rc != ASTC_OK) {
        astc_ctx_destroy(ctx);
        set_error(std::string("context set-up: ") + astc_last_error());
        return rc;
    }
    *out = ctx;
    return ASTC_OK;
}

int astc_ctx_destroy(astc_ctx* ctx) {
    if (!ctx) return ASTC_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->d_tables) (void)hipFree(ctx->d_tables);
    if (ctx->d_dec_tables) (void)hipFree(ctx->d_dec_tables);
    noalloc_teardown(ctx);
    delete ctx;
    return ASTC_OK;
}

int64_t astc_output_size(const astc_image_desc* desc) {
    ImageGeom g;
    int rc = check_desc(desc, &g);
    if (rc != ASTC_OK) return rc;
    return g.nblocks * 16;
}

int astc_encode(astc_ctx* ctx, const astc_image_desc* desc, const void* src_dev, void* dst_dev, void* stream) {
    return guarded([&]() -> int { return encode_one(ctx, desc, src_dev, dst_dev, stream); });
}

int astc_encode_tuned(astc_ctx* ctx, const astc_image_desc* desc, const void* src_dev, void* dst_dev,
                      const astc_plain_tuning* tuning, void* stream) {
    return guarded([&]() -> int {
        if (!tuning) return fail(ASTC_ERR_INVALID_ARG, "tuning is NULL");
        return encode_one(ctx, desc, src_dev, dst_dev, stream, tuning);
    });
}

int astc_plain_autotune(astc_ctx* ctx, const astc_image_desc* desc, const void* src_dev, void* dst_dev, void* stream,
                        int reps, astc_plain_tuning* best) {
    return guarded([&]() -> int {
        if (!ctx || !desc || !best) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
        if (reps < 1) reps = 1;
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipEvent_t t0, t1;
        hipError_t e;
        if ((e = hipSetDevice(ctx->device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
        if ((e = hipEventCreate(&t0)) != hipSuccess) return hip_fail(e, "hipEventCreate");
        if ((e = hipEventCreate(&t1)) != hipSuccess) {
            (void)hipEventDestroy(t0);
            return hip_fail(e, "hipEventCreate");
        }
        float best_ms = 1e30f;
        int rc = ASTC_OK;
        // the flat-alpha test matters only where the four-channel encoder runs
        const bool four = desc->src_format == ASTC_FMT_RGBA8 && !(desc->flags & ASTC_FLAG_NORMAL) &&
                          !(desc->flags & ASTC_FLAG_MAX_ACCUMULATION);
        const bool maxacc = desc->flags & ASTC_FLAG_MAX_ACCUMULATION;
        // lanes 0 stands for one lane with packed texels
        for (int lanes : {1, 2, 0}) {
            if (lanes != 1 && (maxacc || (lanes == 2 && desc->block_dim != 6))) continue;
            for (int bpl : {1, 2, 4, 8})
                for (int pf : {0, 1})
                    for (int flat : {0, 1}) {
                    if (rc != ASTC_OK) break;
                    if (flat && (lanes != 1 || !four)) continue;
                    if (lanes == 0 && pf) continue;
                    const astc_plain_tuning t{lanes == 0 ? 1 : lanes, bpl, pf, flat, lanes == 0 ? 1 : 0};
                    if ((rc = encode_one(ctx, desc, src_dev, dst_dev, stream, &t)) != ASTC_OK) break;  // warm-up
                    if ((e = hipEventRecord(t0, s)) != hipSuccess) {
                        rc = hip_fail(e, "hipEventRecord");
                        break;
                    }
                    for (int r = 0; r < reps && rc == ASTC_OK; ++r) rc = encode_one(ctx, desc, src_dev, dst_dev, stream, &t);
                    if (rc != ASTC_OK) break;
                    if ((e = hipEventRecord(t1, s)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess) {
                        rc = hip_fail(e, "timing event");
                        break;
                    }
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, t0, t1);
                    if (ms < best_ms) {
                        best_ms = ms;
                        *best = t;
                    }
                }
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        return rc;
    });
}

int astc_plain_default_tuning(astc_ctx* ctx, const astc_image_desc* desc, astc_plain_tuning* tuning) {
    return guarded([&]() -> int {
        if (!ctx || !desc || !tuning) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
        ImageGeom g;
        const int rc = check_desc(desc, &g);
        if (rc != ASTC_OK) return rc;
        const PlainTuning t = default_plain_tuning(*desc, g.nblocks, ctx->num_cus);
        tuning->lanes_per_block = t.lanes_per_block;
        tuning->blocks_per_lane = t.blocks_per_lane;
        tuning->prefetch = t.prefetch;
        tuning->flat_alpha = t.flat_alpha;
        tuning->packed_texels = t.packed_texels;
        return ASTC_OK;
    });
}

int astc_decode(astc_ctx* ctx, int width, int height, int block_dim_x, int block_dim_y, const void* blocks_dev,
                void* rgba_dev, void* stream) {
    return guarded([&]() -> int {
        return decode_one(ctx, width, height, block_dim_x, block_dim_y, blocks_dev, rgba_dev, stream);
    });
}

}  // extern "C"

namespace astc {

int encode_one(astc_ctx* ctx, const astc_image_desc* desc, const void* src_dev, void* dst_dev, void* stream,
               const astc_plain_tuning* tuning) {
    if (!ctx) return fail(ASTC_ERR_INVALID_ARG, "ctx is NULL");
    PlainTuning pt{};
    if (tuning) {
        const int bpl = tuning->blocks_per_lane;
        if ((tuning->lanes_per_block != 1 && tuning->lanes_per_block != 2) ||
            !(bpl == 0 || bpl == 1 || bpl == 2 || bpl == 4 || bpl == 8) ||
            (tuning->prefetch != 0 && tuning->prefetch != 1) || (tuning->flat_alpha != 0 && tuning->flat_alpha != 1) ||
            (tuning->packed_texels != 0 && tuning->packed_texels != 1) ||
            (tuning->packed_texels && (tuning->lanes_per_block != 1 || tuning->prefetch || tuning->flat_alpha)))
            return fail(ASTC_ERR_INVALID_ARG, "plain tuning out of range");
        pt.lanes_per_block = tuning->lanes_per_block;
        pt.blocks_per_lane = bpl;
        pt.prefetch = tuning->prefetch;
        pt.flat_alpha = tuning->flat_alpha;
        pt.packed_texels = tuning->packed_texels;
    }
    ImageGeom g;
    int rc = check_desc(desc, &g);
    if (rc != ASTC_OK) return rc;
    if (!src_dev || !dst_dev) return fail(ASTC_ERR_INVALID_ARG, "src/dst is NULL");
    if (reinterpret_cast<uintptr_t>(dst_dev) & 15u) return fail(ASTC_ERR_INVALID_ARG, "dst must be 16-byte aligned");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    e = launch_encode_plain(ctx->d_tables, *desc, g, src_dev, dst_dev, static_cast<hipStream_t>(stream), ctx->num_cus,
                            tuning ? &pt : nullptr);
    if (e != hipSuccess) return hip_fail(e, "encode launch");
    return ASTC_OK;
}

int decode_one(astc_ctx* ctx, int width, int height, int block_dim_x, int block_dim_y, const void* blocks_dev,
               void* rgba_dev, void* stream) {
    if (!ctx || !blocks_dev || !rgba_dev) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
    if (width < 1 || height < 1 || width > (1 << 24) - 1 || height > (1 << 24) - 1)
        return fail(ASTC_ERR_INVALID_ARG, "width/height must be in 1..16777215");
    if (block_dim_x < 4 || block_dim_x > 12 || block_dim_y < 4 || block_dim_y > 12)
        return fail(ASTC_ERR_INVALID_ARG, "block dims must be 4..12");
    if ((reinterpret_cast<uintptr_t>(blocks_dev) & 15u) || (reinterpret_cast<uintptr_t>(rgba_dev) & 3u))
        return fail(ASTC_ERR_INVALID_ARG, "blocks must be 16-byte and rgba 4-byte aligned");
    const int64_t nb = (int64_t)((width + block_dim_x - 1) / block_dim_x) * ((height + block_dim_y - 1) / block_dim_y);
    if (nb > (int64_t)0x7FFFFF00) return fail(ASTC_ERR_UNSUPPORTED, "more than 2^31-256 blocks in one image");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    e = launch_decode(ctx->d_dec_tables, blocks_dev, rgba_dev, width, height, block_dim_x, block_dim_y,
                      static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "decode launch");
    return ASTC_OK;
}

}  // namespace astc

extern "C" {

int astc_encode_each(astc_ctx* ctx, const astc_image_desc* descs, const void* const* srcs, void* const* dsts,
                     int count, void* stream) {
    return guarded([&]() -> int {
        if (!descs || !srcs || !dsts || count < 0) return fail(ASTC_ERR_INVALID_ARG, "NULL arrays");
        for (int i = 0; i < count; ++i) {
            int rc = encode_one(ctx, &descs[i], srcs[i], dsts[i], stream);
            if (rc != ASTC_OK) return rc;
        }
        return ASTC_OK;
    });
}

int astc_write_file(const char* path, int block_dim_x, int block_dim_y, int width, int height,
                    const void* blocks_host, int64_t nbytes) {
    return guarded([&]() -> int {
        if (!path || !blocks_host || nbytes < 0) return fail(ASTC_ERR_INVALID_ARG, "path/blocks invalid");
        if (block_dim_x < 4 || block_dim_x > 12 || block_dim_y < 4 || block_dim_y > 12)
            return fail(ASTC_ERR_INVALID_ARG, "block dims must be 4..12");
        if (width < 1 || height < 1 || width > (1 << 24) - 1 || height > (1 << 24) - 1)
            return fail(ASTC_ERR_INVALID_ARG, "width/height must be in 1..16777215");
        const int64_t expect = 16LL * ((width + block_dim_x - 1) / block_dim_x) * ((height + block_dim_y - 1) / block_dim_y);
        if (nbytes != expect) return fail(ASTC_ERR_INVALID_ARG, "nbytes does not match the block count");
        uint8_t hdr[kAstcHeaderBytes];
        for (int i = 0; i < 4; ++i) hdr[i] = uint8_t(kAstcMagic >> (8 * i));
        hdr[4] = uint8_t(block_dim_x);
        hdr[5] = uint8_t(block_dim_y);
        hdr[6] = 1;
        for (int i = 0; i < 3; ++i) {
            hdr[7 + i] = uint8_t(width >> (8 * i));
            hdr[10 + i] = uint8_t(height >> (8 * i));
        }
        hdr[13] = 1;
        hdr[14] = 0;
        hdr[15] = 0;
        FILE* f = std::fopen(path, "wb");
        if (!f) return fail(ASTC_ERR_IO, std::string("cannot open for writing: ") + path);
        bool ok = std::fwrite(hdr, 1, sizeof(hdr), f) == sizeof(hdr) &&
                  std::fwrite(blocks_host, 1, (size_t)nbytes, f) == (size_t)nbytes;
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return fail(ASTC_ERR_IO, std::string("short write: ") + path);
        return ASTC_OK;
    });
}

int astc_read_file_header(const char* path, int* block_dim_x, int* block_dim_y, int* width, int* height,
                          int64_t* payload_bytes) {
    return guarded([&]() -> int {
        if (!path) return fail(ASTC_ERR_INVALID_ARG, "path is NULL");
        FILE* f = std::fopen(path, "rb");
        if (!f) return fail(ASTC_ERR_IO, std::string("cannot open: ") + path);
        uint8_t h[kAstcHeaderBytes];
        const size_t got = std::fread(h, 1, sizeof(h), f);
        std::fseek(f, 0, SEEK_END);
        const long size = std::ftell(f);
        std::fclose(f);
        if (got != sizeof(h)) return fail(ASTC_ERR_IO, "file shorter than the .astc header");
        uint32_t magic = uint32_t(h[0]) | uint32_t(h[1]) << 8 | uint32_t(h[2]) << 16 | uint32_t(h[3]) << 24;
        if (magic != kAstcMagic) return fail(ASTC_ERR_IO, "bad .astc magic");
        const int bx = h[4], by = h[5];
        const int w = h[7] | h[8] << 8 | h[9] << 16, ht = h[10] | h[11] << 8 | h[12] << 16;
        if (bx < 4 || by < 4 || w < 1 || ht < 1) return fail(ASTC_ERR_IO, "bad .astc header fields");
        const int64_t expect = 16LL * ((w + bx - 1) / bx) * ((ht + by - 1) / by);
        if (size - (long)sizeof(h) != expect) return fail(ASTC_ERR_IO, "payload size does not match the header");
        if (block_dim_x) *block_dim_x = bx;
        if (block_dim_y) *block_dim_y = by;
        if (width) *width = w;
        if (height) *height = ht;
        if (payload_bytes) *payload_bytes = expect;
        return ASTC_OK;
    });
}

int astc_mip_chain_info(const astc_image_desc* base, int max_levels, astc_image_desc* levels, int* count) {
    ImageGeom g;
    int rc = check_desc(base, &g);
    if (rc != ASTC_OK) return rc;
    if (!count) return fail(ASTC_ERR_INVALID_ARG, "count is NULL");
    int w = base->width, h = base->height, n = 0;
    for (;;) {
        if (levels && n < max_levels) {
            levels[n] = *base;
            levels[n].width = w;
            levels[n].height = h;
            levels[n].row_pitch = w * g.bpp;
        }
        ++n;
        if (w < 8 && h < 8) break;
        w = w >= 8 ? w >> 1 : w;
        h = h >= 8 ? h >> 1 : h;
    }
    *count = n;
    return ASTC_OK;
}

int astc_downsample(astc_ctx* ctx, const astc_image_desc* d, const void* src_dev, void* dst_dev, void* stream) {
    if (!ctx) return fail(ASTC_ERR_INVALID_ARG, "ctx is NULL");
    ImageGeom g;
    int rc = check_desc(d, &g);
    if (rc != ASTC_OK) return rc;
    if (d->row_pitch != d->width * g.bpp) return fail(ASTC_ERR_INVALID_ARG, "mip levels must be tightly packed");
    if (!src_dev || !dst_dev) return fail(ASTC_ERR_INVALID_ARG, "src/dst is NULL");
    if (d->width < 8 && d->height < 8) return fail(ASTC_ERR_INVALID_ARG, "no side can halve (both below 8)");
    const int ow = d->width >= 8 ? d->width >> 1 : d->width;
    const int oh = d->height >= 8 ? d->height >> 1 : d->height;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    e = launch_downsample(static_cast<const uint8_t*>(src_dev), d->width, d->height, g.bpp,
                          static_cast<uint8_t*>(dst_dev), ow, oh, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "downsample launch");
    return ASTC_OK;
}

int astc_srgb_table(float out[256]) {
    if (!out) return fail(ASTC_ERR_INVALID_ARG, "out is NULL");
    DeviceTables t;
    build_tables(&t);
    std::memcpy(out, t.srgb, sizeof(t.srgb));
    return ASTC_OK;
}

}  // extern "C"
