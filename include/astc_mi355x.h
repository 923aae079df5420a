/*
 * astc_mi355x.h -- C ABI of the MI355X ASTC block encoder (libastc_mi355x.so).
 *
 * This library replaces the GPU path of niepp/astc_encoder:
 *   encode_astc(ID3D11Device*, ID3D11DeviceContext*, ID3D11Texture2D*, const encode_option&)
 *       (astc_encode.h:87-194) + the MainCS compute shader (ASTC_Encode.hlsl:553-582)
 *   save_astc(path, xdim, ydim, xsize, ysize, buffer, bufsz)       (astc_save.h:52-76)
 * with plain pointers: the caller owns device buffers and streams; the library
 * owns only its context (device tables) and plan objects.
 *
 * Conventions:
 *   - Every function returns ASTC_OK (0) or a negative ASTC_ERR_* code and never
 *     throws; astc_last_error() gives a message for the calling thread.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Calls are asynchronous with respect to the host unless stated otherwise.
 *   - Threads: a context is read-only after astc_ctx_create (except for
 *     astc_encode_batch_noalloc, which serialises its own use of it), so any
 *     number of threads may encode and decode through one context at once.  A batch or a
 *     server is used by one thread at a time (its calls change its plan, graph
 *     and ring).  Destroy a context only after every call using it has
 *     returned, and after its batches and servers are destroyed.
 *   - Output blocks are 16 bytes each in row-major block order
 *     (block (bx, by) at index by * ceil(W/D) + bx), exactly the contents of the
 *     reference's output UAV (RWStructuredBuffer<uint4>, ASTC_Encode.hlsl:81).
 */
#ifndef ASTC_MI355X_H
#define ASTC_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASTC_OK 0
#define ASTC_ERR_INVALID_ARG (-1)
#define ASTC_ERR_HIP (-2)
#define ASTC_ERR_NO_MEMORY (-3)
#define ASTC_ERR_UNSUPPORTED (-4)
#define ASTC_ERR_IO (-5)
#define ASTC_ERR_BUDGET (-6)

/* Source texel formats.  Narrower formats expand like a D3D11 texture Load:
 * RGB8 -> (r, g, b, 255), RG8 -> (r, g, 0, 255). */
#define ASTC_FMT_RGBA8 0
#define ASTC_FMT_RGB8 1
#define ASTC_FMT_RG8 2

/* encode_option (astc_encode.h:14-28) as runtime flags:
 *   ALPHA  has_alpha      -> CEM 12 (LDR RGBA direct), 6-level weights
 *          (otherwise CEM 8, LDR RGB direct, 12-level weights, alpha ignored)
 *   NORMAL is_normal_map  -> blue and alpha forced to 255 after the fetch
 *   SRGB   the source is sRGB-encoded (the reference's R8G8B8A8_UNORM_SRGB
 *          texture): colour channels are linearised on fetch.  The reference's
 *          command line never sets it together with NORMAL (main.cpp:214). */
#define ASTC_FLAG_ALPHA 1u
#define ASTC_FLAG_NORMAL 2u
#define ASTC_FLAG_SRGB 4u
/* Optional endpoint search: max_accumulation_pixel_direction (ASTC_Encode.hlsl:
 * 170-227), which the reference leaves commented out in favour of the principal
 * component analysis (:514).  Per channel, the sum of the texel differences from
 * the mean over the texels whose channel is above the mean; the longest sum is
 * the axis.  Every entry point accepts it; astc_batch_* and astc_server_* encode
 * such images with one plain launch each, after the persistent kernels (in the
 * batch's graph, or on the serving batch's stream). */
#define ASTC_FLAG_MAX_ACCUMULATION 8u

typedef struct astc_image_desc {
    int32_t width;      /* texels, 1 .. 2^24-1 (the .astc header's 24-bit size) */
    int32_t height;     /* texels, 1 .. 2^24-1 */
    int32_t row_pitch;  /* bytes between consecutive source rows, >= width * bytes per texel */
    int32_t src_format; /* ASTC_FMT_* */
    int32_t block_dim;  /* 4 (ASTC 4x4) or 6 (ASTC 6x6) */
    uint32_t flags;     /* ASTC_FLAG_* */
} astc_image_desc;

typedef struct astc_ctx astc_ctx;

/* Message for the last failing call on this thread ("" if none). */
const char* astc_last_error(void);

/* Create a context on HIP device `device` (uploads the sRGB and trit tables).
 * Replaces the reference's per-call shader compile (astc_encode.h:39-85, 92-105). */
int astc_ctx_create(int device, astc_ctx** out);
int astc_ctx_destroy(astc_ctx* ctx);

/* Bytes of encoded output: 16 * ceil(W/D) * ceil(H/D) (astc_encode.h:127-143),
 * or a negative error code for an invalid description. */
int64_t astc_output_size(const astc_image_desc* desc);

/* Encode one texture: replaces encode_astc (astc_encode.h:87-194) and its
 * Dispatch of MainCS (:190).  src_dev: source texels in device memory, row 0
 * first (the reference's texture rows, i.e. after the loader's vertical flip);
 * dst_dev: astc_output_size(desc) bytes of device memory, 16-byte aligned.
 * One kernel launch on `stream`. */
int astc_encode(astc_ctx* ctx, const astc_image_desc* desc, const void* src_dev, void* dst_dev, void* stream);

/* The plain kernel's tuning.  lanes_per_block: 1 (one lane encodes a block)
 * or 2 (two lanes share a 6x6 block, each holding three of its six texel rows;
 * other block sizes and the max-accumulation search run 1).  blocks_per_lane:
 * 1, 2, 4 or 8 (a workgroup covers 256 / lanes_per_block x this many blocks),
 * or 0 for the library's choice.  prefetch: 1 issues the next block's loads
 * before the current block is encoded.  flat_alpha: 1 tests each wave of
 * four-channel blocks (RGBA8 sources, one lane per block) for an alpha that is
 * the same integer-valued byte in every texel of every block, and encodes such
 * waves with the three-channel encoder (faster on opaque textures, about
 * neutral where few waves qualify; astc_encode uses it for four-channel
 * encodes).  packed_texels: 1 keeps each block's texels
 * as the fetched bytes and converts them again in every pass of the encode
 * (fewer registers, more waves per SIMD, more instructions); it needs
 * lanes_per_block 1, prefetch 0 and flat_alpha 0, and applies to the principal
 * component encodes (the max-accumulation search runs the default kernel).
 * Every tuning gives the same blocks. */
typedef struct astc_plain_tuning {
    int32_t lanes_per_block;
    int32_t blocks_per_lane;
    int32_t prefetch;
    int32_t flat_alpha;
    int32_t packed_texels;
} astc_plain_tuning;

/* astc_encode with an explicit plain-kernel tuning (astc_encode uses the
 * library's default for the image).  Same blocks as astc_encode. */
int astc_encode_tuned(astc_ctx* ctx, const astc_image_desc* desc, const void* src_dev, void* dst_dev,
                      const astc_plain_tuning* tuning, void* stream);
/* Time every plain tuning on this image (reps launches each after one warm-up,
 * on `stream`; synchronous; dst_dev is overwritten) and write the fastest to
 * *best.  No reference counterpart (the reference has one fixed dispatch). */
int astc_plain_autotune(astc_ctx* ctx, const astc_image_desc* desc, const void* src_dev, void* dst_dev, void* stream,
                        int reps, astc_plain_tuning* best);
/* The tuning astc_encode uses for this image (blocks_per_lane resolved; fields
 * that do not apply to the image's channel class or method are 0).  No
 * reference counterpart. */
int astc_plain_default_tuning(astc_ctx* ctx, const astc_image_desc* desc, astc_plain_tuning* tuning);

/* Decode blocks (row-major, ceil(W/bx) x ceil(H/by) of them, 16-byte aligned)
 * to a width x height RGBA8 image (4-byte aligned, row pitch width * 4).
 * Single-partition LDR ASTC: block sizes 4..12, any weight grid and
 * quantisation, colour endpoint modes 0/4/8/12, dual plane, LDR void extent;
 * other blocks decode to magenta.  Texel = ((C0*(64-w) + C1*w + 32) >> 6) >> 8
 * with C = endpoint * 257.  No reference counterpart (the reference only
 * encodes); used to verify encoded textures on the device. */
int astc_decode(astc_ctx* ctx, int width, int height, int block_dim_x, int block_dim_y, const void* blocks_dev,
                void* rgba_dev, void* stream);

/* The plain per-texture form over many images: astc_encode for each image in
 * turn (one kernel launch per image, issued from C++), on `stream`. */
int astc_encode_each(astc_ctx* ctx, const astc_image_desc* descs, const void* const* srcs, void* const* dsts,
                     int count, void* stream);

/* Write an .astc file: 16-byte header (magic 0x5CA1AB13, block dims, 24-bit
 * little-endian sizes, z = 1) followed by the blocks.  Replaces save_astc
 * (astc_save.h:52-76); unlike it, a failed open or short write is reported. */
int astc_write_file(const char* path, int block_dim_x, int block_dim_y, int width, int height,
                    const void* blocks_host, int64_t nbytes);

/* Read an .astc file header; returns ASTC_OK and fills the fields. */
int astc_read_file_header(const char* path, int* block_dim_x, int* block_dim_y, int* width, int* height,
                          int64_t* payload_bytes);

/* ---- mip chains ------------------------------------------------------------
 * Level l + 1 of a tightly packed image (row_pitch = width * bytes per texel):
 * a side of at least 8 halves (rounding down), a shorter side is kept; 2x2 box
 * filter in stored space, (a+b+c+d+2)>>2, or (a+b+1)>>1 along the one axis that
 * halves.  The chain ends when neither side can halve (4x4 for power-of-two
 * sides, i.e. max(4, w >> l) x max(4, h >> l)).
 * No reference counterpart (the reference encodes one level, main.cpp:36). */

/* Number of levels (including level 0) and, if `levels` is not NULL, each level's
 * description (block size and flags copied from `base`, row_pitch tight). */
int astc_mip_chain_info(const astc_image_desc* base, int max_levels, astc_image_desc* levels, int* count);
/* Write level l+1 from level l (both tightly packed, device memory) on `stream`. */
int astc_downsample(astc_ctx* ctx, const astc_image_desc* src_desc, const void* src_dev, void* dst_dev,
                    void* stream);

/* ---- batched persistent encoding ----------------------------------------
 * Many images (textures, mip levels; any mix of block sizes, formats and
 * flags) encoded by one persistent kernel fed by a device work list of block
 * ranges, replayed as a captured hipGraph.  No reference counterpart: the
 * reference encodes one texture per Dispatch (astc_encode.h:190); each image
 * here gets exactly the blocks astc_encode would give it. */
typedef struct astc_batch astc_batch;

typedef struct astc_tuning {
    int32_t blocks_per_lane;   /* 1, 2, 4 or 8: a work item is 256 x this many blocks of one image */
    int32_t workgroups_per_cu; /* persistent 256-thread workgroups per compute unit, 1..4 (with
                                  split_modes, of the first kernel; the second runs five) */
    int32_t prefetch;          /* 1: load the next block's texels while encoding the current one */
    int32_t split_modes;       /* 0: one kernel for every item; 1: two kernels with their own
                                  register budgets, one for 6x6 and four-channel 4x4 items and
                                  one, on a second stream, for 4x4 RGB8 and normal-map items */
    int32_t lanes_per_block;   /* lanes per 6x6 block: 1 (0 counts as 1), or 2: two lanes share each
                                  6x6 block (three texel rows each) and the kernel holding 6x6 items
                                  runs three waves per SIMD instead of two.  Same blocks either way */
} astc_tuning;

/* descs/srcs/dsts: `count` images (device pointers as for astc_encode).  The
 * batch keeps its own copy of the descriptions; buffers must outlive it. */
int astc_batch_create(astc_ctx* ctx, const astc_image_desc* descs, const void* const* srcs, void* const* dsts,
                      int count, astc_batch** out);
/* Rebuild the work list and graph for a tuning (astc_batch_create uses defaults). */
int astc_batch_set_tuning(astc_batch* batch, const astc_tuning* tuning);
int astc_batch_get_tuning(const astc_batch* batch, astc_tuning* tuning);
/* Time every candidate tuning (blocks per lane x workgroups per CU x prefetch x
 * split x lanes per 6x6 block; reps graph replays each, on `stream`,
 * synchronous) and keep the fastest; writes it to *best if not NULL. */
int astc_batch_autotune(astc_batch* batch, void* stream, int reps, astc_tuning* best);
/* Encode every image of the batch: one graph replay on `stream`. */
int astc_batch_launch(astc_batch* batch, void* stream);
/* Totals: blocks, real texels (sum of width * height) and work items. */
int astc_batch_info(const astc_batch* batch, int64_t* blocks, int64_t* texels, int32_t* items);
int astc_batch_destroy(astc_batch* batch);

/* Encode `count` images with the batched persistent kernel in one call that
 * allocates no device or host memory: the work list and queue heads live in a
 * fixed array in the library (4.5 MiB per GPU, held until the process ends),
 * fetched by a copy kernel from pinned staging the context allocated at
 * creation.  No
 * counterpart in the reference: this is for callers that cannot allocate
 * (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Images and outputs as for astc_batch_create; each image gets
 * exactly the blocks astc_encode gives it.  Limits: at most 8192 images and
 * 131072 work items (chunks of 256 x blocks_per_lane blocks; images under 256
 * blocks share items); larger batches are refused with ASTC_ERR_INVALID_ARG.
 * tuning: NULL for {4, 2, 1, 0, 1}; split_modes must be 0 (one kernel, one
 * stream).  Launches on `stream`, stream-ordered; calls on one GPU share the
 * array, so calls that overlap on the device (different streams) race, and
 * calls from several threads are serialised on the host. */
int astc_encode_batch_noalloc(astc_ctx* ctx, const astc_image_desc* descs, const void* const* srcs,
                              void* const* dsts, int count, const astc_tuning* tuning, void* stream);

/* ---- serving: continuous batching of compression requests ------------------
 * A request is a set of images (typically one texture's mip chain) whose
 * tightly packed sources lie back to back in device memory, image i at offset
 * sum over j < i of align256(width_j * height_j * bytes per texel); the buffer
 * needs to hold only up to the last image's last byte, and exactly that many
 * bytes are read.  The server
 * owns a device ring of `hbm_budget_bytes`: an admitted request's sources are
 * copied into it (one device-to-device copy) next to space for its blocks; a
 * request waits while the ring lacks room.  Up to `max_inflight` batches run at
 * once, each on its own stream, each one replay of the persistent kernel's
 * graph over every image of every request in it; a new batch is formed from
 * whatever has arrived as soon as a slot is free, so work that arrives while
 * other batches are on the GPU overlaps with them.  No reference counterpart. */
typedef struct astc_server astc_server;

typedef struct astc_server_config {
    int64_t hbm_budget_bytes;   /* ring size (sources + blocks of admitted requests) */
    int32_t max_inflight;       /* batches on the GPU at once, 1..8 */
    int32_t max_batch_requests; /* requests per batch, >= 1 */
    astc_tuning tuning;         /* persistent-kernel tuning */
} astc_server_config;

typedef struct astc_request {
    const astc_image_desc* images; /* nimages descriptions (row_pitch = width * bytes per texel) */
    int32_t nimages;
    int32_t reserved;
    const void* src_dev;           /* the images' texels, laid out as described above */
    void* dst_dev;                 /* NULL, or where the blocks are copied on completion: image after image
                                      with no gaps (image i's 16-byte blocks at the sum of the block bytes
                                      of images j < i), 16-byte aligned, exactly that many bytes written */
    double arrival_s;              /* arrival time, seconds after the start of the trace */
} astc_request;

typedef struct astc_server_stats {
    int64_t requests, images, texels, blocks, batches;
    double wall_s;               /* first arrival to last completion */
    double mean_batch_requests;
    int64_t peak_ring_bytes;     /* most ring bytes held at once */
    int32_t peak_inflight;       /* most batches on the GPU at once */
    int32_t reserved;
    double max_loop_gap_s;       /* longest pass of the serving loop (host stalls show here) */
} astc_server_stats;

/* Bytes the ring holds for one request: its sources and its blocks, each
 * rounded up to a multiple of 256. */
int64_t astc_request_bytes(const astc_image_desc* images, int nimages);
int astc_server_create(astc_ctx* ctx, const astc_server_config* config, astc_server** out);
/* Serve `n` requests (arrival_s non-decreasing) in real time on the calling
 * thread: each request is released at its arrival time, admitted when the ring
 * has room, batched, and completed; latency_s[i] = completion - arrival (host
 * steady clock, completion observed by polling the batch's event). */
int astc_server_run_trace(astc_server* server, const astc_request* requests, int n, double* latency_s,
                          astc_server_stats* stats);
int astc_server_destroy(astc_server* server);

/* ---- online serving with zero-copy admission --------------------------------
 * The caller writes each request's sources straight into ring space the server
 * hands out, so admission copies nothing, and reads the blocks where they were
 * encoded.  The server is driven by astc_server_poll from the caller's own
 * loop (one thread at a time; astc_server_run_trace refuses to run while online
 * requests are outstanding).  No reference counterpart. */

/* Reserve ring space for a request (sources laid out as for astc_request.src_dev,
 * then its blocks).  *src_ring: where the caller writes the sources (device
 * memory); *ticket: the request's id.  ASTC_ERR_BUDGET when the ring has no room
 * now (complete batches with astc_server_poll and release tickets, then retry)
 * or the request exceeds the whole budget. */
int astc_server_reserve(astc_server* server, const astc_image_desc* images, int nimages, void** src_ring,
                        int64_t* ticket);
/* Queue a reserved request for the next batch.  Its sources are taken to be in
 * place once the work queued on `stream` so far has run (NULL: already in place). */
int astc_server_submit(astc_server* server, int64_t ticket, void* stream);
/* Complete finished batches and launch new ones from the submitted requests
 * (up to max_inflight batches of up to max_batch_requests each), without
 * waiting.  Writes up to max_done completed tickets to done[] and their count
 * to *ndone; completions beyond max_done are reported by later calls. */
int astc_server_poll(astc_server* server, int64_t* done, int max_done, int* ndone);
/* A completed request's blocks in the ring, image after image with no gaps
 * (16-byte aligned), valid until the ticket is released. */
int astc_server_blocks(astc_server* server, int64_t ticket, const void** blocks_dev, int64_t* nbytes);
/* Return a completed (or reserved, never submitted) request's ring space. */
int astc_server_release(astc_server* server, int64_t ticket);

/* ---- image loading for the command-line drop-in ----------------------------
 * Replaces load_tex's stbi_load(path, ..., STBI_rgb_alpha) with
 * stbi_set_flip_vertically_on_load(1) (main.cpp:19-56) for PNG input: every PNG
 * colour type and bit depth, expanded to 8-bit RGBA as stb_image does.  *rgba is
 * width * height * 4 bytes, released with astc_free_image. */
int astc_load_image(const char* path, int flip_vertically, int* width, int* height, uint8_t** rgba);
void astc_free_image(uint8_t* rgba);

/* Copy the library's sRGB table (256 floats: float(linear(c/255)) * 255.0f). */
int astc_srgb_table(float out[256]);

#ifdef __cplusplus
}
#endif

#endif /* ASTC_MI355X_H */
