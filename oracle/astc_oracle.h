/*
 * astc_oracle.h -- CPU oracle for the ASTC block-encode path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the HIP encoder
 * is compared against; it is never linked into, loaded by, or called from the
 * product path (astc_encoder_amd/).  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.
 *
 * It restates the reference encoder (niepp/astc_encoder):
 *   ASTC_Encode.hlsl                     -- block encoder (PCA endpoints, weights, block assembly)
 *   ASTC_IntegerSequenceEncoding.hlsl    -- bounded integer sequence encoding (ISE)
 *   ASTC_Table.hlsl                      -- weight scramble table
 *   astc_encode.h:124-134, astc_save.h:5-14,52-76, main.cpp:19-56 -- dispatch grid, .astc header, load
 * and adds an independent ASTC-specification decoder (for PSNR and for
 * decode -> re-pack checks), which the reference does not have.
 *
 * Floating-point schedules (the reference leaves FP association unspecified:
 * it is compiled by FXC without IEEE strictness, which may fuse multiply-adds
 * and fold divisions by constants):
 *   ORACLE_SCHED_CANONICAL  the documented schedule the HIP kernel implements:
 *                           sequential sums, one fused multiply-add chain per
 *                           covariance entry, reciprocal multiplication for 1/N, a
 *                           Newton reciprocal square root for normalize, and
 *                           three exact rearrangements (power iteration on M M
 *                           with one length test, min/max as dot(texel, v) -
 *                           dot(mean, v), weights from texel projections).
 *   ORACLE_SCHED_SOURCE     the HLSL source order taken literally: sequential
 *                           sums, separate multiply and add, IEEE division.
 * Both run in binary32 (float) or binary64 (double, the FP64 twin).
 */
#ifndef ASTC_ORACLE_H
#define ASTC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_SCHED_CANONICAL = 0, ORACLE_SCHED_SOURCE = 1 };
enum { ORACLE_FMT_RGBA8 = 0, ORACLE_FMT_RGB8 = 1, ORACLE_FMT_RG8 = 2 };
enum { ORACLE_FLAG_ALPHA = 1, ORACLE_FLAG_NORMAL = 2, ORACLE_FLAG_SRGB = 4, ORACLE_FLAG_MAX_ACCUMULATION = 8 };
/* Endpoint search: the shader's principal component analysis (:139-168), or the
 * max_accumulation_pixel_direction alternative it leaves commented out (:170-227, :514). */
enum { ORACLE_METHOD_PCA = 0, ORACLE_METHOD_MAX_ACCUMULATION = 1 };

/* Per-block intermediate values, for known-answer tests and FP64 error statements. */
typedef struct {
    double mean[4];
    double axis[4];        /* principal axis returned by the power iteration */
    double e0[4], e1[4];   /* float endpoints after clamp/swap/alpha override */
    uint32_t ep_q[8];      /* quantised endpoints r0,r1,g0,g1,b0,b1,a0,a1 */
    uint32_t w_q[16];      /* quantised weights (linear index, before scramble) */
} oracle_block_trace;

/* Encode one block from fetched texels (texels[k*4+c], k = y*dim+x, values in [0,255]).
 * has_alpha selects CEM 12 / QUANT_6 weights vs CEM 8 / QUANT_12 weights. */
int oracle_encode_block_f32(const float* texels, int dim, int has_alpha, int sched,
                            uint8_t out[16], oracle_block_trace* trace);
int oracle_encode_block_f64(const double* texels, int dim, int has_alpha, int sched,
                            uint8_t out[16], oracle_block_trace* trace);

int oracle_encode_block_method_f32(const float* texels, int dim, int has_alpha, int sched, int method,
                                   uint8_t out[16], oracle_block_trace* trace);
int oracle_encode_block_method_f64(const double* texels, int dim, int has_alpha, int sched, int method,
                                   uint8_t out[16], oracle_block_trace* trace);

/* Fetch one block's texels exactly as MainCS does (ASTC_Encode.hlsl:561-580):
 * out-of-bounds texels are zero, normal maps force b = a = 1, then x255.
 * UNORM channels convert as binary32 c * (1/255) (the hardware conversion)
 * before the x255; sRGB colour channels through oracle_srgb_table. */
/* Binary32 against binary64 runs of the same method over every block of an
 * image, per block: axis error (unit axes compared up to sign; -1 where the
 * binary64 axis is zero), endpoint error (largest difference of the unrounded
 * endpoint channels, 8-bit units), and whether the blocks' bytes differ. */
void oracle_fp64_twin(const uint8_t* src, int width, int height, int row_pitch, int src_format, int dim, int flags,
                      int threads, double* axis_err, double* endpoint_err, uint8_t* differ);

void oracle_fetch_block_f32(const uint8_t* src, int width, int height, int row_pitch, int src_format,
                            int dim, int flags, int bx, int by, float* texels);
void oracle_fetch_block_f64(const uint8_t* src, int width, int height, int row_pitch, int src_format,
                            int dim, int flags, int bx, int by, double* texels);

/* Encode a whole image: output blocks in row-major block order, 16 bytes each.
 * threads <= 0 means all OpenMP threads.  Returns the block count or < 0. */
int64_t oracle_encode_image(const uint8_t* src, int width, int height, int row_pitch, int src_format,
                            int dim, int flags, int sched, int threads, uint8_t* out_blocks);

/* sRGB -> linear x255 table used for SRGB fetches (binary32). */
void oracle_srgb_table(float out[256]);
/* UNORM8 -> float x255 values of the binary32 fetch. */
void oracle_unorm_table(float out[256]);

/* ---- ISE and block assembly restatement (ASTC_IntegerSequenceEncoding.hlsl) ---- */
void oracle_tables(int32_t trits_out[243], int32_t quints_out[125], int32_t scramble_out[12 * 32],
                   int32_t btq_out[21 * 3]);
/* bise_weights / bise_endpoints; returns the spec bit count of the stream. */
int oracle_bise(const uint32_t* numbers, int count, int quant, uint32_t out[4]);
uint32_t oracle_blockmode(int weight_quant);
void oracle_assemble_block(uint32_t blockmode, uint32_t cem, const uint32_t ep_ise[4],
                           const uint32_t wt_ise[4], uint8_t out[16]);
/* Pack quantised endpoints (ep_q[8]) and linear weights (w_q[16]) into a block,
 * exactly as encode_block does after its floating-point stage. */
void oracle_pack_block(const uint32_t ep_q[8], const uint32_t w_q[16], int has_alpha, uint8_t out[16]);

/* ---- independent ASTC specification decoder ---- */
typedef struct {
    int valid;             /* 0 = error block (decoded as magenta) */
    int void_extent;
    int grid_w, grid_h, dual_plane, partitions, cem, weight_quant, endpoint_quant;
    uint32_t ep_q[8];      /* decoded endpoint ISE symbols, unquantised to 0..255 */
    uint32_t w_sym[64];    /* raw weight ISE symbols in grid order */
    uint32_t w_unq[64];    /* unquantised weights 0..64 in grid order */
} oracle_block_info;

/* Decode one block to RGBA8 texels (dim x dim, row-major, 4 bytes each). */
int oracle_decode_block(const uint8_t blk[16], int dim_x, int dim_y, uint8_t* rgba, oracle_block_info* info);
/* As oracle_decode_block, and each texel's infilled weight (0..64, spec C.2.18):
 * plane 0 at texel_w[t * dim_x + s], plane 1 of a dual-plane block after it. */
int oracle_decode_block_weights(const uint8_t blk[16], int dim_x, int dim_y, uint8_t* rgba, int32_t* texel_w);
/* The 6x6 sampling taps the encoder uses (idx_grids / wt_grids, ASTC_Encode.hlsl:268-304):
 * grid point g's four texel indices idx[4g..4g+3] and tap weights wt[4g..4g+3]. */
void oracle_grid6_taps(int32_t idx[64], float wt[64]);
/* Decode a whole image of blocks into a width x height RGBA8 image (row pitch = width*4). */
int oracle_decode_image(const uint8_t* blocks, int width, int height, int dim, uint8_t* rgba);
/* Spec ISE decoding of count values. */
int oracle_ise_decode(const uint8_t* bits_lsb_first, int bit_offset, int count, int quant, uint32_t* out);

/* PSNR of a decoded RGBA8 image against the encoder's fetched texels
 * (mode channels only: RGB, RGBA, or RG for normal maps). */
double oracle_psnr(const uint8_t* src, int width, int height, int row_pitch, int src_format, int flags,
                   const uint8_t* decoded_rgba);

/* Squared error sum and sample count behind oracle_psnr (for aggregating over textures). */
void oracle_sq_error(const uint8_t* src, int width, int height, int row_pitch, int src_format, int flags,
                     const uint8_t* decoded_rgba, double* sse, double* count);

/* The canonical schedule's reciprocal square root (bit-pattern seed + three Newton steps). */
float oracle_rsqrt_nr(float x);

/* Smallest binary32 x with sqrtf(x) >= 1e-5f: "length(v) < SMALL_VALUE" <=> dot(v,v) < this. */
float oracle_length_threshold(void);

#ifdef __cplusplus
}
#endif

#endif /* ASTC_ORACLE_H */
