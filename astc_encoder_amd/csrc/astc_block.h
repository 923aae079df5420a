// astc_block.h -- one ASTC block per lane: fetch, PCA endpoints, weights, ISE, block.
//
// Device code shared by the per-texture kernel and the persistent batched
// kernel.  It computes what ASTC_Encode.hlsl:510-550 (encode_block) computes,
// in a fixed floating-point schedule ("canonical schedule") that is reproduced
// bit for bit by the CPU oracle:
//   * sums over a block run left to right over the texels in row-major order;
//   * sums of products are one fused multiply-add chain over the texels,
//     starting with a plain product;
//   * dot products are fma(a3,b3, fma(a2,b2, fma(a1,b1, a0*b0)));
//   * "x * s + t" expressions of the shader are single fused multiply-adds;
//   * divisions by the constants N and N-1 are multiplications by the
//     correctly rounded reciprocals (what FXC emits for a literal divisor);
//   * normalize(v) = v * rsqrt_nr(dot(v, v)), a reciprocal square root built
//     from IEEE operations only: seed 0x5f375a86 - (bits(x) >> 1), then three
//     Newton steps y = y * fma(-(0.5x * y), y, 1.5) (within 4 ulp; the shader's
//     own rsq is an approximate hardware instruction too);
//   * length(v) < 1e-5 is the equivalent exact test on dot(v, v);
//   * three exact rearrangements of the shader's formulas save a quarter of
//     the arithmetic: after its first iteration the power iteration runs on
//     S = M M without the length test (which can no longer succeed), min/max
//     projections are dot(texel, v) minus dot(mean, v), and grid weights
//     rescale texel projections (dot(k, e0) cancels in the rescaling).  Each is
//     explained where it is used.
// The file must be compiled with -ffp-contract=off so that nothing else fuses.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "astc_format.h"

namespace astc {

enum : uint32_t { FLAG_ALPHA = 1u, FLAG_NORMAL = 2u, FLAG_SRGB = 4u, FLAG_MAX_ACCUMULATION = 8u };
enum : int { FMT_RGBA8 = 0, FMT_RGB8 = 1, FMT_RG8 = 2 };

// Tables staged into LDS by every workgroup.
struct DeviceTables {
    float srgb[256];    // sRGB byte -> float(linear) * 255.0f
    uint8_t trit[256];  // 5-trit tuple index -> packed 8-bit code
};

struct Block128 { uint32_t w[4]; };

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

__device__ __forceinline__ float dot4(const float a[4], const float b[4]) {
    return fma_(a[3], b[3], fma_(a[2], b[2], fma_(a[1], b[1], a[0] * b[0])));
}

// dot4 over the first NC channels.  When the other channels hold zero in one
// operand (the products there are zeros) this is dot4's value: adding a zero
// product changes at most the sign of a zero result.
template <int NC>
__device__ __forceinline__ float dotn(const float a[4], const float b[4]) {
    float r = a[0] * b[0];
#pragma unroll
    for (int c = 1; c < NC; ++c) r = fma_(a[c], b[c], r);
    return r;
}

// 1 / sqrt(x) from IEEE operations only (bit-pattern seed + three Newton steps),
// reproduced bit for bit by the CPU oracle.
__device__ __forceinline__ float rsqrt_nr(float x) {
    float y = __uint_as_float(0x5f375a86u - (__float_as_uint(x) >> 1));
    const float h = 0.5f * x;
    y = y * fma_(-(h * y), y, 1.5f);
    y = y * fma_(-(h * y), y, 1.5f);
    y = y * fma_(-(h * y), y, 1.5f);
    return y;
}

// frexp exponent: x = m * 2^e with 0.5 <= |m| < 1 (x finite, nonzero).
__device__ __forceinline__ int exponent_of(float x) { return __builtin_amdgcn_frexp_expf(x); }
__device__ __forceinline__ float ldexp_(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }

__device__ __forceinline__ bool is_short(float dot) { return dot < __uint_as_float(kLenThresholdBits); }

__device__ __forceinline__ float clamp255(float x) { return fminf(fmaxf(x, 0.0f), 255.0f); }

// Issue priority of the wave (s_setprio; 0 is the default).  The encode raises
// it to 1 for its short serial phases (power iteration; endpoints and weight
// axis; weight quantisation and packing), whose dependent instructions leave
// issue slots that the other wave of the SIMD then fills, and keeps 0 for the
// passes over the texels, which have independent work for every texel.  The
// kernels' loops raise it to 2 from a block's store to the end of the next
// block's conversion (memory and table reads).  Measured 5% faster on 6x6
// textures than the loops' phase alone, and 9% faster than no priorities.
__device__ __forceinline__ void issue_priority(int p) {
    switch (p) {
    case 0: __builtin_amdgcn_s_setprio(0); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    default: __builtin_amdgcn_s_setprio(2); break;
    }
}

// ---------------------------------------------------------------------------
// Fetch (ASTC_Encode.hlsl:561-580 + D3D11 Load): out-of-bounds texels are zero;
// RGB8 expands with alpha 255 and RG8 with blue 0, alpha 255; sRGB converts the
// colour channels through its table, UNORM channels as binary32 c * (1/255)
// times 255 (the hardware conversion, through a table too); normal maps force
// blue = alpha = 255.
// ---------------------------------------------------------------------------
struct FetchArgs {
    const uint8_t* src;
    int width, height, row_pitch, fmt;
    uint32_t flags;
    int aligned;        // src and row_pitch are multiples of 4
    uint32_t or_mask;   // 0xFFFF0000 for normal maps (blue = alpha = 255), else 0
    int xblocks;        // blocks per block row
    uint32_t xb_magic;  // block id / xblocks = mulhi(2 id, xb_magic) >> xb_shift
    int xb_shift;
};

// OR-mask that applies the normal-map override to a packed RGBA8 texel.  Setting
// the bytes to 255 before conversion equals forcing the floats to 255 after it
// (ASTC_Encode.hlsl:575-579): the sRGB and UNORM tables both map 255 to exactly
// 255.0f, and out-of-bounds (zero) texels get blue = alpha = 255 as well.
__host__ __device__ inline uint32_t normal_or_mask(uint32_t flags) { return (flags & FLAG_NORMAL) ? 0xFFFF0000u : 0u; }

// Division of a block id n < 2^31 by d = xblocks without a divide: with
// l = ceil(log2 d) and m = ceil(2^(31+l) / d) (which fits 32 bits),
// floor(n m / 2^(31+l)) = floor(n / d) for every n < 2^31, because
// m d - 2^(31+l) < d <= 2^l keeps the error term below 1/d (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994).
// mulhi(2n, m) >> l is that quotient.
__host__ inline void set_block_row(FetchArgs* f, int xblocks) {
    int l = 0;
    while ((uint64_t(1) << l) < uint64_t(xblocks)) ++l;
    f->xblocks = xblocks;
    f->xb_shift = l;
    f->xb_magic = uint32_t(((uint64_t(1) << (31 + l)) + uint64_t(xblocks) - 1) / uint64_t(xblocks));
}

__device__ __forceinline__ int block_row_of(const FetchArgs& f, int id) {
    return int(__umulhi(uint32_t(id) << 1, f.xb_magic) >> f.xb_shift);
}

// Channel classes: which channels are 255 in every texel of every block.
//   CH_ALL    any content: four channels;
//   CH_RGB8   an RGB8 image tiled exactly by its blocks (no texel outside the
//             image, where alpha would be 0): alpha is constant, three channels;
//   CH_NORMAL normal maps (blue = alpha = 255, outside the image too): two.
enum : int { CH_ALL = 4, CH_RGB8 = 3, CH_NORMAL = 2 };

// A block's texels as the encoder receives them from the fetch, by class:
//   CH_ALL    one packed RGBA8 word per texel;
//   CH_RGB8   each texel row's RGB bytes back to back (the source bytes);
//   CH_NORMAL each texel row's red and green bytes back to back (for an RG8
//             source, the source bytes).
// Texel k of row r = k / DIM, column i = k % DIM, has channel c at byte
// i * BPT + c of its row's WPR words.  The conversion picks the bytes out of the
// words with byte selects, so RGB8 and RG8 rows need no realignment.
template <int DIM, int CH>
struct Layout {
    static constexpr int BPT = CH == CH_RGB8 ? 3 : CH == CH_NORMAL ? 2 : 4;  // bytes per texel
    static constexpr int WPR = (DIM * BPT + 3) / 4;                           // words per row
    static constexpr int WORDS = DIM * WPR;
};

template <int DIM, int CH>
__device__ __forceinline__ uint32_t layout_byte(const uint32_t* w, int k, int c) {
    using L = Layout<DIM, CH>;
    const int j = (k % DIM) * L::BPT + c;
    return (w[(k / DIM) * L::WPR + (j >> 2)] >> (8 * (j & 3))) & 0xFFu;
}

__host__ __device__ inline int bytes_per_texel(int fmt) { return fmt == FMT_RGBA8 ? 4 : fmt == FMT_RGB8 ? 3 : 2; }

// Packed RGBA8 of texel (x, y) with bounds check (slow path).
__device__ __forceinline__ uint32_t load_texel_guarded(const FetchArgs& a, int x, int y) {
    if (x >= a.width || y >= a.height) return 0u;
    const uint8_t* p = a.src + (size_t)y * a.row_pitch + (size_t)x * bytes_per_texel(a.fmt);
    if (a.fmt == FMT_RGBA8) return uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24;
    if (a.fmt == FMT_RGB8) return uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | 0xFF000000u;
    return uint32_t(p[0]) | uint32_t(p[1]) << 8 | 0xFF000000u;
}

// One texel row of an RGB8 or RG8 source as packed RGBA8 words, from
// 4-byte-aligned words around a row that starts at `row`.
template <int DIM>
__device__ __forceinline__ void load_row_expand(const FetchArgs& a, const uint8_t* row, uint32_t out[DIM]) {
    // the row's bytes start `shift` bytes into the first word; the words are
    // realigned with a funnel shift so that byte j sits at a fixed place
    const int bpp = bytes_per_texel(a.fmt);
    const uint32_t shift = uint32_t(reinterpret_cast<uintptr_t>(row) & 3u);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(row - shift);
    constexpr int kWords = (DIM * 3 + 3) / 4 + 1;
    const int nwords = int((DIM * bpp + shift + 3) >> 2);
    uint32_t raw[kWords + 1];
#pragma unroll
    for (int i = 0; i < kWords; ++i) raw[i] = i < nwords ? w[i] : 0u;
    raw[kWords] = 0u;
    uint32_t al[kWords];
#pragma unroll
    for (int i = 0; i < kWords; ++i) al[i] = __builtin_amdgcn_alignbyte(raw[i + 1], raw[i], shift);
    auto byte_at = [&](int j) -> uint32_t { return (al[j >> 2] >> ((j & 3) * 8)) & 0xFFu; };
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        if (a.fmt == FMT_RGB8)
            out[i] = byte_at(3 * i) | byte_at(3 * i + 1) << 8 | byte_at(3 * i + 2) << 16 | 0xFF000000u;
        else
            out[i] = byte_at(2 * i) | byte_at(2 * i + 1) << 8 | 0xFF000000u;
    }
}

// NR texel rows of a fully in-bounds block, the first starting at `row`, in
// the class's layout.  The source format is tested once per block.
template <int DIM, int CH, int NR = DIM>
__device__ __forceinline__ void load_rows_fast(const FetchArgs& a, const uint8_t* row, uint32_t* out) {
    using L = Layout<DIM, CH>;
    const int pitch = a.row_pitch;
    if (CH == CH_RGB8 && DIM == 4) {
        // 12 bytes at 12 bx past an aligned row start: three aligned words
#pragma unroll
        for (int r = 0; r < NR; ++r, row += pitch)
#pragma unroll
            for (int i = 0; i < L::WPR; ++i) out[r * L::WPR + i] = reinterpret_cast<const uint32_t*>(row)[i];
    } else if (CH == CH_RGB8) {
        // 18 bytes at 18 bx: they start 0 or 2 bytes into an aligned word
        const uint32_t shift = uint32_t(reinterpret_cast<uintptr_t>(row) & 3u);
        row -= shift;
#pragma unroll
        for (int r = 0; r < NR; ++r, row += pitch) {
            const uint32_t* w = reinterpret_cast<const uint32_t*>(row);
            uint32_t raw[L::WPR + 1];
#pragma unroll
            for (int i = 0; i < L::WPR; ++i) raw[i] = w[i];
            raw[L::WPR] = 0u;
#pragma unroll
            for (int i = 0; i < L::WPR; ++i) out[r * L::WPR + i] = __builtin_amdgcn_alignbyte(raw[i + 1], raw[i], shift);
        }
    } else if (a.fmt == (CH == CH_NORMAL ? FMT_RG8 : FMT_RGBA8)) {
        // the source bytes are the layout: RGBA8 words, or RG8 pairs of a normal
        // map (8 bytes at 8 bx, 12 at 12 bx), at aligned addresses
#pragma unroll
        for (int r = 0; r < NR; ++r, row += pitch)
#pragma unroll
            for (int i = 0; i < L::WPR; ++i) out[r * L::WPR + i] = reinterpret_cast<const uint32_t*>(row)[i];
    } else {
#pragma unroll
        for (int r = 0; r < NR; ++r, row += pitch) {
            uint32_t t[DIM];
            if (a.fmt == FMT_RGBA8) {
#pragma unroll
                for (int i = 0; i < DIM; ++i) t[i] = reinterpret_cast<const uint32_t*>(row)[i];
            } else {
                load_row_expand<DIM>(a, row, t);
            }
            if (CH == CH_ALL) {
#pragma unroll
                for (int i = 0; i < DIM; ++i) out[r * DIM + i] = t[i];
            } else {
                // normal map in a three- or four-byte format: keep red and green
#pragma unroll
                for (int i = 0; i < DIM; i += 2) out[r * L::WPR + i / 2] = (t[i] & 0xFFFFu) | t[i + 1] << 16;
            }
        }
    }
}

// Load a block's texels in the class's layout (loads only; no conversion), so
// that a caller can issue the next block's loads before encoding this one:
// NR rows from row r0 of the block (all of them by default).
template <int DIM, int CH, int NR = DIM>
__device__ __forceinline__ void load_block(const FetchArgs& a, int bx, int by, uint32_t p[NR * Layout<DIM, CH>::WPR],
                                           int r0 = 0) {
    using L = Layout<DIM, CH>;
    const int x0 = bx * DIM, y0 = by * DIM;
    const int bpp = bytes_per_texel(a.fmt);
    // Fast path: the block is inside the image and the 4-byte words covering its
    // row span end inside the row's texel bytes, so no load reaches past the image.
    const bool fast = a.aligned && (x0 + DIM <= a.width) && (y0 + DIM <= a.height) &&
                      (((x0 + DIM) * bpp + 3) & ~3) <= a.width * bpp;
    if (fast) {
        load_rows_fast<DIM, CH, NR>(
            a, a.src + (size_t)(uint32_t)(y0 + r0) * (uint32_t)a.row_pitch + uint32_t(x0 * bpp), p);
    } else {
        if (CH != CH_ALL) {
#pragma unroll
            for (int i = 0; i < NR * L::WPR; ++i) p[i] = 0u;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
                const uint32_t t = load_texel_guarded(a, x0 + i, y0 + r0 + r);
                if (CH == CH_ALL) {
                    p[r * DIM + i] = t;
                } else {
#pragma unroll
                    for (int c = 0; c < L::BPT; ++c) {
                        const int j = i * L::BPT + c;
                        p[r * L::WPR + (j >> 2)] |= ((t >> (8 * c)) & 0xFFu) << (8 * (j & 3));
                    }
                }
            }
    }
}

// Layout -> floats (x255 scale) through LDS tables: `colour` converts the colour
// channels (sRGB or UNORM), `unorm` the alpha.  Only the class's channels are
// converted (the others are known constants to the encoder, see encode_block).
// or_mask (CH_ALL only) applies the normal-map override to the words.  NR rows
// of texels (all of the block's by default).
template <int DIM, int CH, int NR = DIM>
__device__ __forceinline__ void unpack_block(const uint32_t p[NR * Layout<DIM, CH>::WPR], uint32_t or_mask,
                                             const float* colour, const float* unorm, float x[NR * DIM][4]) {
#pragma unroll
    for (int k = 0; k < NR * DIM; ++k) {
        if (CH == CH_ALL) {
            const uint32_t w = p[k] | or_mask;
#pragma unroll
            for (int c = 0; c < 3; ++c) x[k][c] = colour[(w >> (8 * c)) & 0xFFu];
            x[k][3] = unorm[w >> 24];
        } else {
#pragma unroll
            for (int c = 0; c < CH; ++c) x[k][c] = colour[layout_byte<DIM, CH>(p, k, c)];
        }
    }
}

// A block's texels kept as the fetched words (Layout<DIM, CH>) and converted
// at every read through the LDS tables, as unpack_block converts them (the
// same floats): a lane holds 1-4 bytes per texel instead of four floats.
// encode_block reads the texels in four passes (sums, covariance, extreme
// projections, weight projections); texel_pass between them makes each pass
// convert again instead of keeping the previous pass's floats live.
template <int DIM, int CH>
struct PackedTexels {
    uint32_t p[Layout<DIM, CH>::WORDS];
    const float* colour;  // colour-channel table (sRGB or UNORM)
    const float* unorm;   // alpha table
    struct Texel {
        float v[4];
        __device__ __forceinline__ operator const float*() const { return v; }
    };
    __device__ __forceinline__ Texel operator[](int k) const {
        Texel t;
        if (CH == CH_ALL) {
            const uint32_t w = p[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) t.v[c] = colour[(w >> (8 * c)) & 0xFFu];
            t.v[3] = unorm[w >> 24];
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) t.v[c] = c < CH ? colour[layout_byte<DIM, CH>(p, k, c)] : 255.0f;
        }
        return t;
    }
};

// Between two passes over a block's texels: nothing for converted floats; for
// packed words a compiler memory barrier, so that the next pass reloads the
// table entries rather than reusing (and keeping live) the last pass's.
template <typename TX>
__device__ __forceinline__ void texel_pass(const TX&) {}
template <int DIM, int CH>
__device__ __forceinline__ void texel_pass(const PackedTexels<DIM, CH>&) {
    asm volatile("" ::: "memory");
}

// After texel k of a pass, for packed words, at the end of each texel row: the
// same barrier, with the pass's running values v[0..NV) (sums, extremes,
// projections) taken through it, so that each row's table reads are consumed
// before the next row's are issued (without it a pass's reads are all issued,
// and held, at once).
template <int NV, typename TX>
__device__ __forceinline__ void texel_step(const TX&, int, float*) {}
template <int NV, int DIM, int CH>
__device__ __forceinline__ void texel_step(const PackedTexels<DIM, CH>&, int k, float* v) {
    if (k % DIM == DIM - 1) {
#pragma unroll
        for (int i = 0; i < NV; ++i) asm volatile("" : "+v"(v[i]));
        asm volatile("" ::: "memory");
    }
}

// Per-channel sums of the first NC channels over texels [0, H) and [H, N),
// added (one chain of additions per channel and range, in texel order).
template <int NC, int H, int N>
__device__ __forceinline__ void block_sums(const float (&x)[N][4], float sum[4]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        float r = x[0][c];
#pragma unroll
        for (int k = 1; k < H; ++k) r = r + x[k][c];
        if (H < N) {
            float r2 = x[H][c];
#pragma unroll
            for (int k = H + 1; k < N; ++k) r2 = r2 + x[k][c];
            r = r + r2;
        }
        sum[c] = r;
    }
}
// The same sums texel by texel (each texel converted once per pass).
template <int NC, int H, int N, int DIM, int CH>
__device__ __forceinline__ void block_sums(const PackedTexels<DIM, CH>& x, float sum[4]) {
    float r[4], r2[4];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const typename PackedTexels<DIM, CH>::Texel t = x[k];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (k == 0) r[c] = t.v[c];
            else if (k < H) r[c] = r[c] + t.v[c];
            else if (k == H) r2[c] = t.v[c];
            else r2[c] = r2[c] + t.v[c];
        }
        texel_step<NC>(x, k, k < H ? r : r2);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) sum[c] = H < N ? r[c] + r2[c] : r[c];
}

// ---------------------------------------------------------------------------
// Integer stage: endpoint bytes + weight trits -> 128-bit block
// (ASTC_Encode.hlsl:400-444, 475-508; ASTC_IntegerSequenceEncoding.hlsl:142-176, 207-276)
// ---------------------------------------------------------------------------
// Bits of a 5-value trit group that come from the packed trit code T (8 bits),
// placed for b low bits per value: m0 T[1:0] m1 T[3:2] m2 T[4] m3 T[6:5] m4 T[7].
__host__ __device__ constexpr uint32_t trit_group_spread(uint32_t T, int b) {
    return (T & 3u) << b | ((T >> 2) & 3u) << (2 * b + 2) | ((T >> 4) & 1u) << (3 * b + 4) |
           ((T >> 5) & 3u) << (4 * b + 5) | (T >> 7) << (5 * b + 7);
}

// Pack tables in LDS: for each 5-trit tuple index, the spread code bits of its
// group for QUANT_6 (b = 1) and QUANT_12 (b = 2) weights.
struct PackTables { uint32_t spread[2][256]; };

// wsh[i]: twice weight i's quantised value in the low five bits (the field
// offset into the split tables).
// swap: the endpoints go in the other order (e1 first), which exchanges the two
// bytes of every endpoint pair.
template <bool ALPHA>
__device__ __forceinline__ Block128 pack_block(const uint32_t e0q[4], const uint32_t e1q[4], const uint32_t wsh[16],
                                               const PackTables* pt, bool swap) {
    constexpr WeightSplit split = ALPHA ? kSplitQ6 : kSplitQ12;
    constexpr int b = split.bits;          // bits per weight below the trit
    constexpr int group_bits = 5 * b + 8;  // one 5-value trit group
    const uint32_t* spread = pt->spread[ALPHA ? 0 : 1];
    // weight stream, LSB first; at most 4 groups x 18 bits (the fourth only
    // carries value 15, so it ends below bit 64)
    uint64_t stream = 0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        uint32_t tidx = 0, low = 0;
        constexpr int mpos[5] = {0, b + 2, 2 * b + 4, 3 * b + 5, 4 * b + 7};
        constexpr int pow3[5] = {1, 3, 9, 27, 81};
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const int idx = 5 * g + i;
            if (idx < 16) {
                // v_bfe_u32 reads its offset from the low five bits
                tidx += __builtin_amdgcn_ubfe(split.trit_bits, wsh[idx], 2) * pow3[i];
                low |= __builtin_amdgcn_ubfe(split.low_bits, wsh[idx], 2) << mpos[i];
            }
        }
        const uint64_t grp = uint64_t(spread[tidx] | low);
        stream |= grp << (g * group_bits);
    }
    // endpoints: QUANT_256 raw bytes r0 r1 g0 g1 b0 b1 [a0 a1] from block bit 17
    uint32_t ep_x = e0q[0] | e1q[0] << 8 | e0q[1] << 16 | e1q[1] << 24;
    uint32_t ep_y = ALPHA ? (e0q[2] | e1q[2] << 8 | e0q[3] << 16 | e1q[3] << 24) : (e0q[2] | e1q[2] << 8);
    if (swap) {
        ep_x = (ep_x & 0x00FF00FFu) << 8 | (ep_x >> 8 & 0x00FF00FFu);
        ep_y = (ep_y & 0x00FF00FFu) << 8 | (ep_y >> 8 & 0x00FF00FFu);
    }
    constexpr uint32_t mode = ALPHA ? kModeRGBA : kModeRGB;
    constexpr uint32_t cem = ALPHA ? CEM_LDR_RGBA_DIRECT : CEM_LDR_RGB_DIRECT;
    Block128 blk;
    blk.w[0] = mode | (cem << 13) | (ep_x << 17);
    blk.w[1] = (ep_x >> 15) | (ep_y << 17);
    // weight stream bit k lands on block bit 127 - k
    blk.w[2] = __builtin_bitreverse32(uint32_t(stream >> 32)) | (ep_y >> 15);
    blk.w[3] = __builtin_bitreverse32(uint32_t(stream));
    return blk;
}

// ---------------------------------------------------------------------------
// Floating-point stage (ASTC_Encode.hlsl:93-168, 233-245, 256-393, 510-550)
//
// NC < 4: channels NC..3 hold exactly 255 in every texel of the block (normal
// maps: blue and alpha; an RGB8 block inside the image: alpha).  Such a channel
// has mean exactly 255 (16 and 36 copies of an integer sum and scale back
// exactly), so its differences from the mean, its covariance row, its axis
// component after the first product and its endpoint offsets are all zero, and
// the four-channel computation adds only zero terms for it.  The specialised
// code leaves those terms out; it produces the same bytes.
// ---------------------------------------------------------------------------
// METHOD: METHOD_PCA, the shader's principal component analysis (:139-168), or
// METHOD_MAX_ACCUMULATION, the max_accumulation_pixel_direction search it leaves
// commented out (:170-227, :514).
enum : int { METHOD_PCA = 0, METHOD_MAX_ACCUMULATION = 1 };

// Texels summed by one chain: a 6x6 block's sums run over its two halves
// (texels 0-17, rows 0-2, and 18-35, rows 3-5) separately and add the two, so
// that each half can be summed by its own lane (encode_block_pair); 4x4 blocks
// sum all 16 in one chain.
template <int DIM>
struct SumHalf { static constexpr int value = DIM == 6 ? 18 : DIM * DIM; };

// The power iteration of eigen_vector (:93-106) on the covariance m, from
// v0 = (0.26726, 0.80178, 0.53452, 0).  One iteration is v <- normalize(M (M v)),
// after the test length(M v) < eps, which returns M v.  The first iteration is
// taken as written (M v0, its test, then M (M v0)).  Past it, the test cannot
// succeed: |M v|^2 = dot(v, S v) with S = M M is the Rayleigh quotient of the
// positive semi-definite S at the unit v, and it never decreases along power
// iterates.  The remaining seven iterations use S (formed once,
// S_ij = dot(row i, row j)) and the last one normalises.  Only the direction
// feeds the next product, so the iterate is rescaled by a power of two (exact)
// before products 1, 4 and 7 (RESCALE = 3; 1 rescales before every product,
// for register budgets that need it): three products from a largest component
// in [0.5, 1) stay in the normal range (each grows it by at most
// 4 max|S_ij| < 8e10 and shrinks it by at most half the first test's 1e-10
// bound on the Rayleigh quotient), and scaling by powers of two commutes with
// rounding, so fewer rescales change no bit.
template <int NC, int RESCALE>
__device__ __forceinline__ void principal_axis(const float m[4][4], float v[4]) {
    float s2[4][4];
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = i; j < NC; ++j) s2[i][j] = s2[j][i] = dotn<NC>(m[i], m[j]);
    const float v0[4] = {0.26726f, 0.80178f, 0.53452f, 0.0f};
    float u[4], w[4];
#pragma unroll
    for (int i = 0; i < NC; ++i) u[i] = dotn<NC>(m[i], v0);
    if (is_short(dotn<NC>(u, u))) {
#pragma unroll
        for (int i = 0; i < NC; ++i) v[i] = u[i];
    } else {
#pragma unroll
        for (int i = 0; i < NC; ++i) w[i] = dotn<NC>(m[i], u);
#pragma unroll
        for (int it = 1; it < 8; ++it) {
            if ((it - 1) % RESCALE == 0) {
                float mx = fabsf(w[0]);
#pragma unroll
                for (int i = 1; i < NC; ++i) mx = fmaxf(mx, fabsf(w[i]));
                const int e = exponent_of(mx);
#pragma unroll
                for (int i = 0; i < NC; ++i) v[i] = ldexp_(w[i], -e);
            } else {
#pragma unroll
                for (int i = 0; i < NC; ++i) v[i] = w[i];
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) w[i] = dotn<NC>(s2[i], v);
        }
        const float sc = rsqrt_nr(dotn<NC>(w, w));
#pragma unroll
        for (int i = 0; i < NC; ++i) v[i] = w[i] * sc;
    }
#pragma unroll
    for (int i = NC; i < 4; ++i) v[i] = 0.0f;
}

// The covariance matrix (:149-162) from the upper-triangle sums of products.
template <int NC, int N>
__device__ __forceinline__ void covariance_matrix(const float cov[NC * (NC + 1) / 2], float m[4][4]) {
    constexpr float rcp_n1 = 1.0f / float(N - 1);
    int p = 0;
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
        for (int b2 = a; b2 < NC; ++b2, ++p) m[a][b2] = m[b2][a] = cov[p] * rcp_n1;
}

// Sums of products of the differences from the mean over texels [K0, K1) of x:
// one fused multiply-add chain per upper-triangle entry, starting with a plain
// product.
template <int NC, int K0, int K1, typename TX>
__device__ __forceinline__ void covariance_chain(const TX& x, const float mean[4], float cov[NC * (NC + 1) / 2]) {
#pragma unroll
    for (int k = K0; k < K1; ++k) {
        float d[4];
#pragma unroll
        for (int c = 0; c < NC; ++c) d[c] = x[k][c] - mean[c];
        int p = 0;
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b2 = a; b2 < NC; ++b2, ++p) cov[p] = (k == K0) ? d[a] * d[b2] : fma_(d[a], d[b2], cov[p]);
        texel_step<NC * (NC + 1) / 2>(x, k, cov);
    }
}

// find_min_max's endpoints (:108-137) from the extreme projections lo, hi of
// the texels on v: e = clamp(v * (t - dot(mean, v)) + mean), the alpha override
// of RGB mode, the rounded endpoints (encode_color :233-245) and the swap flag.
// The shader swaps e0 and e1 when e0's rounded RGB sum is the larger
// (:125-130).  Swapping only reverses each endpoint byte pair and negates
// e1 - e0; negation is exact and round-to-nearest is symmetric in sign, so
// every later quantity of the swapped block is the unswapped one with its sign
// flipped.  The swap is carried as that flag instead of moving values.
// With NC < 4 the inactive channels' endpoints are 255, except alpha, which is
// `alpha` (the block's constant alpha for the dynamic three-channel path).
template <bool ALPHA, int NC>
__device__ __forceinline__ bool endpoints(const float v[4], float lo, float hi, const float mean[4], float e0[4],
                                          float e1[4], uint32_t e0q[4], uint32_t e1q[4], float alpha = 255.0f) {
    const float mv = dotn<NC>(mean, v);
    lo = lo - mv;
    hi = hi - mv;
#pragma unroll
    for (int c = 0; c < 3; ++c) e0[c] = e1[c] = 255.0f;
    e0[3] = e1[3] = alpha;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        e0[c] = clamp255(fma_(v[c], lo, mean[c]));
        e1[c] = clamp255(fma_(v[c], hi, mean[c]));
    }
    const float s0 = (rintf(e0[0]) + rintf(e0[1])) + rintf(e0[2]);
    const float s1 = (rintf(e1[0]) + rintf(e1[1])) + rintf(e1[2]);
    if (!ALPHA) {
        e0[3] = 255.0f;
        e1[3] = 255.0f;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        e0q[c] = uint32_t(rintf(e0[c]));
        e1q[c] = uint32_t(rintf(e1[c]));
    }
    return s0 > s1;
}

// The weight axis of calculate_normal_weights (:316-372): k = e1 - e0 over the
// KC channels that differ, scaled to unit length.  The weight of grid point i
// is dot(k, sample_i - e0) rescaled to [0, 1] by its minimum and range over
// the grid, so the constant dot(k, e0) cancels: the projections
// p_i = dot(k, sample_i) are rescaled instead.  |e1 - e0| < 1e-5 gives
// all-zero weights (:323-330); a zero scale gives the same fields without a
// branch: every projection is +-0, the range is 0, so each (p - min) * 1e5 is
// +-0 and rounds to weight 0.  With the swap, the scale is negated (e1 - e0 of
// the swapped block).
template <int KC>
__device__ __forceinline__ void weight_axis(const float e0[4], const float e1[4], bool swap, float k[4]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) k[c] = 0.0f;
#pragma unroll
    for (int c = 0; c < KC; ++c) k[c] = e1[c] - e0[c];
    const float kk = dotn<KC>(k, k);
    float s = is_short(kk) ? 0.0f : rsqrt_nr(kk);
    s = swap ? -s : s;
#pragma unroll
    for (int c = 0; c < KC; ++c) k[c] = k[c] * s;
}

// 6x6 sample projections of grid points [G0, G0 + NG) from texel projections
// pt (texel index minus T0): each sample is a fixed 4-tap combination
// (sample_texel :307-314 with idx_grids / wt_grids :268-304), and so is its
// projection.
template <int G0, int NG, int T0, int NT>
__device__ __forceinline__ void grid6_samples(const float pt[NT], float proj[NG]) {
    constexpr float tap_w[3] = {0.444f, 0.222f, 0.111f};
#pragma unroll
    for (int j = 0; j < NG; ++j) {
        const int i = G0 + j;
        float smp = pt[grid6_tap_index(i, 0) - T0] * tap_w[grid6_tap_far(i, 0)];
        smp = fma_(pt[grid6_tap_index(i, 1) - T0], tap_w[grid6_tap_far(i, 1)], smp);
        smp = fma_(pt[grid6_tap_index(i, 2) - T0], tap_w[grid6_tap_far(i, 2)], smp);
        proj[j] = fma_(pt[grid6_tap_index(i, 3) - T0], tap_w[grid6_tap_far(i, 3)], smp);
    }
}

// quantize_weights (:374-382) of rescaled projections: (w - min) * invlen is at
// most 1 + 2^-22, so round(w * range) never exceeds range and the clamp of
// quantize_weight (:258) never acts.  round-half-even(y) for 0 <= y < 2^22 is
// the low bits of y + 1.5 * 2^23 (the sum's unit in the last place is 1);
// doubled, its low five bits are 2 * round(y).
template <bool ALPHA, int NG>
__device__ __forceinline__ void quantize_weights(const float proj[NG], float minw, float maxw, uint32_t wsh[NG]) {
    constexpr uint32_t range = ALPHA ? 5u : 11u;
    const float invlen = 1.0f / fmaxf(0.00001f, maxw - minw);
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const float wn = (proj[i] - minw) * invlen;
        wsh[i] = __float_as_uint(wn * float(range) + 12582912.0f) << 1;
    }
}

// A four-channel block whose alpha is the same byte c in every texel, with
// UNORM value c itself (130 of the 256 bytes, 0 and 255 among them), encodes to
// the same bytes through the three-channel path: the alpha mean is exactly c
// (sums of 18 and 36 copies of an integer and the scale by fl(1/36) or
// fl(1/16) are exact), so every alpha difference from it is +-0, the alpha
// covariance row and the axis' alpha component are +-0, the alpha endpoints
// are c, and every term the four-channel computation adds for alpha is a
// signed zero, which changes no rounded result.  words: the block's packed
// RGBA8 texels; unorm: the UNORM conversion table.
template <int N>
__device__ __forceinline__ bool flat_alpha(const uint32_t words[N], const float* unorm) {
    // alpha is the top byte, so the largest and smallest words carry the
    // largest and smallest alpha
    uint32_t hi = words[0], lo = words[0];
#pragma unroll
    for (int k = 1; k < N; ++k) {
        hi = max(hi, words[k]);
        lo = min(lo, words[k]);
    }
    const uint32_t a = words[0] >> 24;
    return ((hi ^ lo) >> 24) == 0 && unorm[a] == float(a);
}

// RESCALE: the power iteration's rescale interval, 3 or 1 (every product); both
// give the same blocks (see principal_axis).  alpha: the constant alpha of a
// four-channel block encoded by the three-channel path (NC = 3; see
// flat_alpha), else 255.  x: the texels, float x[DIM * DIM][4] or
// PackedTexels<DIM, CH> (x[k][c] is texel k's channel c, times 255).
template <int DIM, bool ALPHA, int NC = 4, int METHOD = METHOD_PCA, int RESCALE = 3, typename TX>
__device__ __forceinline__ Block128 encode_block(const TX& x, const PackTables* pack, float alpha = 255.0f) {
    static_assert(NC >= 2 && NC <= 4, "two to four active channels");
    constexpr int N = DIM * DIM;
    constexpr int H = SumHalf<DIM>::value;
    constexpr float rcp_n = 1.0f / float(N);

    // mean (principal_component_analysis :142-147)
    float mean[4] = {255.0f, 255.0f, 255.0f, 255.0f};
    {
        float sum[4];
        block_sums<NC, H, N>(x, sum);
#pragma unroll
        for (int c = 0; c < NC; ++c) mean[c] = sum[c] * rcp_n;
    }
    texel_pass(x);

    float v[4];
    if (METHOD == METHOD_MAX_ACCUMULATION) {
        // max_accumulation_pixel_direction (:180-224): acc[c] sums the
        // differences from the mean of the texels whose channel c is above the
        // mean (an fma with a 0/1 factor adds exactly the difference or nothing);
        // the axis is the longest acc[c] (alpha competes only with ALPHA),
        // normalised unless shorter than 1e-5.
        float acc[4][4];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int j = 0; j < NC; ++j) acc[c][j] = 0.0f;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            float d[4];
#pragma unroll
            for (int c = 0; c < NC; ++c) d[c] = x[k][c] - mean[c];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const float on = d[c] > 0.0f ? 1.0f : 0.0f;
#pragma unroll
                for (int j = 0; j < NC; ++j) acc[c][j] = fma_(on, d[j], acc[c][j]);
            }
        }
        float maxdot = dotn<NC>(acc[0], acc[0]);
#pragma unroll
        for (int j = 0; j < NC; ++j) v[j] = acc[0][j];
#pragma unroll
        for (int c = 1; c < NC; ++c) {
            if (c == 3 && !ALPHA) break;
            const float dc = dotn<NC>(acc[c], acc[c]);
            const bool better = dc > maxdot;
            maxdot = better ? dc : maxdot;
#pragma unroll
            for (int j = 0; j < NC; ++j) v[j] = better ? acc[c][j] : v[j];
        }
        if (!is_short(maxdot)) {
            const float sc = rsqrt_nr(maxdot);
#pragma unroll
            for (int j = 0; j < NC; ++j) v[j] = v[j] * sc;
        }
#pragma unroll
        for (int i = NC; i < 4; ++i) v[i] = 0.0f;
    } else {
        // covariance (:149-162): upper triangle of the active channels
        constexpr int NP = NC * (NC + 1) / 2;
        float cov[NP];
        covariance_chain<NC, 0, H>(x, mean, cov);
        if (H < N) {
            float cov2[NP];
            covariance_chain<NC, H, N>(x, mean, cov2);
#pragma unroll
            for (int p = 0; p < NP; ++p) cov[p] = cov[p] + cov2[p];
        }
        float m[4][4];
        covariance_matrix<NC, N>(cov, m);
        issue_priority(1);
        principal_axis<NC, RESCALE>(m, v);
        issue_priority(0);
    }
    texel_pass(x);

    // find_min_max (:108-137): t_k = dot(texel_k - mean, v), taken as
    // dot(texel_k, v) - dot(mean, v); rounding is monotonic, so the minimum and
    // maximum of t_k are the extreme projections minus dot(mean, v).
    float ext[2] = {1e31f, -1e31f};  // lowest and highest projection
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const float t = dotn<NC>(x[k], v);
        ext[0] = fminf(ext[0], t);
        ext[1] = fmaxf(ext[1], t);
        texel_step<2>(x, k, ext);
    }
    const float lo = ext[0], hi = ext[1];
    texel_pass(x);
    issue_priority(1);
    float e0[4], e1[4];
    uint32_t e0q[4], e1q[4];
    const bool swap = endpoints<ALPHA, NC>(v, lo, hi, mean, e0, e1, e0q, e1q, alpha);

    // calculate_normal_weights (:316-372) + quantize_weights (:374-382)
    constexpr int KC = (NC == 4 && !ALPHA) ? 3 : NC;  // RGB mode: endpoint alphas are equal
    float k[4];
    weight_axis<KC>(e0, e1, swap, k);
    float proj[16];
    issue_priority(0);
    if (DIM == 6) {
        float pt[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            pt[j] = dotn<KC>(k, x[j]);
            texel_step<DIM>(x, j, pt + (j / DIM) * DIM);
        }
        grid6_samples<0, 16, 0, N>(pt, proj);
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            proj[i] = dotn<KC>(k, x[i]);
            texel_step<DIM>(x, i, proj + (i / DIM) * DIM);
        }
    }
    float minw = 1e31f, maxw = -1e31f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        minw = fminf(proj[i], minw);
        maxw = fmaxf(proj[i], maxw);
    }
    uint32_t wsh[16];  // 2 * quantised weight in the low five bits
    issue_priority(1);
    quantize_weights<ALPHA, 16>(proj, minw, maxw, wsh);
    const Block128 blk = pack_block<ALPHA>(e0q, e1q, wsh, pack, swap);
    issue_priority(0);
    return blk;
}

// ---------------------------------------------------------------------------
// Two lanes per 6x6 block: lanes 2j and 2j+1 of a wave share one block, lane
// `half` holding texel rows 3 half .. 3 half + 2 (texels 18 half ..
// 18 half + 17).  Each lane sums, projects and samples its own half; the
// halves are combined by one exchange between the two lanes (DPP quad
// permutation [1, 0, 3, 2]):
//   * sums: own + other.  IEEE addition is commutative, so both lanes hold the
//     same bits, and they equal the one-lane schedule's (half 0) + (half 1);
//   * minimum and maximum: exact and order-free;
//   * the grid: grid rows 0-1 sample texel rows 0-2 only and grid rows 2-3
//     texel rows 3-5 only, so each lane samples its own eight grid points, at
//     the same texel offsets within its half; their quantised weights (4 bits
//     each) are exchanged as one word.
// The serial part (power iteration, endpoints, integer packing) runs on both
// lanes with identical inputs.  Blocks equal encode_block's bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pair_other(uint32_t v) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float pair_other(float v) { return __uint_as_float(pair_other(__float_as_uint(v))); }

template <bool ALPHA, int NC = 4, int RESCALE = 3>
__device__ __forceinline__ Block128 encode_block_pair(const float x[18][4], int half, const PackTables* pack) {
    static_assert(NC >= 2 && NC <= 4, "two to four active channels");
    constexpr int N = 36, NH = 18;
    constexpr float rcp_n = 1.0f / float(N);

    float mean[4] = {255.0f, 255.0f, 255.0f, 255.0f};
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        float r = x[0][c];
#pragma unroll
        for (int k = 1; k < NH; ++k) r = r + x[k][c];
        mean[c] = (r + pair_other(r)) * rcp_n;
    }
    constexpr int NP = NC * (NC + 1) / 2;
    float cov[NP];
    covariance_chain<NC, 0, NH>(x, mean, cov);
#pragma unroll
    for (int p = 0; p < NP; ++p) cov[p] = cov[p] + pair_other(cov[p]);
    float m[4][4];
    covariance_matrix<NC, N>(cov, m);
    float v[4];
    issue_priority(1);
    principal_axis<NC, RESCALE>(m, v);
    issue_priority(0);

    float lo = 1e31f, hi = -1e31f;
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        const float t = dotn<NC>(x[k], v);
        lo = fminf(lo, t);
        hi = fmaxf(hi, t);
    }
    lo = fminf(lo, pair_other(lo));
    hi = fmaxf(hi, pair_other(hi));
    issue_priority(1);
    float e0[4], e1[4];
    uint32_t e0q[4], e1q[4];
    const bool swap = endpoints<ALPHA, NC>(v, lo, hi, mean, e0, e1, e0q, e1q);

    constexpr int KC = (NC == 4 && !ALPHA) ? 3 : NC;
    float k[4];
    weight_axis<KC>(e0, e1, swap, k);
    issue_priority(0);
    float pt[NH];
#pragma unroll
    for (int j = 0; j < NH; ++j) pt[j] = dotn<KC>(k, x[j]);
    float proj[8];
    grid6_samples<0, 8, 0, NH>(pt, proj);  // grid rows 0-1 of half 0 = grid rows 2-3 of half 1
    float minw = 1e31f, maxw = -1e31f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        minw = fminf(proj[i], minw);
        maxw = fmaxf(proj[i], maxw);
    }
    minw = fminf(minw, pair_other(minw));
    maxw = fmaxf(maxw, pair_other(maxw));
    uint32_t own[8];
    issue_priority(1);
    quantize_weights<ALPHA, 8>(proj, minw, maxw, own);
    uint32_t packed = 0;  // the eight quantised weights, 4 bits each
#pragma unroll
    for (int i = 0; i < 8; ++i) packed |= ((own[i] >> 1) & 15u) << (4 * i);
    const uint32_t other = pair_other(packed);
    const uint32_t first = half ? other : packed, second = half ? packed : other;
    uint32_t wsh[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        wsh[i] = ((first >> (4 * i)) & 15u) << 1;
        wsh[8 + i] = ((second >> (4 * i)) & 15u) << 1;
    }
    const Block128 blk = pack_block<ALPHA>(e0q, e1q, wsh, pack, swap);
    issue_priority(0);
    return blk;
}

}  // namespace astc
