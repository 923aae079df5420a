// astc_format.h -- ASTC format constants and tables shared by host and device code.
//
// The encoder implements the fixed-mode subset of niepp/astc_encoder
// (ASTC_Encode.hlsl:510-550): one partition, one plane, a 4x4 weight grid for
// both 4x4 and 6x6 blocks, QUANT_256 endpoints, and either
//   RGB mode  : CEM 8  (LDR RGB direct),  QUANT_12 weights (range 0..11)
//   RGBA mode : CEM 12 (LDR RGBA direct), QUANT_6  weights (range 0..5).
// Tables are generated at compile time from the ASTC specification's closed
// forms rather than typed in:
//   - trit packing: the reference's integer_from_trits
//     (ASTC_IntegerSequenceEncoding.hlsl:30-62) is, for every 5-trit tuple, the
//     largest 8-bit code that the specification's trit decoder maps back to it;
//   - weight scramble: ASTC_Table.hlsl:4-66 row m orders the ISE symbols of
//     quantisation level m by their specification-unquantised weight.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define ASTC_HD __host__ __device__
#else
#define ASTC_HD
#endif

namespace astc {

// Quantisation method ids (ASTC_Encode.hlsl:47-68).
enum : int { QUANT_6 = 4, QUANT_12 = 7, QUANT_256 = 20 };
enum : uint32_t { CEM_LDR_RGB_DIRECT = 8, CEM_LDR_RGBA_DIRECT = 12 };

// length(v) < 0.00001f with length = correctly rounded sqrtf(dot(v, v)) holds
// exactly when dot(v, v) < this value (the smallest binary32 whose sqrt
// reaches 1e-5f; sqrt is monotone).
constexpr uint32_t kLenThresholdBits = 0x2edbe6fdu;

// ---- specification trit decoder (8 bits -> 5 trits) ----
constexpr int bits_of(uint32_t v, int hi, int lo) { return int((v >> lo) & ((1u << (hi - lo + 1)) - 1u)); }

struct Trits5 { int t[5]; };
constexpr Trits5 decode_trit_block(uint32_t T) {
    Trits5 r{};
    uint32_t C = 0;
    if (bits_of(T, 4, 2) == 7) {
        C = uint32_t((bits_of(T, 7, 5) << 2) | bits_of(T, 1, 0));
        r.t[4] = 2;
        r.t[3] = 2;
    } else {
        C = uint32_t(bits_of(T, 4, 0));
        if (bits_of(T, 6, 5) == 3) {
            r.t[4] = 2;
            r.t[3] = bits_of(T, 7, 7);
        } else {
            r.t[4] = bits_of(T, 7, 7);
            r.t[3] = bits_of(T, 6, 5);
        }
    }
    const int c4 = bits_of(C, 4, 4), c3 = bits_of(C, 3, 3), c2 = bits_of(C, 2, 2), c1 = bits_of(C, 1, 1),
              c0 = bits_of(C, 0, 0);
    if (c1 == 1 && c0 == 1) {
        r.t[2] = 2;
        r.t[1] = c4;
        r.t[0] = (c3 << 1) | (c2 & (1 - c3));
    } else if (c3 == 1 && c2 == 1) {
        r.t[2] = 2;
        r.t[1] = 2;
        r.t[0] = (c1 << 1) | c0;
    } else {
        r.t[2] = c4;
        r.t[1] = (c3 << 1) | c2;
        r.t[0] = (c1 << 1) | (c0 & (1 - c1));
    }
    return r;
}

// Encoder table indexed by t0 + 3 t1 + 9 t2 + 27 t3 + 81 t4 (243 used of 256).
struct TritEncodeTable { uint8_t code[256]; };
constexpr TritEncodeTable make_trit_encode_table() {
    TritEncodeTable tab{};
    for (uint32_t T = 0; T < 256; ++T) {  // ascending, so the largest code wins
        Trits5 d = decode_trit_block(T);
        tab.code[d.t[0] + 3 * d.t[1] + 9 * d.t[2] + 27 * d.t[3] + 81 * d.t[4]] = uint8_t(T);
    }
    return tab;
}
constexpr TritEncodeTable kTritEncode = make_trit_encode_table();

// ---- specification quint decoder (7 bits -> 3 quints) ----
struct Quints3 { int q[3]; };
constexpr Quints3 decode_quint_block(uint32_t Q) {
    Quints3 r{};
    const int q0b = bits_of(Q, 0, 0);
    if (bits_of(Q, 2, 1) == 3 && bits_of(Q, 6, 5) == 0) {
        r.q[2] = (q0b << 2) | ((bits_of(Q, 4, 4) & (1 - q0b)) << 1) | (bits_of(Q, 3, 3) & (1 - q0b));
        r.q[1] = 4;
        r.q[0] = 4;
        return r;
    }
    int C = 0;
    if (bits_of(Q, 2, 1) == 3) {
        r.q[2] = 4;
        C = (bits_of(Q, 4, 3) << 3) | ((3 - bits_of(Q, 6, 5)) << 1) | q0b;
    } else {
        r.q[2] = bits_of(Q, 6, 5);
        C = bits_of(Q, 4, 0);
    }
    if ((C & 7) == 5) {
        r.q[1] = 4;
        r.q[0] = (C >> 3) & 3;
    } else {
        r.q[1] = (C >> 3) & 3;
        r.q[0] = C & 7;
    }
    return r;
}

// Decoder tables: packed 8-bit trit code -> 5 trits (2 bits each), packed 7-bit
// quint code -> 3 quints (3 bits each).
struct TritQuintDecodeTable { uint16_t trits[256]; uint16_t quints[128]; };
constexpr TritQuintDecodeTable make_decode_tables() {
    TritQuintDecodeTable t{};
    for (uint32_t T = 0; T < 256; ++T) {
        Trits5 d = decode_trit_block(T);
        t.trits[T] = uint16_t(d.t[0] | d.t[1] << 2 | d.t[2] << 4 | d.t[3] << 6 | d.t[4] << 8);
    }
    for (uint32_t Q = 0; Q < 128; ++Q) {
        Quints3 d = decode_quint_block(Q);
        t.quints[Q] = uint16_t(d.q[0] | d.q[1] << 3 | d.q[2] << 6);
    }
    return t;
}
constexpr TritQuintDecodeTable kDecodeTables = make_decode_tables();

// (levels, trits, quints, bits) of the 21 quantisation methods (ASTC_Encode.hlsl:47-68)
struct QuantLevel { uint8_t levels_minus1, trits, quints, bits; };
constexpr QuantLevel kQuantLevels[21] = {
    {1, 0, 0, 1},   {2, 1, 0, 0},   {3, 0, 0, 2},   {4, 0, 1, 0},   {5, 1, 0, 1},   {7, 0, 0, 3},   {9, 0, 1, 1},
    {11, 1, 0, 2},  {15, 0, 0, 4},  {19, 0, 1, 2},  {23, 1, 0, 3},  {31, 0, 0, 5},  {39, 0, 1, 3},  {47, 1, 0, 4},
    {63, 0, 0, 6},  {79, 0, 1, 4},  {95, 1, 0, 5},  {127, 0, 0, 7}, {159, 0, 1, 5}, {191, 1, 0, 6}, {255, 0, 0, 8}};

// ---- weight unquantisation for the two trit-coded weight ranges in use ----
// QUANT_6: 1 trit + 1 bit, C = 50; QUANT_12: 1 trit + 2 bits, C = 23, B = b000b0b
// (ASTC specification, weight unquantisation).
constexpr int unquant_weight_trit(int bits, int sym) {
    const int D = sym >> bits, low = sym & ((1 << bits) - 1);
    const int A = (low & 1) ? 0x7F : 0, b = (low >> 1) & 1;
    const int C = bits == 1 ? 50 : 23;
    const int B = bits == 1 ? 0 : ((b << 6) | (b << 2) | b);
    int v = (A & 0x20) | (((D * C + B) ^ A) >> 2);
    return v > 32 ? v + 1 : v;
}

// Linear weight index q -> (trit, low bits) of its ISE symbol, packed 2 bits per q
// into 32-bit immediates so that the kernel extracts them with one bit-field op.
struct WeightSplit { uint32_t trit_bits, low_bits; int levels, bits; };
constexpr WeightSplit make_weight_split(int levels, int bits) {
    int sym[12] = {};
    for (int s = 0; s < levels; ++s) sym[s] = s;
    for (int i = 1; i < levels; ++i)  // insertion sort by unquantised value (stable)
        for (int j = i; j > 0 && unquant_weight_trit(bits, sym[j - 1]) > unquant_weight_trit(bits, sym[j]); --j) {
            int t = sym[j];
            sym[j] = sym[j - 1];
            sym[j - 1] = t;
        }
    WeightSplit w{0u, 0u, levels, bits};
    for (int q = 0; q < levels; ++q) {
        w.trit_bits |= uint32_t(sym[q] >> bits) << (2 * q);
        w.low_bits |= uint32_t(sym[q] & ((1 << bits) - 1)) << (2 * q);
    }
    return w;
}
constexpr WeightSplit kSplitQ6 = make_weight_split(6, 1);
constexpr WeightSplit kSplitQ12 = make_weight_split(12, 2);

// scramble symbol for linear index q, reconstructed from a split (for host checks)
constexpr int scramble_symbol(const WeightSplit& w, int q) {
    return int(((w.trit_bits >> (2 * q)) & 3u) << w.bits | ((w.low_bits >> (2 * q)) & 3u));
}

// ---- block mode: 2D layout row "D H B B A A R0 0 0 R2 R1", width B+4, height A+2 ----
// (ASTC_Encode.hlsl:446-473): grid 4x4 -> A = 2, B = 0; single plane D = 0;
// weight range R = (q % 6) + 2 with precision bit H = (q >= 6).
constexpr uint32_t block_mode_4x4_grid(int weight_quant) {
    const uint32_t r = uint32_t(weight_quant % 6 + 2), h = weight_quant >= 6 ? 1u : 0u;
    const uint32_t a = 2, b = 0, d = 0;
    return ((r >> 1) & 3u) | ((r & 1u) << 4) | (a << 5) | (b << 7) | (h << 9) | (d << 10);
}
constexpr uint32_t kModeRGBA = block_mode_4x4_grid(QUANT_6);   // 0x043
constexpr uint32_t kModeRGB = block_mode_4x4_grid(QUANT_12);   // 0x251

// ---- 6x6 weight sampling (idx_grids / wt_grids, ASTC_Encode.hlsl:268-304) ----
// Grid point g (0..3) of the 4-point grid over 6 texels lies at texel coordinate
// 5g/3; its bracketing texels are c0 = {0,1,3,4}[g] and c0 + 1, the nearer being
// c0 for even g and c0 + 1 for odd g.  A tap's weight is 0.444 / 0.222 / 0.111
// when it is the nearer texel on both / one / neither axis.
constexpr int grid6_c0(int g) { return g == 0 ? 0 : g == 1 ? 1 : g == 2 ? 3 : 4; }
constexpr int grid6_tap_index(int gridpoint, int tap) {
    return (grid6_c0(gridpoint >> 2) + (tap >> 1)) * 6 + grid6_c0(gridpoint & 3) + (tap & 1);
}
constexpr int grid6_tap_far(int gridpoint, int tap) {
    return ((tap & 1) != ((gridpoint & 3) & 1)) + ((tap >> 1) != ((gridpoint >> 2) & 1));
}

// .astc container (astc_save.h:3-14)
constexpr uint32_t kAstcMagic = 0x5CA1AB13u;
constexpr int kAstcHeaderBytes = 16;

static_assert(kModeRGBA == 0x043 && kModeRGB == 0x251, "block mode");
static_assert(kTritEncode.code[0] == 0 && kTritEncode.code[242] == 254, "trit table");

}  // namespace astc
