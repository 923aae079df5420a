/*
 * astc_oracle.cpp -- CPU oracle for the ASTC block-encode path.
 *
 * TEST INFRASTRUCTURE ONLY (see astc_oracle.h).  The HIP encoder never calls
 * into this file; tests compare the two.
 *
 * Build: strict IEEE, no contraction (-ffp-contract=off, no fast-math).  Fused
 * multiply-adds appear only where written as std::fma.
 *
 * Reference lines restated here (file:line in niepp/astc_encoder):
 *   fetch + addressing   ASTC_Encode.hlsl:561-580, astc_encode.h:124-134
 *   PCA                  ASTC_Encode.hlsl:93-106 (eigen_vector), 108-137 (find_min_max),
 *                        139-168 (principal_component_analysis)
 *   endpoints            ASTC_Encode.hlsl:233-245 (encode_color), 475-489 (endpoint_ise)
 *   weights              ASTC_Encode.hlsl:256-260, 268-393, 491-508
 *   block                ASTC_Encode.hlsl:400-444 (assemble_block), 446-473 (assemble_blockmode),
 *                        510-550 (encode_block)
 *   ISE                  ASTC_IntegerSequenceEncoding.hlsl:5-71 (tables), 98-119 (orbits8_ptr),
 *                        121-137, 142-205 (trits/quints), 207-276 (bise_*)
 *   scramble             ASTC_Table.hlsl:3-66
 *   .astc header         astc_save.h:5-14, 52-76
 */
#include "astc_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ------------------------------------------------------------------------- */
/* Quantisation levels (ASTC_Encode.hlsl:47-68).  Index -> (levels, trits,    */
/* quints, bits), the ASTC specification's range table.                       */
/* ------------------------------------------------------------------------- */
struct QuantInfo { int levels, trits, quints, bits; };
const QuantInfo kQuant[21] = {
    {2, 0, 0, 1},   {3, 1, 0, 0},   {4, 0, 0, 2},   {5, 0, 1, 0},   {6, 1, 0, 1},
    {8, 0, 0, 3},   {10, 0, 1, 1},  {12, 1, 0, 2},  {16, 0, 0, 4},  {20, 0, 1, 2},
    {24, 1, 0, 3},  {32, 0, 0, 5},  {40, 0, 1, 3},  {48, 1, 0, 4},  {64, 0, 0, 6},
    {80, 0, 1, 4},  {96, 1, 0, 5},  {128, 0, 0, 7}, {160, 0, 1, 5}, {192, 1, 0, 6},
    {256, 0, 0, 8}};
constexpr int QUANT_6 = 4, QUANT_12 = 7, QUANT_256 = 20;
constexpr int CEM_LDR_RGB_DIRECT = 8, CEM_LDR_RGBA_DIRECT = 12;

inline uint32_t bitfield(uint32_t v, int hi, int lo) { return (v >> lo) & ((1u << (hi - lo + 1)) - 1u); }

/* Spec trit-block decoder (ASTC specification, "Integer Sequence Encoding"):
 * 8 packed bits T -> 5 trits. */
void spec_decode_trits(uint32_t T, int t[5]) {
    uint32_t C;
    if (bitfield(T, 4, 2) == 7) {
        C = (bitfield(T, 7, 5) << 2) | bitfield(T, 1, 0);
        t[4] = 2; t[3] = 2;
    } else {
        C = bitfield(T, 4, 0);
        if (bitfield(T, 6, 5) == 3) { t[4] = 2; t[3] = (int)bitfield(T, 7, 7); }
        else { t[4] = (int)bitfield(T, 7, 7); t[3] = (int)bitfield(T, 6, 5); }
    }
    if (bitfield(C, 1, 0) == 3) {
        t[2] = 2; t[1] = (int)bitfield(C, 4, 4);
        t[0] = (int)((bitfield(C, 3, 3) << 1) | (bitfield(C, 2, 2) & ~bitfield(C, 3, 3) & 1u));
    } else if (bitfield(C, 3, 2) == 3) {
        t[2] = 2; t[1] = 2; t[0] = (int)bitfield(C, 1, 0);
    } else {
        t[2] = (int)bitfield(C, 4, 4); t[1] = (int)bitfield(C, 3, 2);
        t[0] = (int)((bitfield(C, 1, 1) << 1) | (bitfield(C, 0, 0) & ~bitfield(C, 1, 1) & 1u));
    }
}

/* Spec quint-block decoder: 7 packed bits Q -> 3 quints. */
void spec_decode_quints(uint32_t Q, int q[3]) {
    if (bitfield(Q, 2, 1) == 3 && bitfield(Q, 6, 5) == 0) {
        uint32_t q0b = bitfield(Q, 0, 0);
        q[2] = (int)((q0b << 2) | ((bitfield(Q, 4, 4) & ~q0b & 1u) << 1) | (bitfield(Q, 3, 3) & ~q0b & 1u));
        q[1] = 4; q[0] = 4;
        return;
    }
    uint32_t C;
    if (bitfield(Q, 2, 1) == 3) {
        q[2] = 4;
        C = (bitfield(Q, 4, 3) << 3) | ((~bitfield(Q, 6, 5) & 3u) << 1) | bitfield(Q, 0, 0);
    } else {
        q[2] = (int)bitfield(Q, 6, 5);
        C = bitfield(Q, 4, 0);
    }
    if (bitfield(C, 2, 0) == 5) { q[1] = 4; q[0] = (int)bitfield(C, 4, 3); }
    else { q[1] = (int)bitfield(C, 4, 3); q[0] = (int)bitfield(C, 2, 0); }
}

/* Spec weight unquantisation (ISE symbol -> 0..64), used to derive the scramble order. */
int spec_unquant_weight(int quant, int sym) {
    const QuantInfo& qi = kQuant[quant];
    int b = qi.bits, v;
    if (!qi.trits && !qi.quints) {
        int acc = 0, n = 0;
        while (n < 6) { acc = (acc << b) | sym; n += b; }
        v = acc >> (n - 6);
    } else if (b == 0) {
        v = sym * (qi.trits ? 32 : 16);
        return v;
    } else {
        int D = sym >> b, low = sym & ((1 << b) - 1);
        int A = (low & 1) ? 0x7F : 0, bb = (low >> 1) & 1, cc = (low >> 2) & 1;
        int B = 0, C = 0;
        if (qi.trits) {
            if (b == 1) { C = 50; B = 0; }
            else if (b == 2) { C = 23; B = (bb << 6) | (bb << 2) | bb; }
            else { C = 11; B = (cc << 6) | (bb << 5) | (cc << 1) | bb; }
        } else {
            if (b == 1) { C = 28; B = 0; }
            else { C = 13; B = (bb << 6) | (bb << 1) | bb; }
        }
        int T = D * C + B;
        T ^= A;
        v = (A & 0x20) | (T >> 2);
    }
    if (v > 32) v += 1;
    return v;
}

/* Tables of ASTC_IntegerSequenceEncoding.hlsl:30-71 and ASTC_Table.hlsl:4-66,
 * derived from the specification instead of typed in:
 *  - integer_from_trits[t]  = the largest 8-bit code the spec decoder maps to t
 *  - integer_from_quints[q] = the largest 7-bit code the spec decoder maps to q
 *  - scramble row m         = symbols ordered by spec-unquantised value (argsort). */
struct Tables {
    int trits[243];
    int quints[125];
    int scramble[12 * 32];
    Tables() {
        for (uint32_t code = 0; code < 256; ++code) {
            int t[5];
            spec_decode_trits(code, t);
            trits[t[4] * 81 + t[3] * 27 + t[2] * 9 + t[1] * 3 + t[0]] = (int)code;
        }
        for (uint32_t code = 0; code < 128; ++code) {
            int q[3];
            spec_decode_quints(code, q);
            quints[q[2] * 25 + q[1] * 5 + q[0]] = (int)code;
        }
        std::memset(scramble, 0, sizeof(scramble));
        for (int m = 0; m < 12; ++m) {
            int n = kQuant[m].levels;
            int order[32];
            for (int s = 0; s < n; ++s) order[s] = s;
            std::stable_sort(order, order + n, [&](int a, int b) {
                return spec_unquant_weight(m, a) < spec_unquant_weight(m, b);
            });
            for (int i = 0; i < n; ++i) scramble[m * 32 + i] = order[i];
        }
    }
};
const Tables& tables() { static Tables t; return t; }

/* ------------------------------------------------------------------------- */
/* ISE writer (ASTC_IntegerSequenceEncoding.hlsl:98-276)                      */
/* ------------------------------------------------------------------------- */
/* orbits8_ptr (ISE.hlsl:98-119): OR the low `bitcount` bits of `number` at
 * `bitoffset`, carrying into the next word.  The carry shift only happens
 * when the field straddles a word boundary, so 32 - bit_idx is in 1..31 here
 * (the HLSL evaluates the shift unconditionally, which C++ must not). */
void orbits8(uint32_t out[4], uint32_t& bitoffset, uint32_t number, uint32_t bitcount) {
    uint32_t newpos = bitoffset + bitcount;
    uint32_t nidx = newpos >> 5, uidx = bitoffset >> 5, bit_idx = bitoffset & 31;
    if (uidx < 4) out[uidx] |= number << bit_idx;
    if (nidx > uidx && uidx + 1 < 4) out[uidx + 1] |= number >> (32 - bit_idx);
    bitoffset = newpos;
}

/* split_high_low (ISE.hlsl:121-126) */
void split_high_low(uint32_t n, uint32_t i, int& high, uint32_t& low) {
    low = n & ((1u << i) - 1u);
    high = (int)((n >> i) & 0xFF);
}

/* reverse_byte (ISE.hlsl:131-137) */
uint32_t reverse_byte(uint32_t p) {
    p = ((p & 0xF) << 4) | ((p >> 4) & 0xF);
    p = ((p & 0x33) << 2) | ((p >> 2) & 0x33);
    p = ((p & 0x55) << 1) | ((p >> 1) & 0x55);
    return p;
}

/* encode_trits (ISE.hlsl:142-176) */
void encode_trits(uint32_t bitcount, const uint32_t b[5], uint32_t out[4], uint32_t& pos) {
    int t[5];
    uint32_t m[5];
    for (int i = 0; i < 5; ++i) split_high_low(b[i], bitcount, t[i], m[i]);
    uint32_t packhigh = (uint32_t)tables().trits[t[4] * 81 + t[3] * 27 + t[2] * 9 + t[1] * 3 + t[0]];
    orbits8(out, pos, m[0], bitcount); orbits8(out, pos, packhigh & 3, 2);
    orbits8(out, pos, m[1], bitcount); orbits8(out, pos, (packhigh >> 2) & 3, 2);
    orbits8(out, pos, m[2], bitcount); orbits8(out, pos, (packhigh >> 4) & 1, 1);
    orbits8(out, pos, m[3], bitcount); orbits8(out, pos, (packhigh >> 5) & 3, 2);
    orbits8(out, pos, m[4], bitcount); orbits8(out, pos, (packhigh >> 7) & 1, 1);
}

/* encode_quints (ISE.hlsl:181-205) */
void encode_quints(uint32_t bitcount, const uint32_t b[3], uint32_t out[4], uint32_t& pos) {
    int q[3];
    uint32_t m[3];
    for (int i = 0; i < 3; ++i) split_high_low(b[i], bitcount, q[i], m[i]);
    uint32_t packhigh = (uint32_t)tables().quints[q[2] * 25 + q[1] * 5 + q[0]];
    orbits8(out, pos, m[0], bitcount); orbits8(out, pos, packhigh & 7, 3);
    orbits8(out, pos, m[1], bitcount); orbits8(out, pos, (packhigh >> 3) & 3, 2);
    orbits8(out, pos, m[2], bitcount); orbits8(out, pos, (packhigh >> 5) & 3, 2);
}

/* bise_endpoints / bise_weights (ISE.hlsl:207-276) generalised to `count`
 * values: groups of 5 (trits) or 3 (quints) padded with zeros, else raw bits.
 * Returns the specification's bit count for `count` values. */
int bise(const uint32_t* numbers, int count, int quant, uint32_t out[4]) {
    const QuantInfo& qi = kQuant[quant];
    uint32_t pos = 0;
    if (qi.trits) {
        for (int g = 0; g < count; g += 5) {
            uint32_t b[5];
            for (int i = 0; i < 5; ++i) b[i] = (g + i < count) ? numbers[g + i] : 0u;
            encode_trits((uint32_t)qi.bits, b, out, pos);
        }
        return ((8 + 5 * qi.bits) * count + 4) / 5;
    }
    if (qi.quints) {
        for (int g = 0; g < count; g += 3) {
            uint32_t b[3];
            for (int i = 0; i < 3; ++i) b[i] = (g + i < count) ? numbers[g + i] : 0u;
            encode_quints((uint32_t)qi.bits, b, out, pos);
        }
        return ((7 + 3 * qi.bits) * count + 2) / 3;
    }
    for (int i = 0; i < count; ++i) orbits8(out, pos, numbers[i], (uint32_t)qi.bits);
    return count * qi.bits;
}

/* assemble_blockmode (ASTC_Encode.hlsl:446-473): 4x4 weight grid, single plane. */
uint32_t assemble_blockmode(uint32_t wq) {
    const uint32_t X_GRIDS = 4, Y_GRIDS = 4;
    uint32_t a = (Y_GRIDS - 2) & 3, b = (X_GRIDS - 4) & 3, d = 0;
    uint32_t h = (wq < 6) ? 0 : 1;
    uint32_t r = (wq % 6) + 2;
    uint32_t mode = (r >> 1) & 3;
    mode |= (r & 1) << 4;
    mode |= (a & 3) << 5;
    mode |= (b & 3) << 7;
    mode |= h << 9;
    mode |= d << 10;
    return mode;
}

/* assemble_block (ASTC_Encode.hlsl:400-444), including the plain assignment of
 * phy.y at :438 that discards weight bits taken from wt_ise.z. */
void assemble_block(uint32_t blockmode, uint32_t cem, const uint32_t ep[4], const uint32_t wt[4], uint32_t phy[4]) {
    phy[0] = phy[1] = phy[2] = phy[3] = 0;
    for (int k = 0; k < 3; ++k) {
        uint32_t w = wt[k];
        uint32_t r = (reverse_byte(w & 0xFF) << 24) | (reverse_byte((w >> 8) & 0xFF) << 16) |
                     (reverse_byte((w >> 16) & 0xFF) << 8) | reverse_byte((w >> 24) & 0xFF);
        phy[3 - k] |= r;
    }
    phy[0] = blockmode;
    phy[0] |= (cem & 0xF) << 13;
    phy[0] |= (ep[0] & 0x7FFF) << 17;
    phy[1] = (ep[0] >> 15) & 0x1FFFF;
    phy[1] |= (ep[1] & 0x7FFF) << 17;
    phy[2] |= (ep[1] >> 15) & 0x1FFFF;
}

void store_le(const uint32_t phy[4], uint8_t out[16]) {
    for (int i = 0; i < 4; ++i)
        for (int b = 0; b < 4; ++b) out[i * 4 + b] = (uint8_t)(phy[i] >> (8 * b));
}

/* The integer stage of encode_block (ASTC_Encode.hlsl:528-548, 475-508). */
void pack_block(const uint32_t ep_q_in[8], const uint32_t w_q[16], bool has_alpha, uint8_t out[16]) {
    const uint32_t wq = has_alpha ? QUANT_6 : QUANT_12;
    uint32_t ep_q[8];
    std::memcpy(ep_q, ep_q_in, sizeof(ep_q));
    if (!has_alpha) { ep_q[6] = 0; ep_q[7] = 0; }            /* :480-483 */
    uint32_t ep_ise[4] = {0, 0, 0, 0};
    bise(ep_q, has_alpha ? 8 : 6, QUANT_256, ep_ise);         /* :487, ISE.hlsl:214-218 */
    uint32_t sym[16];
    for (int i = 0; i < 16; ++i) sym[i] = (uint32_t)tables().scramble[wq * 32 + w_q[i]];  /* :498-502 */
    uint32_t wt_ise[4] = {0, 0, 0, 0};
    bise(sym, 16, (int)wq, wt_ise);
    uint32_t phy[4];
    assemble_block(assemble_blockmode(wq), has_alpha ? CEM_LDR_RGBA_DIRECT : CEM_LDR_RGB_DIRECT, ep_ise, wt_ise, phy);
    store_le(phy, out);
}

/* ------------------------------------------------------------------------- */
/* Floating-point stage                                                       */
/* ------------------------------------------------------------------------- */
template <typename T> struct K;
template <> struct K<float> {
    static constexpr float small = 0.00001f;                  /* SMALL_VALUE, :34 */
    static constexpr float big = 1e31f;                       /* :110-111, :333-334 */
    static constexpr float v0[4] = {0.26726f, 0.80178f, 0.53452f, 0.0f};   /* :96 */
    static constexpr float tap[3] = {0.444f, 0.222f, 0.111f}; /* wt_grids values, :287-304 */
};
template <> struct K<double> {
    static constexpr double small = 0.00001;
    static constexpr double big = 1e31;
    static constexpr double v0[4] = {0.26726, 0.80178, 0.53452, 0.0};
    static constexpr double tap[3] = {0.444, 0.222, 0.111};
};
constexpr float K<float>::v0[4];
constexpr float K<float>::tap[3];
constexpr double K<double>::v0[4];
constexpr double K<double>::tap[3];

/* idx_grids / wt_grids (ASTC_Encode.hlsl:268-304): grid point g (0..3) of a
 * 4-point grid over 6 texels sits at texel coordinate 5g/3.  Its two bracketing
 * columns are c0 = {0,1,3,4}[g] and c0+1; the nearer one (round(5g/3)) is c0 for
 * even g and c0+1 for odd g.  Tap weight 0.444 / 0.222 / 0.111 for a tap that is
 * near in both / one / neither axis; taps ordered (r0,c0),(r0,c0+1),(r0+1,c0),(r0+1,c0+1). */
struct Taps { int idx[16][4]; int far_count[16][4]; };
const Taps& taps6() {
    static Taps t = [] {
        Taps r;
        const int c0[4] = {0, 1, 3, 4};
        for (int gy = 0; gy < 4; ++gy)
            for (int gx = 0; gx < 4; ++gx) {
                int g = gy * 4 + gx;
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        int k = dy * 2 + dx;
                        r.idx[g][k] = (c0[gy] + dy) * 6 + (c0[gx] + dx);
                        int near_x = (gx & 1) ? 1 : 0, near_y = (gy & 1) ? 1 : 0;
                        r.far_count[g][k] = (dx != near_x) + (dy != near_y);
                    }
            }
        return r;
    }();
    return t;
}

template <typename T>
struct Arith {
    bool canon;
    T mad(T a, T b, T c) const { return canon ? std::fma(a, b, c) : a * b + c; }
    /* dot (HLSL dp4): canonical = fma chain over channels 0..3 starting with a product. */
    T dot(const T* a, const T* b) const {
        if (canon) return std::fma(a[3], b[3], std::fma(a[2], b[2], std::fma(a[1], b[1], a[0] * b[0])));
        return ((a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]) + a[3] * b[3];
    }
    /* Sum over the block's texels of f(k), left to right over k
     * (ASTC_Encode.hlsl:143-146).  Canonical 6x6: the two halves of the block,
     * rows 0-2 (texels 0-17) and rows 3-5 (texels 18-35), are each summed left
     * to right and the two sums added, so that each half can be summed by its
     * own lane and the halves combined with one exchange. */
    template <typename F> T block_sum(int dim, F f) const {
        const int n = dim * dim, h = (canon && dim == 6) ? n / 2 : n;
        T s = f(0);
        for (int k = 1; k < h; ++k) s = s + f(k);
        if (h == n) return s;
        T s2 = f(h);
        for (int k = h + 1; k < n; ++k) s2 = s2 + f(k);
        return s + s2;
    }
    /* Sum over texels of a[k]*b[k]: canonical is one fma chain starting with a
     * product (for 6x6, one per half as in block_sum, then added); source adds
     * separately rounded products (ASTC_Encode.hlsl:151-161). */
    T block_dot(int dim, const T* a, const T* b, int stride) const {
        const int n = dim * dim;
        if (!canon) {
            T s = 0;
            for (int k = 0; k < n; ++k) s = s + a[k * stride] * b[k * stride];
            return s;
        }
        const int h = dim == 6 ? n / 2 : n;
        T r = a[0] * b[0];
        for (int k = 1; k < h; ++k) r = std::fma(a[k * stride], b[k * stride], r);
        if (h == n) return r;
        T r2 = a[h * stride] * b[h * stride];
        for (int k = h + 1; k < n; ++k) r2 = std::fma(a[k * stride], b[k * stride], r2);
        return r + r2;
    }
    /* x / n for the constant divisors of :147 and :162.  FXC turns a division by
     * a literal into a multiplication by its reciprocal; the canonical schedule
     * does the same, the source schedule divides. */
    T div_const(T x, int n) const { return canon ? x * (T(1) / T(n)) : x / T(n); }
};

template <typename T> T clamp255(T x) { return std::fmin(std::fmax(x, T(0)), T(255)); }

/* length(v) < SMALL_VALUE, with length = sqrt(dot(v, v)). */
template <typename T> bool short_dot(T dot) { return std::sqrt(dot) < K<T>::small; }
template <typename T> bool short_vec(const Arith<T>& ar, const T* v) { return short_dot(ar.dot(v, v)); }
/* Reciprocal square root of the canonical binary32 schedule: seed from the bit
 * pattern, 0x5f375a86 - (bits(x) >> 1), then three Newton steps
 * y = y * fma(-(0.5x * y), y, 1.5).  IEEE operations only, so every machine gets
 * the same bits.  (HLSL's normalize is compiled to the GPU's approximate rsq.) */
float rsqrt_nr(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    b = 0x5f375a86u - (b >> 1);
    float y;
    std::memcpy(&y, &b, 4);
    const float h = 0.5f * x;
    y = y * std::fma(-(h * y), y, 1.5f);
    y = y * std::fma(-(h * y), y, 1.5f);
    y = y * std::fma(-(h * y), y, 1.5f);
    return y;
}

/* normalize(v) = v * rsqrt(dot(v, v)): the Newton rsqrt in the canonical binary32
 * schedule, a correctly rounded 1 / sqrt otherwise (source schedule, FP64 twin). */
template <typename T> void normalize4(const Arith<T>& ar, T* v) {
    const T d = ar.dot(v, v);
    T s;
    if (std::is_same<T, float>::value && ar.canon) s = (T)rsqrt_nr((float)d);
    else s = T(1) / std::sqrt(d);
    for (int c = 0; c < 4; ++c) v[c] = v[c] * s;
}

/* mul(m, v) for the symmetric 4x4 covariance (rows are dot products). */
template <typename T> void matvec(const Arith<T>& ar, const T m[4][4], const T* v, T* out) {
    for (int i = 0; i < 4; ++i) out[i] = ar.dot(m[i], v);
}

/* eigen_vector in the canonical schedule.  One iteration of :97-104 is
 * v <- normalize(M (M v)) after the test length(M v) < eps (which returns M v).
 * The first iteration is taken as written: M v0, its test, M (M v0).  Past it
 * the test cannot succeed: with S = M M (S_ij = dot(row i, row j)), |M v|^2 =
 * dot(v, S v) is the Rayleigh quotient of the positive semi-definite S at the
 * unit v, which never decreases along power iterates.  The other seven
 * iterations multiply by S; the eighth product is normalised.  The iterate is
 * rescaled by a power of two (exact; only the direction feeds the next product)
 * before products 1, 4 and 7: three products from a largest component in
 * [0.5, 1) stay in the normal range (each grows it by at most 4 max|S_ij| < 8e10
 * and shrinks it by at most half the first test's 1e-10 bound on the Rayleigh
 * quotient), and scaling by powers of two commutes with rounding. */
template <typename T> int exponent_of(T x) {
    int e;
    std::frexp(x, &e);
    return e;
}
template <typename T> void eigen_canonical(const Arith<T>& ar, const T m[4][4], T v[4]) {
    T s2[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = i; j < 4; ++j) s2[i][j] = s2[j][i] = ar.dot(m[i], m[j]);
    T u[4], w[4];
    matvec(ar, m, v, u);
    if (short_vec(ar, u)) {
        std::memcpy(v, u, sizeof(u));
        return;
    }
    matvec(ar, m, u, w);
    for (int it = 1; it < 8; ++it) {
        if (it % 3 == 1) {
            const T mx = std::fmax(std::fmax(std::fabs(w[0]), std::fabs(w[1])), std::fmax(std::fabs(w[2]), std::fabs(w[3])));
            const int e = exponent_of(mx);
            for (int i = 0; i < 4; ++i) v[i] = std::ldexp(w[i], -e);
        } else {
            std::memcpy(v, w, sizeof(w));
        }
        matvec(ar, s2, v, w);
    }
    std::memcpy(v, w, sizeof(w));
    normalize4(ar, v);
}

/* calculate_normal_weights (:316-372) in the canonical schedule.  The weight of
 * grid point i is dot(k, sample_i - e0), rescaled to [0, 1] by the minimum and
 * range over the grid, so dot(k, e0) cancels and the projections
 * p_i = dot(k, sample_i) are rescaled instead.  A 6x6 sample is a 4-tap
 * combination of texels (:307-314), and its projection the same combination of
 * texel projections. */
template <typename T>
void weights_canonical(const Arith<T>& ar, const T* tex, int dim, const T axis[4], T projw[16]) {
    if (dim == 6) {
        T pt[36];
        for (int j = 0; j < 36; ++j) pt[j] = ar.dot(axis, &tex[j * 4]);
        const Taps& tp = taps6();
        for (int i = 0; i < 16; ++i) {
            T s = pt[tp.idx[i][0]] * K<T>::tap[tp.far_count[i][0]];
            for (int k = 1; k < 4; ++k) s = std::fma(pt[tp.idx[i][k]], K<T>::tap[tp.far_count[i][k]], s);
            projw[i] = s;
        }
    } else {
        for (int i = 0; i < 16; ++i) projw[i] = ar.dot(axis, &tex[i * 4]);
    }
    T minw = K<T>::big, maxw = -K<T>::big;
    for (int i = 0; i < 16; ++i) {
        minw = std::fmin(projw[i], minw);
        maxw = std::fmax(projw[i], maxw);
    }
    const T invlen = T(1) / std::fmax(K<T>::small, maxw - minw);
    for (int i = 0; i < 16; ++i) projw[i] = (projw[i] - minw) * invlen;
}

/* max_accumulation_pixel_direction (:170-227), the endpoint search the shader
 * leaves commented out (:514): for each channel c, the sum of the texel
 * differences from the mean over the texels whose channel c is above the mean;
 * the axis is the longest of these sums (alpha competes only with has_alpha),
 * normalised unless shorter than 1e-5.  Sums run left to right over the texels
 * in both schedules. */
template <typename T>
void max_accumulation_axis(const Arith<T>& ar, const T* d, int N, bool has_alpha, T v[4]) {
    T sum[4][4] = {};
    for (int k = 0; k < N; ++k)
        for (int c = 0; c < 4; ++c)
            if (d[k * 4 + c] > T(0))
                for (int j = 0; j < 4; ++j) sum[c][j] = sum[c][j] + d[k * 4 + j];
    int best = 0;
    T maxdot = ar.dot(sum[0], sum[0]);
    for (int c = 1; c < (has_alpha ? 4 : 3); ++c) {
        const T dc = ar.dot(sum[c], sum[c]);
        if (dc > maxdot) {
            maxdot = dc;
            best = c;
        }
    }
    std::memcpy(v, sum[best], 4 * sizeof(T));
    if (!short_vec(ar, v)) normalize4(ar, v);
}

template <typename T>
void encode_block_impl(const T* tex /* N*4 */, int dim, bool has_alpha, int sched, uint8_t out[16],
                       oracle_block_trace* trace, int method = ORACLE_METHOD_PCA) {
    const Arith<T> ar{sched == ORACLE_SCHED_CANONICAL};
    const int N = dim * dim;

    /* principal_component_analysis (:139-168) */
    T mean[4];
    for (int c = 0; c < 4; ++c) mean[c] = ar.div_const(ar.block_sum(dim, [&](int k) { return tex[k * 4 + c]; }), N);
    std::vector<T> d((size_t)N * 4);
    for (int k = 0; k < N; ++k)
        for (int c = 0; c < 4; ++c) d[k * 4 + c] = tex[k * 4 + c] - mean[c];
    T cov[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = i; j < 4; ++j) {
            T s = ar.block_dot(dim, &d[i], &d[j], 4);
            cov[i][j] = cov[j][i] = ar.div_const(s, N - 1);
        }

    /* eigen_vector (:93-106): v = M v; return v if short; v = normalize(M v). */
    T v[4] = {K<T>::v0[0], K<T>::v0[1], K<T>::v0[2], K<T>::v0[3]};
    if (method == ORACLE_METHOD_MAX_ACCUMULATION) {
        max_accumulation_axis(ar, d.data(), N, has_alpha, v);
    } else if (ar.canon) {
        eigen_canonical(ar, cov, v);
    } else {
        for (int it = 0; it < 8; ++it) {
            T u[4];
            matvec(ar, cov, v, u);
            std::memcpy(v, u, sizeof(v));
            if (short_vec(ar, v)) break;
            matvec(ar, cov, v, u);
            normalize4(ar, u);
            std::memcpy(v, u, sizeof(v));
        }
    }

    /* find_min_max (:108-137): t_k = dot(texel_k - mean, v).  Canonical: the
     * extreme values of dot(texel_k, v), minus dot(mean, v) (rounding is
     * monotonic, so this equals min/max over k of dot(texel_k, v) - dot(mean, v)). */
    T a = K<T>::big, b = -K<T>::big;
    if (ar.canon) {
        for (int k = 0; k < N; ++k) {
            T t = ar.dot(&tex[k * 4], v);
            a = std::fmin(a, t);
            b = std::fmax(b, t);
        }
        const T mv = ar.dot(mean, v);
        a = a - mv;
        b = b - mv;
    } else {
        for (int k = 0; k < N; ++k) {
            T t = ar.dot(&d[k * 4], v);
            a = std::fmin(a, t);
            b = std::fmax(b, t);
        }
    }
    T e0[4], e1[4];
    for (int c = 0; c < 4; ++c) {
        e0[c] = clamp255(ar.mad(v[c], a, mean[c]));
        e1[c] = clamp255(ar.mad(v[c], b, mean[c]));
    }
    T s0 = (std::rint(e0[0]) + std::rint(e0[1])) + std::rint(e0[2]);
    T s1 = (std::rint(e1[0]) + std::rint(e1[1])) + std::rint(e1[2]);
    if (s0 > s1) std::swap(e0, e1);
    if (!has_alpha) { e0[3] = T(255); e1[3] = T(255); }

    /* encode_color (:233-245): round half to even, channel-interleaved. */
    uint32_t ep_q[8];
    for (int c = 0; c < 4; ++c) {
        ep_q[2 * c] = (uint32_t)std::rint(e0[c]);
        ep_q[2 * c + 1] = (uint32_t)std::rint(e1[c]);
    }

    /* calculate_normal_weights (:316-372) on the float endpoints. */
    T projw[16];
    T axis[4];
    for (int c = 0; c < 4; ++c) axis[c] = e1[c] - e0[c];
    const uint32_t range = has_alpha ? 5u : 11u;          /* weight_range - 1 (:519-521, :540) */
    uint32_t w_q[16];
    if (short_vec(ar, axis)) {
        for (int i = 0; i < 16; ++i) projw[i] = T(0);
    } else if (ar.canon) {
        normalize4(ar, axis);
        weights_canonical(ar, tex, dim, axis, projw);
    } else {
        normalize4(ar, axis);
        T minw = K<T>::big, maxw = -K<T>::big;
        for (int i = 0; i < 16; ++i) {
            T diff[4];
            if (dim == 6) {
                const Taps& tp = taps6();
                for (int c = 0; c < 4; ++c) {
                    T s = tex[tp.idx[i][0] * 4 + c] * K<T>::tap[tp.far_count[i][0]];
                    for (int k = 1; k < 4; ++k) s = ar.mad(tex[tp.idx[i][k] * 4 + c], K<T>::tap[tp.far_count[i][k]], s);
                    diff[c] = s - e0[c];
                }
            } else {
                for (int c = 0; c < 4; ++c) diff[c] = tex[i * 4 + c] - e0[c];
            }
            T w = ar.dot(axis, diff);
            minw = std::fmin(w, minw);
            maxw = std::fmax(w, maxw);
            projw[i] = w;
        }
        T invlen = T(1) / std::fmax(K<T>::small, maxw - minw);
        for (int i = 0; i < 16; ++i) projw[i] = (projw[i] - minw) * invlen;
    }
    /* quantize_weight (:256-260) */
    for (int i = 0; i < 16; ++i) {
        T q = std::rint(projw[i] * T(range));
        uint32_t qi = (uint32_t)q;
        w_q[i] = std::min(qi, range);
    }

    pack_block(ep_q, w_q, has_alpha, out);

    if (trace) {
        for (int c = 0; c < 4; ++c) {
            trace->mean[c] = (double)mean[c];
            trace->axis[c] = (double)v[c];
            trace->e0[c] = (double)e0[c];
            trace->e1[c] = (double)e1[c];
        }
        std::memcpy(trace->ep_q, ep_q, sizeof(ep_q));
        std::memcpy(trace->w_q, w_q, sizeof(w_q));
    }
}

/* ------------------------------------------------------------------------- */
/* Fetch (ASTC_Encode.hlsl:561-580 + D3D11 texture Load semantics)            */
/* ------------------------------------------------------------------------- */
double srgb_to_linear(double u) {
    return u <= 0.04045 ? u / 12.92 : std::pow((u + 0.055) / 1.055, 2.4);
}
/* Binary32 sRGB table: the texture unit returns float(lin(c/255)) and the shader
 * multiplies by 255.0f (:579). */
struct SrgbLut {
    float f[256];
    double d[256];
    SrgbLut() {
        for (int c = 0; c < 256; ++c) {
            double l = srgb_to_linear(c / 255.0);
            f[c] = (float)l * 255.0f;
            d[c] = l * 255.0;
        }
    }
};
const SrgbLut& srgb_lut() { static SrgbLut l; return l; }

/* UNORM8 -> float as D3D hardware converts it, a binary32 multiply by the
 * rounded reciprocal of 255, then the shader's x255 (:579): 126 of the 256
 * values land one ulp off the integer.  (A correctly rounded c / 255 would give
 * every integer back exactly; the reference's golden output shows the rounding
 * noise of an inexact conversion -- blocks whose exact covariance has no
 * colour-alpha coupling still found their alpha axis -- see the leaf tests.)
 * The binary64 twin keeps the ideal value c. */
struct UnormLut {
    float f[256];
    UnormLut() {
        for (int c = 0; c < 256; ++c) f[c] = ((float)c * (1.0f / 255.0f)) * 255.0f;
    }
};
const UnormLut& unorm_lut() { static UnormLut l; return l; }

/* RGBA8 of texel (x, y) with D3D's expansion of narrower formats:
 * RGB8 -> alpha 255; RG8 -> blue 0, alpha 255. */
void load_rgba8(const uint8_t* src, int row_pitch, int fmt, int x, int y, uint8_t rgba[4]) {
    const uint8_t* row = src + (size_t)y * row_pitch;
    if (fmt == ORACLE_FMT_RGBA8) { std::memcpy(rgba, row + 4 * x, 4); }
    else if (fmt == ORACLE_FMT_RGB8) { rgba[0] = row[3 * x]; rgba[1] = row[3 * x + 1]; rgba[2] = row[3 * x + 2]; rgba[3] = 255; }
    else { rgba[0] = row[2 * x]; rgba[1] = row[2 * x + 1]; rgba[2] = 0; rgba[3] = 255; }
}

template <typename T>
void fetch_block(const uint8_t* src, int width, int height, int row_pitch, int fmt, int dim, int flags,
                 int bx, int by, T* texels) {
    const bool srgb = flags & ORACLE_FLAG_SRGB, normal = flags & ORACLE_FLAG_NORMAL;
    for (int k = 0; k < dim * dim; ++k) {
        int x = bx * dim + k % dim, y = by * dim + k / dim;
        T t[4] = {0, 0, 0, 0};                                  /* out of bounds: all zero */
        if (x < width && y < height) {
            uint8_t p[4];
            load_rgba8(src, row_pitch, fmt, x, y, p);
            for (int c = 0; c < 4; ++c) {
                if (srgb && c < 3) t[c] = std::is_same<T, float>::value ? (T)srgb_lut().f[p[c]] : (T)srgb_lut().d[p[c]];
                else t[c] = std::is_same<T, float>::value ? (T)unorm_lut().f[p[c]] : (T)p[c];
            }
        }
        if (normal) { t[2] = T(255); t[3] = T(255); }          /* :575-578, then x255 */
        for (int c = 0; c < 4; ++c) texels[k * 4 + c] = t[c];
    }
}

/* ------------------------------------------------------------------------- */
/* Specification decoder (independent of the encoder above)                  */
/* ------------------------------------------------------------------------- */
uint32_t read_bits(const uint8_t* p, int pos, int n) {
    uint32_t v = 0;
    for (int i = 0; i < n; ++i) v |= (uint32_t)((p[(pos + i) >> 3] >> ((pos + i) & 7)) & 1) << i;
    return v;
}

int ise_bits(int quant, int count) {
    const QuantInfo& q = kQuant[quant];
    if (q.trits) return ((8 + 5 * q.bits) * count + 4) / 5;
    if (q.quints) return ((7 + 3 * q.bits) * count + 2) / 3;
    return q.bits * count;
}

void ise_decode(const uint8_t* p, int pos, int count, int quant, uint32_t* out) {
    const QuantInfo& q = kQuant[quant];
    const int b = q.bits;
    if (q.trits) {
        for (int g = 0; g < count; g += 5) {
            uint32_t m[5] = {0, 0, 0, 0, 0}, T = 0;
            const int tb[5] = {2, 2, 1, 2, 1};
            int tshift = 0;
            for (int i = 0; i < 5; ++i) {
                if (g + i >= count) break;
                m[i] = read_bits(p, pos, b); pos += b;
                T |= read_bits(p, pos, tb[i]) << tshift; pos += tb[i];
                tshift += tb[i];
            }
            int t[5];
            spec_decode_trits(T, t);
            for (int i = 0; i < 5 && g + i < count; ++i) out[g + i] = ((uint32_t)t[i] << b) | m[i];
        }
    } else if (q.quints) {
        for (int g = 0; g < count; g += 3) {
            uint32_t m[3] = {0, 0, 0}, Q = 0;
            const int qb[3] = {3, 2, 2};
            int qshift = 0;
            for (int i = 0; i < 3; ++i) {
                if (g + i >= count) break;
                m[i] = read_bits(p, pos, b); pos += b;
                Q |= read_bits(p, pos, qb[i]) << qshift; pos += qb[i];
                qshift += qb[i];
            }
            int qq[3];
            spec_decode_quints(Q, qq);
            for (int i = 0; i < 3 && g + i < count; ++i) out[g + i] = ((uint32_t)qq[i] << b) | m[i];
        }
    } else {
        for (int i = 0; i < count; ++i) { out[i] = read_bits(p, pos, b); pos += b; }
    }
}

/* Colour endpoint unquantisation (ISE symbol -> 0..255), spec C.2.13. */
uint32_t unquant_color(int quant, uint32_t sym) {
    const QuantInfo& q = kQuant[quant];
    const int b = q.bits;
    if (!q.trits && !q.quints) {
        uint32_t acc = 0; int n = 0;
        while (n < 8) { acc = (acc << b) | sym; n += b; }
        return acc >> (n - 8);
    }
    uint32_t D = sym >> b, low = sym & ((1u << b) - 1);
    uint32_t a = low & 1, bb = (low >> 1) & 1, cc = (low >> 2) & 1, dd = (low >> 3) & 1, ee = (low >> 4) & 1, ff = (low >> 5) & 1;
    uint32_t A = a ? 0x1FF : 0, B = 0, C = 0;
    if (q.trits) {
        switch (b) {
        case 1: C = 204; B = 0; break;
        case 2: C = 93; B = (bb << 8) | (bb << 4) | (bb << 2) | (bb << 1); break;
        case 3: C = 44; B = (cc << 8) | (bb << 7) | (cc << 3) | (bb << 2) | (cc << 1) | bb; break;
        case 4: C = 22; B = (dd << 8) | (cc << 7) | (bb << 6) | (dd << 2) | (cc << 1) | bb; break;
        case 5: C = 11; B = (ee << 8) | (dd << 7) | (cc << 6) | (bb << 5) | (ee << 1) | dd; break;
        default: C = 5; B = (ff << 8) | (ee << 7) | (dd << 6) | (cc << 5) | (bb << 4) | ff; break;
        }
    } else {
        switch (b) {
        case 1: C = 113; B = 0; break;
        case 2: C = 54; B = (bb << 8) | (bb << 3) | (bb << 2); break;
        case 3: C = 26; B = (cc << 8) | (bb << 7) | (cc << 2) | (bb << 1) | cc; break;
        case 4: C = 13; B = (dd << 8) | (cc << 7) | (bb << 6) | (dd << 1) | cc; break;
        default: C = 6; B = (ee << 8) | (dd << 7) | (cc << 6) | (bb << 5) | ee; break;
        }
    }
    uint32_t T = D * C + B;
    T ^= A;
    return (A & 0x80) | (T >> 2);
}

struct BlockMode { bool ok; int gw, gh, dual, quant; };
/* 2D block mode decoding, spec Table C.2.8 + weight range C.2.7. */
BlockMode decode_block_mode(uint32_t m) {
    BlockMode r{false, 0, 0, 0, 0};
    int R, H = (m >> 9) & 1, D = (m >> 10) & 1;
    if (m & 3) {
        R = (int)(((m & 3) << 1) | ((m >> 4) & 1));
        int A = (m >> 5) & 3, B = (m >> 7) & 3;
        switch ((m >> 2) & 3) {
        case 0: r.gw = B + 4; r.gh = A + 2; break;
        case 1: r.gw = B + 8; r.gh = A + 2; break;
        case 2: r.gw = A + 2; r.gh = B + 8; break;
        default:
            if (m & 0x100) { r.gw = (B & 1) + 2; r.gh = A + 2; }
            else { r.gw = A + 2; r.gh = (B & 1) + 6; }
        }
    } else {
        if ((m & 0xF) == 0) return r;
        R = (int)(((m >> 1) & 6) | ((m >> 4) & 1));
        int A = (m >> 5) & 3, B = (m >> 9) & 3;
        switch ((m >> 7) & 3) {
        case 0: r.gw = 12; r.gh = A + 2; break;
        case 1: r.gw = A + 2; r.gh = 12; break;
        case 2: r.gw = A + 6; r.gh = B + 6; D = 0; H = 0; break;
        default: {
            int s = (m >> 5) & 3;
            if (s == 0) { r.gw = 6; r.gh = 10; }
            else if (s == 1) { r.gw = 10; r.gh = 6; }
            else return r;
        }
        }
    }
    if (R < 2) return r;
    r.quant = (R - 2) + 6 * H;
    r.dual = D;
    r.ok = true;
    return r;
}

void fill(uint8_t* rgba, int n, uint8_t r, uint8_t g, uint8_t b, uint8_t a) {
    for (int i = 0; i < n; ++i) { rgba[4 * i] = r; rgba[4 * i + 1] = g; rgba[4 * i + 2] = b; rgba[4 * i + 3] = a; }
}

/* texel_w (optional): each texel's infilled weight 0..64, plane 0 at [t * bw + s]
 * and, for dual-plane blocks, plane 1 at [bw * bh + t * bw + s]. */
int decode_block(const uint8_t blk[16], int bw, int bh, uint8_t* rgba, oracle_block_info* info,
                 int32_t* texel_w = nullptr) {
    oracle_block_info local;
    oracle_block_info& I = info ? *info : local;
    std::memset(&I, 0, sizeof(I));
    const int n = bw * bh;
    uint32_t mode = read_bits(blk, 0, 11);
    if ((mode & 0x1FF) == 0x1FC) {                            /* void extent */
        I.void_extent = 1;
        if (mode & 0x200) { fill(rgba, n, 255, 0, 255, 255); return 0; }  /* HDR: unsupported */
        uint8_t c[4];
        for (int i = 0; i < 4; ++i) c[i] = (uint8_t)(read_bits(blk, 64 + 16 * i, 16) >> 8);
        fill(rgba, n, c[0], c[1], c[2], c[3]);
        I.valid = 1;
        return 1;
    }
    BlockMode bm = decode_block_mode(mode);
    int parts = (int)read_bits(blk, 11, 2) + 1;
    I.partitions = parts;
    if (!bm.ok || parts != 1 || bm.gw > bw || bm.gh > bh) { fill(rgba, n, 255, 0, 255, 255); return 0; }
    int nw = bm.gw * bm.gh * (bm.dual ? 2 : 1);
    int wbits = ise_bits(bm.quant, nw);
    if (nw > 64 || wbits < 24 || wbits > 96) { fill(rgba, n, 255, 0, 255, 255); return 0; }
    int cem = (int)read_bits(blk, 13, 4);
    int nvals = 2 * ((cem >> 2) + 1);
    int avail = 128 - 17 - wbits - (bm.dual ? 2 : 0);
    int equant = -1;
    for (int q = QUANT_256; q >= QUANT_6; --q)
        if (ise_bits(q, nvals) <= avail) { equant = q; break; }
    if (equant < 0 || (cem != 0 && cem != 4 && cem != 8 && cem != 12)) { fill(rgba, n, 255, 0, 255, 255); return 0; }
    I.grid_w = bm.gw; I.grid_h = bm.gh; I.dual_plane = bm.dual; I.cem = cem;
    I.weight_quant = bm.quant; I.endpoint_quant = equant;

    uint32_t ev[8];
    ise_decode(blk, 17, nvals, equant, ev);
    uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < nvals; ++i) v[i] = unquant_color(equant, ev[i]);
    std::memcpy(I.ep_q, v, sizeof(v));
    int e0[4], e1[4];
    auto bc = [](int* e) { e[0] = (e[0] + e[2]) >> 1; e[1] = (e[1] + e[2]) >> 1; };
    if (cem == 0) { for (int c = 0; c < 3; ++c) { e0[c] = (int)v[0]; e1[c] = (int)v[1]; } e0[3] = e1[3] = 255; }
    else if (cem == 4) { for (int c = 0; c < 3; ++c) { e0[c] = (int)v[0]; e1[c] = (int)v[1]; } e0[3] = (int)v[2]; e1[3] = (int)v[3]; }
    else {
        int a0 = cem == 12 ? (int)v[6] : 255, a1 = cem == 12 ? (int)v[7] : 255;
        int s0 = (int)(v[0] + v[2] + v[4]), s1 = (int)(v[1] + v[3] + v[5]);
        if (s1 >= s0) {
            e0[0] = (int)v[0]; e0[1] = (int)v[2]; e0[2] = (int)v[4]; e0[3] = a0;
            e1[0] = (int)v[1]; e1[1] = (int)v[3]; e1[2] = (int)v[5]; e1[3] = a1;
        } else {
            e0[0] = (int)v[1]; e0[1] = (int)v[3]; e0[2] = (int)v[5]; e0[3] = a1;
            e1[0] = (int)v[0]; e1[1] = (int)v[2]; e1[2] = (int)v[4]; e1[3] = a0;
            bc(e0); bc(e1);
        }
    }

    /* Weights: stored bit-reversed from bit 127 downwards. */
    uint8_t rev[16];
    for (int i = 0; i < 16; ++i) rev[i] = (uint8_t)reverse_byte(blk[15 - i]);
    uint32_t wsym[64];
    ise_decode(rev, 0, nw, bm.quant, wsym);
    for (int i = 0; i < nw; ++i) {
        I.w_sym[i] = wsym[i];
        I.w_unq[i] = (uint32_t)spec_unquant_weight(bm.quant, (int)wsym[i]);
    }
    int ccs = bm.dual ? (int)read_bits(blk, 128 - wbits - 2, 2) : -1;

    /* Weight infill, spec C.2.18. */
    int Ds = (1024 + bw / 2) / (bw - 1), Dt = (1024 + bh / 2) / (bh - 1);
    const int planes = bm.dual ? 2 : 1;
    for (int t = 0; t < bh; ++t)
        for (int s = 0; s < bw; ++s) {
            int cs = Ds * s, ct = Dt * t;
            int gs = (cs * (bm.gw - 1) + 32) >> 6, gt = (ct * (bm.gh - 1) + 32) >> 6;
            int js = gs >> 4, fs = gs & 15, jt = gt >> 4, ft = gt & 15;
            int v0 = js + jt * bm.gw;
            int w11 = (fs * ft + 8) >> 4, w10 = ft - w11, w01 = fs - w11, w00 = 16 - fs - ft + w11;
            int wt[2];
            for (int p = 0; p < planes; ++p) {
                auto W = [&](int idx) -> int {
                    int i = idx * planes + p;
                    return i < nw ? (int)I.w_unq[i] : 0;
                };
                int acc = W(v0) * w00;
                if (w01) acc += W(v0 + 1) * w01;
                if (w10) acc += W(v0 + bm.gw) * w10;
                if (w11) acc += W(v0 + bm.gw + 1) * w11;
                wt[p] = (acc + 8) >> 4;
                if (texel_w) texel_w[p * bw * bh + t * bw + s] = wt[p];
            }
            uint8_t* o = rgba + 4 * (t * bw + s);
            for (int c = 0; c < 4; ++c) {
                int w = (c == ccs) ? wt[1] : wt[0];
                int C0 = e0[c] * 257, C1 = e1[c] * 257;
                int C = (C0 * (64 - w) + C1 * w + 32) >> 6;
                o[c] = (uint8_t)(C >> 8);
            }
        }
    I.valid = 1;
    return 1;
}

}  // namespace

/* ------------------------------------------------------------------------- */
/* C API                                                                      */
/* ------------------------------------------------------------------------- */
extern "C" {

int oracle_encode_block_f32(const float* texels, int dim, int has_alpha, int sched, uint8_t out[16],
                            oracle_block_trace* trace) {
    return oracle_encode_block_method_f32(texels, dim, has_alpha, sched, ORACLE_METHOD_PCA, out, trace);
}

int oracle_encode_block_f64(const double* texels, int dim, int has_alpha, int sched, uint8_t out[16],
                            oracle_block_trace* trace) {
    return oracle_encode_block_method_f64(texels, dim, has_alpha, sched, ORACLE_METHOD_PCA, out, trace);
}

int oracle_encode_block_method_f32(const float* texels, int dim, int has_alpha, int sched, int method,
                                   uint8_t out[16], oracle_block_trace* trace) {
    if (dim != 4 && dim != 6) return -1;
    encode_block_impl<float>(texels, dim, has_alpha != 0, sched, out, trace, method);
    return 0;
}

int oracle_encode_block_method_f64(const double* texels, int dim, int has_alpha, int sched, int method,
                                   uint8_t out[16], oracle_block_trace* trace) {
    if (dim != 4 && dim != 6) return -1;
    encode_block_impl<double>(texels, dim, has_alpha != 0, sched, out, trace, method);
    return 0;
}

void oracle_fetch_block_f32(const uint8_t* src, int width, int height, int row_pitch, int src_format, int dim,
                            int flags, int bx, int by, float* texels) {
    fetch_block<float>(src, width, height, row_pitch, src_format, dim, flags, bx, by, texels);
}

void oracle_fetch_block_f64(const uint8_t* src, int width, int height, int row_pitch, int src_format, int dim,
                            int flags, int bx, int by, double* texels) {
    fetch_block<double>(src, width, height, row_pitch, src_format, dim, flags, bx, by, texels);
}

int64_t oracle_encode_image(const uint8_t* src, int width, int height, int row_pitch, int src_format, int dim,
                            int flags, int sched, int threads, uint8_t* out_blocks) {
    if ((dim != 4 && dim != 6) || width <= 0 || height <= 0) return -1;
    const int xb = (width + dim - 1) / dim, yb = (height + dim - 1) / dim;   /* astc_encode.h:127-128 */
    const int64_t nblk = (int64_t)xb * yb;
    const bool alpha = flags & ORACLE_FLAG_ALPHA;
    const int method = (flags & ORACLE_FLAG_MAX_ACCUMULATION) ? ORACLE_METHOD_MAX_ACCUMULATION : ORACLE_METHOD_PCA;
    (void)tables(); (void)taps6(); (void)srgb_lut();
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t id = 0; id < nblk; ++id) {
        float tex[36 * 4];
        int by = (int)(id / xb), bx = (int)(id - (int64_t)by * xb);       /* :568-569 */
        fetch_block<float>(src, width, height, row_pitch, src_format, dim, flags, bx, by, tex);
        encode_block_impl<float>(tex, dim, alpha, sched, out_blocks + 16 * id, nullptr, method);
    }
    return nblk;
}

void oracle_fp64_twin(const uint8_t* src, int width, int height, int row_pitch, int src_format, int dim, int flags,
                      int threads, double* axis_err, double* endpoint_err, uint8_t* differ) {
    /* The same method run in binary32 (canonical schedule, what the GPU runs)
     * and in binary64 on every block of an image (ASTC_Encode.hlsl:510-550):
     * axis error = |a32 - a64| up to sign (-1 where the binary64 axis is zero);
     * endpoint error = largest |e32 - e64| over the unrounded endpoint
     * channels, in 8-bit units; differ = the two blocks' bytes differ. */
    const int xb = (width + dim - 1) / dim, yb = (height + dim - 1) / dim;
    const int64_t nblk = (int64_t)xb * yb;
    const bool alpha = flags & ORACLE_FLAG_ALPHA;
    const int method = (flags & ORACLE_FLAG_MAX_ACCUMULATION) ? ORACLE_METHOD_MAX_ACCUMULATION : ORACLE_METHOD_PCA;
    (void)tables(); (void)taps6(); (void)srgb_lut(); (void)unorm_lut();
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t id = 0; id < nblk; ++id) {
        float t32[36 * 4];
        double t64[36 * 4];
        const int by = (int)(id / xb), bx = (int)(id - (int64_t)by * xb);
        fetch_block<float>(src, width, height, row_pitch, src_format, dim, flags, bx, by, t32);
        fetch_block<double>(src, width, height, row_pitch, src_format, dim, flags, bx, by, t64);
        uint8_t b32[16], b64[16];
        oracle_block_trace r32, r64;
        encode_block_impl<float>(t32, dim, alpha, ORACLE_SCHED_CANONICAL, b32, &r32, method);
        encode_block_impl<double>(t64, dim, alpha, ORACLE_SCHED_CANONICAL, b64, &r64, method);
        double nrm = 0.0, dm = 0.0, dp = 0.0, ep = 0.0;
        for (int c = 0; c < 4; ++c) {
            nrm += r64.axis[c] * r64.axis[c];
            dm += (r32.axis[c] - r64.axis[c]) * (r32.axis[c] - r64.axis[c]);
            dp += (r32.axis[c] + r64.axis[c]) * (r32.axis[c] + r64.axis[c]);
            ep = std::max(ep, std::max(std::fabs(r32.e0[c] - r64.e0[c]), std::fabs(r32.e1[c] - r64.e1[c])));
        }
        axis_err[id] = nrm > 0.0 ? std::sqrt(std::min(dm, dp)) : -1.0;
        endpoint_err[id] = ep;
        differ[id] = std::memcmp(b32, b64, 16) != 0;
    }
}

void oracle_srgb_table(float out[256]) { std::memcpy(out, srgb_lut().f, sizeof(float) * 256); }
void oracle_unorm_table(float out[256]) { std::memcpy(out, unorm_lut().f, sizeof(float) * 256); }

void oracle_tables(int32_t trits_out[243], int32_t quints_out[125], int32_t scramble_out[12 * 32],
                   int32_t btq_out[21 * 3]) {
    const Tables& t = tables();
    for (int i = 0; i < 243; ++i) trits_out[i] = t.trits[i];
    for (int i = 0; i < 125; ++i) quints_out[i] = t.quints[i];
    for (int i = 0; i < 12 * 32; ++i) scramble_out[i] = t.scramble[i];
    for (int i = 0; i < 21; ++i) {
        btq_out[3 * i] = kQuant[i].bits;
        btq_out[3 * i + 1] = kQuant[i].trits;
        btq_out[3 * i + 2] = kQuant[i].quints;
    }
}

int oracle_bise(const uint32_t* numbers, int count, int quant, uint32_t out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0;
    return bise(numbers, count, quant, out);
}

uint32_t oracle_blockmode(int weight_quant) { return assemble_blockmode((uint32_t)weight_quant); }

void oracle_assemble_block(uint32_t blockmode, uint32_t cem, const uint32_t ep_ise[4], const uint32_t wt_ise[4],
                           uint8_t out[16]) {
    uint32_t phy[4];
    assemble_block(blockmode, cem, ep_ise, wt_ise, phy);
    store_le(phy, out);
}

void oracle_pack_block(const uint32_t ep_q[8], const uint32_t w_q[16], int has_alpha, uint8_t out[16]) {
    pack_block(ep_q, w_q, has_alpha != 0, out);
}

int oracle_decode_block(const uint8_t blk[16], int dim_x, int dim_y, uint8_t* rgba, oracle_block_info* info) {
    return decode_block(blk, dim_x, dim_y, rgba, info);
}

int oracle_decode_block_weights(const uint8_t blk[16], int dim_x, int dim_y, uint8_t* rgba, int32_t* texel_w) {
    return decode_block(blk, dim_x, dim_y, rgba, nullptr, texel_w);
}

void oracle_grid6_taps(int32_t idx[64], float wt[64]) {
    const Taps& tp = taps6();
    for (int g = 0; g < 16; ++g)
        for (int k = 0; k < 4; ++k) {
            idx[4 * g + k] = tp.idx[g][k];
            wt[4 * g + k] = K<float>::tap[tp.far_count[g][k]];
        }
}

int oracle_decode_image(const uint8_t* blocks, int width, int height, int dim, uint8_t* rgba) {
    const int xb = (width + dim - 1) / dim, yb = (height + dim - 1) / dim;
    int ok = 1;
#pragma omp parallel for schedule(dynamic, 64) reduction(& : ok)
    for (int64_t id = 0; id < (int64_t)xb * yb; ++id) {
        uint8_t tile[36 * 4];
        int by = (int)(id / xb), bx = (int)(id % xb);
        ok &= decode_block(blocks + 16 * id, dim, dim, tile, nullptr);
        for (int k = 0; k < dim * dim; ++k) {
            int x = bx * dim + k % dim, y = by * dim + k / dim;
            if (x < width && y < height) std::memcpy(rgba + ((size_t)y * width + x) * 4, tile + 4 * k, 4);
        }
    }
    return ok;
}

int oracle_ise_decode(const uint8_t* bits, int bit_offset, int count, int quant, uint32_t* out) {
    if (quant < 0 || quant > 20) return -1;
    ise_decode(bits, bit_offset, count, quant, out);
    return ise_bits(quant, count);
}

void oracle_sq_error(const uint8_t* src, int width, int height, int row_pitch, int src_format, int flags,
                     const uint8_t* dec, double* sse_out, double* count_out) {
    const bool alpha = flags & ORACLE_FLAG_ALPHA, normal = flags & ORACLE_FLAG_NORMAL;
    const int nch = normal ? 2 : (alpha ? 4 : 3);
    double sse = 0.0;
#pragma omp parallel for reduction(+ : sse) schedule(static)
    for (int y = 0; y < height; ++y) {
        double row = 0.0;
        for (int x = 0; x < width; ++x) {
            float t[4];
            uint8_t p[4];
            load_rgba8(src, row_pitch, src_format, x, y, p);
            for (int c = 0; c < 4; ++c)
                t[c] = ((flags & ORACLE_FLAG_SRGB) && c < 3) ? srgb_lut().f[p[c]] : (float)p[c];
            const uint8_t* o = dec + ((size_t)y * width + x) * 4;
            for (int c = 0; c < nch; ++c) {
                double e = (double)t[c] - (double)o[c];
                row += e * e;
            }
        }
        sse += row;
    }
    *sse_out = sse;
    *count_out = (double)width * height * nch;
}

double oracle_psnr(const uint8_t* src, int width, int height, int row_pitch, int src_format, int flags,
                   const uint8_t* dec) {
    double sse, cnt;
    oracle_sq_error(src, width, height, row_pitch, src_format, flags, dec, &sse, &cnt);
    if (sse <= 0.0) return INFINITY;
    return 10.0 * std::log10(255.0 * 255.0 * cnt / sse);
}

float oracle_rsqrt_nr(float x) { return rsqrt_nr(x); }

float oracle_length_threshold(void) {
    /* Smallest x with sqrtf(x) >= SMALL_VALUE; sqrtf is monotone, so search near 1e-10. */
    float x = 1e-10f;
    while (std::sqrt(x) >= K<float>::small) x = std::nextafter(x, 0.0f);
    while (std::sqrt(x) < K<float>::small) x = std::nextafter(x, 1.0f);
    return x;
}

}  /* extern "C" */
