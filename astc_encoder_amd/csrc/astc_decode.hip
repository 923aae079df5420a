// astc_decode.hip -- ASTC LDR decoder kernel (one lane per block).
//
// Decodes the single-partition LDR subset of the ASTC specification to RGBA8:
// any block size up to 12x12, any 2D weight grid up to the block size with the
// specification's bilinear infill, every weight and endpoint quantisation (bits,
// trits, quints), one or two weight planes, colour endpoint modes 0 / 4 / 8 / 12
// (LDR luminance, luminance+alpha, RGB, RGBA direct, with blue contraction), and
// LDR void-extent blocks.  Anything else (multiple partitions, HDR, reserved
// modes) decodes to the error colour magenta.  Each texel is
// ((C0 * (64 - w) + C1 * w + 32) >> 6) >> 8 with C = endpoint * 257.
// It decodes every block the encoder produces and is used to check encoded
// textures at throughput (PSNR) without a round trip to the host.
#include "astc_internal.h"

namespace astc {

namespace {

constexpr int kDecThreads = 256;
constexpr int kMaxWeights = 64;

__device__ __forceinline__ uint32_t get_bits(uint64_t lo, uint64_t hi, int pos, int n) {
    uint64_t v;
    if (pos >= 64) v = hi >> (pos - 64);
    else if (pos == 0) v = lo;
    else v = (lo >> pos) | (hi << (64 - pos));
    return uint32_t(v) & ((1u << n) - 1u);
}

__device__ __forceinline__ int ise_bit_count(int q, int count) {
    const QuantLevel L = kQuantLevels[q];
    if (L.trits) return ((8 + 5 * L.bits) * count + 4) / 5;
    if (L.quints) return ((7 + 3 * L.bits) * count + 2) / 3;
    return L.bits * count;
}

// Decode `count` ISE values of method q starting at bit `pos`; value i goes to
// out[i * kDecThreads] (a per-lane column in LDS).
__device__ void ise_decode(uint64_t lo, uint64_t hi, int pos, int count, int q, const uint16_t* trit_lut,
                           const uint16_t* quint_lut, uint8_t* out) {
    const QuantLevel L = kQuantLevels[q];
    const int b = L.bits;
    if (L.trits) {
        for (int g = 0; g < count; g += 5) {
            uint32_t m[5] = {0, 0, 0, 0, 0}, T = 0;
            int tshift = 0;
            for (int i = 0; i < 5 && g + i < count; ++i) {
                const int tb = (i == 2 || i == 4) ? 1 : 2;
                m[i] = get_bits(lo, hi, pos, b);
                pos += b;
                T |= get_bits(lo, hi, pos, tb) << tshift;
                pos += tb;
                tshift += tb;
            }
            const uint32_t t = trit_lut[T];
            for (int i = 0; i < 5 && g + i < count; ++i)
                out[(g + i) * kDecThreads] = uint8_t((((t >> (2 * i)) & 3u) << b) | m[i]);
        }
    } else if (L.quints) {
        for (int g = 0; g < count; g += 3) {
            uint32_t m[3] = {0, 0, 0}, Q = 0;
            int qshift = 0;
            for (int i = 0; i < 3 && g + i < count; ++i) {
                const int qb = i == 0 ? 3 : 2;
                m[i] = get_bits(lo, hi, pos, b);
                pos += b;
                Q |= get_bits(lo, hi, pos, qb) << qshift;
                pos += qb;
                qshift += qb;
            }
            const uint32_t t = quint_lut[Q];
            for (int i = 0; i < 3 && g + i < count; ++i)
                out[(g + i) * kDecThreads] = uint8_t((((t >> (3 * i)) & 7u) << b) | m[i]);
        }
    } else {
        for (int i = 0; i < count; ++i, pos += b) out[i * kDecThreads] = uint8_t(get_bits(lo, hi, pos, b));
    }
}

// ISE symbol -> 0..255 colour value (specification colour unquantisation)
__device__ uint32_t unquant_color(int q, uint32_t sym) {
    const QuantLevel L = kQuantLevels[q];
    const int b = L.bits;
    if (!L.trits && !L.quints) {  // bit replication to 8 bits
        uint32_t acc = 0;
        int n = 0;
        while (n < 8) {
            acc = (acc << b) | sym;
            n += b;
        }
        return acc >> (n - 8);
    }
    const uint32_t D = sym >> b, low = sym & ((1u << b) - 1u);
    const uint32_t a = low & 1u, bb = (low >> 1) & 1u, cc = (low >> 2) & 1u, dd = (low >> 3) & 1u,
                   ee = (low >> 4) & 1u, ff = (low >> 5) & 1u;
    const uint32_t A = a ? 0x1FFu : 0u;
    uint32_t B = 0, C = 0;
    if (L.trits) {
        switch (b) {
        case 1: C = 204; break;
        case 2: C = 93; B = (bb << 8) | (bb << 4) | (bb << 2) | (bb << 1); break;
        case 3: C = 44; B = (cc << 8) | (bb << 7) | (cc << 3) | (bb << 2) | (cc << 1) | bb; break;
        case 4: C = 22; B = (dd << 8) | (cc << 7) | (bb << 6) | (dd << 2) | (cc << 1) | bb; break;
        case 5: C = 11; B = (ee << 8) | (dd << 7) | (cc << 6) | (bb << 5) | (ee << 1) | dd; break;
        default: C = 5; B = (ff << 8) | (ee << 7) | (dd << 6) | (cc << 5) | (bb << 4) | ff; break;
        }
    } else {
        switch (b) {
        case 1: C = 113; break;
        case 2: C = 54; B = (bb << 8) | (bb << 3) | (bb << 2); break;
        case 3: C = 26; B = (cc << 8) | (bb << 7) | (cc << 2) | (bb << 1) | cc; break;
        case 4: C = 13; B = (dd << 8) | (cc << 7) | (bb << 6) | (dd << 1) | cc; break;
        default: C = 6; B = (ee << 8) | (dd << 7) | (cc << 6) | (bb << 5) | ee; break;
        }
    }
    const uint32_t T = (D * C + B) ^ A;
    return (A & 0x80u) | (T >> 2);
}

// ISE symbol -> 0..64 weight (specification weight unquantisation)
__device__ uint32_t unquant_weight(int q, uint32_t sym) {
    const QuantLevel L = kQuantLevels[q];
    const int b = L.bits;
    uint32_t v;
    if (!L.trits && !L.quints) {
        uint32_t acc = 0;
        int n = 0;
        while (n < 6) {
            acc = (acc << b) | sym;
            n += b;
        }
        v = acc >> (n - 6);
    } else if (b == 0) {
        return sym * (L.trits ? 32u : 16u);
    } else {
        const uint32_t D = sym >> b, low = sym & ((1u << b) - 1u);
        const uint32_t A = (low & 1u) ? 0x7Fu : 0u, bb = (low >> 1) & 1u, cc = (low >> 2) & 1u;
        uint32_t B = 0, C = 0;
        if (L.trits) {
            if (b == 1) C = 50;
            else if (b == 2) { C = 23; B = (bb << 6) | (bb << 2) | bb; }
            else { C = 11; B = (cc << 6) | (bb << 5) | (cc << 1) | bb; }
        } else {
            if (b == 1) C = 28;
            else { C = 13; B = (bb << 6) | (bb << 1) | bb; }
        }
        v = (A & 0x20u) | (((D * C + B) ^ A) >> 2);
    }
    return v > 32 ? v + 1 : v;
}

struct Mode { int ok, gw, gh, dual, quant; };

__device__ Mode decode_mode(uint32_t m) {
    Mode r{0, 0, 0, 0, 0};
    int R, H = (m >> 9) & 1, D = (m >> 10) & 1;
    if (m & 3u) {
        R = int(((m & 3u) << 1) | ((m >> 4) & 1u));
        const int A = (m >> 5) & 3, B = (m >> 7) & 3;
        switch ((m >> 2) & 3) {
        case 0: r.gw = B + 4; r.gh = A + 2; break;
        case 1: r.gw = B + 8; r.gh = A + 2; break;
        case 2: r.gw = A + 2; r.gh = B + 8; break;
        default:
            if (m & 0x100u) { r.gw = (B & 1) + 2; r.gh = A + 2; }
            else { r.gw = A + 2; r.gh = (B & 1) + 6; }
        }
    } else {
        if ((m & 0xFu) == 0) return r;
        R = int(((m >> 1) & 6u) | ((m >> 4) & 1u));
        const int A = (m >> 5) & 3, B = (m >> 9) & 3;
        switch ((m >> 7) & 3) {
        case 0: r.gw = 12; r.gh = A + 2; break;
        case 1: r.gw = A + 2; r.gh = 12; break;
        case 2: r.gw = A + 6; r.gh = B + 6; D = 0; H = 0; break;
        default: {
            const int s = (m >> 5) & 3;
            if (s == 0) { r.gw = 6; r.gh = 10; }
            else if (s == 1) { r.gw = 10; r.gh = 6; }
            else return r;
        }
        }
    }
    if (R < 2) return r;
    r.quant = (R - 2) + 6 * H;
    r.dual = D;
    r.ok = 1;
    return r;
}

struct DecodeArgs {
    const uint4* blocks;
    uint32_t* rgba;  // width * height texels
    int width, height, bw, bh, xblocks, nblocks;
    const TritQuintDecodeTable* tables;
};

__device__ __forceinline__ void fill_block(const DecodeArgs& a, int x0, int y0, uint32_t c) {
    for (int t = 0; t < a.bh; ++t)
        for (int s = 0; s < a.bw; ++s) {
            const int x = x0 + s, y = y0 + t;
            if (x < a.width && y < a.height) a.rgba[(size_t)y * a.width + x] = c;
        }
}

__global__ void __launch_bounds__(kDecThreads) decode_kernel(DecodeArgs a) {
    __shared__ uint16_t trit_lut[256];
    __shared__ uint16_t quint_lut[128];
    __shared__ uint8_t wbuf[kMaxWeights * kDecThreads];
    __shared__ uint8_t ebuf[8 * kDecThreads];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) trit_lut[i] = a.tables->trits[i];
    for (int i = threadIdx.x; i < 128; i += blockDim.x) quint_lut[i] = a.tables->quints[i];
    __syncthreads();
    const int id = blockIdx.x * kDecThreads + threadIdx.x;
    if (id >= a.nblocks) return;
    const int by = id / a.xblocks, bx = id - by * a.xblocks;
    const int x0 = bx * a.bw, y0 = by * a.bh;
    const uint4 blk = a.blocks[id];
    const uint64_t lo = uint64_t(blk.x) | uint64_t(blk.y) << 32, hi = uint64_t(blk.z) | uint64_t(blk.w) << 32;
    const uint32_t kError = 0xFFFF00FFu;  // magenta, alpha 255
    const uint32_t mode = uint32_t(lo) & 0x7FFu;
    if ((mode & 0x1FFu) == 0x1FCu) {
        if (mode & 0x200u) { fill_block(a, x0, y0, kError); return; }
        const uint32_t c = uint32_t((hi >> 8) & 0xFF) | uint32_t((hi >> 24) & 0xFF) << 8 |
                           uint32_t((hi >> 40) & 0xFF) << 16 | uint32_t((hi >> 56) & 0xFF) << 24;
        fill_block(a, x0, y0, c);
        return;
    }
    const Mode m = decode_mode(mode);
    const int parts = int((lo >> 11) & 3) + 1;
    const int nw = m.gw * m.gh * (m.dual ? 2 : 1);
    if (!m.ok || parts != 1 || m.gw > a.bw || m.gh > a.bh || nw > kMaxWeights) { fill_block(a, x0, y0, kError); return; }
    const int wbits = ise_bit_count(m.quant, nw);
    const int cem = int((lo >> 13) & 0xF);
    if (wbits < 24 || wbits > 96 || (cem != 0 && cem != 4 && cem != 8 && cem != 12)) {
        fill_block(a, x0, y0, kError);
        return;
    }
    const int nvals = 2 * ((cem >> 2) + 1);
    const int avail = 128 - 17 - wbits - (m.dual ? 2 : 0);
    int eq = -1;
    for (int q = QUANT_256; q >= QUANT_6; --q)
        if (ise_bit_count(q, nvals) <= avail) { eq = q; break; }
    if (eq < 0) { fill_block(a, x0, y0, kError); return; }

    uint8_t* ecol = ebuf + threadIdx.x;
    ise_decode(lo, hi, 17, nvals, eq, trit_lut, quint_lut, ecol);
    int v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = i < nvals ? int(unquant_color(eq, ecol[i * kDecThreads])) : 0;
    int e0[4], e1[4];
    if (cem == 0) {
        e0[0] = e0[1] = e0[2] = v[0]; e1[0] = e1[1] = e1[2] = v[1]; e0[3] = e1[3] = 255;
    } else if (cem == 4) {
        e0[0] = e0[1] = e0[2] = v[0]; e1[0] = e1[1] = e1[2] = v[1]; e0[3] = v[2]; e1[3] = v[3];
    } else {
        const int a0 = cem == 12 ? v[6] : 255, a1 = cem == 12 ? v[7] : 255;
        if (v[1] + v[3] + v[5] >= v[0] + v[2] + v[4]) {
            e0[0] = v[0]; e0[1] = v[2]; e0[2] = v[4]; e0[3] = a0;
            e1[0] = v[1]; e1[1] = v[3]; e1[2] = v[5]; e1[3] = a1;
        } else {  // blue contraction, endpoints swapped
            e0[0] = (v[1] + v[5]) >> 1; e0[1] = (v[3] + v[5]) >> 1; e0[2] = v[5]; e0[3] = a1;
            e1[0] = (v[0] + v[4]) >> 1; e1[1] = (v[2] + v[4]) >> 1; e1[2] = v[4]; e1[3] = a0;
        }
    }

    // weights: the stream runs downwards from bit 127
    uint8_t* wcol = wbuf + threadIdx.x;
    const uint64_t rlo = __builtin_bitreverse64(hi), rhi = __builtin_bitreverse64(lo);
    ise_decode(rlo, rhi, 0, nw, m.quant, trit_lut, quint_lut, wcol);
    for (int i = 0; i < nw; ++i) wcol[i * kDecThreads] = uint8_t(unquant_weight(m.quant, wcol[i * kDecThreads]));
    const int ccs = m.dual ? int(get_bits(lo, hi, 128 - wbits - 2, 2)) : -1;
    const int planes = m.dual ? 2 : 1;

    // bilinear infill of the weight grid (specification), then interpolation
    const int Ds = (1024 + a.bw / 2) / (a.bw - 1), Dt = (1024 + a.bh / 2) / (a.bh - 1);
    for (int t = 0; t < a.bh; ++t) {
        const int y = y0 + t;
        if (y >= a.height) break;
        const int gt = (Dt * t * (m.gh - 1) + 32) >> 6;
        const int jt = gt >> 4, ft = gt & 15;
        for (int s = 0; s < a.bw; ++s) {
            const int x = x0 + s;
            if (x >= a.width) break;
            const int gs = (Ds * s * (m.gw - 1) + 32) >> 6;
            const int js = gs >> 4, fs = gs & 15;
            const int v0 = js + jt * m.gw;
            const int w11 = (fs * ft + 8) >> 4, w10 = ft - w11, w01 = fs - w11, w00 = 16 - fs - ft + w11;
            int wt[2];
            for (int p = 0; p < planes; ++p) {
                auto W = [&](int idx) -> int {
                    const int i = idx * planes + p;
                    return i < nw ? int(wcol[i * kDecThreads]) : 0;
                };
                int acc = W(v0) * w00;
                if (w01) acc += W(v0 + 1) * w01;
                if (w10) acc += W(v0 + m.gw) * w10;
                if (w11) acc += W(v0 + m.gw + 1) * w11;
                wt[p] = (acc + 8) >> 4;
            }
            uint32_t out = 0;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int w = (c == ccs) ? wt[1] : wt[0];
                const int C = (e0[c] * 257 * (64 - w) + e1[c] * 257 * w + 32) >> 6;
                out |= uint32_t(C >> 8) << (8 * c);
            }
            a.rgba[(size_t)y * a.width + x] = out;
        }
    }
}

}  // namespace

hipError_t launch_decode(const TritQuintDecodeTable* d_tables, const void* blocks, void* rgba, int width, int height,
                         int bw, int bh, hipStream_t stream) {
    DecodeArgs a;
    a.blocks = static_cast<const uint4*>(blocks);
    a.rgba = static_cast<uint32_t*>(rgba);
    a.width = width;
    a.height = height;
    a.bw = bw;
    a.bh = bh;
    a.xblocks = (width + bw - 1) / bw;
    a.nblocks = a.xblocks * ((height + bh - 1) / bh);
    a.tables = d_tables;
    (void)hipGetLastError();  // report this launch's errors only, not an earlier call's
    hipLaunchKernelGGL(decode_kernel, dim3(unsigned((a.nblocks + kDecThreads - 1) / kDecThreads)), dim3(kDecThreads),
                       0, stream, a);
    return hipGetLastError();
}

}  // namespace astc
