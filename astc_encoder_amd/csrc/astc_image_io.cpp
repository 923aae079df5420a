// astc_image_io.cpp -- PNG loading for the command-line drop-in (astc_load_image).
//
// The reference loads its input with stb_image as 8-bit RGBA, flipped
// vertically (main.cpp:24-25).  This is a PNG decoder over zlib written for this
// library: every colour type (grey, grey+alpha, RGB, RGBA, palette with tRNS),
// bit depths 1/2/4/8/16 (16-bit samples keep their high byte, as stb_image's
// 16-to-8 conversion does), the five scanline filters and Adam7 interlacing.
// Expansion to RGBA follows stb_image: grey -> (g, g, g, 255), grey+alpha ->
// (g, g, g, a), RGB -> alpha 255, palette -> the entry with its tRNS alpha, and a
// tRNS colour key in grey/RGB images makes matching pixels transparent.
#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "astc_internal.h"

using namespace astc;

namespace {

uint32_t be32(const uint8_t* p) { return uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

struct Png {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> plte, trns, idat;
};

int channels_of(int ctype) { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }

// Unfilter one pass (w x h) from `in` (starting at *pos); writes raw scanlines to out.
bool unfilter(const std::vector<uint8_t>& in, size_t* pos, uint32_t w, uint32_t h, int bpp_bits,
              std::vector<uint8_t>* out) {
    const size_t stride = (size_t(w) * bpp_bits + 7) / 8;
    const int bpp = std::max(1, bpp_bits / 8);
    out->assign(stride * h, 0);
    std::vector<uint8_t> prev(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        if (*pos + 1 + stride > in.size()) return false;
        const int ft = in[*pos];
        const uint8_t* src = &in[*pos + 1];
        uint8_t* cur = &(*out)[y * stride];
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            int v = src[i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: return false;
            }
            cur[i] = uint8_t(v);
        }
        std::memcpy(prev.data(), cur, stride);
        *pos += 1 + stride;
    }
    return true;
}

// Sample k (channel-interleaved) of a raw scanline at `depth` bits, as 8 bits.
uint8_t sample8(const uint8_t* row, size_t k, int depth, bool index) {
    if (depth == 8) return row[k];
    if (depth == 16) return row[2 * k];  // high byte
    const int per = 8 / depth;
    const int v = (row[k / per] >> (8 - depth * (int(k % per) + 1))) & ((1 << depth) - 1);
    if (index) return uint8_t(v);
    return uint8_t(v * 255 / ((1 << depth) - 1));  // scale low-depth grey to 0..255
}

uint16_t sample16(const uint8_t* row, size_t k, int depth) {
    if (depth == 16) return uint16_t(row[2 * k] << 8 | row[2 * k + 1]);
    if (depth == 8) return row[k];
    const int per = 8 / depth;
    return uint16_t((row[k / per] >> (8 - depth * (int(k % per) + 1))) & ((1 << depth) - 1));
}

void put_pixel(const Png& p, const uint8_t* row, uint32_t x, uint8_t* dst) {
    const int ch = channels_of(p.ctype);
    if (p.ctype == 3) {
        const int idx = sample8(row, x, p.depth, true);
        const uint8_t* e = idx * 3 + 2 < (int)p.plte.size() ? &p.plte[idx * 3] : nullptr;
        dst[0] = e ? e[0] : 0;
        dst[1] = e ? e[1] : 0;
        dst[2] = e ? e[2] : 0;
        dst[3] = idx < (int)p.trns.size() ? p.trns[idx] : 255;
        return;
    }
    uint8_t s[4];
    for (int c = 0; c < ch; ++c) s[c] = sample8(row, size_t(x) * ch + c, p.depth, false);
    bool key = false;  // tRNS colour key (compared at the stored bit depth)
    if (p.ctype == 0 && p.trns.size() >= 2)
        key = sample16(row, x, p.depth) == (p.trns[0] << 8 | p.trns[1]);
    if (p.ctype == 2 && p.trns.size() >= 6) {
        key = true;
        for (int c = 0; c < 3; ++c)
            key = key && sample16(row, size_t(x) * 3 + c, p.depth) == (p.trns[2 * c] << 8 | p.trns[2 * c + 1]);
    }
    switch (p.ctype) {
    case 0: dst[0] = dst[1] = dst[2] = s[0]; dst[3] = key ? 0 : 255; break;
    case 4: dst[0] = dst[1] = dst[2] = s[0]; dst[3] = s[1]; break;
    case 2: dst[0] = s[0]; dst[1] = s[1]; dst[2] = s[2]; dst[3] = key ? 0 : 255; break;
    default: dst[0] = s[0]; dst[1] = s[1]; dst[2] = s[2]; dst[3] = s[3]; break;
    }
}

int load_png(const char* path, std::vector<uint8_t>* rgba, int* wo, int* ho) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(ASTC_ERR_IO, std::string("can't fopen: ") + path);
    std::vector<uint8_t> file;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) file.insert(file.end(), buf, buf + n);
    std::fclose(f);
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (file.size() < 8 || std::memcmp(file.data(), sig, 8) != 0)
        return fail(ASTC_ERR_UNSUPPORTED, std::string("not a PNG file: ") + path);
    Png p;
    size_t pos = 8;
    bool have_ihdr = false, have_iend = false;
    while (pos + 12 <= file.size() && !have_iend) {
        const uint32_t len = be32(&file[pos]);
        if (pos + 12 + (size_t)len > file.size()) return fail(ASTC_ERR_IO, "truncated PNG chunk");
        const uint8_t* type = &file[pos + 4];
        const uint8_t* data = &file[pos + 8];
        if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
            p.w = be32(data);
            p.h = be32(data + 4);
            p.depth = data[8];
            p.ctype = data[9];
            p.interlace = data[12];
            have_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            p.plte.assign(data, data + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            p.trns.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            p.idat.insert(p.idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            have_iend = true;
        }
        pos += 12 + len;
    }
    if (!have_ihdr || p.w == 0 || p.h == 0 || p.w > (1u << 24) - 1 || p.h > (1u << 24) - 1)
        return fail(ASTC_ERR_UNSUPPORTED, "bad PNG header");
    const bool ok_depth = (p.ctype == 0 && (p.depth == 1 || p.depth == 2 || p.depth == 4 || p.depth == 8 ||
                                            p.depth == 16)) ||
                          (p.ctype == 3 && (p.depth == 1 || p.depth == 2 || p.depth == 4 || p.depth == 8)) ||
                          ((p.ctype == 2 || p.ctype == 4 || p.ctype == 6) && (p.depth == 8 || p.depth == 16));
    if (!ok_depth || p.interlace > 1) return fail(ASTC_ERR_UNSUPPORTED, "unsupported PNG colour type / depth");
    if (p.ctype == 3 && p.plte.empty()) return fail(ASTC_ERR_UNSUPPORTED, "palette PNG without PLTE");
    const int bpp_bits = channels_of(p.ctype) * p.depth;

    // filtered scanline bytes the header implies (every Adam7 pass when interlaced)
    static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
    uint64_t expected = 0;
    for (int k = 0; k < (p.interlace ? 7 : 1); ++k) {
        const uint64_t dx = p.interlace ? adx[k] : 1, dy = p.interlace ? ady[k] : 1;
        const uint64_t ox = p.interlace ? ax0[k] : 0, oy = p.interlace ? ay0[k] : 0;
        const uint64_t pw = p.w > ox ? (p.w - ox + dx - 1) / dx : 0, ph = p.h > oy ? (p.h - oy + dy - 1) / dy : 0;
        if (pw && ph) expected += ph * (1 + (pw * uint64_t(bpp_bits) + 7) / 8);
    }

    // inflate (never past what the header implies)
    std::vector<uint8_t> raw;
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return fail(ASTC_ERR_NO_MEMORY, "inflateInit failed");
    zs.next_in = p.idat.data();
    zs.avail_in = (uInt)p.idat.size();
    int zr;
    do {
        zs.next_out = buf;
        zs.avail_out = sizeof(buf);
        zr = inflate(&zs, Z_NO_FLUSH);
        if (zr == Z_BUF_ERROR) break;  // no progress possible: the input is used up
        if (zr != Z_OK && zr != Z_STREAM_END) {
            inflateEnd(&zs);
            return fail(ASTC_ERR_IO, "corrupt PNG image data");
        }
        raw.insert(raw.end(), buf, buf + (sizeof(buf) - zs.avail_out));
        if (raw.size() > expected) {
            inflateEnd(&zs);
            return fail(ASTC_ERR_IO, "PNG image data larger than its header allows");
        }
        // a full output buffer may leave decoded bytes pending even with no input left
    } while (zr != Z_STREAM_END && (zs.avail_in > 0 || zs.avail_out == 0));
    inflateEnd(&zs);
    // Every scanline the header implies must be there before the image is
    // allocated: the header alone must not size an allocation (a few bytes can
    // claim 2^24 x 2^24 texels).
    if (raw.size() < expected) return fail(ASTC_ERR_IO, "PNG image data shorter than its header requires");

    rgba->assign(size_t(p.w) * p.h * 4, 0);
    size_t rp = 0;
    std::vector<uint8_t> pass;
    if (!p.interlace) {
        if (!unfilter(raw, &rp, p.w, p.h, bpp_bits, &pass)) return fail(ASTC_ERR_IO, "corrupt PNG scanlines");
        const size_t stride = (size_t(p.w) * bpp_bits + 7) / 8;
        for (uint32_t y = 0; y < p.h; ++y)
            for (uint32_t x = 0; x < p.w; ++x) put_pixel(p, &pass[y * stride], x, &(*rgba)[(size_t(y) * p.w + x) * 4]);
    } else {
        static const int x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
        static const int dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
        for (int k = 0; k < 7; ++k) {
            const uint32_t pw = (p.w - x0[k] + dx[k] - 1) / dx[k] * (p.w > (uint32_t)x0[k]);
            const uint32_t ph = (p.h - y0[k] + dy[k] - 1) / dy[k] * (p.h > (uint32_t)y0[k]);
            if (!pw || !ph) continue;
            if (!unfilter(raw, &rp, pw, ph, bpp_bits, &pass)) return fail(ASTC_ERR_IO, "corrupt PNG scanlines");
            const size_t stride = (size_t(pw) * bpp_bits + 7) / 8;
            for (uint32_t y = 0; y < ph; ++y)
                for (uint32_t x = 0; x < pw; ++x) {
                    const size_t X = x0[k] + size_t(x) * dx[k], Y = y0[k] + size_t(y) * dy[k];
                    put_pixel(p, &pass[y * stride], x, &(*rgba)[(Y * p.w + X) * 4]);
                }
        }
    }
    *wo = (int)p.w;
    *ho = (int)p.h;
    return ASTC_OK;
}

}  // namespace

extern "C" {

int astc_load_image(const char* path, int flip_vertically, int* width, int* height, uint8_t** rgba) {
    return guarded([&]() -> int {
        if (!path || !width || !height || !rgba) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
        *rgba = nullptr;
        std::vector<uint8_t> img;
        int w = 0, h = 0;
        int rc = load_png(path, &img, &w, &h);
        if (rc != ASTC_OK) return rc;
        uint8_t* out = static_cast<uint8_t*>(std::malloc(img.size()));
        if (!out) return fail(ASTC_ERR_NO_MEMORY, "host allocation failed");
        const size_t row = size_t(w) * 4;
        for (int y = 0; y < h; ++y)
            std::memcpy(out + size_t(y) * row, img.data() + size_t(flip_vertically ? h - 1 - y : y) * row, row);
        *width = w;
        *height = h;
        *rgba = out;
        return ASTC_OK;
    });
}

void astc_free_image(uint8_t* rgba) { std::free(rgba); }

}  // extern "C"
