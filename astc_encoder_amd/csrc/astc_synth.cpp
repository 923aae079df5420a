// astc_synth.cpp -- seeded synthetic textures for the encoder's workloads
// (libastc_synth.so, host only).
//
// Every value is produced from a SplitMix64 stream and IEEE double arithmetic
// only (no libm transcendental calls; the sine is a fixed polynomial), built
// with -ffp-contract=off, so every x86-64 host produces identical bytes.
//
// Workloads:
//   sinusoid RGB8   sum of 3 low-frequency sinusoid colour gradients (periods
//                   64-512 px log-uniform, amplitude 0-200 per channel) plus
//                   uniform noise +-8, clamped to 0..255
//   noise RGBA8     4-octave value noise (lattice 256 -> 32 px, persistence
//                   0.5) with hard-edged solid rectangles covering >= 10%;
//                   alpha = radial gradient x cut-out mask, 30% of texels
//                   exactly 0 and 20% exactly 255
//   normal RG8      5-octave fBm heightfield (lattice 256 -> 16 px), central
//                   differences, n = normalize(-s dh/dx, -s dh/dy, 1),
//                   RG = round((n.xy * 0.5 + 0.5) * 255)
//   mip chain       2x2 box filter in stored space, (a+b+c+d+2)>>2, or
//                   (a+b+1)>>1 along the one axis that still halves; a side
//                   halves while it is at least 8 (power-of-two sides end at 4)
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

uint64_t mix(uint64_t a, uint64_t b) {
    SplitMix64 m(a * 0x100000001B3ull ^ b);
    return m.next();
}

// lattice value in [0, 1) for integer lattice point (ix, iy) of one noise layer
double lattice(uint64_t key, int64_t ix, int64_t iy) {
    uint64_t h = mix(key, (uint64_t(uint32_t(ix)) << 32) | uint32_t(iy));
    return double(h >> 11) * (1.0 / 9007199254740992.0);
}

double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

double value_noise(uint64_t key, double x, double y, double cell) {
    const double fx = x / cell, fy = y / cell;
    const double flx = std::floor(fx), fly = std::floor(fy);
    const int64_t ix = int64_t(flx), iy = int64_t(fly);
    const double tx = smooth(fx - flx), ty = smooth(fy - fly);
    const double a = lattice(key, ix, iy), b = lattice(key, ix + 1, iy);
    const double c = lattice(key, ix, iy + 1), d = lattice(key, ix + 1, iy + 1);
    const double top = a + (b - a) * tx, bot = c + (d - c) * tx;
    return top + (bot - top) * ty;
}

double fbm(uint64_t key, double x, double y, double cell0, int octaves) {
    double sum = 0.0, amp = 1.0, norm = 0.0, cell = cell0;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * value_noise(mix(key, uint64_t(o) + 1), x, y, cell);
        norm += amp;
        amp *= 0.5;
        cell *= 0.5;
    }
    return sum / norm;
}

// sin(x) for any finite x: reduce to [-pi, pi] and use a degree-15 odd Taylor
// polynomial in Horner form (|error| < 1e-6 there, ample for 8-bit texels).
double poly_sin(double x) {
    const double two_pi = 6.283185307179586;
    const double pi = 3.141592653589793;
    double r = x - two_pi * std::floor(x / two_pi);  // [0, 2pi)
    if (r > pi) r -= two_pi;
    const double r2 = r * r;
    double p = -1.0 / 1307674368000.0;
    p = p * r2 + 1.0 / 6227020800.0;
    p = p * r2 - 1.0 / 39916800.0;
    p = p * r2 + 1.0 / 362880.0;
    p = p * r2 - 1.0 / 5040.0;
    p = p * r2 + 1.0 / 120.0;
    p = p * r2 - 1.0 / 6.0;
    p = p * r2 + 1.0;
    return r * p;
}
double poly_cos(double x) { return poly_sin(x + 1.5707963267948966); }

// 2^t for t >= 0: 2^floor(t) (exact) times a degree-12 Taylor series of e^(f ln 2).
double poly_exp2(double t) {
    const double fl = std::floor(t);
    const double y = (t - fl) * 0.6931471805599453;
    double p = 1.0, term = 1.0;
    for (int k = 1; k <= 12; ++k) {
        term = term * y / k;
        p = p + term;
    }
    return std::ldexp(p, int(fl));
}

uint8_t to_u8(double v) {
    v = std::min(255.0, std::max(0.0, v));
    return uint8_t(std::nearbyint(v));
}

}  // namespace

extern "C" {

// Sinusoid RGB8, row pitch = width * 3.
void astc_synth_sinusoid_rgb8(uint64_t seed, int width, int height, uint8_t* out) {
    SplitMix64 rng(mix(seed, 0xC0FFEE));
    double period[3], ct[3], st[3], phase[3], amp[3][3];
    for (int k = 0; k < 3; ++k) {
        period[k] = 64.0 * poly_exp2(3.0 * rng.uniform());  // log-uniform in [64, 512)
        const double th = 6.283185307179586 * rng.uniform();
        ct[k] = poly_cos(th);
        st[k] = poly_sin(th);
        phase[k] = 6.283185307179586 * rng.uniform();
        for (int c = 0; c < 3; ++c) amp[k][c] = 200.0 * rng.uniform();
    }
    const uint64_t noise_key = mix(seed, 0xBEEF);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y) {
        SplitMix64 nr(mix(noise_key, uint64_t(y)));
        uint8_t* row = out + (size_t)y * width * 3;
        for (int x = 0; x < width; ++x) {
            double s[3];
            for (int k = 0; k < 3; ++k)
                s[k] = 0.5 + 0.5 * poly_sin(6.283185307179586 * (x * ct[k] + y * st[k]) / period[k] + phase[k]);
            for (int c = 0; c < 3; ++c) {
                const double v = amp[0][c] * s[0] + amp[1][c] * s[1] + amp[2][c] * s[2] + (16.0 * nr.uniform() - 8.0);
                row[3 * x + c] = to_u8(v);
            }
        }
    }
}

// Value noise RGBA8 with rectangles and a cut-out alpha, row pitch = width * 4.
void astc_synth_noise_rgba8(uint64_t seed, int width, int height, uint8_t* out) {
    const uint64_t key = mix(seed, 0x5EED);
    std::vector<float> mask((size_t)width * height);
    const double cx = 0.5 * width, cy = 0.5 * height, rmax = std::sqrt(cx * cx + cy * cy);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y) {
        uint8_t* row = out + (size_t)y * width * 4;
        for (int x = 0; x < width; ++x) {
            for (int c = 0; c < 3; ++c) row[4 * x + c] = to_u8(255.0 * fbm(mix(key, uint64_t(c) + 10), x, y, 256.0, 4));
            mask[(size_t)y * width + x] = float(value_noise(mix(key, 99), x, y, 512.0));
        }
    }
    // hard-edged solid rectangles until >= 10% of the texels are covered
    SplitMix64 rng(mix(key, 7));
    std::vector<uint8_t> covered((size_t)width * height, 0);
    size_t count = 0, target = ((size_t)width * height + 9) / 10;
    const int max_side = std::max(2, std::min(width, height) / 4), min_side = std::max(1, max_side / 16);
    while (count < target) {
        const int rw = min_side + int(rng.uniform() * (max_side - min_side + 1));
        const int rh = min_side + int(rng.uniform() * (max_side - min_side + 1));
        const int x0 = int(rng.uniform() * width), y0 = int(rng.uniform() * height);
        uint8_t col[3];
        for (int c = 0; c < 3; ++c) col[c] = uint8_t(rng.next() >> 56);
        for (int y = y0; y < std::min(height, y0 + rh); ++y)
            for (int x = x0; x < std::min(width, x0 + rw); ++x) {
                uint8_t* p = out + ((size_t)y * width + x) * 4;
                p[0] = col[0]; p[1] = col[1]; p[2] = col[2];
                uint8_t& cv = covered[(size_t)y * width + x];
                if (!cv) { cv = 1; ++count; }
            }
    }
    // alpha: lowest 30% of the mask -> 0, highest 20% -> 255, the rest a radial gradient in 1..254
    std::vector<float> sorted(mask);
    const size_t n = sorted.size();
    const size_t i30 = n * 3 / 10, i80 = n * 8 / 10;
    std::nth_element(sorted.begin(), sorted.begin() + i30, sorted.end());
    const float t30 = sorted[i30];
    std::nth_element(sorted.begin(), sorted.begin() + i80, sorted.end());
    const float t80 = sorted[i80];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const float m = mask[(size_t)y * width + x];
            uint8_t a;
            if (m < t30) a = 0;
            else if (m >= t80) a = 255;
            else {
                const double dx = x + 0.5 - cx, dy = y + 0.5 - cy;
                const double g = 1.0 - std::sqrt(dx * dx + dy * dy) / rmax;
                a = uint8_t(std::min(254.0, std::max(1.0, std::nearbyint(255.0 * g))));
            }
            out[((size_t)y * width + x) * 4 + 3] = a;
        }
}

// Normal map RG8 from an fBm heightfield, row pitch = width * 2.
void astc_synth_normal_rg8(uint64_t seed, int width, int height, uint8_t* out) {
    const uint64_t key = mix(seed, 0x4E4F);
    std::vector<double> h((size_t)width * height);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) h[(size_t)y * width + x] = fbm(key, x, y, 256.0, 5);
    const double s = 400.0;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const int xl = std::max(0, x - 1), xr = std::min(width - 1, x + 1);
            const int yu = std::max(0, y - 1), yd = std::min(height - 1, y + 1);
            const double dx = (h[(size_t)y * width + xr] - h[(size_t)y * width + xl]) * 0.5;
            const double dy = (h[(size_t)yd * width + x] - h[(size_t)yu * width + x]) * 0.5;
            const double nx = -s * dx, ny = -s * dy, len = std::sqrt(nx * nx + ny * ny + 1.0);
            out[((size_t)y * width + x) * 2] = to_u8((nx / len * 0.5 + 0.5) * 255.0);
            out[((size_t)y * width + x) * 2 + 1] = to_u8((ny / len * 0.5 + 0.5) * 255.0);
        }
}

// Next mip level of a tightly packed image with `bpp` bytes per texel.
// Returns the new size through *nw, *nh: a side of at least 8 halves (rounding
// down), a shorter side is kept, so the filter never reads past the image.
void astc_synth_downsample(const uint8_t* src, int w, int h, int bpp, uint8_t* dst, int* nw, int* nh) {
    const int ow = w >= 8 ? w >> 1 : w, oh = h >= 8 ? h >> 1 : h;
    const bool hx = ow < w, hy = oh < h;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x)
            for (int c = 0; c < bpp; ++c) {
                const int sx = hx ? 2 * x : x, sy = hy ? 2 * y : y;
                auto at = [&](int xx, int yy) { return int(src[((size_t)yy * w + xx) * bpp + c]); };
                int v;
                if (hx && hy) v = (at(sx, sy) + at(sx + 1, sy) + at(sx, sy + 1) + at(sx + 1, sy + 1) + 2) >> 2;
                else if (hx) v = (at(sx, sy) + at(sx + 1, sy) + 1) >> 1;
                else if (hy) v = (at(sx, sy) + at(sx, sy + 1) + 1) >> 1;
                else v = at(sx, sy);
                dst[((size_t)y * ow + x) * bpp + c] = uint8_t(v);
            }
    *nw = ow;
    *nh = oh;
}

}  // extern "C"
