// astc_internal.h -- host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "../../include/astc_mi355x.h"
#include "astc_block.h"

struct astc_ctx {
    int device = 0;
    astc::DeviceTables* d_tables = nullptr;  // device copy of the sRGB / trit tables
    astc::DeviceTables h_tables{};           // host copy
    astc::TritQuintDecodeTable* d_dec_tables = nullptr;  // decoder tables
    std::mutex mu;                           // guards lazily created objects
};

namespace astc {

// thread-local last error
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Host-side description checks shared by every entry point.
struct ImageGeom {
    int xblocks, yblocks;
    int64_t nblocks;
    int bpp;
};
int check_desc(const astc_image_desc* d, ImageGeom* g);

// Host construction of the tables (identical on every host: sRGB via double pow,
// rounded to binary32, times 255.0f; trits from the specification decoder).
void build_tables(DeviceTables* t);

constexpr int kPlainThreads = 256;  // threads per workgroup of the encode kernels

// ---- batched persistent encoder: device work list ----
enum : int { MODE_4x4_RGB = 0, MODE_4x4_RGBA = 1, MODE_6x6_RGB = 2, MODE_6x6_RGBA = 3 };
constexpr int kMaxSegs = 16;  // images packed into one work item

struct BatchLevel {   // one image (texture or mip level)
    FetchArgs f;
    uint32_t* dst;
    int xblocks, nblocks;
};
struct BatchSeg { int level, first_block, count, pad; };
struct BatchItem { int mode, seg_begin, nseg, pad; };
struct BatchArgs {
    const BatchLevel* levels;
    const BatchSeg* segs;
    const BatchItem* items;
    const int* nitems;  // read on the device, so a captured graph can be re-fed
    int* counter;       // work-queue head, reset by a kernel node before each run
    const DeviceTables* tables;
    int prefetch;
};
// Launches the counter reset and the persistent kernel (grid workgroups).
hipError_t launch_encode_batch(const BatchArgs& a, int grid, hipStream_t stream);
inline int mode_of(int block_dim, uint32_t flags) {
    return (block_dim == 6 ? 2 : 0) + ((flags & FLAG_ALPHA) ? 1 : 0);
}

hipError_t launch_downsample(const uint8_t* src, int w, int h, int bpp, uint8_t* dst, int ow, int oh,
                             hipStream_t stream);

hipError_t launch_decode(const TritQuintDecodeTable* d_tables, const void* blocks, void* rgba, int width, int height,
                         int bw, int bh, hipStream_t stream);

// Plain per-texture launch (one lane per block).
// bpl_override: blocks per lane (1, 2, 4 or 8); 0 chooses from the block count.
hipError_t launch_encode_plain(const DeviceTables* d_tables, const astc_image_desc& d, const ImageGeom& g,
                               const void* src, void* dst, hipStream_t stream, int bpl_override = 0);
int plain_blocks_per_lane(int64_t nblocks, int dim);

}  // namespace astc
