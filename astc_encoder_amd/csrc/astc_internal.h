// astc_internal.h -- host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/astc_mi355x.h"
#include "astc_block.h"

namespace astc {
struct NoAllocPlan;
// State of astc_encode_batch_noalloc, set up with the context: the device
// address of the library's scratch array on the context's device, pinned host
// staging of the same layout, the event after which the staging may be reused,
// and host work arrays.  The mutex serialises the entry point's use of them.
struct NoAllocState {
    void* scratch = nullptr;
    NoAllocPlan* staging = nullptr;
    const void* staging_dev = nullptr;  // the staging's device address (read by the copy kernel)
    hipEvent_t staged = nullptr;
    std::mutex mu;
    uint64_t keys[2 * 8192];
    int modes[8192];
};
}  // namespace astc

struct astc_ctx {
    int device = 0;
    astc::DeviceTables* d_tables = nullptr;  // device copy of the sRGB / trit tables
    astc::DeviceTables h_tables{};           // host copy
    astc::TritQuintDecodeTable* d_dec_tables = nullptr;  // decoder tables
    int num_cus = 256;                       // compute units of the device
    astc::NoAllocState noalloc;              // astc_encode_batch_noalloc only (it locks noalloc.mu)
    // Otherwise written only by astc_ctx_create / astc_ctx_destroy: every other
    // call only reads it, so threads may share one context.
};

namespace astc {

// thread-local last error
// (set_error and fail copy the message into a fixed per-thread buffer and never
// allocate)
void set_error(const char* msg);
void set_error(const std::string& msg);
int fail(int code, const char* msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Host-side description checks shared by every entry point.
struct ImageGeom {
    int xblocks, yblocks;
    int64_t nblocks;
    int bpp;
};
int check_desc(const astc_image_desc* d, ImageGeom* g);

// Runs an entry point's body so that no C++ exception crosses the C ABI: a
// failed host allocation becomes ASTC_ERR_NO_MEMORY.  The handlers record their
// message without allocating.
template <typename F>
auto guarded(F&& f) -> decltype(f()) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(ASTC_ERR_NO_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        set_error(e.what());
        return ASTC_ERR_INVALID_ARG;
    } catch (...) {
        return fail(ASTC_ERR_INVALID_ARG, "unexpected error");
    }
}

// Bodies of astc_encode / astc_decode (the entry points run them under guarded).
int encode_one(astc_ctx* ctx, const astc_image_desc* desc, const void* src_dev, void* dst_dev, void* stream,
               const astc_plain_tuning* tuning = nullptr);
int decode_one(astc_ctx* ctx, int width, int height, int block_dim_x, int block_dim_y, const void* blocks_dev,
               void* rgba_dev, void* stream);

// Host construction of the tables (identical on every host: sRGB via double pow,
// rounded to binary32, times 255.0f; trits from the specification decoder).
void build_tables(DeviceTables* t);

constexpr int kPlainThreads = 256;  // threads per workgroup of the encode kernels

// ---- batched persistent encoder: device work list ----
// work-item mode: block size x alpha, the sRGB fetch bit and the channel class
// (which channels are 255 in every block: alpha of an RGB8 image tiled exactly
// by its blocks, blue and alpha of a normal map)
enum : int {
    MODE_4x4_RGB = 0, MODE_4x4_RGBA = 1, MODE_6x6_RGB = 2, MODE_6x6_RGBA = 3, MODE_SRGB = 4,
    MODE_CH_ALL = 0, MODE_CH_RGB8 = 8, MODE_CH_NORMAL = 16, MODE_CH_MASK = 24, MODE_COUNT = 32
};
constexpr int kMaxSegs = 16;  // images packed into one work item

struct BatchLevel {   // one image (texture or mip level)
    FetchArgs f;
    uint32_t* dst;
    int nblocks, pad;
};
struct BatchSeg { int level, first_block, count, pad; };
struct BatchItem { int mode, seg_begin, nseg, pad; };
struct BatchArgs {
    const BatchLevel* levels;
    const BatchSeg* segs;
    const BatchItem* items;
    const int* scalars;  // [0] item count, [1] index of the first 4x4 item; read on
                         // the device, so a captured graph can be re-fed
    int* counters;       // work-queue heads [all or 6x6, 4x4], reset by a kernel node
    const DeviceTables* tables;
    int prefetch;
};
struct BatchLaunch {
    int lanes = 1; // lanes per 6x6 block (1 or 2)
    int grid;      // workgroups of the (6x6 or only) persistent kernel
    int grid4;     // workgroups of the 4x4 kernel when split
    int split;     // 0: one kernel for every item; 1: 6x6 and 4x4 kernels
    hipStream_t side;            // second stream for the 4x4 kernel (split only)
    hipEvent_t fork, join;       // ordering events between the two streams
};
// Launches the counter reset and the persistent kernel(s) on `stream`.
hipError_t launch_encode_batch(const BatchArgs& a, const BatchLaunch& l, hipStream_t stream);
// A batch tuning within range (lanes_per_block 0 counts as 1).
inline bool valid_batch_tuning(const astc_tuning& t) {
    const int bpl = t.blocks_per_lane;
    return (bpl == 1 || bpl == 2 || bpl == 4 || bpl == 8) && t.workgroups_per_cu >= 1 && t.workgroups_per_cu <= 4 &&
           (t.prefetch == 0 || t.prefetch == 1) && (t.split_modes == 0 || t.split_modes == 1) &&
           (t.lanes_per_block >= 0 && t.lanes_per_block <= 2);
}
inline int batch_lanes(const astc_tuning& t) { return t.lanes_per_block == 2 ? 2 : 1; }
inline int mode_of(int block_dim, int src_format, uint32_t flags, int width, int height) {
    const bool tiled = width % block_dim == 0 && height % block_dim == 0;
    const int ch = (flags & FLAG_NORMAL) ? MODE_CH_NORMAL
                   : (src_format == FMT_RGB8 && tiled) ? MODE_CH_RGB8 : MODE_CH_ALL;
    return (block_dim == 6 ? 2 : 0) + ((flags & FLAG_ALPHA) ? 1 : 0) + ((flags & FLAG_SRGB) ? MODE_SRGB : 0) + ch;
}
// relative cost order of modes (most expensive first) for the work list: block
// size, then active channels, alpha and sRGB
inline int mode_cost_rank(int mode) {
    const int ch = (mode & MODE_CH_MASK) == MODE_CH_NORMAL ? 0 : (mode & MODE_CH_MASK) == MODE_CH_RGB8 ? 1 : 2;
    return ((mode & 2) ? 16 : 0) + ch * 4 + (mode & 1) * 2 + ((mode & MODE_SRGB) ? 1 : 0);
}

// Work-list planner (astc_batch.cpp) into caller arrays, allocation-free;
// false if the segments or items exceed their capacity.  plan_work sizes
// vectors for it.
bool plan_work_into(const BatchLevel* levels, const int* modes, int nlevels, int blocks_per_lane, uint64_t* keys,
                    BatchSeg* segs, int seg_cap, BatchItem* items, int item_cap, int* nseg, int* nitem,
                    int* first_light);

// astc_encode_batch_noalloc's limits and the layout of its scratch array (and
// of the pinned host staging it is uploaded from).
constexpr int kNoAllocMaxImages = 8192;
constexpr int kNoAllocMaxItems = 131072;
struct NoAllocPlan {
    BatchLevel levels[kNoAllocMaxImages];
    BatchSeg segs[kNoAllocMaxItems];
    BatchItem items[kNoAllocMaxItems];
    int scalars[4];  // item count, first 4x4 item, two queue heads
};
constexpr size_t kNoAllocScratchBytes = (sizeof(NoAllocPlan) + 255) / 256 * 256;
static_assert(sizeof(BatchLevel) == 64 && sizeof(BatchSeg) == 16 && sizeof(BatchItem) == 16, "work-list layout");
static_assert(sizeof(NoAllocState::keys) / sizeof(uint64_t) == 2 * kNoAllocMaxImages &&
                  sizeof(NoAllocState::modes) / sizeof(int) == kNoAllocMaxImages,
              "no-alloc host arrays");
// Set up / release the no-alloc state of a context (astc_noalloc.hip), with the
// context's device current.
int noalloc_setup(astc_ctx* ctx);
void noalloc_teardown(astc_ctx* ctx);

// Device-to-device copy of `bytes` on `stream` with the library's copy kernel.
hipError_t launch_ring_copy(void* dst, const void* src, int64_t bytes, hipStream_t stream);
hipError_t launch_downsample(const uint8_t* src, int w, int h, int bpp, uint8_t* dst, int ow, int oh,
                             hipStream_t stream);

hipError_t launch_decode(const TritQuintDecodeTable* d_tables, const void* blocks, void* rgba, int width, int height,
                         int bw, int bh, hipStream_t stream);

// Plain per-texture launch on a device with num_cus CUs.  tuning: NULL for
// the default (default_plain_tuning); blocks_per_lane 0 picks the count.
struct PlainTuning {
    int lanes_per_block;  // 1, or 2 (6x6 principal-component encodes only; others run 1)
    int blocks_per_lane;  // blocks per lane per workgroup: 1, 2, 4, 8, or 0 for the library's choice
    int prefetch;         // 1: the next block's loads are issued before the current block is encoded
    int flat_alpha;       // 1: one-lane four-channel encodes test each wave for a flat alpha
    int packed_texels;    // 1: one-lane encodes hold the fetched words and convert them in every pass
};
hipError_t launch_encode_plain(const DeviceTables* d_tables, const astc_image_desc& d, const ImageGeom& g,
                               const void* src, void* dst, hipStream_t stream, int num_cus,
                               const PlainTuning* tuning = nullptr);
int plain_blocks_per_lane(int64_t nblocks, int dim, int lanes_per_block, int num_cus, bool packed = false);
PlainTuning default_plain_tuning(const astc_image_desc& d, int64_t nblocks, int num_cus);

}  // namespace astc
