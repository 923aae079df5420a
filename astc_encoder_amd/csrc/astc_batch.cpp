// astc_batch.cpp -- batched persistent encoding: work-list planning, hipGraph
// capture and the autotuner (include/astc_mi355x.h, astc_batch_*).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "astc_internal.h"

using namespace astc;

struct astc_batch {
    astc_ctx* ctx = nullptr;
    int device = 0;
    int num_cus = 0;
    std::vector<BatchLevel> levels;  // host copies
    std::vector<int> modes;          // kNoMode for images encoded by plain launches
    // images with the max-accumulation search: one plain launch each, captured in
    // the same graph after the persistent kernels (the work list has no such mode)
    struct PlainImage {
        astc_image_desc desc;
        ImageGeom geom;
        const void* src;
        void* dst;
    };
    std::vector<PlainImage> plain;
    int64_t blocks = 0, texels = 0;
    astc_tuning tuning{4, 2, 1, 1, 1};
    // device work list
    BatchLevel* d_levels = nullptr;
    BatchSeg* d_segs = nullptr;
    BatchItem* d_items = nullptr;
    int* d_scalars = nullptr;  // [0] item count, [1] first 4x4 item, [2..3] queue heads
    int nitems = 0;
    hipStream_t capture_stream = nullptr, side_stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace astc {

// Mode code of an image that is not in the work list (no code the planner visits).
constexpr int kNoMode = -1;

// Image index bits of a planner sort key.
constexpr int kKeyIndexBits = 27;

// Sort n unsigned keys (distinct, so the result is one order) in place with a
// least-significant-digit radix sort over 8-bit digits; tmp holds n keys.  A
// digit that is the same in every key is skipped.
void radix_sort_keys(uint64_t* keys, uint64_t* tmp, int n) {
    uint64_t* src = keys;
    uint64_t* dst = tmp;
    for (int shift = 0; shift < 64; shift += 8) {
        int count[257] = {0};
        for (int i = 0; i < n; ++i) ++count[((src[i] >> shift) & 0xFF) + 1];
        if (n == 0 || count[((src[0] >> shift) & 0xFF) + 1] == n) continue;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (int i = 0; i < n; ++i) dst[count[(src[i] >> shift) & 0xFF]++] = src[i];
        std::swap(src, dst);
    }
    if (src != keys) std::memcpy(keys, src, size_t(n) * sizeof(uint64_t));
}

// Position of each mode code in the planner's visiting order: most expensive
// first (block size, then active channels, alpha, sRGB), ties by code.
int mode_position(int mode) {
    int pos = 0;
    for (int m = 0; m < MODE_COUNT; ++m)
        pos += mode_cost_rank(m) > mode_cost_rank(mode) || (mode_cost_rank(m) == mode_cost_rank(mode) && m < mode);
    return pos;
}

// Work list for a tuning.  Modes are ordered by cost (6x6 RGBA first) so the
// longest items start first; within a mode, images by size (large first, then
// by index), large ones split into chunks of 256 x blocks_per_lane blocks,
// images under 256 blocks packed up to kMaxSegs per item (one block per
// thread).  The split point (`first_light`) is the first item of the first
// 4x4 reduced-channel mode in that order (4x4 four-channel items rank above
// them and go with the 6x6 items).  Codes with both class bits set are not
// modes (no image has them); images with a mode outside 0..MODE_COUNT-1 are
// left out.  Writes into the caller's arrays (keys: 2 nlevels of room) and
// allocates nothing; returns false if segs or items would exceed their capacity
// or nlevels is 2^27 or more.
bool plan_work_into(const BatchLevel* levels, const int* modes, int nlevels, int blocks_per_lane, uint64_t* keys,
                    BatchSeg* segs, int seg_cap, BatchItem* items, int item_cap, int* nseg, int* nitem,
                    int* first_light) {
    if (nlevels < 0 || nlevels >= (1 << kKeyIndexBits)) return false;
    int pos[MODE_COUNT];
    int light_pos = MODE_COUNT;
    for (int m = 0; m < MODE_COUNT; ++m) {
        pos[m] = mode_position(m);
        const int ch = m & MODE_CH_MASK;
        if (!(m & 2) && (ch == MODE_CH_RGB8 || ch == MODE_CH_NORMAL)) light_pos = std::min(light_pos, pos[m]);
    }
    // visiting order: mode position, then blocks (large first), then index, as
    // one unsigned key per image (position | complement of the block count |
    // index), sorted by a least-significant-digit radix sort
    int n = 0;
    for (int i = 0; i < nlevels; ++i)
        if (modes[i] >= 0 && modes[i] < MODE_COUNT)
            keys[n++] = uint64_t(pos[modes[i]]) << (kKeyIndexBits + 31) |
                        uint64_t(0x7FFFFFFF - std::max(0, levels[i].nblocks)) << kKeyIndexBits | uint64_t(i);
    radix_sort_keys(keys, keys + n, n);
    const uint64_t index_mask = (uint64_t(1) << kKeyIndexBits) - 1;
    const int chunk = kPlainThreads * blocks_per_lane;
    int ns = 0, ni = 0;
    *first_light = -1;
    BatchItem pack{0, 0, 0, 0};
    int pack_blocks = 0, cur_mode = -1;
    auto flush = [&]() -> bool {
        if (pack.nseg) {
            if (ni >= item_cap) return false;
            items[ni++] = pack;
        }
        pack = BatchItem{cur_mode, ns, 0, 0};
        pack_blocks = 0;
        return true;
    };
    for (int k = 0; k < n; ++k) {
        const int li = int(keys[k] & index_mask), mode = modes[li];
        if (mode != cur_mode) {
            if (!flush()) return false;
            cur_mode = mode;
            pack = BatchItem{mode, ns, 0, 0};
        }
        if (*first_light < 0 && pos[mode] >= light_pos) *first_light = ni;
        const int nb = levels[li].nblocks;
        if (nb >= kPlainThreads) {
            for (int first = 0; first < nb; first += chunk) {
                if (ni >= item_cap || ns >= seg_cap) return false;
                items[ni++] = BatchItem{mode, ns, 1, 0};
                segs[ns++] = BatchSeg{li, first, std::min(chunk, nb - first), 0};
            }
            pack.seg_begin = ns;
            continue;
        }
        if (pack_blocks + nb > kPlainThreads || pack.nseg == kMaxSegs) {
            if (!flush()) return false;  // packs stay contiguous in the segment array
        }
        if (ns >= seg_cap) return false;
        segs[ns++] = BatchSeg{li, 0, nb, 0};
        ++pack.nseg;
        pack_blocks += nb;
    }
    if (!flush()) return false;
    if (*first_light < 0) *first_light = ni;
    *nseg = ns;
    *nitem = ni;
    return true;
}

void plan_work(const std::vector<BatchLevel>& levels, const std::vector<int>& modes, int blocks_per_lane,
               std::vector<BatchSeg>* segs, std::vector<BatchItem>* items, int* first_light) {
    // every image yields at most ceil(nblocks / chunk) segments, and every item
    // holds at least one segment
    const int64_t chunk = int64_t(kPlainThreads) * blocks_per_lane;
    int64_t cap = 1;
    for (const BatchLevel& L : levels) cap += (L.nblocks + chunk - 1) / chunk;
    std::vector<uint64_t> keys(2 * levels.size());
    segs->resize(size_t(cap));
    items->resize(size_t(cap));
    int ns = 0, ni = 0;
    if (!plan_work_into(levels.data(), modes.data(), (int)levels.size(), blocks_per_lane, keys.data(), segs->data(),
                        (int)cap, items->data(), (int)cap, &ns, &ni, first_light))
        throw std::length_error("work list: too many images");
    segs->resize(size_t(ns));
    items->resize(size_t(ni));
}

}  // namespace astc

namespace {

void free_worklist(astc_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (void* p : {(void*)b->d_segs, (void*)b->d_items})
        if (p) (void)hipFree(p);
    b->d_segs = nullptr;
    b->d_items = nullptr;
}

int build(astc_batch* b) {
    free_worklist(b);
    std::vector<BatchSeg> segs;
    std::vector<BatchItem> items;
    int first_light = 0;
    plan_work(b->levels, b->modes, b->tuning.blocks_per_lane, &segs, &items, &first_light);
    hipError_t e;
    if ((e = hipMalloc(&b->d_segs, std::max<size_t>(1, segs.size()) * sizeof(BatchSeg))) != hipSuccess)
        return hip_fail(e, "hipMalloc(segs)");
    if ((e = hipMalloc(&b->d_items, std::max<size_t>(1, items.size()) * sizeof(BatchItem))) != hipSuccess)
        return hip_fail(e, "hipMalloc(items)");
    if (!segs.empty() &&
        (e = hipMemcpy(b->d_segs, segs.data(), segs.size() * sizeof(BatchSeg), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(segs)");
    if (!items.empty() && (e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BatchItem),
                                         hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(items)");
    b->nitems = (int)items.size();
    int scal[4] = {b->nitems, first_light, 0, 0};
    if ((e = hipMemcpy(b->d_scalars, scal, sizeof(scal), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(scalars)");

    BatchArgs a;
    a.levels = b->d_levels;
    a.segs = b->d_segs;
    a.items = b->d_items;
    a.scalars = b->d_scalars;
    a.counters = b->d_scalars + 2;
    a.tables = b->ctx->d_tables;
    a.prefetch = b->tuning.prefetch;
    BatchLaunch l;
    l.lanes = batch_lanes(b->tuning);
    l.grid = std::max(1, b->num_cus * b->tuning.workgroups_per_cu);
    // the 4x4 kernel's own occupancy (five waves per SIMD), whatever the 6x6
    // kernel's: with one 6x6 workgroup per CU, three 4x4 waves share each SIMD
    // with the 6x6 wave from the start and cover its latency
    l.grid4 = std::max(1, 5 * b->num_cus);
    l.split = b->tuning.split_modes;
    l.side = b->side_stream;
    l.fork = b->fork;
    l.join = b->join;
    if ((e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_fail(e, "hipStreamBeginCapture");
    hipError_t le = launch_encode_batch(a, l, b->capture_stream);
    for (size_t i = 0; i < b->plain.size() && le == hipSuccess; ++i) {
        const astc_batch::PlainImage& p = b->plain[i];
        le = launch_encode_plain(b->ctx->d_tables, p.desc, p.geom, p.src, p.dst, b->capture_stream, b->num_cus);
    }
    e = hipStreamEndCapture(b->capture_stream, &b->graph);
    if (le != hipSuccess) return hip_fail(le, "batch launch capture");
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess)
        return hip_fail(e, "hipGraphInstantiate");
    return ASTC_OK;
}

bool valid_tuning(const astc_tuning* t) { return valid_batch_tuning(*t); }

}  // namespace

extern "C" {

int astc_batch_create(astc_ctx* ctx, const astc_image_desc* descs, const void* const* srcs, void* const* dsts,
                      int count, astc_batch** out) {
    return guarded([&]() -> int {
        if (!out) return fail(ASTC_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        if (!ctx || !descs || !srcs || !dsts || count < 1) return fail(ASTC_ERR_INVALID_ARG, "empty or NULL batch");
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        astc_batch* b = new (std::nothrow) astc_batch();
        if (!b) return fail(ASTC_ERR_NO_MEMORY, "host allocation failed");
        b->ctx = ctx;
        b->device = ctx->device;
        b->levels.resize(count);
        b->modes.resize(count);
        for (int i = 0; i < count; ++i) {
            ImageGeom g;
            int rc = check_desc(&descs[i], &g);
            if (rc != ASTC_OK || !srcs[i] || !dsts[i] || (reinterpret_cast<uintptr_t>(dsts[i]) & 15u)) {
                delete b;
                if (rc == ASTC_OK) rc = fail(ASTC_ERR_INVALID_ARG, "image " + std::to_string(i) + ": bad src/dst pointer");
                else set_error("image " + std::to_string(i) + ": " + astc_last_error());
                return rc;
            }
            BatchLevel& L = b->levels[i];
            L.f.src = static_cast<const uint8_t*>(srcs[i]);
            L.f.width = descs[i].width;
            L.f.height = descs[i].height;
            L.f.row_pitch = descs[i].row_pitch;
            L.f.fmt = descs[i].src_format;
            L.f.flags = descs[i].flags;
            L.f.aligned = ((reinterpret_cast<uintptr_t>(srcs[i]) & 3u) == 0) && ((descs[i].row_pitch & 3) == 0);
            L.f.or_mask = normal_or_mask(descs[i].flags);
            set_block_row(&L.f, g.xblocks);
            L.dst = static_cast<uint32_t*>(dsts[i]);
            L.nblocks = (int)g.nblocks;
            b->modes[i] = mode_of(descs[i].block_dim, descs[i].src_format, descs[i].flags, descs[i].width, descs[i].height);
            if (descs[i].flags & ASTC_FLAG_MAX_ACCUMULATION) {
                b->modes[i] = kNoMode;
                b->plain.push_back(astc_batch::PlainImage{descs[i], g, srcs[i], dsts[i]});
            }
            b->blocks += g.nblocks;
            b->texels += (int64_t)descs[i].width * descs[i].height;
        }
        int rc = ASTC_OK;
        if ((e = hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount, ctx->device)) != hipSuccess)
            rc = hip_fail(e, "hipDeviceGetAttribute");
        else if ((e = hipMalloc(&b->d_levels, count * sizeof(BatchLevel))) != hipSuccess)
            rc = hip_fail(e, "hipMalloc(levels)");
        else if ((e = hipMemcpy(b->d_levels, b->levels.data(), count * sizeof(BatchLevel), hipMemcpyHostToDevice)) !=
                 hipSuccess)
            rc = hip_fail(e, "hipMemcpy(levels)");
        else if ((e = hipMalloc(&b->d_scalars, 4 * sizeof(int))) != hipSuccess)
            rc = hip_fail(e, "hipMalloc(scalars)");
        else if ((e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking)) != hipSuccess ||
                 (e = hipStreamCreateWithFlags(&b->side_stream, hipStreamNonBlocking)) != hipSuccess ||
                 (e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess ||
                 (e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming)) != hipSuccess)
            rc = hip_fail(e, "stream/event creation");
        else
            rc = build(b);
        if (rc != ASTC_OK) {
            astc_batch_destroy(b);
            return rc;
        }
        *out = b;
        return ASTC_OK;
    });
}

int astc_batch_set_tuning(astc_batch* b, const astc_tuning* t) {
    return guarded([&]() -> int {
        if (!b || !t) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
        if (!valid_tuning(t)) return fail(ASTC_ERR_INVALID_ARG, "tuning out of range");
        hipError_t e = hipSetDevice(b->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        b->tuning = *t;
        return build(b);
    });
}

int astc_batch_get_tuning(const astc_batch* b, astc_tuning* t) {
    if (!b || !t) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
    *t = b->tuning;
    return ASTC_OK;
}

int astc_batch_launch(astc_batch* b, void* stream) {
    if (!b || !b->exec) return fail(ASTC_ERR_INVALID_ARG, "batch is NULL or not built");
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream))) != hipSuccess)
        return hip_fail(e, "hipGraphLaunch");
    return ASTC_OK;
}

int astc_batch_autotune(astc_batch* b, void* stream, int reps, astc_tuning* best_out) {
    return guarded([&]() -> int {
        if (!b) return fail(ASTC_ERR_INVALID_ARG, "batch is NULL");
        if (reps < 1) reps = 1;
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipEvent_t t0, t1;
        hipError_t e;
        if ((e = hipEventCreate(&t0)) != hipSuccess) return hip_fail(e, "hipEventCreate");
        if ((e = hipEventCreate(&t1)) != hipSuccess) {
            (void)hipEventDestroy(t0);
            return hip_fail(e, "hipEventCreate");
        }
        astc_tuning best = b->tuning;
        float best_ms = 1e30f;
        std::vector<astc_tuning> candidates;
        for (int lanes : {1, 2})
            for (int split : {0, 1})
                for (int bpl : {1, 2, 4, 8})
                    for (int wpc : {1, 2})
                        for (int pf : {0, 1}) {
                            if (bpl == 1 && pf == 1) continue;  // nothing to prefetch within an item
                            candidates.push_back(astc_tuning{bpl, wpc, pf, split, lanes});
                        }
        int rc = ASTC_OK;
        for (const astc_tuning& t : candidates) {
            if ((rc = astc_batch_set_tuning(b, &t)) != ASTC_OK) break;
            if ((rc = astc_batch_launch(b, stream)) != ASTC_OK) break;  // warm-up
            if ((e = hipEventRecord(t0, s)) != hipSuccess) {
                rc = hip_fail(e, "hipEventRecord");
                break;
            }
            for (int r = 0; r < reps && rc == ASTC_OK; ++r) rc = astc_batch_launch(b, stream);
            if (rc != ASTC_OK) break;
            if ((e = hipEventRecord(t1, s)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess) {
                rc = hip_fail(e, "timing event");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, t0, t1);
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        if (rc != ASTC_OK) return rc;
        if ((rc = astc_batch_set_tuning(b, &best)) != ASTC_OK) return rc;
        if (best_out) *best_out = best;
        return ASTC_OK;
    });
}

int astc_batch_info(const astc_batch* b, int64_t* blocks, int64_t* texels, int32_t* items) {
    if (!b) return fail(ASTC_ERR_INVALID_ARG, "batch is NULL");
    if (blocks) *blocks = b->blocks;
    if (texels) *texels = b->texels;
    if (items) *items = b->nitems;
    return ASTC_OK;
}

int astc_batch_destroy(astc_batch* b) {
    if (!b) return ASTC_OK;
    (void)hipSetDevice(b->device);
    free_worklist(b);
    if (b->d_levels) (void)hipFree(b->d_levels);
    if (b->d_scalars) (void)hipFree(b->d_scalars);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    if (b->side_stream) (void)hipStreamDestroy(b->side_stream);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    delete b;
    return ASTC_OK;
}

}  // extern "C"
