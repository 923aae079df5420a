// astc_noalloc.hip -- astc_encode_batch_noalloc: the batched persistent encode
// with its work list in a fixed device array, for callers that cannot allocate
// device memory (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// The persistent kernel reads a work list (image descriptions, block ranges,
// items) and two queue heads: device memory beyond the inputs and outputs.
// astc_batch_create allocates it per batch; this entry point instead keeps it
// in `scratch`, a file-scope device array of kNoAllocScratchBytes that every
// GPU the library runs on holds until the process ends.  Its address on each
// device is taken once, in astc_ctx_create (noalloc_setup, with the device
// current), and only that address is used on the host.  The plan is written
// into pinned host staging that the context allocated at creation, copied into
// scratch on the caller's stream by the library's copy kernel (which reads the
// staging through its device address), and the staging is reused only after
// that copy has completed (an event).  The queue heads are cleared by a kernel on
// the same stream before the persistent kernel runs.
//
// Every call on one GPU uses that GPU's scratch: calls on one stream follow
// each other, but calls that overlap on the device (different streams, or other
// threads' streams) race on it.
#include "astc_internal.h"

namespace astc {

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

int noalloc_setup(astc_ctx* ctx) {
    hipError_t e = hipGetSymbolAddress(&ctx->noalloc.scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        ctx->noalloc.scratch = nullptr;
        return hip_fail(e, "hipGetSymbolAddress(scratch)");
    }
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&ctx->noalloc.staging), sizeof(NoAllocPlan), 0)) != hipSuccess) {
        ctx->noalloc.staging = nullptr;
        return hip_fail(e, "hipHostMalloc(no-alloc staging)");
    }
    if ((e = hipHostGetDevicePointer(const_cast<void**>(&ctx->noalloc.staging_dev), ctx->noalloc.staging, 0)) !=
        hipSuccess) {
        ctx->noalloc.staging_dev = nullptr;
        return hip_fail(e, "hipHostGetDevicePointer(no-alloc staging)");
    }
    if ((e = hipEventCreateWithFlags(&ctx->noalloc.staged, hipEventDisableTiming)) != hipSuccess) {
        ctx->noalloc.staged = nullptr;
        return hip_fail(e, "hipEventCreate(no-alloc staging)");
    }
    return ASTC_OK;
}

void noalloc_teardown(astc_ctx* ctx) {
    if (ctx->noalloc.staged) {
        (void)hipEventSynchronize(ctx->noalloc.staged);
        (void)hipEventDestroy(ctx->noalloc.staged);
    }
    if (ctx->noalloc.staging) (void)hipHostFree(ctx->noalloc.staging);
    ctx->noalloc.staged = nullptr;
    ctx->noalloc.staging = nullptr;
    ctx->noalloc.staging_dev = nullptr;
    ctx->noalloc.scratch = nullptr;
}

}  // namespace astc

using namespace astc;

extern "C" int astc_encode_batch_noalloc(astc_ctx* ctx, const astc_image_desc* descs, const void* const* srcs,
                                         void* const* dsts, int count, const astc_tuning* tuning, void* stream) {
    return guarded([&]() -> int {
        if (!ctx || !descs || !srcs || !dsts || count < 1) return fail(ASTC_ERR_INVALID_ARG, "empty or NULL batch");
        if (!ctx->noalloc.scratch || !ctx->noalloc.staging || !ctx->noalloc.staging_dev || !ctx->noalloc.staged)
            return fail(ASTC_ERR_INVALID_ARG, "context has no no-alloc scratch");
        if (count > kNoAllocMaxImages) return fail(ASTC_ERR_INVALID_ARG, "more images than the no-alloc maximum (8192)");
        const astc_tuning t = tuning ? *tuning : astc_tuning{4, 2, 1, 0, 1};
        const int bpl = t.blocks_per_lane;
        if (!valid_batch_tuning(t) || t.split_modes != 0)
            return fail(ASTC_ERR_INVALID_ARG, "tuning out of range (the no-alloc form runs one kernel: split_modes 0)");
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        hipStream_t s = static_cast<hipStream_t>(stream);

        std::lock_guard<std::mutex> lock(ctx->noalloc.mu);
        // the previous call's upload has been read out of the staging
        if ((e = hipEventSynchronize(ctx->noalloc.staged)) != hipSuccess) return hip_fail(e, "no-alloc staging event");
        NoAllocPlan& P = *ctx->noalloc.staging;
        int* modes = ctx->noalloc.modes;
        int nplain = 0;
        for (int i = 0; i < count; ++i) {
            ImageGeom g;
            if (check_desc(&descs[i], &g) != ASTC_OK) return ASTC_ERR_INVALID_ARG;  // message set by check_desc
            if (!srcs[i] || !dsts[i] || (reinterpret_cast<uintptr_t>(dsts[i]) & 15u))
                return fail(ASTC_ERR_INVALID_ARG, "bad src/dst pointer (dst must be 16-byte aligned)");
            BatchLevel& L = P.levels[i];
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
            L.pad = 0;
            modes[i] = mode_of(descs[i].block_dim, descs[i].src_format, descs[i].flags, descs[i].width,
                               descs[i].height);
            if (descs[i].flags & ASTC_FLAG_MAX_ACCUMULATION) {
                modes[i] = -1;  // one plain launch after the persistent kernel
                ++nplain;
            }
        }
        int nseg = 0, nitem = 0, first_light = 0;
        if (!plan_work_into(P.levels, modes, count, bpl, ctx->noalloc.keys, P.segs, kNoAllocMaxItems, P.items,
                            kNoAllocMaxItems, &nseg, &nitem, &first_light))
            return fail(ASTC_ERR_INVALID_ARG, "the batch needs more work items than the no-alloc maximum (131072)");
        P.scalars[0] = nitem;
        P.scalars[1] = first_light;
        P.scalars[2] = P.scalars[3] = 0;

        uint8_t* base = static_cast<uint8_t*>(ctx->noalloc.scratch);
        BatchArgs a;
        a.levels = reinterpret_cast<const BatchLevel*>(base + offsetof(NoAllocPlan, levels));
        a.segs = reinterpret_cast<const BatchSeg*>(base + offsetof(NoAllocPlan, segs));
        a.items = reinterpret_cast<const BatchItem*>(base + offsetof(NoAllocPlan, items));
        a.scalars = reinterpret_cast<const int*>(base + offsetof(NoAllocPlan, scalars));
        a.counters = reinterpret_cast<int*>(base + offsetof(NoAllocPlan, scalars)) + 2;
        a.tables = ctx->d_tables;
        a.prefetch = t.prefetch;
        const struct { size_t off, bytes; } parts[4] = {
            {offsetof(NoAllocPlan, levels), size_t(count) * sizeof(BatchLevel)},
            {offsetof(NoAllocPlan, segs), size_t(nseg) * sizeof(BatchSeg)},
            {offsetof(NoAllocPlan, items), size_t(nitem) * sizeof(BatchItem)},
            {offsetof(NoAllocPlan, scalars), sizeof(P.scalars)}};
        // fetched by the library's copy kernel through the staging's device address
        // (no DMA-engine copy: those stalled the host for milliseconds now and then)
        const uint8_t* staged_dev = static_cast<const uint8_t*>(ctx->noalloc.staging_dev);
        for (const auto& part : parts) {
            if (!part.bytes) continue;
            if ((e = launch_ring_copy(base + part.off, staged_dev + part.off, int64_t(part.bytes), s)) != hipSuccess)
                return hip_fail(e, "no-alloc work-list upload");
        }
        if ((e = hipEventRecord(ctx->noalloc.staged, s)) != hipSuccess) return hip_fail(e, "no-alloc staging event");
        BatchLaunch l{};
        l.lanes = batch_lanes(t);
        l.grid = std::max(1, ctx->num_cus * t.workgroups_per_cu);
        l.split = 0;
        if ((e = launch_encode_batch(a, l, s)) != hipSuccess) return hip_fail(e, "no-alloc batch launch");
        for (int i = 0; i < count && nplain; ++i) {
            if (!(descs[i].flags & ASTC_FLAG_MAX_ACCUMULATION)) continue;
            ImageGeom g;
            (void)check_desc(&descs[i], &g);
            if ((e = launch_encode_plain(ctx->d_tables, descs[i], g, srcs[i], dsts[i], s, ctx->num_cus)) != hipSuccess)
                return hip_fail(e, "no-alloc plain launch");
        }
        return ASTC_OK;
    });
}
