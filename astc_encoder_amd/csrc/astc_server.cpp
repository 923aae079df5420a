// astc_server.cpp -- continuous batching of compression requests
// (include/astc_mi355x.h, astc_server_*).
//
// The server owns a device ring (the HBM budget) and up to max_inflight batch
// slots, each with its own stream and captured graph of the persistent kernel.
// A batch is: waits for the requests' sources (an event per request submitted
// from a caller's stream), device-to-device copies of sources into the ring
// for requests that were handed over elsewhere, one host-to-device copy of the
// batch's work list into the slot's buffers, one replay of the slot's graph
// (counter reset + persistent kernel), optional copies of the blocks out, and
// an event.  Slots use separate streams, so a batch formed while others run
// overlaps with them on the device.
//
// Two ways to feed it: astc_server_run_trace replays a trace of requests whose
// sources lie elsewhere in device memory (copied into the ring on admission),
// in real time on the calling thread; the online calls (reserve, submit, poll,
// blocks, release) hand the caller ring space to write its sources into, so
// admission needs no copy, and leave the blocks in the ring until released.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <new>
#include <unordered_map>
#include <thread>
#include <vector>

#include "astc_internal.h"

using namespace astc;

namespace astc {
void plan_work(const std::vector<BatchLevel>& levels, const std::vector<int>& modes, int blocks_per_lane,
               std::vector<BatchSeg>* segs, std::vector<BatchItem>* items, int* first_light);
}

namespace {

constexpr int64_t kAlign = 256;
int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// FIFO ring allocator over [0, size): allocations are made at the head and
// released in any order; space is reclaimed from the oldest live allocation.
class Ring {
  public:
    explicit Ring(int64_t size) : size_(size) {}
    bool alloc(int64_t bytes, int64_t* off) {
        if (bytes > size_) return false;
        int64_t at;
        if (live_.empty()) {
            at = 0;
        } else {
            const int64_t tail = live_.front().off;
            const bool wrapped = live_.back().off < tail;  // newest sits before the oldest
            if (wrapped) {
                if (head_ + bytes > tail) return false;
                at = head_;
            } else if (head_ + bytes <= size_) {
                at = head_;
            } else if (bytes <= tail) {
                at = 0;
            } else {
                return false;
            }
        }
        head_ = at + bytes;
        live_.push_back({at, bytes, false});
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        *off = at;
        return true;
    }
    void release(int64_t off) {
        for (auto& a : live_)
            if (a.off == off && !a.freed) {
                a.freed = true;
                used_ -= a.bytes;
                break;
            }
        while (!live_.empty() && live_.front().freed) live_.pop_front();
        if (live_.empty()) head_ = 0;
    }
    int64_t peak() const { return peak_; }
    bool empty() const { return live_.empty(); }
    void reset_peak() { peak_ = used_; }

  private:
    struct A { int64_t off, bytes; bool freed; };
    int64_t size_, head_ = 0, used_ = 0, peak_ = 0;
    std::deque<A> live_;
};

struct Slot {
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t done = nullptr, fork = nullptr, join = nullptr;
    BatchLevel* d_levels = nullptr;
    BatchSeg* d_segs = nullptr;
    BatchItem* d_items = nullptr;
    int* d_scalars = nullptr;  // [0] items, [1] first 4x4 item, [2..3] queue heads
    BatchLevel* h_levels = nullptr;  // pinned staging
    BatchSeg* h_segs = nullptr;
    BatchItem* h_items = nullptr;
    int* h_scalars = nullptr;
    // device addresses of the pinned staging: the work list is fetched into the
    // device buffers by the library's copy kernel on the slot's stream (the
    // runtime's DMA path for these uploads stalled the serving loop for 5-9 ms
    // now and then)
    const void* dv_levels = nullptr;
    const void* dv_segs = nullptr;
    const void* dv_items = nullptr;
    const void* dv_scalars = nullptr;
    size_t cap_levels = 0, cap_segs = 0, cap_items = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool busy = false;
    std::vector<int> reqs;        // run_trace: indices of the batch's requests
    std::vector<int64_t> offs;    // run_trace: their ring offsets
    std::vector<int64_t> tickets; // online: the batch's tickets
};

struct ReqGeom {
    std::vector<BatchLevel> levels;  // pointers relative to the request's ring base
    std::vector<int> modes;          // -1 (no work-list mode) for plain-launched images
    // images with the max-accumulation search, encoded by one plain launch each
    // after the batch's graph: index into levels, description, geometry
    std::vector<int> plain;
    std::vector<astc_image_desc> plain_desc;
    std::vector<ImageGeom> plain_geom;
    // src_copy: bytes the caller's source buffer holds (the last image ends
    // there); out_bytes: the blocks, image after image with no gaps (16-byte
    // aligned since every image's blocks are); src_bytes: the ring's room for
    // the sources (src_copy rounded up, so the blocks start 256-byte aligned)
    int64_t src_copy = 0, src_bytes = 0, out_bytes = 0, texels = 0, blocks = 0;
};

}  // namespace

namespace {
// An online request: its geometry, ring space, state and the event after which
// its sources are in place (NULL: in place when submitted).
struct Online {
    ReqGeom geo;
    int64_t off = 0;
    enum State { RESERVED, SUBMITTED, IN_FLIGHT, DONE } state = RESERVED;
    hipEvent_t ready = nullptr;
};

// One request of a batch being formed: its geometry and ring offset, the event
// to wait for, a source to copy into the ring first and a destination to copy
// the blocks to afterwards (each optional).
struct BatchEntry {
    const ReqGeom* geo;
    int64_t off;
    hipEvent_t ready;
    const void* copy_src;
    void* copy_dst;
};
}  // namespace

struct astc_server {
    astc_ctx* ctx = nullptr;
    astc_server_config cfg{};
    int num_cus = 0;
    uint8_t* ring_base = nullptr;
    std::vector<Slot> slots;
    Ring ring{0};
    // online requests
    std::unordered_map<int64_t, Online> online;
    std::deque<int64_t> submitted;
    std::deque<int64_t> completed;  // completed tickets not yet reported by poll
    int64_t next_ticket = 1;
    std::vector<hipEvent_t> free_events;
    // batch formation scratch
    std::vector<BatchLevel> levels;
    std::vector<int> modes;
    std::vector<BatchSeg> segs;
    std::vector<BatchItem> items;
    std::vector<BatchEntry> entries;
};

namespace {

void free_slot_buffers(Slot& s) {
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    if (s.graph) (void)hipGraphDestroy(s.graph);
    s.exec = nullptr;
    s.graph = nullptr;
    for (void* p : {(void*)s.d_levels, (void*)s.d_segs, (void*)s.d_items}) if (p) (void)hipFree(p);
    for (void* p : {(void*)s.h_levels, (void*)s.h_segs, (void*)s.h_items}) if (p) (void)hipHostFree(p);
    s.d_levels = nullptr; s.d_segs = nullptr; s.d_items = nullptr;
    s.h_levels = nullptr; s.h_segs = nullptr; s.h_items = nullptr;
    s.dv_levels = nullptr; s.dv_segs = nullptr; s.dv_items = nullptr;
    s.cap_levels = s.cap_segs = s.cap_items = 0;
}

// Make the slot's buffers hold at least the given counts; (re)capture its graph.
int ensure_capacity(astc_server* sv, Slot& s, size_t nlev, size_t nseg, size_t nitem) {
    if (nlev <= s.cap_levels && nseg <= s.cap_segs && nitem <= s.cap_items && s.exec) return ASTC_OK;
    free_slot_buffers(s);
    s.cap_levels = std::max<size_t>(std::max<size_t>(nlev, 64), 2 * nlev);
    s.cap_segs = std::max<size_t>(std::max<size_t>(nseg, 256), 2 * nseg);
    s.cap_items = std::max<size_t>(std::max<size_t>(nitem, 256), 2 * nitem);
    hipError_t e;
    if ((e = hipMalloc(&s.d_levels, s.cap_levels * sizeof(BatchLevel))) != hipSuccess ||
        (e = hipMalloc(&s.d_segs, s.cap_segs * sizeof(BatchSeg))) != hipSuccess ||
        (e = hipMalloc(&s.d_items, s.cap_items * sizeof(BatchItem))) != hipSuccess)
        return hip_fail(e, "hipMalloc(work list)");
    if ((e = hipHostMalloc(&s.h_levels, s.cap_levels * sizeof(BatchLevel), 0)) != hipSuccess ||
        (e = hipHostMalloc(&s.h_segs, s.cap_segs * sizeof(BatchSeg), 0)) != hipSuccess ||
        (e = hipHostMalloc(&s.h_items, s.cap_items * sizeof(BatchItem), 0)) != hipSuccess)
        return hip_fail(e, "hipHostMalloc(staging)");
    if ((e = hipHostGetDevicePointer(const_cast<void**>(&s.dv_levels), s.h_levels, 0)) != hipSuccess ||
        (e = hipHostGetDevicePointer(const_cast<void**>(&s.dv_segs), s.h_segs, 0)) != hipSuccess ||
        (e = hipHostGetDevicePointer(const_cast<void**>(&s.dv_items), s.h_items, 0)) != hipSuccess)
        return hip_fail(e, "hipHostGetDevicePointer(staging)");
    BatchArgs a;
    a.levels = s.d_levels;
    a.segs = s.d_segs;
    a.items = s.d_items;
    a.scalars = s.d_scalars;
    a.counters = s.d_scalars + 2;
    a.tables = sv->ctx->d_tables;
    a.prefetch = sv->cfg.tuning.prefetch;
    BatchLaunch l;
    l.lanes = batch_lanes(sv->cfg.tuning);
    l.grid = std::max(1, sv->num_cus * sv->cfg.tuning.workgroups_per_cu);
    // the 4x4 kernel's own occupancy (five waves per SIMD), whatever the 6x6
    // kernel's: with one 6x6 workgroup per CU, three 4x4 waves share each SIMD
    // with the 6x6 wave from the start and cover its latency
    l.grid4 = std::max(1, 5 * sv->num_cus);
    l.split = sv->cfg.tuning.split_modes;
    l.side = s.side;
    l.fork = s.fork;
    l.join = s.join;
    if ((e = hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_fail(e, "hipStreamBeginCapture");
    hipError_t le = launch_encode_batch(a, l, s.stream);
    e = hipStreamEndCapture(s.stream, &s.graph);
    if (le != hipSuccess) return hip_fail(le, "capture of the batch kernel");
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    if ((e = hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0)) != hipSuccess)
        return hip_fail(e, "hipGraphInstantiate");
    return ASTC_OK;
}

int geometry(const astc_request& r, ReqGeom* g) {
    if (!r.images || r.nimages < 1 || !r.src_dev) return fail(ASTC_ERR_INVALID_ARG, "request without images");
    g->levels.resize(r.nimages);
    g->modes.resize(r.nimages);
    int64_t so = 0, oo = 0;
    for (int i = 0; i < r.nimages; ++i) {
        ImageGeom ig;
        int rc = check_desc(&r.images[i], &ig);
        if (rc != ASTC_OK) return rc;
        if (r.images[i].row_pitch != r.images[i].width * ig.bpp)
            return fail(ASTC_ERR_INVALID_ARG, "request images must be tightly packed");
        BatchLevel& L = g->levels[i];
        L.f.src = reinterpret_cast<const uint8_t*>(so);  // offsets until placed in the ring
        L.f.width = r.images[i].width;
        L.f.height = r.images[i].height;
        L.f.row_pitch = r.images[i].row_pitch;
        L.f.fmt = r.images[i].src_format;
        L.f.flags = r.images[i].flags;
        L.f.aligned = (r.images[i].row_pitch & 3) == 0;
        L.f.or_mask = normal_or_mask(r.images[i].flags);
        set_block_row(&L.f, ig.xblocks);
        L.dst = reinterpret_cast<uint32_t*>(oo);
        L.nblocks = (int)ig.nblocks;
        g->modes[i] = mode_of(r.images[i].block_dim, r.images[i].src_format, r.images[i].flags, r.images[i].width, r.images[i].height);
        if (r.images[i].flags & ASTC_FLAG_MAX_ACCUMULATION) {
            g->modes[i] = -1;  // the planner visits codes 0..MODE_COUNT-1 only
            g->plain.push_back(i);
            g->plain_desc.push_back(r.images[i]);
            g->plain_geom.push_back(ig);
        }
        const int64_t bytes = (int64_t)r.images[i].row_pitch * r.images[i].height;
        g->src_copy = so + bytes;
        so += align_up(bytes);
        oo += ig.nblocks * 16;
        g->texels += (int64_t)r.images[i].width * r.images[i].height;
        g->blocks += ig.nblocks;
    }
    g->src_bytes = align_up(g->src_copy);
    g->out_bytes = oo;
    return ASTC_OK;
}

int64_t ring_bytes(const ReqGeom& g) { return g.src_bytes + align_up(g.out_bytes); }

// Launch one batch of sv->entries on slot s (idle).  On success the slot is
// busy until its `done` event completes.
int launch_batch(astc_server* sv, Slot& s) {
    std::vector<BatchLevel>& levels = sv->levels;
    std::vector<int>& modes = sv->modes;
    levels.clear();
    modes.clear();
    for (const BatchEntry& en : sv->entries) {
        uint8_t* base = sv->ring_base + en.off;
        for (size_t i = 0; i < en.geo->levels.size(); ++i) {
            BatchLevel L = en.geo->levels[i];
            L.f.src = base + reinterpret_cast<uintptr_t>(L.f.src);
            L.dst = reinterpret_cast<uint32_t*>(base + en.geo->src_bytes + reinterpret_cast<uintptr_t>(L.dst));
            levels.push_back(L);
            modes.push_back(en.geo->modes[i]);
        }
    }
    int first_light = 0;
    plan_work(levels, modes, sv->cfg.tuning.blocks_per_lane, &sv->segs, &sv->items, &first_light);
    int rc = ASTC_OK;
    if (levels.size() > s.cap_levels || sv->segs.size() > s.cap_segs || sv->items.size() > s.cap_items) {
        // grow every idle slot at once: regrowing (allocation + graph capture)
        // costs milliseconds, so it should not recur slot by slot mid-stream
        for (Slot& o : sv->slots)
            if (!o.busy && rc == ASTC_OK)
                rc = ensure_capacity(sv, o, std::max(levels.size(), o.cap_levels),
                                     std::max(sv->segs.size(), o.cap_segs), std::max(sv->items.size(), o.cap_items));
        if (rc != ASTC_OK) return rc;
    }
    std::memcpy(s.h_levels, levels.data(), levels.size() * sizeof(BatchLevel));
    std::memcpy(s.h_segs, sv->segs.data(), sv->segs.size() * sizeof(BatchSeg));
    std::memcpy(s.h_items, sv->items.data(), sv->items.size() * sizeof(BatchItem));
    s.h_scalars[0] = (int)sv->items.size();
    s.h_scalars[1] = first_light;
    s.h_scalars[2] = 0;
    s.h_scalars[3] = 0;
    hipError_t e = hipSuccess;
    hipEvent_t last_wait = nullptr;
    for (const BatchEntry& en : sv->entries) {
        if (e == hipSuccess && en.ready && en.ready != last_wait) {
            e = hipStreamWaitEvent(s.stream, en.ready, 0);
            last_wait = en.ready;
        }
        if (e == hipSuccess && en.copy_src)
            e = launch_ring_copy(sv->ring_base + en.off, en.copy_src, en.geo->src_copy, s.stream);
    }
    if (e == hipSuccess) e = launch_ring_copy(s.d_levels, s.dv_levels, levels.size() * sizeof(BatchLevel), s.stream);
    if (e == hipSuccess) e = launch_ring_copy(s.d_segs, s.dv_segs, sv->segs.size() * sizeof(BatchSeg), s.stream);
    if (e == hipSuccess) e = launch_ring_copy(s.d_items, s.dv_items, sv->items.size() * sizeof(BatchItem), s.stream);
    if (e == hipSuccess) e = launch_ring_copy(s.d_scalars, s.dv_scalars, 4 * sizeof(int), s.stream);
    if (e == hipSuccess) e = hipGraphLaunch(s.exec, s.stream);
    for (const BatchEntry& en : sv->entries) {
        const ReqGeom& q = *en.geo;
        uint8_t* base = sv->ring_base + en.off;
        for (size_t j = 0; j < q.plain.size() && e == hipSuccess; ++j) {
            const BatchLevel& L = q.levels[q.plain[j]];
            e = launch_encode_plain(sv->ctx->d_tables, q.plain_desc[j], q.plain_geom[j],
                                    base + reinterpret_cast<uintptr_t>(L.f.src),
                                    base + q.src_bytes + reinterpret_cast<uintptr_t>(L.dst), s.stream, sv->num_cus);
        }
    }
    for (const BatchEntry& en : sv->entries)
        if (e == hipSuccess && en.copy_dst)
            e = launch_ring_copy(en.copy_dst, sv->ring_base + en.off + en.geo->src_bytes, en.geo->out_bytes, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    if (e != hipSuccess) return hip_fail(e, "batch submission");
    s.busy = true;
    return ASTC_OK;
}

hipError_t take_event(astc_server* sv, hipEvent_t* ev) {
    if (!sv->free_events.empty()) {
        *ev = sv->free_events.back();
        sv->free_events.pop_back();
        return hipSuccess;
    }
    return hipEventCreateWithFlags(ev, hipEventDisableTiming);
}

void give_event(astc_server* sv, hipEvent_t ev) {
    if (ev) sv->free_events.push_back(ev);
}

}  // namespace

extern "C" {

int64_t astc_request_bytes(const astc_image_desc* images, int nimages) {
    return guarded([&]() -> int64_t {
        astc_request r{};
        r.images = images;
        r.nimages = nimages;
        r.src_dev = reinterpret_cast<const void*>(1);
        ReqGeom g;
        int rc = geometry(r, &g);
        if (rc != ASTC_OK) return rc;
        return g.src_bytes + align_up(g.out_bytes);
    });
}

int astc_server_create(astc_ctx* ctx, const astc_server_config* cfg, astc_server** out) {
    return guarded([&]() -> int {
        if (!out) return fail(ASTC_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        if (!ctx || !cfg) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
        const astc_tuning& t = cfg->tuning;
        if (cfg->hbm_budget_bytes < kAlign || cfg->max_inflight < 1 || cfg->max_inflight > 8 ||
            cfg->max_batch_requests < 1 || !valid_batch_tuning(t))
            return fail(ASTC_ERR_INVALID_ARG, "server configuration out of range");
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        astc_server* sv = new (std::nothrow) astc_server();
        if (!sv) return fail(ASTC_ERR_NO_MEMORY, "host allocation failed");
        sv->ctx = ctx;
        sv->cfg = *cfg;
        int rc = ASTC_OK;
        if ((e = hipDeviceGetAttribute(&sv->num_cus, hipDeviceAttributeMultiprocessorCount, ctx->device)) != hipSuccess)
            rc = hip_fail(e, "hipDeviceGetAttribute");
        else if ((e = hipMalloc(&sv->ring_base, cfg->hbm_budget_bytes)) != hipSuccess)
            rc = fail(ASTC_ERR_NO_MEMORY, std::string("ring allocation failed: ") + hipGetErrorString(e));
        sv->ring = Ring(cfg->hbm_budget_bytes);
        if (rc == ASTC_OK) {
            sv->slots.resize(cfg->max_inflight);
            for (Slot& s : sv->slots) {
                if ((e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)) != hipSuccess ||
                    (e = hipStreamCreateWithFlags(&s.side, hipStreamNonBlocking)) != hipSuccess ||
                    (e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess ||
                    (e = hipEventCreateWithFlags(&s.fork, hipEventDisableTiming)) != hipSuccess ||
                    (e = hipEventCreateWithFlags(&s.join, hipEventDisableTiming)) != hipSuccess ||
                    (e = hipMalloc(&s.d_scalars, 4 * sizeof(int))) != hipSuccess ||
                    (e = hipHostMalloc(&s.h_scalars, 4 * sizeof(int), 0)) != hipSuccess ||
                    (e = hipHostGetDevicePointer(const_cast<void**>(&s.dv_scalars), s.h_scalars, 0)) != hipSuccess) {
                    rc = hip_fail(e, "slot setup");
                    break;
                }
                // work-list room for a full batch of mip-chained textures (16 levels
                // per request, at most 65536 levels up front) and ~1 G texels, so
                // that regrowing (reallocation and graph capture, milliseconds)
                // stays rare; larger batches still grow on demand
                const size_t levels_room =
                    std::min<size_t>(65536, std::max<size_t>(1024, 16 * size_t(cfg->max_batch_requests)));
                if ((rc = ensure_capacity(sv, s, levels_room, 65536, 65536)) != ASTC_OK) break;
            }
        }
        // Replay every slot's graph once on an empty work list (the kernels exit at
        // once), with the work-list fetch, a ring-to-ring copy and an event on the
        // slot's stream, so that the runtime's first-use work (queue assignment,
        // code-object loading) happens here rather than under the first requests.
        for (Slot& s : sv->slots) {
            if (rc != ASTC_OK) break;
            s.h_scalars[0] = s.h_scalars[1] = s.h_scalars[2] = s.h_scalars[3] = 0;
            if ((e = launch_ring_copy(s.d_scalars, s.dv_scalars, 4 * sizeof(int), s.stream)) != hipSuccess ||
                (e = hipGraphLaunch(s.exec, s.stream)) != hipSuccess ||
                (e = launch_ring_copy(sv->ring_base + 128, sv->ring_base, 128, s.stream)) != hipSuccess ||
                (e = hipEventRecord(s.done, s.stream)) != hipSuccess || (e = hipEventSynchronize(s.done)) != hipSuccess)
                rc = hip_fail(e, "slot warm-up");
        }
        if (rc != ASTC_OK) {
            astc_server_destroy(sv);
            return rc;
        }
        *out = sv;
        return ASTC_OK;
    });
}

int astc_server_destroy(astc_server* sv) {
    if (!sv) return ASTC_OK;
    (void)hipSetDevice(sv->ctx->device);
    for (Slot& s : sv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        free_slot_buffers(s);
        if (s.d_scalars) (void)hipFree(s.d_scalars);
        if (s.h_scalars) (void)hipHostFree(s.h_scalars);
        if (s.side) (void)hipStreamSynchronize(s.side);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.fork) (void)hipEventDestroy(s.fork);
        if (s.join) (void)hipEventDestroy(s.join);
        if (s.side) (void)hipStreamDestroy(s.side);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    for (auto& kv : sv->online)
        if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
    for (hipEvent_t ev : sv->free_events) (void)hipEventDestroy(ev);
    if (sv->ring_base) (void)hipFree(sv->ring_base);
    delete sv;
    return ASTC_OK;
}

int astc_server_run_trace(astc_server* sv, const astc_request* reqs, int n, double* latency_s,
                          astc_server_stats* stats) {
    return guarded([&]() -> int {
        if (!sv || (!reqs && n > 0) || !latency_s || n < 0) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
        hipError_t e = hipSetDevice(sv->ctx->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        std::vector<ReqGeom> geo(n);
        astc_server_stats st{};
        for (int i = 0; i < n; ++i) {
            int rc = geometry(reqs[i], &geo[i]);
            if (rc != ASTC_OK) return fail(rc, "request " + std::to_string(i) + ": " + astc_last_error());
            if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s)
                return fail(ASTC_ERR_INVALID_ARG, "arrival times must be non-decreasing");
            if (geo[i].src_bytes + align_up(geo[i].out_bytes) > sv->cfg.hbm_budget_bytes)
                return fail(ASTC_ERR_BUDGET, "request " + std::to_string(i) + " is larger than the HBM budget");
            st.images += reqs[i].nimages;
            st.texels += geo[i].texels;
            st.blocks += geo[i].blocks;
        }
        st.requests = n;
        if (!sv->online.empty()) return fail(ASTC_ERR_INVALID_ARG, "online requests are outstanding on this server");
        Ring& ring = sv->ring;
        ring.reset_peak();
        std::deque<int> waiting;
        int next = 0, done = 0, inflight = 0;
        int64_t batched_reqs = 0;
        using clk = std::chrono::steady_clock;
        const auto t0 = clk::now();
        auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
        double last_done = 0.0, last_loop = 0.0;
        int rc = ASTC_OK;

        while (done < n && rc == ASTC_OK) {
            bool progressed = false;
            // completions
            for (Slot& s : sv->slots) {
                if (!s.busy) continue;
                e = hipEventQuery(s.done);
                if (e == hipErrorNotReady) continue;
                if (e != hipSuccess) {
                    rc = hip_fail(e, "batch event");
                    break;
                }
                const double t = now();
                for (size_t k = 0; k < s.reqs.size(); ++k) {
                    latency_s[s.reqs[k]] = t - reqs[s.reqs[k]].arrival_s;
                    ring.release(s.offs[k]);
                }
                done += (int)s.reqs.size();
                last_done = t;
                s.busy = false;
                --inflight;
                progressed = true;
            }
            if (rc != ASTC_OK) break;
            // arrivals
            const double t = now();
            while (next < n && reqs[next].arrival_s <= t) waiting.push_back(next++);
            // launches
            for (Slot& s : sv->slots) {
                if (s.busy || waiting.empty()) continue;
                s.reqs.clear();
                s.offs.clear();
                sv->entries.clear();
                while (!waiting.empty() && (int)s.reqs.size() < sv->cfg.max_batch_requests) {
                    const int r = waiting.front();
                    int64_t off;
                    if (!ring.alloc(ring_bytes(geo[r]), &off)) break;
                    waiting.pop_front();
                    s.reqs.push_back(r);
                    s.offs.push_back(off);
                    sv->entries.push_back(BatchEntry{&geo[r], off, nullptr, reqs[r].src_dev, reqs[r].dst_dev});
                }
                if (s.reqs.empty()) break;  // the ring is full: wait for completions
                if ((rc = launch_batch(sv, s)) != ASTC_OK) break;
                ++inflight;
                ++st.batches;
                batched_reqs += (int64_t)s.reqs.size();
                st.peak_inflight = std::max(st.peak_inflight, inflight);
                progressed = true;
            }
            if (rc != ASTC_OK) break;
            // Idle until the next arrival when nothing is on the GPU: a sleep can
            // overshoot by milliseconds, so only long gaps sleep (waking 2 ms early)
            // and the rest is polled.
            const double tl = now();
            st.max_loop_gap_s = std::max(st.max_loop_gap_s, tl - last_loop);
            if (!progressed && inflight == 0 && waiting.empty() && next < n) {
                const double gap = reqs[next].arrival_s - tl;
                if (gap > 3e-3) std::this_thread::sleep_for(std::chrono::duration<double>(gap - 2e-3));
            }
            last_loop = now();
        }
        for (Slot& s : sv->slots) (void)hipStreamSynchronize(s.stream);
        if (rc != ASTC_OK) {
            // the trace owned the whole ring (no online requests): start it afresh
            ring = Ring(sv->cfg.hbm_budget_bytes);
            for (Slot& s : sv->slots) s.busy = false;
            return rc;
        }
        st.wall_s = n ? last_done - reqs[0].arrival_s : 0.0;
        st.mean_batch_requests = st.batches ? double(batched_reqs) / st.batches : 0.0;
        st.peak_ring_bytes = ring.peak();
        if (stats) *stats = st;
        return ASTC_OK;
    });
}

int astc_server_reserve(astc_server* sv, const astc_image_desc* images, int nimages, void** src_ring,
                        int64_t* ticket) {
    return guarded([&]() -> int {
        if (!sv || !src_ring || !ticket) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
        astc_request r{};
        r.images = images;
        r.nimages = nimages;
        r.src_dev = sv->ring_base;  // placed in the ring below
        Online rec;
        int rc = geometry(r, &rec.geo);
        if (rc != ASTC_OK) return rc;
        if (ring_bytes(rec.geo) > sv->cfg.hbm_budget_bytes)
            return fail(ASTC_ERR_BUDGET, "the request is larger than the HBM budget");
        if (!sv->ring.alloc(ring_bytes(rec.geo), &rec.off))
            return fail(ASTC_ERR_BUDGET, "no ring space now: complete batches (astc_server_poll) and retry");
        const int64_t t = sv->next_ticket++;
        *src_ring = sv->ring_base + rec.off;
        *ticket = t;
        sv->online.emplace(t, std::move(rec));
        return ASTC_OK;
    });
}

int astc_server_submit(astc_server* sv, int64_t ticket, void* stream) {
    return guarded([&]() -> int {
        if (!sv) return fail(ASTC_ERR_INVALID_ARG, "server is NULL");
        auto it = sv->online.find(ticket);
        if (it == sv->online.end() || it->second.state != Online::RESERVED)
            return fail(ASTC_ERR_INVALID_ARG, "not a reserved ticket");
        hipError_t e = hipSetDevice(sv->ctx->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        if (stream) {
            hipEvent_t ev = nullptr;
            if ((e = take_event(sv, &ev)) != hipSuccess) return hip_fail(e, "hipEventCreate");
            if ((e = hipEventRecord(ev, static_cast<hipStream_t>(stream))) != hipSuccess) {
                give_event(sv, ev);
                return hip_fail(e, "hipEventRecord");
            }
            it->second.ready = ev;
        }
        it->second.state = Online::SUBMITTED;
        sv->submitted.push_back(ticket);
        return ASTC_OK;
    });
}

int astc_server_poll(astc_server* sv, int64_t* done, int max_done, int* ndone) {
    return guarded([&]() -> int {
        if (!sv || !ndone || max_done < 0 || (max_done > 0 && !done))
            return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
        *ndone = 0;
        hipError_t e = hipSetDevice(sv->ctx->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        for (Slot& s : sv->slots) {
            if (!s.busy || s.tickets.empty()) continue;
            e = hipEventQuery(s.done);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return hip_fail(e, "batch event");
            for (int64_t t : s.tickets) {
                sv->online[t].state = Online::DONE;
                sv->completed.push_back(t);
            }
            s.tickets.clear();
            s.busy = false;
        }
        for (Slot& s : sv->slots) {
            if (s.busy || sv->submitted.empty()) continue;
            sv->entries.clear();
            s.tickets.clear();
            while (!sv->submitted.empty() && (int)s.tickets.size() < sv->cfg.max_batch_requests) {
                const int64_t t = sv->submitted.front();
                sv->submitted.pop_front();
                Online& rec = sv->online[t];
                sv->entries.push_back(BatchEntry{&rec.geo, rec.off, rec.ready, nullptr, nullptr});
                s.tickets.push_back(t);
            }
            int rc = launch_batch(sv, s);
            if (rc != ASTC_OK) {
                // put the batch back in front of the queue, in order
                for (auto it = s.tickets.rbegin(); it != s.tickets.rend(); ++it) sv->submitted.push_front(*it);
                s.tickets.clear();
                return rc;
            }
            for (int64_t t : s.tickets) {
                Online& rec = sv->online[t];
                rec.state = Online::IN_FLIGHT;
                give_event(sv, rec.ready);  // the wait is enqueued: the event may be recorded again
                rec.ready = nullptr;
            }
        }
        while (*ndone < max_done && !sv->completed.empty()) {
            done[(*ndone)++] = sv->completed.front();
            sv->completed.pop_front();
        }
        return ASTC_OK;
    });
}

int astc_server_blocks(astc_server* sv, int64_t ticket, const void** blocks_dev, int64_t* nbytes) {
    if (!sv || !blocks_dev) return fail(ASTC_ERR_INVALID_ARG, "NULL argument");
    auto it = sv->online.find(ticket);
    if (it == sv->online.end() || it->second.state != Online::DONE)
        return fail(ASTC_ERR_INVALID_ARG, "not a completed ticket");
    *blocks_dev = sv->ring_base + it->second.off + it->second.geo.src_bytes;
    if (nbytes) *nbytes = it->second.geo.out_bytes;
    return ASTC_OK;
}

int astc_server_release(astc_server* sv, int64_t ticket) {
    return guarded([&]() -> int {
        if (!sv) return fail(ASTC_ERR_INVALID_ARG, "server is NULL");
        auto it = sv->online.find(ticket);
        if (it == sv->online.end() ||
            (it->second.state != Online::RESERVED && it->second.state != Online::DONE))
            return fail(ASTC_ERR_INVALID_ARG, "only a reserved (not submitted) or completed ticket can be released");
        sv->ring.release(it->second.off);
        give_event(sv, it->second.ready);
        sv->online.erase(it);
        return ASTC_OK;
    });
}

}  // extern "C"
