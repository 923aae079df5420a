// astc_kernels.hip -- gfx950 kernels of the ASTC encoder.
//
// encode_plain_kernel: the per-texture launch that replaces MainCS
// (ASTC_Encode.hlsl:553-582).  One lane encodes one block: with ~1-2.5k
// floating-point operations per block and hundreds of thousands of blocks per
// texture, keeping each block in one lane avoids cross-lane reductions and
// replicated serial work (the 16-step power iteration), so every lane issues
// useful VALU work.  A workgroup of 256 lanes covers BPL x 256 consecutive
// blocks in row-major block order: in step k lane t encodes block
// base + 256 k + t, so the lanes of a wave read adjacent block rows of the
// texture and write 16-byte blocks to consecutive addresses.  By default a
// lane issues its next block's loads as soon as the current block is
// converted (tiled RGB8 sources: after the store); the 6x6 encoder holds 144 texel floats
// per lane and runs two waves per SIMD.  encode_plain_pair_kernel is the
// two-lanes-per-block 6x6 form (astc_plain_tuning.lanes_per_block = 2), and
// encode_plain_packed_kernel the form that holds the fetched bytes instead of
// floats (astc_plain_tuning.packed_texels = 1).
#include "astc_internal.h"

namespace astc {

struct PlainArgs {
    FetchArgs f;
    int nblocks;
    uint32_t* dst;
    const DeviceTables* tables;
};

// LDS tables: lut[0..255] sRGB -> float(linear) * 255, lut[256..511] UNORM8 ->
// float * 255 as the hardware converts it (binary32 c * (1/255), then x255; 126
// values land one ulp off the integer), and the trit groups' code bits for both
// weight quantisations.
__device__ __forceinline__ void stage_tables(const DeviceTables* t, float* lut, PackTables* pack) {
    const int tid = threadIdx.x;
    for (int i = tid; i < 256; i += blockDim.x) {
        lut[i] = t->srgb[i];
        lut[256 + i] = (float(i) * (1.0f / 255.0f)) * 255.0f;
        const uint32_t T = t->trit[i];
        pack->spread[0][i] = trit_group_spread(T, kSplitQ6.bits);
        pack->spread[1][i] = trit_group_spread(T, kSplitQ12.bits);
    }
}

template <int DIM, int CH>
__device__ __forceinline__ void load_block_id(const FetchArgs& f, int id, uint32_t p[Layout<DIM, CH>::WORDS]) {
    const int by = block_row_of(f, id);
    load_block<DIM, CH>(f, id - by * f.xblocks, by, p);
}

__host__ __device__ inline int channel_class(int fmt, uint32_t flags, int width, int height, int dim) {
    if (flags & FLAG_NORMAL) return CH_NORMAL;
    return (fmt == FMT_RGB8 && width % dim == 0 && height % dim == 0) ? CH_RGB8 : CH_ALL;
}

// SRGB: 1 / 0 fixes the fetch at compile time, -1 selects the table per image
// (sRGB or identity, no branch).  The channel class CH is the number of active
// channels the encoder works on.  OVERRIDE applies the normal-map override
// (blue = alpha = 255) to the texel words.  Only an encoder that reads blue or
// alpha of a normal map needs it: the principal component encoders take every
// normal map through the two-channel class, which reads red and green only, and
// the mask is zero for every other image, so only the four-channel
// max-accumulation search sets it.
template <int DIM, int SRGB, int CH, bool OVERRIDE = false, int NR = DIM>
__device__ __forceinline__ void unpack_texels(const uint32_t p[NR * Layout<DIM, CH>::WPR], const FetchArgs& f,
                                              const float* lut, float x[NR * DIM][4]) {
    const float* colour = SRGB < 0 ? lut + ((f.flags & FLAG_SRGB) ? 0 : 256) : SRGB == 1 ? lut : lut + 256;
    unpack_block<DIM, CH, NR>(p, OVERRIDE ? f.or_mask : 0u, colour, lut + 256, x);
}

__device__ __forceinline__ void store_block(uint32_t* dst, int id, const Block128& b) {
    reinterpret_cast<uint4*>(dst)[id] = make_uint4(b.w[0], b.w[1], b.w[2], b.w[3]);
}

// Power-iteration rescale interval: the batch kernels' 6x6 path (SRGB < 0: the
// table is chosen per image) rescales before every product, which keeps the
// kernels that hold it at 224 VGPRs; every other path rescales before three of
// the seven (fewer instructions).  Both give the same blocks.  At 224 VGPRs
// (allocated in steps of 8) a 6x6 wave leaves 288 of a SIMD's 512 for three
// waves of the 4x4 batch kernel, which run beside it when the split form runs
// one 6x6 workgroup per CU; one register more leaves room for only two.
template <int DIM, int SRGB>
struct BatchRescale { static constexpr int value = (SRGB < 0 && DIM == 6) ? 1 : 3; };

// Encode blocks [first, first + count) of one image with the workgroup's
// threads: step k, thread t takes block first + k * blockDim.x + t.  `cur`
// holds the packed texels of the thread's first block (the caller issues those
// loads).  With prefetch the next step's texels are loaded as soon as the
// current ones are converted (into the same registers), so the loads are in
// flight while the block is encoded; without, after the block is stored.
// From the store of one block to the end of the next block's conversion (and,
// with prefetch, the issue of the following block's loads) the wave runs at
// issue priority 2 (issue_priority): that phase is a chain of dependent memory
// and LDS operations with little arithmetic, and giving it precedence over the
// other wave's encode keeps the two waves of a SIMD out of phase.
// FLAT: a wave whose blocks all have a flat alpha (flat_alpha) is encoded by
// the three-channel path (same bytes); a separately compiled kernel, which the
// default plain tuning uses for four-channel encodes (the batch kernels do not
// test for flat alpha).
template <int DIM, bool ALPHA, int SRGB, int CH, int METHOD = METHOD_PCA, bool FLAT = false>
__device__ __forceinline__ void encode_loaded(const FetchArgs& f, uint32_t* dst, int first, int count, bool prefetch,
                                              const float* lut, const PackTables* trit,
                                              uint32_t cur[Layout<DIM, CH>::WORDS]) {
    const int stride = blockDim.x;
    constexpr int RESCALE = BatchRescale<DIM, SRGB>::value;
#pragma unroll 1
    for (int o = threadIdx.x;; o += stride) {
        const bool has_next = o + stride < count;
        constexpr bool kFlat = FLAT && CH == CH_ALL && METHOD == METHOD_PCA;
        const bool flat = kFlat && __all(flat_alpha<DIM * DIM>(cur, lut + 256));
        float x[DIM * DIM][4];
        unpack_texels<DIM, SRGB, CH, METHOD == METHOD_MAX_ACCUMULATION>(cur, f, lut, x);
        // the packed words are dead once converted: the next block's loads reuse them
        if (has_next && prefetch) load_block_id<DIM, CH>(f, first + o + stride, cur);
        issue_priority(0);
        Block128 b;
        if (kFlat && flat) {
            // keeps the two encoders' code apart: merged, their live values exceed the register file
            __builtin_amdgcn_sched_barrier(0);
            b = encode_block<DIM, ALPHA, CH_RGB8, METHOD, RESCALE>(x, trit, x[0][3]);
        } else {
            b = encode_block<DIM, ALPHA, CH, METHOD, RESCALE>(x, trit);
        }
        issue_priority(2);
        store_block(dst, first + o, b);
        if (!has_next) break;
        if (!prefetch) load_block_id<DIM, CH>(f, first + o + stride, cur);
    }
    issue_priority(0);
}

template <int DIM, bool ALPHA, int SRGB, int CH, int METHOD = METHOD_PCA, bool FLAT = false>
__device__ __forceinline__ void encode_chunk(const FetchArgs& f, uint32_t* dst, int first, int count, bool prefetch,
                                             const float* lut, const PackTables* trit) {
    if (int(threadIdx.x) >= count) return;
    uint32_t cur[Layout<DIM, CH>::WORDS];
    issue_priority(2);  // the first block's fetch, like every later one (encode_loaded)
    load_block_id<DIM, CH>(f, first + threadIdx.x, cur);
    encode_loaded<DIM, ALPHA, SRGB, CH, METHOD, FLAT>(f, dst, first, count, prefetch, lut, trit, cur);
}

// Minimum waves per SIMD the register budget must allow: 4x4 encoders fit in
// 128 VGPRs (the reduced-channel ones use 68-90 and run at five to seven waves);
// 6x6 blocks hold 144 texel floats and run at two.
template <int DIM>
struct PlainWaves { static constexpr int value = DIM == 4 ? 4 : 2; };

template <int DIM, bool ALPHA, bool SRGB, int CH, int METHOD, bool FLAT = false>
__global__ void __launch_bounds__(kPlainThreads)
__attribute__((amdgpu_waves_per_eu(PlainWaves<DIM>::value))) encode_plain_kernel(PlainArgs a, int bpl, int prefetch) {
    __shared__ float lut[512];
    __shared__ PackTables trit;
    stage_tables(a.tables, lut, &trit);
    __syncthreads();
    const int chunk = kPlainThreads * bpl;
    const int first = blockIdx.x * chunk;
    const int count = min(chunk, a.nblocks - first);
    encode_chunk<DIM, ALPHA, SRGB ? 1 : 0, CH, METHOD, FLAT>(a.f, a.dst, first, count, prefetch != 0, lut, &trit);
}

// ---------------------------------------------------------------------------
// One lane per block with the texels held as the fetched words (PackedTexels,
// astc_plain_tuning.packed_texels = 1): each pass of the encode converts them
// again through the LDS tables, which trades VALU and LDS instructions for
// registers: a 6x6 four-channel block needs ~170 VGPRs instead of 218 (three
// waves per SIMD instead of two), a 4x4 one ~96 (five instead of four).  The
// words are live through the encode, so the next block is loaded after the
// store (no prefetch).  Principal component encodes only.
// ---------------------------------------------------------------------------
template <int DIM>
struct PackedWaves { static constexpr int value = DIM == 4 ? 5 : 3; };

template <int DIM, bool ALPHA, bool SRGB, int CH>
__global__ void __launch_bounds__(kPlainThreads) __attribute__((amdgpu_waves_per_eu(PackedWaves<DIM>::value)))
encode_plain_packed_kernel(PlainArgs a, int bpl) {
    __shared__ float lut[512];
    __shared__ PackTables trit;
    stage_tables(a.tables, lut, &trit);
    __syncthreads();
    const int chunk = kPlainThreads * bpl;
    const int first = blockIdx.x * chunk;
    const int count = min(chunk, a.nblocks - first);
    if (int(threadIdx.x) >= count) return;
    PackedTexels<DIM, CH> x;
    x.colour = SRGB ? lut : lut + 256;
    x.unorm = lut + 256;
    load_block_id<DIM, CH>(a.f, first + threadIdx.x, x.p);
#pragma unroll 1
    for (int o = threadIdx.x;; o += kPlainThreads) {
        const bool has_next = o + kPlainThreads < count;
        store_block(a.dst, first + o, encode_block<DIM, ALPHA, CH, METHOD_PCA, 3>(x, &trit));
        if (!has_next) break;
        load_block_id<DIM, CH>(a.f, first + o + kPlainThreads, x.p);
    }
}

// ---------------------------------------------------------------------------
// Two lanes per 6x6 block (encode_block_pair): lanes 2j and 2j+1 load, convert
// and encode rows 0-2 and 3-5 of one block and each store 8 of its 16 bytes.
// A workgroup of 256 lanes covers 128 blocks per step: in step k, lane pair j
// encodes block first + 128 k + j.  Half the texels per lane lowers the
// register budget (a third wave per SIMD) and halves the work of one wave
// iteration (finer balance over the SIMDs); the serial part of the encode runs
// on both lanes.
// ---------------------------------------------------------------------------
template <int CH>
__device__ __forceinline__ void load_half_block(const FetchArgs& f, int id, int half,
                                                uint32_t p[3 * Layout<6, CH>::WPR]) {
    const int by = block_row_of(f, id);
    load_block<6, CH, 3>(f, id - by * f.xblocks, by, p, 3 * half);
}

__device__ __forceinline__ void store_half_block(uint32_t* dst, int id, int half, const Block128& b) {
    reinterpret_cast<uint2*>(dst)[2 * id + half] = half ? make_uint2(b.w[2], b.w[3]) : make_uint2(b.w[0], b.w[1]);
}

template <bool ALPHA, int SRGB, int CH, int RESCALE>
__device__ __forceinline__ void encode_chunk_pair(const FetchArgs& f, uint32_t* dst, int first, int count,
                                                  bool prefetch, const float* lut, const PackTables* trit) {
    const int half = threadIdx.x & 1;
    const int stride = blockDim.x >> 1;
    int o = threadIdx.x >> 1;
    if (o >= count) return;
    uint32_t cur[3 * Layout<6, CH>::WPR];
    load_half_block<CH>(f, first + o, half, cur);
#pragma unroll 1
    for (;; o += stride) {
        const bool has_next = o + stride < count;
        float x[18][4];
        unpack_texels<6, SRGB, CH, false, 3>(cur, f, lut, x);
        issue_priority(0);
        if (has_next && prefetch) load_half_block<CH>(f, first + o + stride, half, cur);
        const Block128 b = encode_block_pair<ALPHA, CH, RESCALE>(x, half, trit);
        issue_priority(2);  // as in encode_loaded
        store_half_block(dst, first + o, half, b);
        if (!has_next) break;
        if (!prefetch) load_half_block<CH>(f, first + o + stride, half, cur);
    }
    issue_priority(0);
}

template <bool ALPHA, bool SRGB, int CH>
__global__ void __launch_bounds__(kPlainThreads) __attribute__((amdgpu_waves_per_eu(3)))
encode_plain_pair_kernel(PlainArgs a, int bpl, int prefetch) {
    __shared__ float lut[512];
    __shared__ PackTables trit;
    stage_tables(a.tables, lut, &trit);
    __syncthreads();
    const int chunk = (kPlainThreads / 2) * bpl;
    const int first = blockIdx.x * chunk;
    const int count = min(chunk, a.nblocks - first);
    encode_chunk_pair<ALPHA, SRGB ? 1 : 0, CH, 3>(a.f, a.dst, first, count, prefetch != 0, lut, &trit);
}

// ---------------------------------------------------------------------------
// Batched persistent kernel: workgroups take work items from a device queue
// (atomic counter) until it is empty.  An item is either a chunk of one image
// (up to 256 x bpl blocks) or a pack of up to kMaxSegs small images (at most
// one block per thread).  Items are uniform in mode (block size x alpha), so
// each workgroup runs one encoder specialisation per item.
// ---------------------------------------------------------------------------
__global__ void reset_counter_kernel(int* counters) {
    if (threadIdx.x < 2 && blockIdx.x == 0) counters[threadIdx.x] = 0;
}

// LPB = 2: 6x6 items with two lanes per block (encode_block_pair): a chunk
// runs encode_chunk_pair; a pack of small images (at most 256 blocks) takes
// two rounds of 128 lane pairs.
template <int DIM, bool ALPHA, int CH, int LPB = 1>
__device__ __forceinline__ void process_item(const BatchArgs& a, const BatchItem& item, const float* lut,
                                             const PackTables* trit) {
    if (DIM == 6 && LPB == 2) {
        if (item.nseg == 1) {
            const BatchSeg sg = a.segs[item.seg_begin];
            const BatchLevel& L = a.levels[sg.level];
            encode_chunk_pair<ALPHA, -1, CH, 3>(L.f, L.dst, sg.first_block, sg.count, a.prefetch != 0, lut, trit);
            return;
        }
        const int half = threadIdx.x & 1, pairs = blockDim.x >> 1;
        for (int r = 0; r < 2; ++r) {
            int o = r * pairs + (threadIdx.x >> 1);
            for (int s = 0; s < item.nseg; ++s) {
                const BatchSeg sg = a.segs[item.seg_begin + s];
                if (o < sg.count) {
                    const BatchLevel& L = a.levels[sg.level];
                    uint32_t p[3 * Layout<6, CH>::WPR];
                    load_half_block<CH>(L.f, sg.first_block + o, half, p);
                    float x[18][4];
                    unpack_texels<6, -1, CH, false, 3>(p, L.f, lut, x);
                    store_half_block(L.dst, sg.first_block + o, half, encode_block_pair<ALPHA, CH, 3>(x, half, trit));
                    break;
                }
                o -= sg.count;
            }
        }
        return;
    }
    if (item.nseg == 1) {
        const BatchSeg sg = a.segs[item.seg_begin];
        const BatchLevel& L = a.levels[sg.level];
        encode_chunk<DIM, ALPHA, -1, CH>(L.f, L.dst, sg.first_block, sg.count, a.prefetch != 0, lut, trit);
        return;
    }
    int o = threadIdx.x;
    for (int s = 0; s < item.nseg; ++s) {
        const BatchSeg sg = a.segs[item.seg_begin + s];
        if (o < sg.count) {
            const BatchLevel& L = a.levels[sg.level];
            uint32_t p[Layout<DIM, CH>::WORDS];
            issue_priority(2);  // fetch and conversion, as in encode_loaded
            load_block_id<DIM, CH>(L.f, sg.first_block + o, p);
            float x[DIM * DIM][4];
            unpack_texels<DIM, -1, CH>(p, L.f, lut, x);
            issue_priority(0);
            store_block(L.dst, sg.first_block + o, encode_block<DIM, ALPHA, CH, METHOD_PCA, BatchRescale<DIM, -1>::value>(x, trit));
            return;
        }
        o -= sg.count;
    }
}

template <int DIM, bool ALPHA, int LPB>
__device__ __forceinline__ void process_item_ch(const BatchArgs& a, const BatchItem& item, const float* lut,
                                                const PackTables* trit) {
    switch (item.mode & MODE_CH_MASK) {
    case MODE_CH_NORMAL: process_item<DIM, ALPHA, CH_NORMAL, LPB>(a, item, lut, trit); break;
    case MODE_CH_RGB8: process_item<DIM, ALPHA, CH_RGB8, LPB>(a, item, lut, trit); break;
    default: process_item<DIM, ALPHA, CH_ALL, LPB>(a, item, lut, trit); break;
    }
}

// SEL = 0: every item (one kernel holds every path and the 6x6 register
// budget, two waves per SIMD).  SEL = 6 / 4 split the list at the first 4x4
// item of a reduced-channel class (RGB8 tiles, normal maps): the 6x6 kernel
// takes the front (6x6 items, then 4x4 four-channel items, which fit its
// budget), the 4x4 kernel the rest, with the smaller budget of those two
// classes (five waves per SIMD).  LPB = 2 (6x6 items with two lanes per
// block) lowers the budget of the kernels that hold 6x6 items to three waves.
template <int SEL, int LPB>
struct BatchWaves { static constexpr int value = SEL == 4 ? 5 : LPB == 2 ? 3 : 2; };

template <int SEL, bool ALPHA>
__device__ __forceinline__ void process_item_4x4(const BatchArgs& a, const BatchItem& item, const float* lut,
                                                 const PackTables* trit) {
    switch (item.mode & MODE_CH_MASK) {
    case MODE_CH_NORMAL:
        if (SEL != 6) process_item<4, ALPHA, CH_NORMAL>(a, item, lut, trit);
        break;
    case MODE_CH_RGB8:
        if (SEL != 6) process_item<4, ALPHA, CH_RGB8>(a, item, lut, trit);
        break;
    default:
        if (SEL != 4) process_item<4, ALPHA, CH_ALL>(a, item, lut, trit);
        break;
    }
}

template <int SEL, int LPB = 1>
__global__ void __launch_bounds__(kPlainThreads) __attribute__((amdgpu_waves_per_eu(BatchWaves<SEL, LPB>::value)))
encode_batch_kernel(BatchArgs a) {
    __shared__ float lut[512];
    __shared__ PackTables trit;
    __shared__ int next_item;
    stage_tables(a.tables, lut, &trit);
    const int nitems = a.scalars[0], first_light = a.scalars[1];
    const int begin = SEL == 4 ? first_light : 0;
    const int end = SEL == 6 ? first_light : nitems;
    int* counter = a.counters + (SEL == 4 ? 1 : 0);
    // Thread 0 claims the workgroup's next item while the current one is being
    // encoded, so the atomic's round trip is off the critical path.
    int claimed = 0;
    if (threadIdx.x == 0) claimed = atomicAdd(counter, 1);
    for (;;) {
        __syncthreads();  // tables staged / everyone done with the previous item index
        if (threadIdx.x == 0) next_item = begin + claimed;
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(next_item);
        if (it >= end) break;
        if (threadIdx.x == 0) claimed = atomicAdd(counter, 1);
        const BatchItem item = a.items[it];
        // one code path per block size x alpha x channel class; linear and sRGB
        // fetches share it through the table choice (separate specialisations
        // would not fit the register budget)
        switch (item.mode & 3) {
        case MODE_4x4_RGB:
            process_item_4x4<SEL, false>(a, item, lut, &trit);
            break;
        case MODE_4x4_RGBA:
            process_item_4x4<SEL, true>(a, item, lut, &trit);
            break;
        case MODE_6x6_RGB:
            if (SEL != 4) process_item_ch<6, false, LPB>(a, item, lut, &trit);
            break;
        default:
            if (SEL != 4) process_item_ch<6, true, LPB>(a, item, lut, &trit);
            break;
        }
    }
}

hipError_t launch_encode_batch(const BatchArgs& a, const BatchLaunch& l, hipStream_t stream) {
    (void)hipGetLastError();  // report this launch's errors only, not an earlier call's
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, a.counters);
    if (!l.split) {
        if (l.lanes == 2)
            hipLaunchKernelGGL((encode_batch_kernel<0, 2>), dim3(unsigned(l.grid)), dim3(kPlainThreads), 0, stream, a);
        else
            hipLaunchKernelGGL((encode_batch_kernel<0, 1>), dim3(unsigned(l.grid)), dim3(kPlainThreads), 0, stream, a);
        return hipGetLastError();
    }
    // 6x6 kernel first on `stream`; the 4x4 kernel on the side stream picks up
    // compute units as 6x6 workgroups finish; joined back into `stream`.
    hipError_t e;
    if ((e = hipEventRecord(l.fork, stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(l.side, l.fork, 0)) != hipSuccess) return e;
    if (l.lanes == 2)
        hipLaunchKernelGGL((encode_batch_kernel<6, 2>), dim3(unsigned(l.grid)), dim3(kPlainThreads), 0, stream, a);
    else
        hipLaunchKernelGGL((encode_batch_kernel<6, 1>), dim3(unsigned(l.grid)), dim3(kPlainThreads), 0, stream, a);
    hipLaunchKernelGGL((encode_batch_kernel<4, 1>), dim3(unsigned(l.grid4)), dim3(kPlainThreads), 0, l.side, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(l.join, l.side)) != hipSuccess) return e;
    return hipStreamWaitEvent(stream, l.join, 0);
}

// ---------------------------------------------------------------------------
// Mip generation: next level of a tightly packed 8-bit image; a side of at
// least 8 halves (rounding down) and a shorter side is kept, so 2 * out <= in
// and every tap is inside the source.  2x2 box filter in stored space,
// (a+b+c+d+2)>>2, or (a+b+1)>>1 along the one axis that halves.  One thread per
// output byte.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) downsample_kernel(const uint8_t* __restrict__ src, int w, int h, int bpp,
                                                        uint8_t* __restrict__ dst, int ow, int oh) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t total = int64_t(ow) * oh * bpp;
    if (i >= total) return;
    const int c = int(i % bpp);
    const int64_t t = i / bpp;
    const int x = int(t % ow), y = int(t / ow);
    const bool hx = ow < w, hy = oh < h;
    const int sx = hx ? 2 * x : x, sy = hy ? 2 * y : y;
    auto at = [&](int xx, int yy) { return int(src[(int64_t(yy) * w + xx) * bpp + c]); };
    int v;
    if (hx && hy) v = (at(sx, sy) + at(sx + 1, sy) + at(sx, sy + 1) + at(sx + 1, sy + 1) + 2) >> 2;
    else if (hx) v = (at(sx, sy) + at(sx + 1, sy) + 1) >> 1;
    else if (hy) v = (at(sx, sy) + at(sx, sy + 1) + 1) >> 1;
    else v = at(sx, sy);
    dst[i] = uint8_t(v);
}

// Device-to-device copy for the server's ring (sources in, blocks out): 16-byte
// vectors when both ends and the size allow, bytes otherwise; grid-stride.  A
// kernel of the library's own keeps these copies on the compute queue of the
// slot's stream.
__global__ void __launch_bounds__(256) ring_copy_kernel(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src,
                                                        int64_t bytes) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | uintptr_t(bytes)) & 15u) == 0) {
        const int64_t n = bytes >> 4;
        for (int64_t i = t; i < n; i += stride)
            reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
    } else {
        for (int64_t i = t; i < bytes; i += stride) dst[i] = src[i];
    }
}

hipError_t launch_ring_copy(void* dst, const void* src, int64_t bytes, hipStream_t stream) {
    if (bytes <= 0) return hipSuccess;
    const int64_t per_block = 256 * 16 * 4;  // four vectors per thread per pass
    const unsigned blocks = unsigned(std::min<int64_t>(2048, (bytes + per_block - 1) / per_block));
    hipLaunchKernelGGL(ring_copy_kernel, dim3(blocks), dim3(256), 0, stream, static_cast<uint8_t*>(dst),
                       static_cast<const uint8_t*>(src), bytes);
    return hipGetLastError();
}

hipError_t launch_downsample(const uint8_t* src, int w, int h, int bpp, uint8_t* dst, int ow, int oh,
                             hipStream_t stream) {
    const int64_t total = int64_t(ow) * oh * bpp;
    (void)hipGetLastError();  // report this launch's errors only, not an earlier call's
    hipLaunchKernelGGL(downsample_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, stream, src, w, h, bpp,
                       dst, ow, oh);
    return hipGetLastError();
}

template <int DIM, bool ALPHA, int CH, int METHOD = METHOD_PCA>
static void launch_plain_ch(const PlainArgs& a, int64_t nblocks, const PlainTuning& t, hipStream_t stream) {
    const bool srgb = a.f.flags & FLAG_SRGB;
    if (DIM == 6 && METHOD == METHOD_PCA && t.lanes_per_block == 2) {
        const int chunk = (kPlainThreads / 2) * t.blocks_per_lane;
        const dim3 grid(unsigned((nblocks + chunk - 1) / chunk));
        if (srgb)
            hipLaunchKernelGGL((encode_plain_pair_kernel<ALPHA, true, CH>), grid, dim3(kPlainThreads), 0, stream, a,
                               t.blocks_per_lane, t.prefetch);
        else
            hipLaunchKernelGGL((encode_plain_pair_kernel<ALPHA, false, CH>), grid, dim3(kPlainThreads), 0, stream, a,
                               t.blocks_per_lane, t.prefetch);
        return;
    }
    const int chunk = kPlainThreads * t.blocks_per_lane;
    const dim3 grid(unsigned((nblocks + chunk - 1) / chunk));
    if (METHOD == METHOD_PCA && t.packed_texels) {
        if (srgb)
            hipLaunchKernelGGL((encode_plain_packed_kernel<DIM, ALPHA, true, CH>), grid, dim3(kPlainThreads), 0, stream,
                               a, t.blocks_per_lane);
        else
            hipLaunchKernelGGL((encode_plain_packed_kernel<DIM, ALPHA, false, CH>), grid, dim3(kPlainThreads), 0,
                               stream, a, t.blocks_per_lane);
        return;
    }
    constexpr bool kFlatCapable = CH == CH_ALL && METHOD == METHOD_PCA;
    if (kFlatCapable && t.flat_alpha) {
        if (srgb)
            hipLaunchKernelGGL((encode_plain_kernel<DIM, ALPHA, true, CH, METHOD, kFlatCapable>), grid,
                               dim3(kPlainThreads), 0, stream, a, t.blocks_per_lane, t.prefetch);
        else
            hipLaunchKernelGGL((encode_plain_kernel<DIM, ALPHA, false, CH, METHOD, kFlatCapable>), grid,
                               dim3(kPlainThreads), 0, stream, a, t.blocks_per_lane, t.prefetch);
        return;
    }
    if (srgb)
        hipLaunchKernelGGL((encode_plain_kernel<DIM, ALPHA, true, CH, METHOD>), grid, dim3(kPlainThreads), 0, stream, a,
                           t.blocks_per_lane, t.prefetch);
    else
        hipLaunchKernelGGL((encode_plain_kernel<DIM, ALPHA, false, CH, METHOD>), grid, dim3(kPlainThreads), 0, stream,
                           a, t.blocks_per_lane, t.prefetch);
}

template <int DIM, bool ALPHA>
static void launch_plain_t(const PlainArgs& a, int64_t nblocks, const PlainTuning& t, hipStream_t stream) {
    if (a.f.flags & FLAG_MAX_ACCUMULATION) {  // optional search: four-channel, one-lane encoder only
        PlainTuning one = t;
        one.lanes_per_block = 1;
        launch_plain_ch<DIM, ALPHA, CH_ALL, METHOD_MAX_ACCUMULATION>(a, nblocks, one, stream);
        return;
    }
    switch (channel_class(a.f.fmt, a.f.flags, a.f.width, a.f.height, DIM)) {
    case CH_NORMAL: launch_plain_ch<DIM, ALPHA, CH_NORMAL>(a, nblocks, t, stream); break;
    case CH_RGB8: launch_plain_ch<DIM, ALPHA, CH_RGB8>(a, nblocks, t, stream); break;
    default: launch_plain_ch<DIM, ALPHA, CH_ALL>(a, nblocks, t, stream); break;
    }
}

// Blocks per lane for a plain launch: the fewest (1, 2, 4 or 8) that let every
// workgroup be resident at once (one-lane 6x6: two workgroups of four waves per
// CU; two-lane or packed-texel 6x6: three; 4x4: four, packed five), so that no
// second round of workgroup starts (table staging, first fetch) trails the first.
int plain_blocks_per_lane(int64_t nblocks, int dim, int lanes_per_block, int num_cus, bool packed) {
    // 4x4 one-lane encodes: one block per lane, whatever the count; the
    // workgroups that do not fit at first start as others finish, which
    // balances the end of the launch (4096x4096 RGBA8: 21.4 us against 23.0 us
    // at the resident-workgroup rule below)
    if (dim == 4 && !packed) return 1;
    const int per_cu = dim == 4 ? (packed ? PackedWaves<4>::value : PlainWaves<4>::value)
                                : (lanes_per_block == 2 ? 3 : packed ? PackedWaves<6>::value : PlainWaves<6>::value);
    const int64_t slots = int64_t(num_cus) * per_cu;
    const int64_t per_step = kPlainThreads / lanes_per_block;
    int bpl = 1;
    while (bpl < 8 && (nblocks + per_step * bpl - 1) / (per_step * bpl) > slots) bpl *= 2;
    return bpl;
}

// Default plain tuning: one lane per block; the next block's loads issued
// before the current block is encoded except for tiled RGB8 sources (measured
// with the issue priorities: 1-4% faster on RGBA8 and normal maps, 4% slower on
// 4096x4096 RGB8 6x6); the flat-alpha test for four-channel principal component
// encodes (with prefetch: 3% faster on the 4096x4096 noise RGBA8 6x6 texture,
// 5-9% on opaque textures, within 1% slower on 4x4 noise).
PlainTuning default_plain_tuning(const astc_image_desc& d, int64_t nblocks, int num_cus) {
    PlainTuning t;
    t.lanes_per_block = 1;
    const int ch = channel_class(d.src_format, d.flags, d.width, d.height, d.block_dim);
    t.prefetch = ch == CH_RGB8 ? 0 : 1;
    t.flat_alpha = (ch == CH_ALL && !(d.flags & FLAG_MAX_ACCUMULATION)) ? 1 : 0;
    t.packed_texels = 0;
    t.blocks_per_lane = plain_blocks_per_lane(nblocks, d.block_dim, t.lanes_per_block, num_cus);
    return t;
}

hipError_t launch_encode_plain(const DeviceTables* d_tables, const astc_image_desc& d, const ImageGeom& g,
                               const void* src, void* dst, hipStream_t stream, int num_cus, const PlainTuning* tuning) {
    PlainArgs a;
    a.f.src = static_cast<const uint8_t*>(src);
    a.f.width = d.width;
    a.f.height = d.height;
    a.f.row_pitch = d.row_pitch;
    a.f.fmt = d.src_format;
    a.f.flags = d.flags;
    a.f.aligned = ((reinterpret_cast<uintptr_t>(src) & 3u) == 0) && ((d.row_pitch & 3) == 0);
    a.f.or_mask = normal_or_mask(d.flags);
    set_block_row(&a.f, g.xblocks);
    a.nblocks = int(g.nblocks);
    a.dst = static_cast<uint32_t*>(dst);
    a.tables = d_tables;
    PlainTuning t = tuning ? *tuning : default_plain_tuning(d, g.nblocks, num_cus);
    (void)hipGetLastError();  // report this launch's errors only, not an earlier call's
    if (d.block_dim != 6) t.lanes_per_block = 1;
    if (t.blocks_per_lane == 0)
        t.blocks_per_lane = plain_blocks_per_lane(g.nblocks, d.block_dim, t.lanes_per_block, num_cus,
                                                  t.packed_texels && !(d.flags & FLAG_MAX_ACCUMULATION));
    const bool alpha = d.flags & FLAG_ALPHA;
    if (d.block_dim == 4) {
        if (alpha) launch_plain_t<4, true>(a, g.nblocks, t, stream);
        else launch_plain_t<4, false>(a, g.nblocks, t, stream);
    } else {
        if (alpha) launch_plain_t<6, true>(a, g.nblocks, t, stream);
        else launch_plain_t<6, false>(a, g.nblocks, t, stream);
    }
    return hipGetLastError();
}

}  // namespace astc
