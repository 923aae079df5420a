// astc_kernels.hip -- gfx950 kernels of the ASTC encoder.
//
// encode_plain_kernel: the per-texture launch that replaces MainCS
// (ASTC_Encode.hlsl:553-582).  One lane encodes one block: with ~1-2.5k
// floating-point operations per block and hundreds of thousands of blocks per
// texture, keeping each block in one lane avoids cross-lane reductions and
// replicated serial work (the 16-step power iteration), so every lane issues
// useful VALU work.  A workgroup of 256 lanes covers 256 consecutive blocks in
// row-major block order, so the lanes of a wave read adjacent block rows of the
// texture and write 16-byte blocks to consecutive addresses.
#include "astc_internal.h"

namespace astc {

struct PlainArgs {
    FetchArgs f;
    int xblocks;
    int nblocks;
    uint32_t* dst;
    const DeviceTables* tables;
};

constexpr int kPlainThreads = 256;

__device__ __forceinline__ void stage_tables(const DeviceTables* t, float* lut, uint8_t* trit) {
    const int tid = threadIdx.x;
    for (int i = tid; i < 256; i += blockDim.x) {
        lut[i] = t->srgb[i];
        trit[i] = t->trit[i];
    }
}

template <int DIM, bool ALPHA>
__global__ void __launch_bounds__(kPlainThreads) encode_plain_kernel(PlainArgs a) {
    __shared__ float lut[256];
    __shared__ uint8_t trit[256];
    stage_tables(a.tables, lut, trit);
    __syncthreads();
    const int id = blockIdx.x * kPlainThreads + threadIdx.x;
    if (id >= a.nblocks) return;
    const int by = id / a.xblocks;
    const int bx = id - by * a.xblocks;
    float x[DIM * DIM][4];
    fetch_block<DIM>(a.f, lut, bx, by, x);
    const Block128 blk = encode_block<DIM, ALPHA>(x, trit);
    reinterpret_cast<uint4*>(a.dst)[id] = make_uint4(blk.w[0], blk.w[1], blk.w[2], blk.w[3]);
}

hipError_t launch_encode_plain(const DeviceTables* d_tables, const astc_image_desc& d, const ImageGeom& g,
                               const void* src, void* dst, hipStream_t stream) {
    PlainArgs a;
    a.f.src = static_cast<const uint8_t*>(src);
    a.f.width = d.width;
    a.f.height = d.height;
    a.f.row_pitch = d.row_pitch;
    a.f.fmt = d.src_format;
    a.f.flags = d.flags;
    a.f.aligned = ((reinterpret_cast<uintptr_t>(src) & 3u) == 0) && ((d.row_pitch & 3) == 0);
    a.xblocks = g.xblocks;
    a.nblocks = int(g.nblocks);
    a.dst = static_cast<uint32_t*>(dst);
    a.tables = d_tables;
    const dim3 grid(unsigned((g.nblocks + kPlainThreads - 1) / kPlainThreads)), block(kPlainThreads);
    const bool alpha = d.flags & FLAG_ALPHA;
    if (d.block_dim == 4) {
        if (alpha) hipLaunchKernelGGL((encode_plain_kernel<4, true>), grid, block, 0, stream, a);
        else hipLaunchKernelGGL((encode_plain_kernel<4, false>), grid, block, 0, stream, a);
    } else {
        if (alpha) hipLaunchKernelGGL((encode_plain_kernel<6, true>), grid, block, 0, stream, a);
        else hipLaunchKernelGGL((encode_plain_kernel<6, false>), grid, block, 0, stream, a);
    }
    return hipGetLastError();
}

}  // namespace astc
