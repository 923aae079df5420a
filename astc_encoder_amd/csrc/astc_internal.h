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

// Plain per-texture launch (one lane per block).
hipError_t launch_encode_plain(const DeviceTables* d_tables, const astc_image_desc& d, const ImageGeom& g,
                               const void* src, void* dst, hipStream_t stream);

}  // namespace astc
