// astc_cs_enc -- command-line drop-in for the reference tool (main.cpp:140-258).
//
//   astc_cs_enc <input.png> [-4x4] [-6x6] [-alpha] [-norm] [-srgb]
//
// Same behaviour as the reference, on the MI355X encoder:
//   * options are booleans set by their presence from argv[2] on; unknown
//     arguments are ignored (main.cpp:140-178); fewer than 2 arguments prints
//     "wrong args count" and exits with -1 (main.cpp:182-185);
//   * the image is loaded as 8-bit RGBA and flipped vertically (main.cpp:24-25);
//     it is an sRGB texture only with -srgb and without -norm (main.cpp:214);
//   * the output is the input path with everything from its last '.' removed,
//     plus ".astc" (strrchr over the whole path, main.cpp:130-138, 244-246).
// One deliberate difference: -6x6 selects 6x6 blocks.  In the reference it sets
// a field nothing reads, so its output is always 4x4 (astc_encode.h:21,59,124).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/astc_mi355x.h"

namespace {

struct Options {
    bool is4x4 = true, is6x6 = false, is_normal_map = false, has_alpha = false, srgb = false;
};

void parse(int argc, char** argv, Options& o) {
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-4x4") o.is4x4 = true;
        else if (a == "-6x6") o.is6x6 = true;
        else if (a == "-norm") o.is_normal_map = true;
        else if (a == "-srgb") o.srgb = true;
        else if (a == "-alpha") o.has_alpha = true;
    }
}

std::string output_path(const std::string& in) {
    std::string out = in;
    const size_t dot = out.rfind('.');
    if (dot != std::string::npos) out.resize(dot);
    return out + ".astc";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cout << "wrong args count" << std::endl;
        return -1;
    }
    Options opt;
    parse(argc, argv, opt);
    std::cout << "encode option setting:\n"
              << "has_alpha\t" << std::boolalpha << opt.has_alpha << std::endl
              << "is 4x4 block\t" << !opt.is6x6 << std::endl
              << "normal map\t" << opt.is_normal_map << std::endl
              << "encode in gamma color space\t" << opt.srgb << std::endl;

    const std::string src_path = argv[1];
    int w = 0, h = 0;
    uint8_t* rgba = nullptr;
    if (astc_load_image(src_path.c_str(), 1, &w, &h, &rgba) != ASTC_OK) {
        std::printf("Failed to load image %s\nReason: %s\n", src_path.c_str(), astc_last_error());
        std::cout << "load source texture failed! [" << src_path << "]" << std::endl;
        return -1;
    }
    const int dim = opt.is6x6 ? 6 : 4;
    astc_image_desc desc;
    desc.width = w;
    desc.height = h;
    desc.row_pitch = w * 4;
    desc.src_format = ASTC_FMT_RGBA8;
    desc.block_dim = dim;
    desc.flags = (opt.has_alpha ? ASTC_FLAG_ALPHA : 0u) | (opt.is_normal_map ? ASTC_FLAG_NORMAL : 0u) |
                 ((opt.srgb && !opt.is_normal_map) ? ASTC_FLAG_SRGB : 0u);
    const int64_t out_bytes = astc_output_size(&desc);

    astc_ctx* ctx = nullptr;
    void* src_dev = nullptr;
    void* dst_dev = nullptr;
    std::vector<uint8_t> blocks;
    bool ok = out_bytes > 0 && astc_ctx_create(0, &ctx) == ASTC_OK;
    if (ok) ok = hipMalloc(&src_dev, size_t(w) * h * 4) == hipSuccess && hipMalloc(&dst_dev, out_bytes) == hipSuccess;
    if (ok) ok = hipMemcpy(src_dev, rgba, size_t(w) * h * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = astc_encode(ctx, &desc, src_dev, dst_dev, nullptr) == ASTC_OK;
    if (ok) {
        blocks.resize(out_bytes);
        ok = hipMemcpy(blocks.data(), dst_dev, out_bytes, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (!ok) std::cout << "encode astc failed! " << astc_last_error() << std::endl;
    if (src_dev) (void)hipFree(src_dev);
    if (dst_dev) (void)hipFree(dst_dev);
    astc_ctx_destroy(ctx);
    astc_free_image(rgba);
    if (!ok) return -1;

    const std::string dst_path = output_path(src_path);
    if (astc_write_file(dst_path.c_str(), dim, dim, w, h, blocks.data(), out_bytes) != ASTC_OK) {
        std::cout << "save astc failed! " << astc_last_error() << std::endl;
        return -1;
    }
    std::cout << "save astc to:" << dst_path << std::endl;
    return 0;
}
