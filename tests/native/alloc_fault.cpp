// Host-only check that no C++ exception crosses the C ABI when host
// allocations fail.  This program replaces the global operator new (which the
// library's calls resolve to as well) with one that throws std::bad_alloc while
// `g_armed` is set, then calls entry points whose error or body paths allocate.
// Prints one "name rc message" line per call; never launches a kernel.
#include <cstdio>
#include <cstdlib>
#include <new>

#include "astc_mi355x.h"

static bool g_armed = false;

void* operator new(std::size_t n) {
    if (g_armed) throw std::bad_alloc();
    if (void* p = std::malloc(n ? n : 1)) return p;
    throw std::bad_alloc();
}
void* operator new[](std::size_t n) { return operator new(n); }
void operator delete(void* p) noexcept { std::free(p); }
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete(void* p, std::size_t) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

static void report(const char* name, long long rc) { std::printf("%s %lld %s\n", name, rc, astc_last_error()); }

int main() {
    astc_image_desc bad = {0, 16, 64, ASTC_FMT_RGBA8, 4, 0};        // width 0: rejected with a long message
    astc_image_desc good = {64, 32, 256, ASTC_FMT_RGBA8, 4, 0};
    astc_ctx* fake = reinterpret_cast<astc_ctx*>(16);               // never dereferenced before the checks
    g_armed = true;
    report("encode_bad_desc", astc_encode(fake, &bad, (const void*)16, (void*)16, nullptr));
    report("decode_bad_size", astc_decode(fake, 0, 16, 4, 4, (const void*)16, (void*)16, nullptr));
    report("output_size_bad_desc", astc_output_size(&bad));
    report("request_bytes", astc_request_bytes(&good, 1));           // body allocates: caught inside
    report("write_file_bad_path", astc_write_file("/nonexistent-dir/x.astc", 4, 4, 4, 4, "0123456789abcdef", 16));
    g_armed = false;
    report("request_bytes_unarmed", astc_request_bytes(&good, 1));
    return 0;
}
