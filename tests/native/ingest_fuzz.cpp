// Loads every file named on the command line through the C ABI's file
// loaders (axr_image_load for .png/.jpg/.ppm, axr_obj_load for .obj) and
// prints "ok" with the sizes or "err" per file.  Built with AddressSanitizer
// and UBSan by tests/test_ingest_fuzz.py to run malformed inputs through
// axiomr_amd/csrc/image.cpp, jpeg.cpp and ingest.cpp.
#include <cstdio>
#include <cstring>
#include <vector>

#include "axiomr.h"

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        const char* p = argv[i];
        const size_t n = std::strlen(p);
        if (n > 4 && std::strcmp(p + n - 4, ".obj") == 0) {
            axr_obj* o = nullptr;
            if (axr_obj_load(p, &o) == AXR_OK) {
                uint64_t nv = 0, nf = 0;
                uint32_t ng = 0, nm = 0;
                axr_obj_counts(o, &nv, &nf, &ng, &nm);
                std::vector<float> v(nv * 14 + 1);
                std::vector<uint32_t> idx(nf * 3 + 1);
                axr_obj_arrays(o, v.data(), idx.data());
                for (uint64_t k = 0; k < nf * 3; ++k)
                    if (idx[k] >= nv) { std::printf("index out of range\n"); return 3; }
                axr_obj_destroy(o);
                std::printf("ok obj %llu %llu\n", (unsigned long long)nv, (unsigned long long)nf);
            } else {
                std::printf("err %s\n", axr_obj_last_error());
            }
        } else {
            int w = 0, h = 0;
            uint8_t* px = nullptr;
            if (axr_image_load(p, &w, &h, &px) == AXR_OK) {
                unsigned sum = 0;   // touch every byte
                for (size_t k = 0; k < (size_t)w * h * 4; ++k) sum += px[k];
                axr_image_free(px);
                std::printf("ok image %dx%d %u\n", w, h, sum);
            } else {
                std::printf("err %s\n", axr_image_last_error());
            }
        }
    }
    return 0;
}
