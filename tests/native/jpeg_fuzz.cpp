// Decodes every file named on the command line with axr::decode_jpeg
// (axiomr_amd/csrc/jpeg.cpp) and prints "ok WxH" or "err <message>" per
// file.  Built with AddressSanitizer and UBSan by tests/test_jpeg.py to run
// truncated and corrupted JPEGs through the decoder.
#include <cstdio>
#include <string>
#include <vector>

#include "jpeg.hpp"

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        FILE* f = std::fopen(argv[i], "rb");
        if (!f) { std::printf("missing %s\n", argv[i]); return 2; }
        std::vector<uint8_t> d;
        int c;
        while ((c = std::fgetc(f)) != EOF) d.push_back((uint8_t)c);
        std::fclose(f);
        int w = 0, h = 0;
        std::vector<uint8_t> rgba;
        std::string err;
        if (axr::decode_jpeg(d.data(), d.size(), &w, &h, rgba, err)) {
            if (rgba.size() != (size_t)w * h * 4) { std::printf("size mismatch\n"); return 3; }
            std::printf("ok %dx%d\n", w, h);
        } else {
            std::printf("err %s\n", err.c_str());
        }
    }
    return 0;
}
