// jpeg.hpp - JPEG decoding (jpeg.cpp) for the texture loader.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace axr {

// Decode a whole JPEG file in memory to RGBA8 (top row first, alpha 255).
// False with a message in err for a malformed or unsupported file.
bool decode_jpeg(const uint8_t* data, size_t n, int* w, int* h, std::vector<uint8_t>& rgba, std::string& err);

}  // namespace axr
