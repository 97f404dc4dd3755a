// image.cpp - texture decoding and framebuffer image output, host C++ over
// zlib.
//
// Decoding gives what the reference's Texture gets from
// stbi_load(path, ..., STBI_rgb_alpha) (src/texture.cpp:21-35): RGBA8, top
// row first, alpha 255 where the file has none.  Formats: PNG (colour types
// gray, RGB, palette, gray+alpha, RGBA; bit depths 1-16, 16-bit samples
// reduced to their high byte; tRNS; Adam7 interlacing), JPEG (jpeg.cpp) and
// binary PPM/PGM (P6/P5, maxval 255).  The other stb_image formats are not
// decoded (AXR_ERR_INVALID).
//
// Output: PNG, RGBA8, no filter, zlib level 6.  A framebuffer (BGRA, row 0 at
// the bottom) is written top row first, as the reference presents it
// (src/window.cpp:70-84 hands the bottom-up BGRA buffer to the display).
#include <zlib.h>

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/axiomr.h"
#include "jpeg.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}

bool read_file(const char* path, std::vector<uint8_t>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) { std::fclose(f); return false; }
    out.resize((size_t)n);
    const size_t got = n ? std::fread(out.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Undo the per-row filters of one (sub)image in place; rows are
// 1 + rowbytes long in `data`, the output rows are rowbytes long.
bool unfilter(const uint8_t* data, size_t rows, size_t rowbytes, size_t bpp, std::vector<uint8_t>& out) {
    out.assign(rows * rowbytes, 0);
    for (size_t y = 0; y < rows; ++y) {
        const uint8_t ft = data[y * (rowbytes + 1)];
        const uint8_t* in = data + y * (rowbytes + 1) + 1;
        uint8_t* cur = out.data() + y * rowbytes;
        const uint8_t* prev = y ? cur - rowbytes : nullptr;
        for (size_t i = 0; i < rowbytes; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int v = in[i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: return false;
            }
            cur[i] = (uint8_t)v;
        }
    }
    return true;
}

struct PngInfo {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> plte;       // 3 per entry
    std::vector<uint8_t> plte_alpha; // tRNS for palettes
    bool has_key = false;
    uint16_t key[3] = {0, 0, 0};     // tRNS colour key (gray or RGB), in file bit depth
};

int channels_of(int ctype) {
    switch (ctype) {
    case 0: return 1;
    case 2: return 3;
    case 3: return 1;
    case 4: return 2;
    case 6: return 4;
    default: return 0;
    }
}

// Sample s (0-based) of a row at the file's bit depth, as a 16-bit value for
// depth 16 and as the raw value otherwise.
uint32_t sample_at(const uint8_t* row, size_t s, int depth) {
    switch (depth) {
    case 16: return (uint32_t)row[2 * s] << 8 | row[2 * s + 1];
    case 8: return row[s];
    default: {
        const size_t bit = s * (size_t)depth;
        const int shift = 8 - depth - (int)(bit & 7);
        return (row[bit >> 3] >> shift) & ((1u << depth) - 1u);
    }
    }
}

// Gray value scaled to 8 bits the way stb_image expands low bit depths.
uint8_t gray8(uint32_t v, int depth) {
    switch (depth) {
    case 1: return (uint8_t)(v * 0xFF);
    case 2: return (uint8_t)(v * 0x55);
    case 4: return (uint8_t)(v * 0x11);
    case 16: return (uint8_t)(v >> 8);
    default: return (uint8_t)v;
    }
}

// Write the pixels of one unfiltered (sub)image row into the RGBA output.
void emit_row(const PngInfo& P, const uint8_t* row, uint32_t n, uint8_t* rgba, size_t x0, size_t dx, size_t stride_px) {
    const int ch = channels_of(P.ctype);
    for (uint32_t i = 0; i < n; ++i) {
        uint8_t* o = rgba + 4 * (x0 + (size_t)i * dx);
        (void)stride_px;
        uint32_t s[4] = {0, 0, 0, 0};
        for (int c = 0; c < ch; ++c) s[c] = sample_at(row, (size_t)i * ch + c, P.depth);
        switch (P.ctype) {
        case 0: {
            const uint8_t g = gray8(s[0], P.depth);
            o[0] = o[1] = o[2] = g;
            o[3] = (P.has_key && s[0] == P.key[0]) ? 0 : 255;
            break;
        }
        case 2:
            for (int c = 0; c < 3; ++c) o[c] = P.depth == 16 ? (uint8_t)(s[c] >> 8) : (uint8_t)s[c];
            o[3] = (P.has_key && s[0] == P.key[0] && s[1] == P.key[1] && s[2] == P.key[2]) ? 0 : 255;
            break;
        case 3: {
            const uint32_t k = s[0];
            if (3 * (size_t)k + 2 < P.plte.size()) {
                o[0] = P.plte[3 * k];
                o[1] = P.plte[3 * k + 1];
                o[2] = P.plte[3 * k + 2];
            } else {
                o[0] = o[1] = o[2] = 0;
            }
            o[3] = k < P.plte_alpha.size() ? P.plte_alpha[k] : 255;
            break;
        }
        case 4: {
            const uint8_t g = gray8(s[0], P.depth);
            o[0] = o[1] = o[2] = g;
            o[3] = gray8(s[1], P.depth);
            break;
        }
        default:
            for (int c = 0; c < 4; ++c) o[c] = P.depth == 16 ? (uint8_t)(s[c] >> 8) : (uint8_t)s[c];
            break;
        }
    }
}

int decode_png(const std::vector<uint8_t>& f, int* w, int* h, std::vector<uint8_t>& rgba) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return fail(AXR_ERR_INVALID, "not a PNG file");
    PngInfo P;
    std::vector<uint8_t> idat;
    size_t pos = 8;
    bool ihdr = false, iend = false;
    while (pos + 12 <= f.size() && !iend) {
        const uint32_t len = be32(&f[pos]);
        if (len > f.size() - pos - 12) return fail(AXR_ERR_INVALID, "PNG chunk runs past the end of the file");
        const char* type = (const char*)&f[pos + 4];
        const uint8_t* d = &f[pos + 8];
        const uint32_t crc = be32(&f[pos + 8 + len]);
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), (const Bytef*)type, 4 + len) != crc)
            return fail(AXR_ERR_INVALID, "PNG chunk CRC mismatch");
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) return fail(AXR_ERR_INVALID, "bad IHDR");
            P.w = be32(d);
            P.h = be32(d + 4);
            P.depth = d[8];
            P.ctype = d[9];
            P.interlace = d[12];
            if (d[10] != 0 || d[11] != 0 || P.interlace > 1) return fail(AXR_ERR_INVALID, "unsupported PNG compression, filter or interlace method");
            const int ch = channels_of(P.ctype);
            const bool ok_depth = (P.ctype == 0 && (P.depth == 1 || P.depth == 2 || P.depth == 4 || P.depth == 8 || P.depth == 16)) ||
                                  (P.ctype == 3 && (P.depth == 1 || P.depth == 2 || P.depth == 4 || P.depth == 8)) ||
                                  ((P.ctype == 2 || P.ctype == 4 || P.ctype == 6) && (P.depth == 8 || P.depth == 16));
            if (!ch || !ok_depth) return fail(AXR_ERR_INVALID, "unsupported PNG colour type / bit depth");
            if (P.w == 0 || P.h == 0 || (uint64_t)P.w * P.h > (1ull << 28)) return fail(AXR_ERR_INVALID, "PNG size out of range");
            ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            P.plte.assign(d, d + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            if (P.ctype == 3) {
                P.plte_alpha.assign(d, d + len);
            } else if (P.ctype == 0 && len >= 2) {
                P.has_key = true;
                P.key[0] = (uint16_t)(d[0] << 8 | d[1]);
            } else if (P.ctype == 2 && len >= 6) {
                P.has_key = true;
                for (int c = 0; c < 3; ++c) P.key[c] = (uint16_t)(d[2 * c] << 8 | d[2 * c + 1]);
            }
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            iend = true;
        }
        pos += 12 + len;
    }
    if (!ihdr || idat.empty()) return fail(AXR_ERR_INVALID, "PNG without IHDR or IDAT");
    if (P.ctype == 3 && P.plte.empty()) return fail(AXR_ERR_INVALID, "palette PNG without PLTE");
    const int ch = channels_of(P.ctype);
    const size_t bits_pp = (size_t)ch * P.depth;
    const size_t bpp = (bits_pp + 7) / 8;
    // Adam7 passes: x0, y0, dx, dy (one pass of step 1 when not interlaced)
    static const int A7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int npass = P.interlace ? 7 : 1;
    size_t need = 0;
    uint32_t pw[7], ph[7];
    for (int k = 0; k < npass; ++k) {
        const int x0 = P.interlace ? A7[k][0] : 0, y0 = P.interlace ? A7[k][1] : 0;
        const int dx = P.interlace ? A7[k][2] : 1, dy = P.interlace ? A7[k][3] : 1;
        pw[k] = P.w > (uint32_t)x0 ? (P.w - x0 + dx - 1) / dx : 0;
        ph[k] = P.h > (uint32_t)y0 ? (P.h - y0 + dy - 1) / dy : 0;
        if (pw[k] && ph[k]) need += (size_t)ph[k] * (1 + (pw[k] * bits_pp + 7) / 8);
    }
    std::vector<uint8_t> raw(need);
    uLongf got = (uLongf)need;
    if (uncompress(raw.data(), &got, idat.data(), (uLong)idat.size()) != Z_OK || got != need)
        return fail(AXR_ERR_INVALID, "PNG image data does not inflate to the expected size");
    rgba.assign((size_t)P.w * P.h * 4, 0);
    size_t off = 0;
    std::vector<uint8_t> rows;
    for (int k = 0; k < npass; ++k) {
        if (!pw[k] || !ph[k]) continue;
        const size_t rowbytes = (pw[k] * bits_pp + 7) / 8;
        if (!unfilter(raw.data() + off, ph[k], rowbytes, bpp, rows)) return fail(AXR_ERR_INVALID, "bad PNG filter type");
        off += (size_t)ph[k] * (rowbytes + 1);
        const int x0 = P.interlace ? A7[k][0] : 0, y0 = P.interlace ? A7[k][1] : 0;
        const int dx = P.interlace ? A7[k][2] : 1, dy = P.interlace ? A7[k][3] : 1;
        for (uint32_t y = 0; y < ph[k]; ++y) {
            const size_t oy = (size_t)y0 + (size_t)y * dy;
            emit_row(P, rows.data() + y * rowbytes, pw[k], rgba.data() + oy * P.w * 4, (size_t)x0, (size_t)dx, 0);
        }
    }
    *w = (int)P.w;
    *h = (int)P.h;
    return AXR_OK;
}

// Binary PPM (P6) / PGM (P5) with maxval 255.
int decode_pnm(const std::vector<uint8_t>& f, int* w, int* h, std::vector<uint8_t>& rgba) {
    size_t p = 2;
    auto skip = [&]() {
        for (;;) {
            while (p < f.size() && std::isspace(f[p])) ++p;
            if (p < f.size() && f[p] == '#') {
                while (p < f.size() && f[p] != '\n') ++p;
            } else {
                break;
            }
        }
    };
    auto num = [&](long& v) {
        skip();
        v = 0;
        size_t s = p;
        while (p < f.size() && std::isdigit(f[p])) v = v * 10 + (f[p++] - '0');
        return p > s && v < (1l << 20);
    };
    long W, H, M;
    if (!num(W) || !num(H) || !num(M) || M != 255 || W <= 0 || H <= 0) return fail(AXR_ERR_INVALID, "unsupported PPM/PGM header");
    ++p;   // the single whitespace byte before the raster
    const int ch = f[1] == '6' ? 3 : 1;
    if (f.size() < p + (size_t)W * H * ch) return fail(AXR_ERR_INVALID, "PPM/PGM raster truncated");
    rgba.resize((size_t)W * H * 4);
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const uint8_t* s = &f[p + i * ch];
        uint8_t* o = &rgba[4 * i];
        o[0] = s[0];
        o[1] = ch == 3 ? s[1] : s[0];
        o[2] = ch == 3 ? s[2] : s[0];
        o[3] = 255;
    }
    *w = (int)W;
    *h = (int)H;
    return AXR_OK;
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* d, size_t n) {
    put32(out, (uint32_t)n);
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    if (n) out.insert(out.end(), d, d + n);
    put32(out, (uint32_t)crc32(crc32(0L, Z_NULL, 0), out.data() + at, (uInt)(4 + n)));
}

}  // namespace

extern "C" {

const char* axr_image_last_error(void) { return g_err.c_str(); }

int axr_image_load(const char* path, int* w, int* h, uint8_t** rgba) {
    if (!path || !w || !h || !rgba) return fail(AXR_ERR_INVALID, "null argument");
    *rgba = nullptr;
    std::vector<uint8_t> f;
    if (!read_file(path, f)) return fail(AXR_ERR_INVALID, std::string("cannot read ") + path);
    std::vector<uint8_t> px;
    int r;
    if (f.size() >= 8 && f[0] == 137 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G') r = decode_png(f, w, h, px);
    else if (f.size() >= 2 && f[0] == 'P' && (f[1] == '6' || f[1] == '5')) r = decode_pnm(f, w, h, px);
    else if (f.size() >= 3 && f[0] == 0xFF && f[1] == 0xD8 && f[2] == 0xFF) {
        std::string e;
        r = axr::decode_jpeg(f.data(), f.size(), w, h, px, e) ? AXR_OK : fail(AXR_ERR_INVALID, "JPEG: " + e);
    }
    else r = fail(AXR_ERR_INVALID, std::string("unsupported image format: ") + path);
    if (r) return r;
    *rgba = (uint8_t*)std::malloc(px.size());
    if (!*rgba) return fail(AXR_ERR_OOM, "out of memory");
    std::memcpy(*rgba, px.data(), px.size());
    return AXR_OK;
}

void axr_image_free(uint8_t* rgba) { std::free(rgba); }

int axr_image_write_png(const char* path, int w, int h, const uint8_t* rgba, int flip_rows) {
    if (!path || !rgba || w <= 0 || h <= 0) return fail(AXR_ERR_INVALID, "bad image");
    std::vector<uint8_t> raw((size_t)h * (1 + (size_t)w * 4));
    for (int y = 0; y < h; ++y) {
        const int sy = flip_rows ? h - 1 - y : y;
        uint8_t* row = raw.data() + (size_t)y * (1 + (size_t)w * 4);
        row[0] = 0;
        std::memcpy(row + 1, rgba + (size_t)sy * w * 4, (size_t)w * 4);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) return fail(AXR_ERR_OOM, "deflate failed");
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    uint8_t ihdr[13];
    const uint32_t W = (uint32_t)w, H = (uint32_t)h;
    for (int k = 0; k < 4; ++k) {
        ihdr[k] = (uint8_t)(W >> (24 - 8 * k));
        ihdr[4 + k] = (uint8_t)(H >> (24 - 8 * k));
    }
    ihdr[8] = 8;    // bit depth
    ihdr[9] = 6;    // RGBA
    ihdr[10] = ihdr[11] = ihdr[12] = 0;
    chunk(out, "IHDR", ihdr, 13);
    chunk(out, "IDAT", z.data(), zn);
    chunk(out, "IEND", nullptr, 0);
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(AXR_ERR_INVALID, std::string("cannot write ") + path);
    const size_t n = std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    return n == out.size() ? AXR_OK : fail(AXR_ERR_INVALID, std::string("short write to ") + path);
}

}  // extern "C"
