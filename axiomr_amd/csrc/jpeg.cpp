// jpeg.cpp - JPEG decoding to RGBA8 as the reference's Texture receives it
// from stbi_load(path, ..., STBI_rgb_alpha) (src/texture.cpp:21-35; the
// reference vendors stb_image v2.25, external/stb_image/stb_image.h).
//
// The steps follow that decoder so that the pixels are the same bytes: 8-bit
// baseline, extended and progressive frames with 1, 3 or 4 components and
// sampling factors 1..4; Huffman decoding with restart intervals (an
// interval that does not end in RSTn ends the scan, as there);
// dequantization into 16-bit coefficients (products wrap to 16 bits);
// the islow integer IDCT in the form the reference's x86-64 build runs (its
// SSE2 kernel: 16-bit sums of coefficient pairs, each pass's result
// saturated to 16 bits, constants rounded to 12-bit fixed point); the
// reference's chroma upsampling (triangle filters for 2x, nearest
// otherwise, centred on the sample rows); YCbCr->RGB in its fixed-point
// form; JFIF/Adobe colour-transform detection, RGB components, Adobe CMYK
// and YCCK through the same rounded 8x8-bit product.  Alpha is 255.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "jpeg.hpp"

namespace {

struct Reader {   // marker segments
    const uint8_t* p;
    const uint8_t* end;
    int get8() { return p < end ? *p++ : 0; }
    int get16() { int a = get8(); return (a << 8) | get8(); }
    void skip(int n) { p = (end - p < n) ? end : p + n; }
    bool eof() const { return p >= end; }
};

// Canonical Huffman table (JPEG Annex C), decoded through a 9-bit lookahead.
struct Huffman {
    uint8_t nsym = 0;
    uint8_t values[256] = {};
    uint8_t size[257] = {};
    uint16_t code[256] = {};
    uint32_t maxcode[18] = {};   // largest code + 1 of each length, left-aligned to 16 bits
    int delta[17] = {};          // symbol index = code + delta[length]
    uint8_t fast[512] = {};      // 255: code longer than 9 bits
    bool ok = false;
};

bool build_huffman(Huffman& h, const int* counts) {
    int k = 0;
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < counts[i]; ++j) {
            if (k >= 256) return false;
            h.size[k++] = (uint8_t)(i + 1);
        }
    h.size[k] = 0;
    h.nsym = (uint8_t)k;
    uint32_t code = 0;
    k = 0;
    for (int len = 1; len <= 16; ++len) {
        h.delta[len] = k - (int)code;
        const int first = k;
        while (h.size[k] == len) h.code[k++] = (uint16_t)code++;
        if (k > first && code - 1 >= (1u << len)) return false;   // more codes than the length holds
        h.maxcode[len] = code << (16 - len);
        code <<= 1;
    }
    h.maxcode[17] = 0xFFFFFFFFu;
    std::memset(h.fast, 255, sizeof(h.fast));
    for (int i = 0; i < k; ++i) {
        const int s = h.size[i];
        if (s <= 9) {
            const int c = h.code[i] << (9 - s);
            for (int j = 0; j < (1 << (9 - s)); ++j) h.fast[c + j] = (uint8_t)i;
        }
    }
    h.ok = true;
    return true;
}

struct Component {
    int id = 0, h = 1, v = 1, tq = 0, hd = 0, ha = 0;
    int dc_pred = 0;
    int x = 0, y = 0;        // samples of this component
    int w2 = 0, h2 = 0;      // padded to whole MCUs
    std::vector<uint8_t> data;
    std::vector<int16_t> coeff;   // progressive: coefficients of every block
    int coeff_w = 0;
};

const uint8_t kDezigzag[64 + 15] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
    30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63,
    // corrupt streams may run past the end
    63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

inline int16_t wrap16(int32_t v) { return (int16_t)(uint16_t)(uint32_t)v; }
inline int16_t sat16(int32_t v) { return v > 32767 ? (int16_t)32767 : v < -32768 ? (int16_t)-32768 : (int16_t)v; }
inline int32_t add32(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
inline int32_t sub32(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
inline uint8_t clamp8(int v) { return v < 0 ? 0 : v > 255 ? 255 : (uint8_t)v; }

// 12-bit fixed-point constants of the islow IDCT, rounded as the reference
// rounds them: (int)(c * 4096 + 0.5) with c a float.
inline int f2f(float c) { return (int)(c * 4096 + 0.5); }

struct IdctConst {
    int16_t r00, r01, r10, r11, a0, a1, b0, b1, c0, c1, d0, d1, e0, e1, g0, g1;
    IdctConst() {
        const int k0541 = f2f(0.5411961f), kn1847 = f2f(-1.847759065f), k0765 = f2f(0.765366865f);
        const int k1175 = f2f(1.175875602f), kn0899 = f2f(-0.899976223f), kn2562 = f2f(-2.562915447f);
        const int kn1961 = f2f(-1.961570560f), k0298 = f2f(0.298631336f), k3072 = f2f(3.072711026f);
        const int kn0390 = f2f(-0.390180644f), k2053 = f2f(2.053119869f), k1501 = f2f(1.501321110f);
        r00 = (int16_t)k0541; r01 = (int16_t)(k0541 + kn1847);           // t2 from (s2, s6)
        r10 = (int16_t)(k0541 + k0765); r11 = (int16_t)k0541;            // t3 from (s2, s6)
        a0 = (int16_t)(kn1961 + k0298); a1 = (int16_t)kn1961;            // from (s7, s3)
        b0 = (int16_t)kn1961; b1 = (int16_t)(kn1961 + k3072);            // from (s7, s3)
        c0 = (int16_t)(kn0390 + k2053); c1 = (int16_t)kn0390;            // from (s5, s1)
        d0 = (int16_t)kn0390; d1 = (int16_t)(kn0390 + k1501);            // from (s5, s1)
        e0 = (int16_t)(k1175 + kn0899); e1 = (int16_t)k1175;             // from (s1+s7, s3+s5)
        g0 = (int16_t)k1175; g1 = (int16_t)(k1175 + kn2562);             // from (s1+s7, s3+s5)
    }
};
const IdctConst kC;

inline int32_t dot2(int16_t x, int16_t y, int16_t cx, int16_t cy) {
    return add32((int32_t)x * cx, (int32_t)y * cy);
}

// One 1-D pass on 8 samples: results = (even +- odd + bias) >> shift,
// saturated to 16 bits.
inline void idct_1d(const int16_t s[8], int16_t o[8], int32_t bias, int shift) {
    const int32_t t2 = dot2(s[2], s[6], kC.r00, kC.r01);
    const int32_t t3 = dot2(s[2], s[6], kC.r10, kC.r11);
    const int32_t t0 = (int32_t)wrap16((int32_t)s[0] + s[4]) * 4096;
    const int32_t t1 = (int32_t)wrap16((int32_t)s[0] - s[4]) * 4096;
    const int32_t x0 = add32(t0, t3), x3 = sub32(t0, t3), x1 = add32(t1, t2), x2 = sub32(t1, t2);
    const int32_t y0 = dot2(s[7], s[3], kC.a0, kC.a1);
    const int32_t y2 = dot2(s[7], s[3], kC.b0, kC.b1);
    const int32_t y1 = dot2(s[5], s[1], kC.c0, kC.c1);
    const int32_t y3 = dot2(s[5], s[1], kC.d0, kC.d1);
    const int16_t s17 = wrap16((int32_t)s[1] + s[7]), s35 = wrap16((int32_t)s[3] + s[5]);
    const int32_t y4 = dot2(s17, s35, kC.e0, kC.e1);
    const int32_t y5 = dot2(s17, s35, kC.g0, kC.g1);
    const int32_t x4 = add32(y0, y4), x5 = add32(y1, y5), x6 = add32(y2, y5), x7 = add32(y3, y4);
    auto bfly = [&](int32_t a, int32_t b, int16_t& lo, int16_t& hi) {
        const int32_t ab = add32(a, bias);
        lo = sat16(add32(ab, b) >> shift);
        hi = sat16(sub32(ab, b) >> shift);
    };
    bfly(x0, x7, o[0], o[7]);
    bfly(x1, x6, o[1], o[6]);
    bfly(x2, x5, o[2], o[5]);
    bfly(x3, x4, o[3], o[4]);
}

// 8x8 block (row-major, dequantized) to 8-bit samples: columns, then rows.
void idct_block(uint8_t* out, int stride, const int16_t* d) {
    int16_t m[64], col[8], res[8];
    for (int c = 0; c < 8; ++c) {
        for (int k = 0; k < 8; ++k) col[k] = d[k * 8 + c];
        idct_1d(col, res, 512, 10);
        for (int k = 0; k < 8; ++k) m[k * 8 + c] = res[k];
    }
    for (int r = 0; r < 8; ++r) {
        idct_1d(m + r * 8, res, 65536 + (128 << 17), 17);
        for (int k = 0; k < 8; ++k) out[r * stride + k] = clamp8(res[k]);
    }
}

class Decoder {
public:
    Decoder(const uint8_t* data, size_t n) { r_.p = data; r_.end = data + n; }

    bool decode(int* w, int* h, std::vector<uint8_t>& rgba, std::string& err);

private:
    Reader r_;
    std::string err_;
    Huffman hdc_[4], hac_[4];
    uint16_t dequant_[4][64] = {};
    Component comp_[4];
    int ncomp_ = 0, width_ = 0, height_ = 0;
    int hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0;
    bool progressive_ = false, jfif_ = false;
    int app14_ = -1, rgb_ = 0;
    int restart_ = 0;
    // scan state
    int scan_n_ = 0, order_[4] = {};
    int ss_ = 0, se_ = 63, ah_ = 0, al_ = 0;
    int eob_run_ = 0, todo_ = 0;
    // entropy-coded data
    uint32_t buf_ = 0;
    int bits_ = 0;
    int marker_ = -1;     // marker met in (or right after) the entropy-coded data
    bool nomore_ = false;

    bool fail(const char* m) { err_ = m; return false; }

    int next_marker() {   // the pending marker, else the next one in the stream (-1: none here)
        if (marker_ >= 0) { const int m = marker_; marker_ = -1; return m; }
        if (r_.get8() != 0xFF) return -1;
        int m = r_.get8();
        while (m == 0xFF) m = r_.get8();
        return m;
    }

    void fill() {
        do {
            uint32_t b = nomore_ ? 0 : (uint32_t)r_.get8();
            if (b == 0xFF) {
                int c = r_.get8();
                while (c == 0xFF) c = r_.get8();
                if (c != 0) { marker_ = c; nomore_ = true; return; }
            }
            buf_ |= b << (24 - bits_);
            bits_ += 8;
        } while (bits_ <= 24);
    }

    int huff(const Huffman& hf) {
        if (bits_ < 16) fill();
        const int k = hf.fast[buf_ >> 23];
        if (k < 255) {
            const int s = hf.size[k];
            if (s > bits_) return -1;
            buf_ <<= s;
            bits_ -= s;
            return hf.values[k];
        }
        const uint32_t t = buf_ >> 16;
        int len = 10;
        while (len < 17 && t >= hf.maxcode[len]) ++len;
        if (len == 17) { bits_ -= 16; return -1; }
        if (len > bits_) return -1;
        const int c = (int)((buf_ >> (32 - len)) & ((1u << len) - 1)) + hf.delta[len];
        if (c < 0 || c >= hf.nsym) return -1;
        bits_ -= len;
        buf_ <<= len;
        return hf.values[c];
    }

    int receive_extend(int n) {   // n in 1..16
        if (bits_ < n) fill();
        const uint32_t v = buf_ >> (32 - n);
        buf_ <<= n;
        bits_ -= n;
        return (v & (1u << (n - 1))) ? (int)v : (int)v - (1 << n) + 1;
    }

    int get_bits(int n) {
        if (n == 0) return 0;
        if (bits_ < n) fill();
        const uint32_t v = buf_ >> (32 - n);
        buf_ <<= n;
        bits_ -= n;
        return (int)v;
    }

    int get_bit() {
        if (bits_ < 1) fill();
        const uint32_t v = buf_ >> 31;
        buf_ <<= 1;
        --bits_;
        return (int)v;
    }

    void reset_entropy() {
        buf_ = 0;
        bits_ = 0;
        nomore_ = false;
        marker_ = -1;
        for (auto& c : comp_) c.dc_pred = 0;
        todo_ = restart_ ? restart_ : 0x7FFFFFFF;
        eob_run_ = 0;
    }

    bool marker_segment(int m);
    bool frame_header(int m);
    bool scan_header();
    bool entropy_data();
    bool block_baseline(int16_t* data, Component& c);
    bool block_prog_dc(int16_t* data, Component& c);
    bool block_prog_ac(int16_t* data, Component& c);
    bool restart_due();
    void finish_progressive();
    void to_rgba(std::vector<uint8_t>& rgba);
};

bool Decoder::marker_segment(int m) {
    if (m == 0xDD) {   // DRI
        if (r_.get16() != 4) return fail("bad DRI length");
        restart_ = r_.get16();
        return true;
    }
    if (m == 0xDB) {   // DQT
        int L = r_.get16() - 2;
        while (L > 0) {
            const int q = r_.get8(), p = q >> 4, t = q & 15;
            if (p > 1 || t > 3) return fail("bad DQT");
            for (int i = 0; i < 64; ++i) dequant_[t][kDezigzag[i]] = (uint16_t)(p ? r_.get16() : r_.get8());
            L -= p ? 129 : 65;
        }
        return L == 0 ? true : fail("bad DQT length");
    }
    if (m == 0xC4) {   // DHT
        int L = r_.get16() - 2;
        while (L > 0) {
            const int q = r_.get8(), tc = q >> 4, th = q & 15;
            if (tc > 1 || th > 3) return fail("bad DHT");
            int counts[16], n = 0;
            for (int i = 0; i < 16; ++i) { counts[i] = r_.get8(); n += counts[i]; }
            L -= 17;
            Huffman& hf = tc ? hac_[th] : hdc_[th];
            if (n > 256 || !build_huffman(hf, counts)) return fail("bad Huffman code lengths");
            for (int i = 0; i < n; ++i) hf.values[i] = (uint8_t)r_.get8();
            L -= n;
        }
        return L == 0 ? true : fail("bad DHT length");
    }
    if ((m >= 0xE0 && m <= 0xEF) || m == 0xFE) {   // APPn, COM
        int L = r_.get16();
        if (L < 2) return fail("bad segment length");
        L -= 2;
        if (m == 0xE0 && L >= 5) {
            const uint8_t tag[5] = {'J', 'F', 'I', 'F', 0};
            bool ok = true;
            for (int i = 0; i < 5; ++i) ok &= r_.get8() == tag[i];
            L -= 5;
            if (ok) jfif_ = true;
        } else if (m == 0xEE && L >= 12) {
            const uint8_t tag[6] = {'A', 'd', 'o', 'b', 'e', 0};
            bool ok = true;
            for (int i = 0; i < 6; ++i) ok &= r_.get8() == tag[i];
            L -= 6;
            if (ok) {
                r_.get8();   // version
                r_.get16();  // flags0
                r_.get16();  // flags1
                app14_ = r_.get8();
                L -= 6;
            }
        }
        r_.skip(L);
        return true;
    }
    return fail("unknown marker");
}

bool Decoder::frame_header(int m) {
    progressive_ = m == 0xC2;
    const int Lf = r_.get16();
    if (Lf < 11) return fail("bad SOF length");
    if (r_.get8() != 8) return fail("only 8-bit JPEG is supported");
    height_ = r_.get16();
    width_ = r_.get16();
    if (height_ == 0) return fail("JPEG with delayed height is not supported");
    if (width_ == 0) return fail("zero width");
    ncomp_ = r_.get8();
    if (ncomp_ != 1 && ncomp_ != 3 && ncomp_ != 4) return fail("bad component count");
    if (Lf != 8 + 3 * ncomp_) return fail("bad SOF length");
    rgb_ = 0;
    for (int i = 0; i < ncomp_; ++i) {
        Component& c = comp_[i];
        c.id = r_.get8();
        if (ncomp_ == 3 && c.id == "RGB"[i]) ++rgb_;
        const int q = r_.get8();
        c.h = q >> 4;
        c.v = q & 15;
        c.tq = r_.get8();
        if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) return fail("bad sampling factors");
        if (c.tq > 3) return fail("bad quantization table index");
    }
    if ((uint64_t)width_ * height_ * ncomp_ > (1ull << 30)) return fail("image too large");
    for (int i = 0; i < ncomp_; ++i) {
        if (comp_[i].h > hmax_) hmax_ = comp_[i].h;
        if (comp_[i].v > vmax_) vmax_ = comp_[i].v;
    }
    mcux_ = (width_ + hmax_ * 8 - 1) / (hmax_ * 8);
    mcuy_ = (height_ + vmax_ * 8 - 1) / (vmax_ * 8);
    for (int i = 0; i < ncomp_; ++i) {
        Component& c = comp_[i];
        c.x = (width_ * c.h + hmax_ - 1) / hmax_;
        c.y = (height_ * c.v + vmax_ - 1) / vmax_;
        c.w2 = mcux_ * c.h * 8;
        c.h2 = mcuy_ * c.v * 8;
        c.data.assign((size_t)c.w2 * c.h2, 0);
        if (progressive_) {
            c.coeff_w = c.w2 / 8;
            c.coeff.assign((size_t)c.w2 * c.h2, 0);
        }
    }
    return true;
}

bool Decoder::scan_header() {
    const int Ls = r_.get16();
    scan_n_ = r_.get8();
    if (scan_n_ < 1 || scan_n_ > 4 || scan_n_ > ncomp_) return fail("bad SOS component count");
    if (Ls != 6 + 2 * scan_n_) return fail("bad SOS length");
    for (int i = 0; i < scan_n_; ++i) {
        const int id = r_.get8(), q = r_.get8();
        int which = 0;
        while (which < ncomp_ && comp_[which].id != id) ++which;
        if (which == ncomp_) return fail("SOS names an unknown component");
        comp_[which].hd = q >> 4;
        comp_[which].ha = q & 15;
        if (comp_[which].hd > 3 || comp_[which].ha > 3) return fail("bad Huffman table index");
        order_[i] = which;
    }
    ss_ = r_.get8();
    se_ = r_.get8();
    const int a = r_.get8();
    ah_ = a >> 4;
    al_ = a & 15;
    if (progressive_) {
        if (ss_ > 63 || se_ > 63 || ss_ > se_ || ah_ > 13 || al_ > 13) return fail("bad SOS");
    } else {
        if (ss_ != 0 || ah_ != 0 || al_ != 0) return fail("bad SOS");
        se_ = 63;
    }
    return true;
}

bool Decoder::block_baseline(int16_t* data, Component& c) {
    if (!hdc_[c.hd].ok || !hac_[c.ha].ok) return fail("missing Huffman table");
    const uint16_t* dq = dequant_[c.tq];
    const int t = huff(hdc_[c.hd]);
    if (t < 0 || t > 16) return fail("bad Huffman code");
    std::memset(data, 0, 64 * sizeof(int16_t));
    const int diff = t ? receive_extend(t) : 0;
    c.dc_pred += diff;
    data[0] = wrap16(c.dc_pred * (int)dq[0]);
    int k = 1;
    do {
        const int rs = huff(hac_[c.ha]);
        if (rs < 0) return fail("bad Huffman code");
        const int s = rs & 15, r = rs >> 4;
        if (s == 0) {
            if (rs != 0xF0) break;   // end of block
            k += 16;
        } else {
            k += r;
            const int z = kDezigzag[k++];
            data[z] = wrap16(receive_extend(s) * (int)dq[z]);
        }
    } while (k < 64);
    return true;
}

bool Decoder::block_prog_dc(int16_t* data, Component& c) {
    if (se_ != 0) return fail("DC and AC in one progressive scan");
    if (ah_ == 0) {
        if (!hdc_[c.hd].ok) return fail("missing Huffman table");
        std::memset(data, 0, 64 * sizeof(int16_t));
        const int t = huff(hdc_[c.hd]);
        if (t < 0 || t > 16) return fail("bad Huffman code");
        const int diff = t ? receive_extend(t) : 0;
        c.dc_pred += diff;
        data[0] = wrap16((int32_t)((uint32_t)c.dc_pred << al_));
    } else if (get_bit()) {
        data[0] = wrap16(data[0] + (1 << al_));
    }
    return true;
}

bool Decoder::block_prog_ac(int16_t* data, Component& c) {
    if (ss_ == 0) return fail("DC and AC in one progressive scan");
    const Huffman& hf = hac_[c.ha];
    if (!hf.ok) return fail("missing Huffman table");
    if (ah_ == 0) {   // first pass over these coefficients
        if (eob_run_) { --eob_run_; return true; }
        int k = ss_;
        do {
            const int rs = huff(hf);
            if (rs < 0) return fail("bad Huffman code");
            const int s = rs & 15, r = rs >> 4;
            if (s == 0) {
                if (r < 15) {
                    eob_run_ = (1 << r) - 1 + get_bits(r);
                    break;
                }
                k += 16;
            } else {
                k += r;
                const int z = kDezigzag[k++];
                data[z] = wrap16((int32_t)((uint32_t)receive_extend(s) << al_));
            }
        } while (k <= se_);
        return true;
    }
    // refinement of coefficients already non-zero, new ones of +-1 << al
    const int16_t bit = (int16_t)(1 << al_);
    auto refine = [&](int16_t* p) {
        if (get_bit() && (*p & bit) == 0) *p = wrap16(*p > 0 ? *p + bit : *p - bit);
    };
    if (eob_run_) {
        --eob_run_;
        for (int k = ss_; k <= se_; ++k) {
            int16_t* p = &data[kDezigzag[k]];
            if (*p != 0) refine(p);
        }
        return true;
    }
    int k = ss_;
    do {
        const int rs = huff(hf);
        if (rs < 0) return fail("bad Huffman code");
        int s = rs & 15, r = rs >> 4;
        if (s == 0) {
            if (r < 15) {
                eob_run_ = (1 << r) - 1 + get_bits(r);
                r = 64;   // the rest of the band only refines
            }
        } else {
            if (s != 1) return fail("bad Huffman code");
            s = get_bit() ? bit : -bit;
        }
        while (k <= se_) {
            int16_t* p = &data[kDezigzag[k++]];
            if (*p != 0) {
                refine(p);
            } else {
                if (r == 0) { *p = (int16_t)s; break; }
                --r;
            }
        }
    } while (k <= se_);
    return true;
}

// After an MCU: when the restart interval runs out, a RSTn marker restarts
// the entropy decoder; anything else ends the scan.
bool Decoder::restart_due() {
    if (--todo_ > 0) return false;
    if (bits_ < 24) fill();
    if (marker_ < 0xD0 || marker_ > 0xD7) return true;
    reset_entropy();
    return false;
}

bool Decoder::entropy_data() {
    reset_entropy();
    alignas(16) int16_t blk[64];
    if (scan_n_ == 1) {   // one component: its blocks in raster order
        Component& c = comp_[order_[0]];
        const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
        for (int j = 0; j < bh; ++j)
            for (int i = 0; i < bw; ++i) {
                if (!progressive_) {
                    if (!block_baseline(blk, c)) return false;
                    idct_block(c.data.data() + (size_t)c.w2 * j * 8 + i * 8, c.w2, blk);
                } else {
                    int16_t* d = c.coeff.data() + 64 * ((size_t)i + (size_t)j * c.coeff_w);
                    if (!(ss_ == 0 ? block_prog_dc(d, c) : block_prog_ac(d, c))) return false;
                }
                if (restart_due()) return true;
            }
        return true;
    }
    for (int j = 0; j < mcuy_; ++j)   // interleaved MCUs
        for (int i = 0; i < mcux_; ++i) {
            for (int k = 0; k < scan_n_; ++k) {
                Component& c = comp_[order_[k]];
                for (int y = 0; y < c.v; ++y)
                    for (int x = 0; x < c.h; ++x) {
                        const int bx = i * c.h + x, by = j * c.v + y;
                        if (!progressive_) {
                            if (!block_baseline(blk, c)) return false;
                            idct_block(c.data.data() + (size_t)c.w2 * by * 8 + bx * 8, c.w2, blk);
                        } else {
                            int16_t* d = c.coeff.data() + 64 * ((size_t)bx + (size_t)by * c.coeff_w);
                            if (!block_prog_dc(d, c)) return false;
                        }
                    }
            }
            if (restart_due()) return true;
        }
    return true;
}

void Decoder::finish_progressive() {
    alignas(16) int16_t blk[64];
    for (int n = 0; n < ncomp_; ++n) {
        Component& c = comp_[n];
        const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
        const uint16_t* dq = dequant_[c.tq];
        for (int j = 0; j < bh; ++j)
            for (int i = 0; i < bw; ++i) {
                const int16_t* d = c.coeff.data() + 64 * ((size_t)i + (size_t)j * c.coeff_w);
                for (int k = 0; k < 64; ++k) blk[k] = wrap16((int32_t)d[k] * dq[k]);
                idct_block(c.data.data() + (size_t)c.w2 * j * 8 + i * 8, c.w2, blk);
            }
    }
}

// Upsampling of one output row from the two nearest sample rows.
const uint8_t* upsample(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int hs, int vs) {
    if (hs == 1 && vs == 1) return near;
    if (hs == 1 && vs == 2) {
        for (int i = 0; i < w; ++i) out[i] = (uint8_t)((3 * near[i] + far[i] + 2) >> 2);
        return out;
    }
    if (hs == 2 && vs == 1) {
        if (w == 1) { out[0] = out[1] = near[0]; return out; }
        out[0] = near[0];
        out[1] = (uint8_t)((near[0] * 3 + near[1] + 2) >> 2);
        int i = 1;
        for (; i < w - 1; ++i) {
            const int n = 3 * near[i] + 2;
            out[2 * i] = (uint8_t)((n + near[i - 1]) >> 2);
            out[2 * i + 1] = (uint8_t)((n + near[i + 1]) >> 2);
        }
        out[2 * i] = (uint8_t)((near[w - 2] * 3 + near[w - 1] + 2) >> 2);
        out[2 * i + 1] = near[w - 1];
        return out;
    }
    if (hs == 2 && vs == 2) {
        if (w == 1) { out[0] = out[1] = (uint8_t)((3 * near[0] + far[0] + 2) >> 2); return out; }
        int t1 = 3 * near[0] + far[0];
        out[0] = (uint8_t)((t1 + 2) >> 2);
        for (int i = 1; i < w; ++i) {
            const int t0 = t1;
            t1 = 3 * near[i] + far[i];
            out[2 * i - 1] = (uint8_t)((3 * t0 + t1 + 8) >> 4);
            out[2 * i] = (uint8_t)((3 * t1 + t0 + 8) >> 4);
        }
        out[2 * w - 1] = (uint8_t)((t1 + 2) >> 2);
        return out;
    }
    for (int i = 0; i < w; ++i)   // other factors: nearest
        for (int j = 0; j < hs; ++j) out[i * hs + j] = near[i];
    return out;
}

// YCbCr -> RGBA, four bytes a pixel, in the reference's 32-bit fixed-point
// form (its SSE2 form, high halves of 16x16-bit products, gives the same
// bytes for all 2^24 inputs).
void ycbcr_row(uint8_t* out, const uint8_t* y, const uint8_t* cb, const uint8_t* cr, int count) {
    auto fx = [](float c) { return ((int)(c * 4096.0f + 0.5f)) << 8; };
    const int kr = fx(1.40200f), kg = -fx(0.71414f), kgb = -fx(0.34414f), kb = fx(1.77200f);
    for (int i = 0; i < count; ++i) {
        const int yf = (y[i] << 20) + (1 << 19);
        const int crv = cr[i] - 128, cbv = cb[i] - 128;
        const int r = (yf + crv * kr) >> 20;
        const int g = (yf + crv * kg + (int)((uint32_t)(cbv * kgb) & 0xFFFF0000u)) >> 20;
        const int b = (yf + cbv * kb) >> 20;
        out[4 * i] = clamp8(r);
        out[4 * i + 1] = clamp8(g);
        out[4 * i + 2] = clamp8(b);
        out[4 * i + 3] = 255;
    }
}

// x * y / 255, rounded (the reference's CMYK product).
inline uint8_t mul8(uint8_t x, uint8_t y) {
    const unsigned t = (unsigned)x * y + 128;
    return (uint8_t)((t + (t >> 8)) >> 8);
}

void Decoder::to_rgba(std::vector<uint8_t>& rgba) {
    const bool is_rgb = ncomp_ == 3 && (rgb_ == 3 || (app14_ == 0 && !jfif_));
    struct Resample { int hs, vs, ystep, ypos, wlo; const uint8_t* line0; const uint8_t* line1; };
    Resample rs[4];
    std::vector<uint8_t> linebuf[4];
    for (int k = 0; k < ncomp_; ++k) {
        Component& c = comp_[k];
        Resample& r = rs[k];
        r.hs = hmax_ / c.h;
        r.vs = vmax_ / c.v;
        r.ystep = r.vs >> 1;
        r.wlo = (width_ + r.hs - 1) / r.hs;
        r.ypos = 0;
        r.line0 = r.line1 = c.data.data();
        linebuf[k].assign((size_t)width_ + 3, 0);
    }
    rgba.assign((size_t)width_ * height_ * 4, 0);
    const uint8_t* row[4] = {};
    for (int j = 0; j < height_; ++j) {
        uint8_t* out = rgba.data() + (size_t)4 * width_ * j;
        for (int k = 0; k < ncomp_; ++k) {
            Resample& r = rs[k];
            const bool bottom = r.ystep >= (r.vs >> 1);
            row[k] = upsample(linebuf[k].data(), bottom ? r.line1 : r.line0, bottom ? r.line0 : r.line1, r.wlo, r.hs,
                              r.vs);
            if (++r.ystep >= r.vs) {
                r.ystep = 0;
                r.line0 = r.line1;
                if (++r.ypos < comp_[k].y) r.line1 += comp_[k].w2;
            }
        }
        if (ncomp_ == 3) {
            if (is_rgb) {
                for (int i = 0; i < width_; ++i) {
                    out[4 * i] = row[0][i];
                    out[4 * i + 1] = row[1][i];
                    out[4 * i + 2] = row[2][i];
                    out[4 * i + 3] = 255;
                }
            } else {
                ycbcr_row(out, row[0], row[1], row[2], width_);
            }
        } else if (ncomp_ == 4) {
            if (app14_ == 0) {   // CMYK
                for (int i = 0; i < width_; ++i) {
                    const uint8_t m = row[3][i];
                    out[4 * i] = mul8(row[0][i], m);
                    out[4 * i + 1] = mul8(row[1][i], m);
                    out[4 * i + 2] = mul8(row[2][i], m);
                    out[4 * i + 3] = 255;
                }
            } else if (app14_ == 2) {   // YCCK
                ycbcr_row(out, row[0], row[1], row[2], width_);
                for (int i = 0; i < width_; ++i) {
                    const uint8_t m = row[3][i];
                    out[4 * i] = mul8((uint8_t)(255 - out[4 * i]), m);
                    out[4 * i + 1] = mul8((uint8_t)(255 - out[4 * i + 1]), m);
                    out[4 * i + 2] = mul8((uint8_t)(255 - out[4 * i + 2]), m);
                }
            } else {   // YCbCr plus a fourth channel, which is ignored
                ycbcr_row(out, row[0], row[1], row[2], width_);
            }
        } else {
            for (int i = 0; i < width_; ++i) {
                out[4 * i] = out[4 * i + 1] = out[4 * i + 2] = row[0][i];
                out[4 * i + 3] = 255;
            }
        }
    }
}

bool Decoder::decode(int* w, int* h, std::vector<uint8_t>& rgba, std::string& err) {
    auto done = [&](bool ok) {
        if (!ok) err = err_.empty() ? "corrupt JPEG" : err_;
        return ok;
    };
    int m = next_marker();
    if (m != 0xD8) return done(fail("no SOI marker"));
    m = next_marker();
    while (m != 0xC0 && m != 0xC1 && m != 0xC2) {   // tables and APP segments before the frame
        if (!marker_segment(m)) return done(false);
        m = next_marker();
        while (m < 0) {
            if (r_.eof()) return done(fail("no SOF marker"));
            m = next_marker();
        }
    }
    if (!frame_header(m)) return done(false);
    m = next_marker();
    while (m != 0xD9) {   // EOI
        if (m == 0xDA) {
            if (!scan_header() || !entropy_data()) return done(false);
            if (marker_ < 0) {
                // trailing bytes after the entropy-coded data: take the
                // byte after the next 0xFF as the marker
                while (!r_.eof()) {
                    if (r_.get8() == 0xFF) { marker_ = r_.get8(); break; }
                }
            }
        } else if (m == 0xDC) {   // DNL
            if (r_.get16() != 4) return done(fail("bad DNL length"));
            if (r_.get16() != height_) return done(fail("bad DNL height"));
        } else if (m < 0) {
            return done(fail("expected a marker"));
        } else if (!marker_segment(m)) {
            return done(false);
        }
        m = next_marker();
    }
    if (progressive_) finish_progressive();
    to_rgba(rgba);
    *w = width_;
    *h = height_;
    return true;
}

}  // namespace

namespace axr {

bool decode_jpeg(const uint8_t* data, size_t n, int* w, int* h, std::vector<uint8_t>& rgba, std::string& err) {
    Decoder d(data, n);
    return d.decode(w, h, rgba, err);
}

}  // namespace axr
