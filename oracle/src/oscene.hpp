// ORACLE - test infrastructure only (see omath.hpp).
//
// Resource types of the restated pipeline: Vertex, Texture, Material, Mesh,
// Framebuffer.
//   Vertex       reference: include/mesh.hpp:9-18   (14 floats = 56 bytes)
//   Material     reference: include/mesh.hpp:20-34  (textures + specular exponent)
//   MaterialGroup reference: include/mesh.hpp:35-39 (contiguous face ranges)
//   Texture      reference: include/texture.hpp:12-33 (nearest sample, v flipped)
//   Framebuffer  reference: include/framebuffer.hpp:14-48, src/framebuffer.cpp:26-42
//                (BGRA8 colour, FP32 depth cleared to +inf)
// Additions with no reference code: a bilinear filter mode on Texture, and a
// per-face RGBA8 colour stream on Mesh (used by FlatColorShader).
#pragma once
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include "omath.hpp"

namespace orc {

struct Vertex {
    vec3 position;
    vec2 uv;
    vec3 normal;
    vec3 tangent;
    vec3 bitangent;
};
static_assert(sizeof(Vertex) == 56, "Vertex must be 14 floats");

enum class Filter { Nearest = 0, Bilinear = 1 };

class Texture {
public:
    Texture(const uint8_t* rgba, int w, int h, Filter f) : m_W(w), m_H(h), m_Filter(f) {
        m_Data.assign(rgba, rgba + (size_t)w * h * 4);
    }
    int width() const { return m_W; }
    int height() const { return m_H; }

    vec4 sample(const vec2& uv) const {
        if (m_Data.empty()) return vec4(0, 0, 0, 1);
        return m_Filter == Filter::Bilinear ? sampleBilinear(uv) : sampleNearest(uv);
    }
    // Reference texture.hpp:18-33: x = (int)(u*(W-1)), clamp, row flipped.
    vec4 sampleNearest(const vec2& uv) const {
        int x = cvt_trunc_i32(uv.x * (float)(m_W - 1));
        int y = cvt_trunc_i32(uv.y * (float)(m_H - 1));
        x = std::max(0, std::min(x, m_W - 1));
        y = std::max(0, std::min(y, m_H - 1));
        y = m_H - 1 - y;
        return texel(x, y);
    }
    // Bilinear filter on the same texel grid as sampleNearest: texel centres
    // at integer positions of (u*(W-1), v*(H-1)), clamp-to-edge, rows flipped.
    //   c = mix(mix(t00, t10, fx), mix(t01, t11, fx), fy)   per channel
    vec4 sampleBilinear(const vec2& uv) const {
        float sx = uv.x * (float)(m_W - 1);
        float sy = uv.y * (float)(m_H - 1);
        float flx = std::floor(sx), fly = std::floor(sy);
        float fx = sx - flx, fy = sy - fly;
        int x0 = cvt_trunc_i32(flx), y0 = cvt_trunc_i32(fly);
        int x1 = (x0 == INT32_MAX) ? x0 : x0 + 1;
        int y1 = (y0 == INT32_MAX) ? y0 : y0 + 1;
        x0 = std::max(0, std::min(x0, m_W - 1)); x1 = std::max(0, std::min(x1, m_W - 1));
        y0 = std::max(0, std::min(y0, m_H - 1)); y1 = std::max(0, std::min(y1, m_H - 1));
        int r0 = m_H - 1 - y0, r1 = m_H - 1 - y1;
        vec4 t00 = texel(x0, r0), t10 = texel(x1, r0), t01 = texel(x0, r1), t11 = texel(x1, r1);
        vec4 top = mix(t00, t10, fx);
        vec4 bot = mix(t01, t11, fx);
        return mix(top, bot, fy);
    }
    vec4 texel(int x, int y) const {
        const uint8_t* p = &m_Data[((size_t)y * m_W + x) * 4];
        const float inv255 = 1.0f / 255.0f;
        return vec4(p[0] * inv255, p[1] * inv255, p[2] * inv255, p[3] * inv255);
    }

private:
    std::vector<uint8_t> m_Data;
    int m_W, m_H;
    Filter m_Filter;
};

struct Material {
    const Texture* diffuseTexture = nullptr;
    const Texture* metallicTexture = nullptr;
    const Texture* bumpTexture = nullptr;
    const Texture* roughnessTexture = nullptr;
    const Texture* aoTexture = nullptr;
    float specularExponent = 1.0f;
};

struct MaterialGroup {
    uint32_t material;      // index into Mesh::materials
    size_t startIndex;
    size_t faceCount;
};

struct Mesh {
    std::vector<Vertex> vertices;
    std::vector<uint32_t> indices;      // 3 per face
    std::vector<MaterialGroup> groups;
    std::vector<Material> materials;
    std::vector<uint32_t> faceColors;   // optional RGBA8 per face (r in the low byte)
};

class Framebuffer {
public:
    Framebuffer(int w, int h) : m_W(w), m_H(h), m_Color((size_t)w * h * 4, 0),
        m_Depth((size_t)w * h, std::numeric_limits<float>::infinity()), m_Ids((size_t)w * h, 0xFFFFFFFFu) {}
    int width() const { return m_W; }
    int height() const { return m_H; }
    // Reference framebuffer.cpp:26-36: packed a<<24 | r<<16 | g<<8 | b.
    void clearColor(uint32_t packedBGRA) {
        uint32_t* p = reinterpret_cast<uint32_t*>(m_Color.data());
        std::fill(p, p + (size_t)m_W * m_H, packedBGRA);
        std::fill(m_Ids.begin(), m_Ids.end(), 0xFFFFFFFFu);
    }
    void clearDepth(float d) { std::fill(m_Depth.begin(), m_Depth.end(), d); }
    uint8_t* color() { return m_Color.data(); }
    float* depth() { return m_Depth.data(); }
    // Winner primitive key per pixel (face submission rank * 8 + sub-triangle);
    // a debugging channel of the restatement, the reference has none.
    uint32_t* ids() { return m_Ids.data(); }

private:
    int m_W, m_H;
    std::vector<uint8_t> m_Color;
    std::vector<float> m_Depth;
    std::vector<uint32_t> m_Ids;
};

}  // namespace orc
