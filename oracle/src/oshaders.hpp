// ORACLE - test infrastructure only (see omath.hpp).
//
// The shader plug-in interface and the shaders, restated.
//   VertexOutput / VSTransformedTriangle  reference: include/IShader.hpp:11-28
//   IShader (vertex/fragment + uniforms)  reference: include/IShader.hpp:30-45
//   FlatShader   reference: include/shaders/shaders.hpp:19-62
//   PhongShader  reference: include/shaders/shaders.hpp:140-247
//   PBRShader    reference: include/shaders/shaders.hpp:252-423
// Shaders with no reference code (the workloads ask for them):
//   GouraudShader     per-vertex Lambert intensity, interpolated, times a texture
//   PhongPointShader  PhongShader's normal-mapped model summed over point lights
//   FlatColorShader   one RGBA8 colour per primitive
// Extension of the interface: VSTransformedTriangle carries the primitive's
// face index, and the uniforms carry the mesh's per-face colour stream.
#pragma once
#include <algorithm>
#include <memory>
#include "oscene.hpp"

namespace orc {

struct VertexOutput {
    vec3 normal;
    float zNDC = 1.0f;
    vec2 uv;
    vec3 worldPos;
    mat3 tbn;
};

struct VSTransformedTriangle {
    VertexOutput vertices[3];
    uint32_t primitiveId = 0;
    const VertexOutput& operator[](int i) const { return vertices[i]; }
};

struct IShader {
    virtual ~IShader() = default;
    virtual VertexOutput vertex(const Vertex& v, int indexInsideFace) = 0;
    virtual bool fragment(vec3& bar, vec4& color, const VSTransformedTriangle& tri) = 0;
    virtual std::unique_ptr<IShader> clone() const = 0;

    mat4 mvp, model, viewProj, viewportMat;
    vec3 cameraPosition;
    const Material* material = nullptr;
    const uint32_t* faceColors = nullptr;
};

template <class D> struct ShaderBase : IShader {
    std::unique_ptr<IShader> clone() const override { return std::make_unique<D>(static_cast<const D&>(*this)); }
};

// Barycentric blend written left to right: (b.x*v0 + b.y*v1) + b.z*v2.
template <class T> inline T bary3(const vec3& b, const T& v0, const T& v1, const T& v2) {
    return (b.x * v0 + b.y * v1) + b.z * v2;
}
inline float bary3f(const vec3& b, float v0, float v1, float v2) { return (b.x * v0 + b.y * v1) + b.z * v2; }

// world-space TBN of one vertex, shared by the Phong-family vertex stages
// (reference shaders.hpp:156-166).
inline void world_tbn(const mat4& model, const Vertex& v, VertexOutput& o) {
    o.uv = v.uv;
    o.worldPos = xyz(model * vec4(v.position, 1.0f));
    vec3 T = normalize(xyz(model * vec4(v.tangent, 0.0f)));
    vec3 B = normalize(xyz(model * vec4(v.bitangent, 0.0f)));
    vec3 N = normalize(xyz(model * vec4(v.normal, 0.0f)));
    o.tbn[0] = T; o.tbn[1] = B; o.tbn[2] = N;
}

struct FlatShader : ShaderBase<FlatShader> {
    vec3 lightDirection;
    VertexOutput vertex(const Vertex& v, int) override {
        VertexOutput o;
        mat3 nm = to_mat3(transpose(inverse(model)));
        o.normal = nm * v.normal;
        return o;
    }
    bool fragment(vec3& bar, vec4& color, const VSTransformedTriangle& t) override {
        vec3 n = normalize(bary3(bar, t[0].normal, t[1].normal, t[2].normal));
        float intensity = sclamp(dot(-lightDirection, n), 0.0f, 1.0f);
        color = vec4(1.0f * intensity, 1.0f * intensity, 1.0f * intensity, 1.0f * intensity);
        return false;
    }
};

// Normal-mapped surface normal and albedo shared by PhongShader and
// PhongPointShader (reference shaders.hpp:172-213).
inline void phong_surface(const vec3& bar, const VSTransformedTriangle& t, const Material* m,
                          vec3& normal, vec4& albedo, vec3& fragPos) {
    vec2 uv = bary3(bar, t[0].uv, t[1].uv, t[2].uv);
    vec4 nmv = m->bumpTexture->sample(uv);
    vec3 nms = normalize(xyz(nmv) * 2.0f - vec3(1.0f));
    mat3 itbn = (bar.x * t[0].tbn + bar.y * t[1].tbn) + bar.z * t[2].tbn;
    vec3 iT = itbn[0];
    vec3 iN = normalize(itbn[2]);
    vec3 tN = normalize(iT - iN * dot(iN, iT));
    vec3 bN = cross(iN, tN);
    mat3 ftbn; ftbn[0] = tN; ftbn[1] = bN; ftbn[2] = iN;
    albedo = m->diffuseTexture->sample(uv);
    normal = normalize(ftbn * nms);
    fragPos = bary3(bar, t[0].worldPos, t[1].worldPos, t[2].worldPos);
}

struct PhongShader : ShaderBase<PhongShader> {
    vec3 lightDirection, lightColor;
    VertexOutput vertex(const Vertex& v, int) override { VertexOutput o; world_tbn(model, v, o); return o; }
    bool fragment(vec3& bar, vec4& color, const VSTransformedTriangle& t) override {
        vec3 normal, fragPos; vec4 albedo;
        phong_surface(bar, t, material, normal, albedo, fragPos);
        vec3 viewDir = normalize(cameraPosition - fragPos);
        vec3 lightDir = -lightDirection;
        vec3 ambient = 0.1f * lightColor;
        float diff = smax(dot(normal, lightDir), 0.0f);
        vec3 diffuse = diff * lightColor;
        vec3 reflectDir = reflect(-lightDir, normal);
        float spec = std::pow(smax(dot(viewDir, reflectDir), 0.0f), material->specularExponent * 50.0f);
        vec3 specular = (0.5f * spec) * lightColor;
        vec3 fc = ((ambient + diffuse) + specular) * xyz(albedo);
        color = vec4(fc, 1.0f);
        return false;
    }
};

struct PhongPointShader : ShaderBase<PhongPointShader> {
    int numLights = 0;
    vec3 lightPos[4], lightColor[4];
    VertexOutput vertex(const Vertex& v, int) override { VertexOutput o; world_tbn(model, v, o); return o; }
    bool fragment(vec3& bar, vec4& color, const VSTransformedTriangle& t) override {
        vec3 normal, fragPos; vec4 albedo;
        phong_surface(bar, t, material, normal, albedo, fragPos);
        vec3 viewDir = normalize(cameraPosition - fragPos);
        vec3 sum(0.0f);
        for (int l = 0; l < numLights; ++l) {
            vec3 L = normalize(lightPos[l] - fragPos);
            vec3 ambient = 0.1f * lightColor[l];
            float diff = smax(dot(normal, L), 0.0f);
            vec3 diffuse = diff * lightColor[l];
            vec3 reflectDir = reflect(-L, normal);
            float spec = std::pow(smax(dot(viewDir, reflectDir), 0.0f), material->specularExponent * 50.0f);
            vec3 specular = (0.5f * spec) * lightColor[l];
            sum = sum + ((ambient + diffuse) + specular);
        }
        color = vec4(sum * xyz(albedo), 1.0f);
        return false;
    }
};

struct PBRShader : ShaderBase<PBRShader> {
    vec3 lightDirection, lightColor;
    VertexOutput vertex(const Vertex& v, int) override { VertexOutput o; world_tbn(model, v, o); return o; }
    bool fragment(vec3& bar, vec4& color, const VSTransformedTriangle& t) override {
        const float PI = 3.14159265358979323846f;
        vec2 uv = bary3(bar, t[0].uv, t[1].uv, t[2].uv);
        vec4 nmv = material->bumpTexture->sample(uv);
        vec3 nms = normalize((xyz(nmv) * 2.0f) - 1.0f);
        mat3 itbn = (bar.x * t[0].tbn + bar.y * t[1].tbn) + bar.z * t[2].tbn;
        vec3 iT = itbn[0];
        vec3 iN = normalize(itbn[2]);
        vec3 nt = iN * dot(iN, iT);
        vec3 tN = normalize(iT - nt);
        vec3 bN = normalize(cross(iN, tN));
        (void)bN;  // the reference re-derives N from T and B but never uses it
        vec3 albedo = xyz(material->diffuseTexture->sample(uv));
        albedo = vpow(albedo, vec3(2.2f));
        float metallic = material->metallicTexture->sample(uv).x;
        float roughness = material->roughnessTexture->sample(uv).x;
        float ao = material->aoTexture->sample(uv).x;
        float r2 = roughness * roughness;
        float r4 = r2 * r2;
        float omm = 1.0f - metallic;
        vec3 normal = normalize(itbn * nms);
        vec3 fragPos = bary3(bar, t[0].worldPos, t[1].worldPos, t[2].worldPos);
        vec3 viewDir = normalize(cameraPosition - fragPos);
        vec3 lightDir = -lightDirection;
        vec3 halfway = normalize(lightDir + viewDir);
        float NdotH = smax(dot(normal, halfway), 0.0f);
        float NdotV = smax(dot(normal, viewDir), 0.0f);
        float NdotL = smax(dot(normal, lightDir), 0.0f);
        vec3 F0 = vec3(0.04f) + metallic * (albedo - vec3(0.04f));
        float NdotH2 = NdotH * NdotH;
        float dp = NdotH2 * (r4 - 1.0f) + 1.0f;
        float NDF = r4 / ((PI * dp) * dp);
        float r = roughness + 1.0f;
        float k = (r * r) / 8.0f;
        float nvk = NdotV * (1.0f - k) + k;
        float nlk = NdotL * (1.0f - k) + k;
        float G = (NdotV / nvk) * (NdotL / nlk);
        float cosT = smax(dot(halfway, normal), 0.0f);
        float omc = 1.0f - cosT;
        float term = omc * omc; term = term * term; term = term * omc;
        vec3 F = F0 + (vec3(1.0f) - F0) * term;
        vec3 numerator = (F * NDF) * G;
        float denom = ((4.0f * NdotV) * NdotL) + 0.0001f;
        vec3 specular = numerator / denom;
        vec3 kD = (vec3(1.0f) - F) * omm;
        vec3 diffuse = (kD * albedo) * (1.0f / PI);
        vec3 ambient = (vec3(0.03f) * albedo) * ao;
        vec3 fc = ambient + (((diffuse + specular) * lightColor) * NdotL);
        fc = fc / (fc + vec3(1.0f));
        fc = vpow(fc, vec3(1.0f / 2.2f));
        color = vec4(fc, 1.0f);
        return false;
    }
};

struct GouraudShader : ShaderBase<GouraudShader> {
    vec3 lightDirection;
    VertexOutput vertex(const Vertex& v, int) override {
        VertexOutput o;
        vec3 N = normalize(xyz(model * vec4(v.normal, 0.0f)));
        float intensity = sclamp(dot(-lightDirection, N), 0.0f, 1.0f);
        o.normal = vec3(intensity, 0.0f, 0.0f);
        o.uv = v.uv;
        return o;
    }
    bool fragment(vec3& bar, vec4& color, const VSTransformedTriangle& t) override {
        float i = bary3f(bar, t[0].normal.x, t[1].normal.x, t[2].normal.x);
        vec2 uv = bary3(bar, t[0].uv, t[1].uv, t[2].uv);
        vec4 albedo = material->diffuseTexture->sample(uv);
        color = vec4(xyz(albedo) * (0.1f + i), 1.0f);
        return false;
    }
};

struct FlatColorShader : ShaderBase<FlatColorShader> {
    VertexOutput vertex(const Vertex&, int) override { return VertexOutput{}; }
    bool fragment(vec3&, vec4& color, const VSTransformedTriangle& t) override {
        uint32_t c = faceColors ? faceColors[t.primitiveId] : 0xFFFFFFFFu;
        const float inv255 = 1.0f / 255.0f;
        color = vec4(((float)(c & 255u) + 0.5f) * inv255, ((float)((c >> 8) & 255u) + 0.5f) * inv255,
                     ((float)((c >> 16) & 255u) + 0.5f) * inv255, ((float)(c >> 24) + 0.5f) * inv255);
        return false;
    }
};

}  // namespace orc
