// shade.hpp - the fragment-stage shaders as device functors.
//
// Each shader is a struct with
//   Var vs(const Vertex56&, const Uniforms&)         the vertex stage
//   vec4 fs(vec3 bar, const Var v[3], const ShadeCtx&)  the fragment stage
// and the raster kernel instantiates one resolve loop per shader, so the
// shader choice costs one uniform branch per tile.
//
//   FlatShader        reference: include/shaders/shaders.hpp:19-62
//   PhongShader       reference: include/shaders/shaders.hpp:140-247
//   PBRShader         reference: include/shaders/shaders.hpp:252-423
//   GouraudShader     no reference code: per-vertex Lambert intensity, times a
//                     texture, colour = albedo * (0.1 + i)
//   PhongPointShader  no reference code: PhongShader's normal-mapped model,
//                     light terms summed over up to 4 point lights
//   FlatColorShader   no reference code: per-face RGBA8, written back exactly
// Texture::sample (reference include/texture.hpp:12-33) is restated as
// sample_nearest; sample_bilinear is the added filter on the same texel grid.
#pragma once
#include "axr_types.hpp"

#pragma clang fp contract(off)

namespace axr {

// float -> int32 with x86 cvttss2si semantics: truncation, NaN and
// out-of-range give INT_MIN (the reference's (int) casts on its x86 target).
AXR_HD int cvt_i32(float f) {
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : INT32_MIN;
}

// Texel (x, y) with 0 <= x < w, 0 <= y < h; w * h <= 2^30, so the index
// fits 32 bits.
AXR_HD uint32_t texel_word(const DevTexture& t, int x, int y) {
    // a 32-bit byte offset: with the texel pointer in scalar registers the
    // load takes a scalar base and a 32-bit vector offset
    const uint32_t off = ((uint32_t)y * (uint32_t)t.w + (uint32_t)x) * 4u;
    return gld(reinterpret_cast<const uint32_t*>(reinterpret_cast<const unsigned char*>(t.texels) + off), 0u);
}

AXR_HD vec4 texel(const DevTexture& t, int x, int y) {
    uint32_t p = texel_word(t, x, y);
    const float inv255 = 1.0f / 255.0f;
    return vec4{(float)(p & 255u) * inv255, (float)((p >> 8) & 255u) * inv255,
                (float)((p >> 16) & 255u) * inv255, (float)(p >> 24) * inv255};
}

// A texel's four bytes as floats 0..255 (the bilinear filter blends these
// and scales by 1/255 once).
AXR_HD vec4 texel_bytes(const DevTexture& t, int x, int y) {
    uint32_t p = texel_word(t, x, y);
    return vec4{(float)(p & 255u), (float)((p >> 8) & 255u), (float)((p >> 16) & 255u), (float)(p >> 24)};
}

// Colour-only arithmetic with fused multiply-adds.  A fused multiply-add
// differs from the separately rounded pair by at most one rounding of the
// product (~6e-8 relative); through the shading terms that stays orders of
// magnitude below one 8-bit step (1/255), which is the tolerance colour is
// held to.  Barycentrics, texture coordinates, coverage and depth never go
// through these.
AXR_HD float cdot(vec3 a, vec3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
AXR_HD vec3 cmad(vec3 a, float s, vec3 c) { return make3(fmaf(a.x, s, c.x), fmaf(a.y, s, c.y), fmaf(a.z, s, c.z)); }
AXR_HD vec4 cmad4(vec4 a, float s, vec4 c) {
    return make4(fmaf(a.x, s, c.x), fmaf(a.y, s, c.y), fmaf(a.z, s, c.z), fmaf(a.w, s, c.w));
}
AXR_HD vec3 cbary3(vec3 b, vec3 v0, vec3 v1, vec3 v2) { return cmad(v2, b.z, cmad(v1, b.y, v0 * b.x)); }
AXR_HD float cbary1(vec3 b, float v0, float v1, float v2) { return fmaf(b.z, v2, fmaf(b.y, v1, b.x * v0)); }
AXR_HD vec3 cmul(const mat3& m, vec3 v) { return cmad(m.c[2], v.z, cmad(m.c[1], v.y, m.c[0] * v.x)); }
AXR_HD vec3 cdir(const mat4& m, float x, float y, float z) {
    return cmad(xyz(m.c[2]), z, cmad(xyz(m.c[1]), y, xyz(m.c[0]) * x));
}
AXR_HD vec3 cpoint(const mat4& m, float x, float y, float z) {
    return cmad(xyz(m.c[2]), z, cmad(xyz(m.c[1]), y, cmad(xyz(m.c[0]), x, xyz(m.c[3]))));
}
AXR_HD vec3 creflect(vec3 i, vec3 n) { return cmad(n, -2.0f * cdot(n, i), i); }
AXR_HD vec4 cmix(vec4 x, vec4 y, float a) { return cmad4(y - x, a, x); }

AXR_HD vec4 sample_nearest(const DevTexture& t, vec2 uv) {
    int x = cvt_i32(uv.x * (float)(t.w - 1));
    int y = cvt_i32(uv.y * (float)(t.h - 1));
    x = imax(0, imin(x, t.w - 1));
    y = imax(0, imin(y, t.h - 1));
    return texel(t, x, t.h - 1 - y);
}

AXR_HD vec4 sample_bilinear(const DevTexture& t, vec2 uv) {
    float sx = uv.x * (float)(t.w - 1);
    float sy = uv.y * (float)(t.h - 1);
    float flx = floorf(sx), fly = floorf(sy);
    float fx = sx - flx, fy = sy - fly;
    int x0 = cvt_i32(flx), y0 = cvt_i32(fly);
    int x1 = (x0 == INT32_MAX) ? x0 : x0 + 1;
    int y1 = (y0 == INT32_MAX) ? y0 : y0 + 1;
    x0 = imax(0, imin(x0, t.w - 1)); x1 = imax(0, imin(x1, t.w - 1));
    y0 = imax(0, imin(y0, t.h - 1)); y1 = imax(0, imin(y1, t.h - 1));
    int r0 = t.h - 1 - y0, r1 = t.h - 1 - y1;
    vec4 t00 = texel_bytes(t, x0, r0), t10 = texel_bytes(t, x1, r0);
    vec4 t01 = texel_bytes(t, x0, r1), t11 = texel_bytes(t, x1, r1);
    vec4 top = cmix(t00, t10, fx);
    vec4 bot = cmix(t01, t11, fx);
    return cmix(top, bot, fy) * (1.0f / 255.0f);
}

AXR_HD vec4 sample(const DevTexture& t, vec2 uv) {
    if (!t.texels) return vec4{0.0f, 0.0f, 0.0f, 1.0f};
    return t.filter ? sample_bilinear(t, uv) : sample_nearest(t, uv);
}

// Shading-only math.  Colour is compared with the reference to within one
// 8-bit step, not bit for bit, so the lighting terms use the hardware's
// ~1-ulp reciprocal square root, log2 and exp2 (v_rsq_f32, v_log_f32,
// v_exp_f32) instead of correctly rounded sqrt, divide and powf: their error
// is about 1e-7 relative, four orders of magnitude below 1/255.  Texture
// coordinates and barycentrics stay exact (nearest-texel sampling is
// discontinuous in uv), and so does everything that decides coverage, depth
// and winners.
#if defined(__HIP_DEVICE_COMPILE__)
AXR_HD vec3 fnormalize(vec3 v) { return v * __builtin_amdgcn_rsqf(cdot(v, v)); }
AXR_HD float fpow(float x, float e) {
    // x >= 0 at every call site; x^0 = 1 (also for x = 0, as powf)
    return e == 0.0f ? 1.0f : __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}
#else
AXR_HD vec3 fnormalize(vec3 v) { return normalize(v); }
AXR_HD float fpow(float x, float e) { return powf(x, e); }
#endif

struct ShadeCtx {
    const Uniforms* u;
    const DevMaterial* m;
    const uint32_t* face_colors;
    uint32_t prim;     // mesh face index of the primitive
};

AXR_HD vec2 bary2(vec3 b, vec2 v0, vec2 v1, vec2 v2) { return (b.x * v0 + b.y * v1) + b.z * v2; }

struct FlatShader {
    struct Var { vec3 n; };
    static AXR_HD Var vs(const Vertex56& v, const Uniforms& u) {
        return Var{cmul(u.normal_mat, make3(v.nx, v.ny, v.nz))};
    }
    static AXR_HD vec4 fs(vec3 bar, const Var* t, const ShadeCtx& c) {
        vec3 n = fnormalize(cbary3(bar, t[0].n, t[1].n, t[2].n));
        float i = sclamp(cdot(-c.u->light_dir, n), 0.0f, 1.0f);
        return vec4{1.0f * i, 1.0f * i, 1.0f * i, 1.0f * i};
    }
};

// Varyings of the normal-mapped shaders (reference shaders.hpp:147-168).
struct TbnVar { vec2 uv; vec3 wp; vec3 t, b, n; };
AXR_HD TbnVar tbn_vs(const Vertex56& v, const Uniforms& u) {
    TbnVar o;
    o.uv = make2(v.u, v.v);
    o.wp = cpoint(u.model, v.px, v.py, v.pz);
    o.t = fnormalize(cdir(u.model, v.tx, v.ty, v.tz));
    o.b = fnormalize(cdir(u.model, v.bx, v.by, v.bz));
    o.n = fnormalize(cdir(u.model, v.nx, v.ny, v.nz));
    return o;
}
// Interpolated TBN columns: (b.x*M0 + b.y*M1) + b.z*M2, column by column.
AXR_HD void tbn_interp(vec3 bar, const TbnVar* t, vec3& iT, vec3& iB, vec3& iN) {
    iT = cbary3(bar, t[0].t, t[1].t, t[2].t);
    iB = cbary3(bar, t[0].b, t[1].b, t[2].b);
    iN = cbary3(bar, t[0].n, t[1].n, t[2].n);
}
// The Phong shaders' varyings: the same without the bitangent, which their
// fragment stage never reads (it rebuilds B as N x T), so the batch's vertex
// pass stores 44 bytes a vertex instead of 56 and the resolve loads as much
// less per winning pixel.
struct TnVar { vec2 uv; vec3 wp; vec3 t, n; };
AXR_HD TnVar tn_vs(const Vertex56& v, const Uniforms& u) {
    TnVar o;
    o.uv = make2(v.u, v.v);
    o.wp = cpoint(u.model, v.px, v.py, v.pz);
    o.t = fnormalize(cdir(u.model, v.tx, v.ty, v.tz));
    o.n = fnormalize(cdir(u.model, v.nx, v.ny, v.nz));
    return o;
}
// Normal-mapped world normal, albedo and position (reference shaders.hpp:172-218).
AXR_HD void phong_surface(vec3 bar, const TnVar* t, const DevMaterial& m, vec3& normal, vec4& albedo, vec3& fragPos) {
    vec2 uv = bary2(bar, t[0].uv, t[1].uv, t[2].uv);
    vec4 nmv = sample(m.bump, uv);
    vec3 nms = fnormalize(cmad(xyz(nmv), 2.0f, make3(-1.0f, -1.0f, -1.0f)));
    const vec3 iT = cbary3(bar, t[0].t, t[1].t, t[2].t);
    const vec3 iN = cbary3(bar, t[0].n, t[1].n, t[2].n);
    vec3 N = fnormalize(iN);
    vec3 T = fnormalize(cmad(N, -cdot(N, iT), iT));
    vec3 B = cross(N, T);
    mat3 f{{T, B, N}};
    albedo = sample(m.diffuse, uv);
    normal = fnormalize(cmul(f, nms));
    fragPos = cbary3(bar, t[0].wp, t[1].wp, t[2].wp);
}

struct PhongShader {
    using Var = TnVar;
    static AXR_HD Var vs(const Vertex56& v, const Uniforms& u) { return tn_vs(v, u); }
    static AXR_HD vec4 fs(vec3 bar, const Var* t, const ShadeCtx& c) {
        vec3 normal, fragPos; vec4 albedo;
        phong_surface(bar, t, *c.m, normal, albedo, fragPos);
        const Uniforms& u = *c.u;
        vec3 viewDir = fnormalize(u.camera_pos - fragPos);
        vec3 lightDir = -u.light_dir;
        float diff = smax(cdot(normal, lightDir), 0.0f);
        vec3 reflectDir = creflect(-lightDir, normal);
        float spec = fpow(smax(cdot(viewDir, reflectDir), 0.0f), c.m->specular_exponent * 50.0f);
        // ambient + diffuse + specular = light_color * (0.1 + diff + 0.5 spec)
        return make4((u.light_color * fmaf(0.5f, spec, 0.1f + diff)) * xyz(albedo), 1.0f);
    }
};

struct PhongPointShader {
    using Var = TnVar;
    static AXR_HD Var vs(const Vertex56& v, const Uniforms& u) { return tn_vs(v, u); }
    static AXR_HD vec4 fs(vec3 bar, const Var* t, const ShadeCtx& c) {
        vec3 normal, fragPos; vec4 albedo;
        phong_surface(bar, t, *c.m, normal, albedo, fragPos);
        const Uniforms& u = *c.u;
        vec3 viewDir = fnormalize(u.camera_pos - fragPos);
        const float e = c.m->specular_exponent * 50.0f;
        vec3 sum = make3(0.0f, 0.0f, 0.0f);
        for (int l = 0; l < u.num_lights; ++l) {
            vec3 L = fnormalize(u.light_pos[l] - fragPos);
            float diff = smax(cdot(normal, L), 0.0f);
            vec3 reflectDir = creflect(-L, normal);
            float spec = fpow(smax(cdot(viewDir, reflectDir), 0.0f), e);
            // ambient + diffuse + specular = colour * (0.1 + diff + 0.5 spec)
            sum = cmad(u.light_colors[l], fmaf(0.5f, spec, 0.1f + diff), sum);
        }
        return make4(sum * xyz(albedo), 1.0f);
    }
};

AXR_HD vec3 vpow(vec3 a, float e) { return make3(fpow(a.x, e), fpow(a.y, e), fpow(a.z, e)); }

struct PBRShader {
    using Var = TbnVar;
    static AXR_HD Var vs(const Vertex56& v, const Uniforms& u) { return tbn_vs(v, u); }
    static AXR_HD vec4 fs(vec3 bar, const Var* t, const ShadeCtx& c) {
        const float PI = 3.14159265358979323846f;
        const DevMaterial& m = *c.m;
        const Uniforms& u = *c.u;
        vec2 uv = bary2(bar, t[0].uv, t[1].uv, t[2].uv);
        vec4 nmv = sample(m.bump, uv);
        vec3 nms = fnormalize(cmad(xyz(nmv), 2.0f, make3(-1.0f, -1.0f, -1.0f)));
        vec3 iT, iB, iN;
        tbn_interp(bar, t, iT, iB, iN);
        mat3 itbn{{iT, iB, iN}};
        vec3 albedo = vpow(xyz(sample(m.diffuse, uv)), 2.2f);
        float metallic = sample(m.metallic, uv).x;
        float roughness = sample(m.roughness, uv).x;
        float ao = sample(m.ao, uv).x;
        float r2 = roughness * roughness;
        float r4 = r2 * r2;
        float omm = 1.0f - metallic;
        // The reference re-orthogonalises the frame but lights with the
        // interpolated one (shaders.hpp:316-343); only the latter matters.
        vec3 normal = fnormalize(cmul(itbn, nms));
        vec3 fragPos = cbary3(bar, t[0].wp, t[1].wp, t[2].wp);
        vec3 viewDir = fnormalize(u.camera_pos - fragPos);
        vec3 lightDir = -u.light_dir;
        vec3 halfway = fnormalize(lightDir + viewDir);
        float NdotH = smax(cdot(normal, halfway), 0.0f);
        float NdotV = smax(cdot(normal, viewDir), 0.0f);
        float NdotL = smax(cdot(normal, lightDir), 0.0f);
        vec3 c04 = make3(0.04f, 0.04f, 0.04f);
        vec3 F0 = c04 + metallic * (albedo - c04);
        float NdotH2 = NdotH * NdotH;
        float dp = NdotH2 * (r4 - 1.0f) + 1.0f;
        float NDF = r4 / ((PI * dp) * dp);
        float r = roughness + 1.0f;
        float k = (r * r) / 8.0f;
        float nvk = NdotV * (1.0f - k) + k;
        float nlk = NdotL * (1.0f - k) + k;
        float G = (NdotV / nvk) * (NdotL / nlk);
        float cosT = smax(cdot(halfway, normal), 0.0f);
        float omc = 1.0f - cosT;
        float term = omc * omc; term = term * term; term = term * omc;
        vec3 one = make3(1.0f, 1.0f, 1.0f);
        vec3 F = F0 + (one - F0) * term;
        vec3 numerator = (F * NDF) * G;
        float denom = ((4.0f * NdotV) * NdotL) + 0.0001f;
        vec3 specular = numerator / denom;
        vec3 kD = (one - F) * omm;
        vec3 diffuse = (kD * albedo) * (1.0f / PI);
        vec3 ambient = (make3(0.03f, 0.03f, 0.03f) * albedo) * ao;
        vec3 fc = ambient + (((diffuse + specular) * u.light_color) * NdotL);
        fc = fc / (fc + one);
        fc = vpow(fc, 1.0f / 2.2f);
        return make4(fc, 1.0f);
    }
};

struct GouraudShader {
    struct Var { float i; vec2 uv; };
    static AXR_HD Var vs(const Vertex56& v, const Uniforms& u) {
        vec3 N = fnormalize(cdir(u.model, v.nx, v.ny, v.nz));
        return Var{sclamp(cdot(-u.light_dir, N), 0.0f, 1.0f), make2(v.u, v.v)};
    }
    static AXR_HD vec4 fs(vec3 bar, const Var* t, const ShadeCtx& c) {
        float i = cbary1(bar, t[0].i, t[1].i, t[2].i);
        vec2 uv = bary2(bar, t[0].uv, t[1].uv, t[2].uv);
        vec4 albedo = sample(c.m->diffuse, uv);
        return make4(xyz(albedo) * (0.1f + i), 1.0f);
    }
};

struct FlatColorShader {
    struct Var { };
    static AXR_HD Var vs(const Vertex56&, const Uniforms&) { return Var{}; }
    static AXR_HD vec4 fs(vec3, const Var*, const ShadeCtx& c) {
        uint32_t k = c.face_colors ? gld(c.face_colors, c.prim) : 0xFFFFFFFFu;
        const float inv255 = 1.0f / 255.0f;
        return vec4{((float)(k & 255u) + 0.5f) * inv255, ((float)((k >> 8) & 255u) + 0.5f) * inv255,
                    ((float)((k >> 16) & 255u) + 0.5f) * inv255, ((float)(k >> 24) + 0.5f) * inv255};
    }
};

// (uint8_t)(clamp(c, 0, 1) * 255) with truncation (reference
// tiled_pipeline.cpp:577-580); NaN converts to 0.
AXR_HD uint32_t to_u8(float c) { return (uint32_t)cvt_i32(sclamp(c, 0.0f, 1.0f) * 255.0f) & 255u; }
// Colour in the framebuffer's byte order B, G, R, A (reference
// tiled_pipeline.cpp:1171-1174).
AXR_HD uint32_t pack_bgra(vec4 c) {
    return to_u8(c.z) | (to_u8(c.y) << 8) | (to_u8(c.x) << 16) | (to_u8(c.w) << 24);
}

}  // namespace axr
