// ORACLE - test infrastructure only (see omath.hpp).
//
// C entry points over the restated pipeline, loaded with ctypes by the test
// suite, the smoke check and the CPU-baseline leg of bench.py.  Also exports
// the clipper and the triangle setup on their own, so that tests can check
// them against closed forms and invariants.
#include <cstring>
#include <memory>
#include <new>
#include "otiled.hpp"

using namespace orc;

extern "C" {

struct orc_draw_params {
    float model[16];
    float view_proj[16];
    float viewport[16];
    float camera_pos[3];
    int shader;  // 0 flat, 1 phong, 2 pbr, 3 gouraud, 4 phong_point, 5 flat_color
    float light_dir[3];
    float light_color[3];
    int num_lights;
    float light_pos[4][3];
    float light_colors[4][3];
};

struct orc_material_desc {
    void* diffuse;
    void* metallic;
    void* bump;
    void* roughness;
    void* ao;
    float specular_exponent;
};

void* orc_texture_create(const uint8_t* rgba, int w, int h, int filter) {
    return new (std::nothrow) Texture(rgba, w, h, filter ? Filter::Bilinear : Filter::Nearest);
}
void orc_texture_destroy(void* t) { delete static_cast<Texture*>(t); }

// groups: ngroups triples (material index, first face, face count)
void* orc_mesh_create(const float* verts56, size_t nverts, const uint32_t* idx, size_t nfaces,
                      const uint32_t* groups, int ngroups, const orc_material_desc* mats, int nmats,
                      const uint32_t* face_colors) {
    Mesh* m = new (std::nothrow) Mesh();
    if (!m) return nullptr;
    m->vertices.resize(nverts);
    std::memcpy(m->vertices.data(), verts56, nverts * sizeof(Vertex));
    m->indices.assign(idx, idx + nfaces * 3);
    for (int g = 0; g < ngroups; ++g)
        m->groups.push_back(MaterialGroup{groups[3 * g], groups[3 * g + 1], groups[3 * g + 2]});
    for (int i = 0; i < nmats; ++i) {
        Material mt;
        mt.diffuseTexture = static_cast<const Texture*>(mats[i].diffuse);
        mt.metallicTexture = static_cast<const Texture*>(mats[i].metallic);
        mt.bumpTexture = static_cast<const Texture*>(mats[i].bump);
        mt.roughnessTexture = static_cast<const Texture*>(mats[i].roughness);
        mt.aoTexture = static_cast<const Texture*>(mats[i].ao);
        mt.specularExponent = mats[i].specular_exponent;
        m->materials.push_back(mt);
    }
    if (face_colors) m->faceColors.assign(face_colors, face_colors + nfaces);
    return m;
}
void orc_mesh_destroy(void* m) { delete static_cast<Mesh*>(m); }

void* orc_fb_create(int w, int h) { return new (std::nothrow) Framebuffer(w, h); }
void orc_fb_destroy(void* fb) { delete static_cast<Framebuffer*>(fb); }
void orc_fb_clear(void* fb, uint32_t bgra, float depth) {
    auto* f = static_cast<Framebuffer*>(fb);
    f->clearColor(bgra);
    f->clearDepth(depth);
}
void orc_fb_read(void* fb, uint8_t* bgra, float* depth, uint32_t* ids) {
    auto* f = static_cast<Framebuffer*>(fb);
    size_t n = (size_t)f->width() * f->height();
    if (bgra) std::memcpy(bgra, f->color(), n * 4);
    if (depth) std::memcpy(depth, f->depth(), n * 4);
    if (ids) std::memcpy(ids, f->ids(), n * 4);
}

struct OracleCtx {
    std::unique_ptr<TiledPipeline> pipe;
};

void* orc_pipeline_create(int threads) {
    auto* c = new (std::nothrow) OracleCtx();
    if (!c) return nullptr;
    c->pipe = std::make_unique<TiledPipeline>((size_t)(threads > 0 ? threads : 1), nullptr);
    return c;
}
void orc_pipeline_destroy(void* p) { delete static_cast<OracleCtx*>(p); }

static vec3 v3(const float* f) { return vec3(f[0], f[1], f[2]); }

// Returns 0 on success, -1 on an unknown shader, -2 on a missing material.
int orc_draw(void* p, void* fb, void* mesh, const orc_draw_params* dp) {
    auto* c = static_cast<OracleCtx*>(p);
    std::unique_ptr<IShader> sh;
    switch (dp->shader) {
    case 0: { auto s = std::make_unique<FlatShader>(); s->lightDirection = v3(dp->light_dir); sh = std::move(s); break; }
    case 1: { auto s = std::make_unique<PhongShader>(); s->lightDirection = v3(dp->light_dir);
              s->lightColor = v3(dp->light_color); sh = std::move(s); break; }
    case 2: { auto s = std::make_unique<PBRShader>(); s->lightDirection = v3(dp->light_dir);
              s->lightColor = v3(dp->light_color); sh = std::move(s); break; }
    case 3: { auto s = std::make_unique<GouraudShader>(); s->lightDirection = v3(dp->light_dir); sh = std::move(s); break; }
    case 4: { auto s = std::make_unique<PhongPointShader>(); s->numLights = dp->num_lights < 4 ? dp->num_lights : 4;
              for (int l = 0; l < s->numLights; ++l) { s->lightPos[l] = v3(dp->light_pos[l]); s->lightColor[l] = v3(dp->light_colors[l]); }
              sh = std::move(s); break; }
    case 5: sh = std::make_unique<FlatColorShader>(); break;
    default: return -1;
    }
    c->pipe->setFramebuffer(static_cast<Framebuffer*>(fb));
    c->pipe->setShader(sh.get());
    c->pipe->setViewProj(mat4::from_cols(dp->view_proj), mat4::from_cols(dp->viewport), v3(dp->camera_pos));
    try {
        c->pipe->drawMesh(mat4::from_cols(dp->model), *static_cast<Mesh*>(mesh));
    } catch (const std::out_of_range&) {
        c->pipe->setShader(nullptr);
        return -2;
    }
    c->pipe->setShader(nullptr);
    return 0;
}

// clipping, culling, setup phase / init tiles / binning / tile prep / raster / merge
void orc_phase_times(void* p, double* out6) {
    const PhaseTimes& t = static_cast<OracleCtx*>(p)->pipe->times();
    out6[0] = t.clipping; out6[1] = t.initTiles; out6[2] = t.binning;
    out6[3] = t.tilePrep; out6[4] = t.raster; out6[5] = t.merging;
}
size_t orc_last_triangle_count(void* p) { return static_cast<OracleCtx*>(p)->pipe->lastTriangleCount(); }

// ---- unit-level entry points ----------------------------------------------
// A clipped vertex is 18 floats: the 14 vertex floats, then clip x, y, z, w.

static void load_cv(const float* f, ClippedVertex& v) {
    std::memcpy(&v.vertex, f, sizeof(Vertex));
    v.clipPos = vec4(f[14], f[15], f[16], f[17]);
}
static void store_cv(const ClippedVertex& v, float* f) {
    std::memcpy(f, &v.vertex, sizeof(Vertex));
    f[14] = v.clipPos.x; f[15] = v.clipPos.y; f[16] = v.clipPos.z; f[17] = v.clipPos.w;
}

// Clips one triangle (3 x 18 floats) against the six frustum planes; writes up
// to 24 vertices (8 triangles) and returns the number of vertices.
int orc_clip_triangle(const float* in3, float* out24) {
    std::array<ClippedVertex, MAX_CLIPPED_VERTS> t{};
    for (int i = 0; i < 3; ++i) load_cv(in3 + 18 * i, t[i]);
    size_t n = 3;
    clipTriangle(n, t);
    for (size_t i = 0; i < n; ++i) store_cv(t[i], out24 + 18 * i);
    return (int)n;
}

// One plane of the clipper: returns 0, 3 or 4 and writes v0..v3 (4 x 18 floats).
int orc_clip_single_plane(int plane, const float* in3, float* out4) {
    ClippedVertex v[4];
    for (int i = 0; i < 3; ++i) load_cv(in3 + 18 * i, v[i]);
    int r = clipTriangleSinglePlane(plane, v[0], v[1], v[2], v[3]);
    for (int i = 0; i < 4; ++i) store_cv(v[i], out4 + 18 * i);
    return r;
}

// Triangle setup from three clip positions (12 floats).  out13 = screen x,y
// of the three corners, ndc z of the three corners, minX, minY, maxX, maxY.
// Returns 1 when the triangle is back-facing.
int orc_setup_triangle(const float* clip12, int W, int H, float* out13) {
    ClippedVertex c[3];
    for (int i = 0; i < 3; ++i) c[i].clipPos = vec4(clip12[4 * i], clip12[4 * i + 1], clip12[4 * i + 2], clip12[4 * i + 3]);
    Triangle t(c[0], c[1], c[2], W, H, nullptr);
    for (int i = 0; i < 3; ++i) { out13[2 * i] = t.screenPos[i].x; out13[2 * i + 1] = t.screenPos[i].y; }
    out13[6] = t.ndcZ.x; out13[7] = t.ndcZ.y; out13[8] = t.ndcZ.z;
    out13[9] = t.minX; out13[10] = t.minY; out13[11] = t.maxX; out13[12] = t.maxY;
    return Triangle::isBackface(c[0], c[1], c[2], W, H) ? 1 : 0;
}

// Texture sample (filter 0 nearest, 1 bilinear) for the sampling tests.
void orc_texture_sample(void* t, float u, float v, float* out4) {
    vec4 c = static_cast<Texture*>(t)->sample(vec2(u, v));
    out4[0] = c.x; out4[1] = c.y; out4[2] = c.z; out4[3] = c.w;
}

// 4x4 helpers so tests can pin the matrix evaluation orders.
void orc_mat4_mul(const float* a, const float* b, float* out) {
    mat4 r = mat4::from_cols(a) * mat4::from_cols(b);
    for (int c = 0; c < 4; ++c) { out[4 * c] = r[c].x; out[4 * c + 1] = r[c].y; out[4 * c + 2] = r[c].z; out[4 * c + 3] = r[c].w; }
}
void orc_mat4_vec4(const float* m, const float* v, float* out) {
    vec4 r = mat4::from_cols(m) * vec4(v[0], v[1], v[2], v[3]);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

}  // extern "C"
