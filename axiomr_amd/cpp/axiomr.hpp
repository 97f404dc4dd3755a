// axiomr.hpp - C++ mirror of the reference renderer's pipeline interface over
// the C ABI of include/axiomr.h (header only; link libaxiomr_amd.so).
//
// Names and argument meaning follow the reference:
//   Pipeline / TiledPipeline   reference include/pipeline.hpp:18-48,
//                              include/tiled_pipeline.hpp:58-62
//   IShader and the shaders    reference include/IShader.hpp:30-46,
//                              include/shaders/shaders.hpp
//   Vertex, Material, MaterialGroup, Mesh   reference include/mesh.hpp
//   Texture, Framebuffer, Camera, Color     reference include/texture.hpp,
//                              include/framebuffer.hpp, include/camera.hpp,
//                              include/math.hpp:17-19
// Differences: matrices are column-major float[16] (the reference passes
// glm::mat4, whose memory layout is the same); resources live on the GPU and
// belong to a Context; drawMesh runs on the GPU and returns when the frame is
// drawn.  Like the reference, drawMesh does nothing without a shader, camera
// or framebuffer; failures of the device path throw AR::Error.
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "axiomr.h"

namespace AR {

using mat4 = std::array<float, 16>;   // column-major
struct vec2 { float x, y; };
struct vec3 { float x, y, z; };

inline mat4 identity() { return mat4{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; }

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

class Context {
public:
    explicit Context(int device = 0, uint64_t hbmBudget = 0) {
        int r = axr_ctx_create(device, hbmBudget, &m_Ctx);
        if (r != AXR_OK) throw Error(r, "axr_ctx_create failed");
    }
    ~Context() { if (m_Ctx) axr_ctx_destroy(m_Ctx); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    axr_ctx* handle() const { return m_Ctx; }
    void check(int r) const { if (r != AXR_OK) throw Error(r, axr_last_error(m_Ctx)); }
    void sync() const { check(axr_sync(m_Ctx)); }
private:
    axr_ctx* m_Ctx = nullptr;
};

struct Color { unsigned char r, g, b, a; };

class Texture {
public:
    enum Filter { Nearest = AXR_FILTER_NEAREST, Bilinear = AXR_FILTER_BILINEAR };
    // rgba: width * height * 4 bytes, row 0 first (what stb_image returns)
    Texture(Context& ctx, const uint8_t* rgba, int width, int height, Filter f = Nearest) : m_Ctx(&ctx) {
        ctx.check(axr_texture_create(ctx.handle(), rgba, width, height, f, &m_Tex));
        m_W = width;
        m_H = height;
    }
    // reference Texture(const std::string& filePath) (src/texture.cpp:6-35):
    // PNG or binary PPM/PGM, decoded to RGBA8 with row 0 first
    Texture(Context& ctx, const std::string& filePath, Filter f = Nearest) : m_Ctx(&ctx) {
        uint8_t* rgba = nullptr;
        int r = axr_image_load(filePath.c_str(), &m_W, &m_H, &rgba);
        if (r != AXR_OK) throw Error(r, axr_image_last_error());
        r = axr_texture_create(ctx.handle(), rgba, m_W, m_H, f, &m_Tex);
        axr_image_free(rgba);
        ctx.check(r);
    }
    ~Texture() { if (m_Tex) axr_texture_destroy(m_Tex); }
    Texture(const Texture&) = delete;
    Texture& operator=(const Texture&) = delete;
    int getWidth() const { return m_W; }
    int getHeight() const { return m_H; }
    axr_texture* handle() const { return m_Tex; }
private:
    Context* m_Ctx;
    axr_texture* m_Tex = nullptr;
    int m_W = 0, m_H = 0;
};

struct Vertex {                      // reference include/mesh.hpp:9-18
    float position[3];
    float uv[2];
    float normal[3];
    float tangent[3];
    float bitangent[3];
};
static_assert(sizeof(Vertex) == 56, "Vertex is 14 floats");

struct Material {                    // reference include/mesh.hpp:20-34
    const Texture* diffuseTexture = nullptr;
    const Texture* metallicTexture = nullptr;
    const Texture* bumpTexture = nullptr;
    const Texture* roughnessTexture = nullptr;
    const Texture* aoTexture = nullptr;
    float specularExponent = 1.0f;
};

struct MaterialGroup {               // reference include/mesh.hpp:35-39
    uint32_t material;               // index into the mesh's materials
    size_t startIndex;
    size_t faceCount;
};

class Mesh {
public:
    Mesh(Context& ctx, const std::vector<Vertex>& vertices, const std::vector<uint32_t>& indices,
         const std::vector<MaterialGroup>& groups, const std::vector<Material>& materials,
         const std::vector<uint32_t>& faceColors = {}) {
        std::vector<axr_group> g;
        for (const auto& x : groups) g.push_back(axr_group{x.material, (uint32_t)x.startIndex, (uint32_t)x.faceCount});
        std::vector<axr_material> m;
        auto h = [](const Texture* t) { return t ? t->handle() : nullptr; };
        for (const auto& x : materials)
            m.push_back(axr_material{h(x.diffuseTexture), h(x.metallicTexture), h(x.bumpTexture), h(x.roughnessTexture),
                                     h(x.aoTexture), x.specularExponent});
        ctx.check(axr_mesh_create(ctx.handle(), reinterpret_cast<const float*>(vertices.data()), vertices.size(),
                                  indices.data(), indices.size() / 3, g.data(), (uint32_t)g.size(), m.data(),
                                  (uint32_t)m.size(), faceColors.empty() ? nullptr : faceColors.data(), &m_Mesh));
        m_Faces = indices.size() / 3;
    }
    // reference Mesh(const std::string& filePath) (src/mesh.cpp:8-27): the OBJ,
    // the materials of <dir>/<basename>.mtl and their textures (owned by the
    // mesh).  A group naming a material the .mtl lacks throws, as the
    // reference's getMaterial (.at()) does; a file without usemtl draws nothing.
    Mesh(Context& ctx, const std::string& filePath, Texture::Filter f = Texture::Nearest) {
        axr_obj* obj = nullptr;
        int r = axr_obj_load(filePath.c_str(), &obj);
        if (r != AXR_OK) throw Error(r, axr_obj_last_error());
        struct Guard { axr_obj* o; ~Guard() { axr_obj_destroy(o); } } guard{obj};
        uint64_t nv = 0, nf = 0;
        uint32_t ng = 0, nm = 0;
        axr_obj_counts(obj, &nv, &nf, &ng, &nm);
        std::vector<float> v(nv * 14);
        std::vector<uint32_t> idx(nf * 3);
        axr_obj_arrays(obj, v.data(), idx.data());
        std::vector<axr_material> mats;
        for (uint32_t i = 0; i < nm; ++i) {
            axr_obj_material om;
            axr_obj_get_material(obj, i, &om);
            auto tex = [&](const char* path) -> axr_texture* {
                if (!path[0]) return nullptr;
                m_Textures.emplace_back(new Texture(ctx, std::string(path), f));
                return m_Textures.back()->handle();
            };
            mats.push_back(axr_material{tex(om.diffuse_map), tex(om.metallic_map), tex(om.bump_map),
                                        tex(om.roughness_map), tex(om.ao_map), om.specular_exponent});
        }
        std::vector<axr_group> groups;
        for (uint32_t i = 0; i < ng; ++i) {
            axr_obj_group og;
            axr_obj_get_group(obj, i, &og);
            if (og.material_index == UINT32_MAX)
                throw Error(AXR_ERR_INVALID, std::string("material '") + og.material + "' is not in the .mtl file");
            groups.push_back(axr_group{og.material_index, (uint32_t)og.first_face, (uint32_t)og.face_count});
        }
        if (mats.empty()) mats.push_back(axr_material{nullptr, nullptr, nullptr, nullptr, nullptr, 1.0f});
        if (groups.empty()) groups.push_back(axr_group{0, 0, 0});
        ctx.check(axr_mesh_create(ctx.handle(), v.data(), nv, idx.data(), nf, groups.data(), (uint32_t)groups.size(),
                                  mats.data(), (uint32_t)mats.size(), nullptr, &m_Mesh));
        m_Faces = nf;
    }
    ~Mesh() { if (m_Mesh) axr_mesh_destroy(m_Mesh); }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    axr_mesh* handle() const { return m_Mesh; }
    size_t faceCount() const { return m_Faces; }
private:
    axr_mesh* m_Mesh = nullptr;
    size_t m_Faces = 0;
    std::vector<std::unique_ptr<Texture>> m_Textures;
};

class Framebuffer {
public:
    Framebuffer(Context& ctx, int width, int height, bool withIds = false) : m_Ctx(&ctx), m_W(width), m_H(height) {
        ctx.check(axr_fb_create(ctx.handle(), width, height, withIds ? 1 : 0, &m_Fb));
    }
    ~Framebuffer() { if (m_Fb) axr_fb_destroy(m_Fb); }
    Framebuffer(const Framebuffer&) = delete;
    Framebuffer& operator=(const Framebuffer&) = delete;
    // reference src/framebuffer.cpp:26-42
    void clearColor(const Color& c) {
        m_Clear = (uint32_t(c.a) << 24) | (uint32_t(c.r) << 16) | (uint32_t(c.g) << 8) | uint32_t(c.b);
        m_Ctx->check(axr_fb_clear(m_Fb, m_Clear, m_Depth));
    }
    void clearDepth(float d = __builtin_huge_valf()) {
        m_Depth = d;
        m_Ctx->check(axr_fb_clear(m_Fb, m_Clear, m_Depth));
    }
    uint32_t getWidth() const { return (uint32_t)m_W; }
    uint32_t getHeight() const { return (uint32_t)m_H; }
    // BGRA8 (4 bytes per pixel) and FP32 depth, row 0 = bottom; either may be null
    void read(uint8_t* bgra, float* depth, uint32_t* ids = nullptr) const {
        m_Ctx->check(axr_fb_read(m_Fb, bgra, depth, ids));
    }
    // the colour plane as an RGBA PNG, top row first
    void writePNG(const std::string& path) const { m_Ctx->check(axr_fb_write_png(m_Fb, path.c_str())); }
    axr_fb* handle() const { return m_Fb; }
    Context& context() const { return *m_Ctx; }
private:
    Context* m_Ctx;
    axr_fb* m_Fb = nullptr;
    int m_W, m_H;
    uint32_t m_Clear = 0;
    float m_Depth = __builtin_huge_valf();
};

// What the reference's Window reports to Camera::handleInput
// (src/camera.cpp:100-127).  Camera::handleInput takes this or any type with
// the same three calls, such as the reference's own Window.
struct InputState {
    vec2 mouse{0, 0};
    bool rightButton = false;
    std::string keys;                // held keys among "WSAD"
    bool isMouseButtonDown(int button) const { return button == 1 && rightButton; }
    vec2 getMousePos() const { return mouse; }
    bool isKeyDown(char key) const { return keys.find(key) != std::string::npos; }
};

// reference Camera (include/camera.hpp, src/camera.cpp) over the native
// axr_camera record.  As in the reference, the view is identity until the
// first update(), and setFov/setAspectRatio do not rebuild the projection.
class Camera {
public:
    Camera(vec3 position = {0, 0, 0}, vec3 target = {0, 0, -1}, float fov = 60.0f, float aspectRatio = 16.0f / 9.0f) {
        float p[3] = {position.x, position.y, position.z}, t[3] = {target.x, target.y, target.z};
        if (axr_camera_init(&m_C, p, t, fov, aspectRatio) != AXR_OK)
            throw Error(AXR_ERR_INVALID, "Camera: position and target must be finite");
    }
    void setPosition(vec3 p) { float v[3] = {p.x, p.y, p.z}; axr_camera_set_position(&m_C, v); }
    void setTarget(vec3 t) { float v[3] = {t.x, t.y, t.z}; axr_camera_set_target(&m_C, v); }
    void setFov(float fov) { axr_camera_set_fov(&m_C, fov); }
    void setAspectRatio(float a) { axr_camera_set_aspect_ratio(&m_C, a); }
    void setViewport(int x, int y, int w, int h) { axr_camera_set_viewport(&m_C, x, y, w, h); }
    template <class Window>
    void handleInput(const Window& window) {
        axr_camera_input in{};
        auto m = window.getMousePos();
        in.mouse_x = m.x;
        in.mouse_y = m.y;
        in.right_button = window.isMouseButtonDown(1) ? 1 : 0;
        in.keys = (window.isKeyDown('W') ? AXR_KEY_W : 0u) | (window.isKeyDown('S') ? AXR_KEY_S : 0u) |
                  (window.isKeyDown('A') ? AXR_KEY_A : 0u) | (window.isKeyDown('D') ? AXR_KEY_D : 0u);
        axr_camera_handle_input(&m_C, &in);
    }
    void update(float deltaTime) { axr_camera_update(&m_C, deltaTime); }
    mat4 getViewProjectionMatrix() const { return load(m_C.view_proj); }
    mat4 getViewMatrix() const { return load(m_C.view); }
    mat4 getProjectionMatrix() const { return load(m_C.proj); }
    mat4 getViewportMatrix() const { mat4 m; axr_camera_viewport_matrix(&m_C, m.data()); return m; }
    vec3 getPosition() const { return {m_C.position[0], m_C.position[1], m_C.position[2]}; }
    const axr_camera& state() const { return m_C; }
private:
    static mat4 load(const float* f) { mat4 m; std::copy(f, f + 16, m.data()); return m; }
    axr_camera m_C{};
};

// reference IShader: the concrete type selects the device shader; the public
// fields become the draw's uniforms.
struct IShader {
    virtual ~IShader() = default;
    virtual void fill(axr_draw_params& p) const = 0;
};
inline void put3(float* d, vec3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

struct FlatShader : IShader {        // reference shaders.hpp:19-62
    vec3 lightDirection{0, -1, 0};
    void fill(axr_draw_params& p) const override { p.shader = AXR_SHADER_FLAT; put3(p.light_dir, lightDirection); }
};
struct PhongShader : IShader {       // reference shaders.hpp:140-247
    vec3 lightDirection{0, -1, 0}, lightColor{0.6f, 0.6f, 0.6f};
    void fill(axr_draw_params& p) const override {
        p.shader = AXR_SHADER_PHONG; put3(p.light_dir, lightDirection); put3(p.light_color, lightColor);
    }
};
struct PBRShader : IShader {         // reference shaders.hpp:252-423
    vec3 lightDirection{0, -1, 0}, lightColor{5.6f, 5.6f, 5.6f};
    void fill(axr_draw_params& p) const override {
        p.shader = AXR_SHADER_PBR; put3(p.light_dir, lightDirection); put3(p.light_color, lightColor);
    }
};
struct GouraudShader : IShader {     // no reference counterpart
    vec3 lightDirection{0, -1, 0};
    void fill(axr_draw_params& p) const override { p.shader = AXR_SHADER_GOURAUD; put3(p.light_dir, lightDirection); }
};
struct PhongPointShader : IShader {  // no reference counterpart
    struct Light { vec3 position, color; };
    std::vector<Light> lights;       // up to 4
    void fill(axr_draw_params& p) const override {
        p.shader = AXR_SHADER_PHONG_POINT;
        p.num_lights = (int32_t)std::min<size_t>(lights.size(), 4);
        for (int i = 0; i < p.num_lights; ++i) { put3(p.light_pos[i], lights[i].position); put3(p.light_colors[i], lights[i].color); }
    }
};
struct FlatColorShader : IShader {   // no reference counterpart
    void fill(axr_draw_params& p) const override { p.shader = AXR_SHADER_FLAT_COLOR; }
};

class Pipeline {                     // reference include/pipeline.hpp:18-48
public:
    Pipeline(Camera* cam, Framebuffer* fb) : m_Framebuffer(fb), m_Camera(cam) {}
    virtual ~Pipeline() = default;
    void setShader(IShader* shader) { m_Shader = shader; }
    void setCamera(Camera* cam) { m_Camera = cam; }
    void setFramebuffer(Framebuffer* fb) { m_Framebuffer = fb; }
    virtual void drawMesh(const mat4& modelMatrix, const Mesh& mesh) = 0;
    mat4 getViewportMat() { return m_Camera ? m_Camera->getViewportMatrix() : identity(); }
protected:
    IShader* m_Shader = nullptr;
    Framebuffer* m_Framebuffer = nullptr;
    const Camera* m_Camera = nullptr;
};

// The GPU drop-in for the reference's TiledPipeline::drawMesh.
class GpuTiledPipeline : public Pipeline {
public:
    GpuTiledPipeline(Camera* cam, Framebuffer* fb) : Pipeline(cam, fb) {}
    void drawMesh(const mat4& modelMatrix, const Mesh& mesh) override {
        if (!m_Shader || !m_Camera || !m_Framebuffer) return;   // reference tiled_pipeline.cpp:146
        axr_draw_params p;
        std::memset(&p, 0, sizeof(p));
        std::memcpy(p.model, modelMatrix.data(), 64);
        std::memcpy(p.view_proj, m_Camera->getViewProjectionMatrix().data(), 64);
        mat4 vp = m_Camera->getViewportMatrix();
        std::memcpy(p.viewport, vp.data(), 64);
        put3(p.camera_pos, m_Camera->getPosition());
        m_Shader->fill(p);
        Context& ctx = m_Framebuffer->context();
        ctx.check(axr_draw(ctx.handle(), m_Framebuffer->handle(), mesh.handle(), &p));
    }
};

}  // namespace AR
