// example_drop_in.cpp - the reference renderer's frame loop (clear, drawMesh,
// read back) written against axiomr.hpp instead of the reference's
// TiledPipeline: an 800x600 UV sphere (100 x 100) with Gouraud shading over a
// bilinear 256^2 checker, camera at (0, 0, 3).  Prints the covered pixel
// count and an FNV-1a hash of the BGRA and depth planes.
//
//   example_drop_in --obj model.obj frame.png
// loads a model the way the reference's Renderer does (Mesh(path), its .mtl
// and textures; src/renderer.cpp:58-94), draws it with the Phong shader
// (light (0, -1, 0)) from (0, 0, 3) at 800x600 and writes the frame as PNG.
//
//   example_drop_in --fly N
// runs the camera half of the reference's frame loop (handleInput, then
// update(dt); src/renderer.cpp:122-123) for N frames of a scripted input
// (no GPU needed) and prints the final position and projection*view.
//
//   g++ -std=c++17 -O2 -I include -I axiomr_amd/cpp axiomr_amd/cpp/example_drop_in.cpp \
//       -L axiomr_amd -laxiomr_amd -Wl,-rpath,<dir of libaxiomr_amd.so> -o example_drop_in
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "axiomr.hpp"

static uint64_t fnv(const std::vector<uint8_t>& bgra, const std::vector<float>& depth, size_t* covered) {
    uint64_t h = 1469598103934665603ull;
    auto mixb = [&](const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    };
    mixb(bgra.data(), bgra.size());
    mixb(depth.data(), depth.size() * 4);
    *covered = 0;
    for (float d : depth) *covered += std::isfinite(d);
    return h;
}

static int draw_obj(const char* path, const char* png) {
    const int W = 800, H = 600;
    try {
        AR::Context ctx(0);
        AR::Mesh mesh(ctx, path);
        AR::Framebuffer fb(ctx, W, H);
        AR::Camera cam({0, 0, 3}, {0, 0, 0}, 60.0f, (float)W / H);
        cam.setViewport(0, 0, W, H);
        cam.update(0.0f);                // the frame loop's update builds the view
        AR::GpuTiledPipeline pipe(&cam, &fb);
        AR::PhongShader shader;
        shader.lightDirection = {0, -1, 0};
        pipe.setShader(&shader);
        fb.clearColor(AR::Color{0, 0, 0, 255});
        fb.clearDepth();
        pipe.drawMesh(AR::identity(), mesh);
        fb.writePNG(png);
        std::vector<uint8_t> bgra(W * H * 4);
        std::vector<float> depth(W * H);
        fb.read(bgra.data(), depth.data());
        size_t covered;
        const uint64_t h = fnv(bgra, depth, &covered);
        std::printf("covered %zu hash %016llx\n", covered, (unsigned long long)h);
    } catch (const AR::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}

// Frame i of the scripted input: the mouse drifts right and up, the right
// button is held except on every fifth frame, and W, W+D, A cycle.
static AR::InputState scripted_input(int i) {
    AR::InputState in;
    in.mouse = {100.0f + 3.0f * i, 200.0f - 2.0f * i};
    in.rightButton = (i % 5) != 0;
    in.keys = (i % 3 == 0) ? "W" : (i % 3 == 1) ? "WD" : "A";
    return in;
}

static int fly(int frames) {
    AR::Camera cam({0, 1, 5}, {0, 0, 0}, 60.0f, 800.0f / 600.0f);
    cam.setViewport(0, 0, 800, 600);
    for (int i = 0; i < frames; ++i) {
        cam.handleInput(scripted_input(i));
        cam.update(1.0f / 60.0f);
    }
    AR::vec3 p = cam.getPosition();
    std::printf("position %.9g %.9g %.9g\nview_proj", p.x, p.y, p.z);
    for (float v : cam.getViewProjectionMatrix()) std::printf(" %.9g", v);
    std::printf("\n");
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 3 && std::string(argv[1]) == "--obj") return draw_obj(argv[2], argv[3]);
    if (argc > 2 && std::string(argv[1]) == "--fly") return fly(std::atoi(argv[2]));
    int frames = argc > 1 ? std::atoi(argv[1]) : 1;
    const int W = 800, H = 600, stacks = 100, slices = 100;
    std::vector<AR::Vertex> verts((stacks + 1) * (slices + 1));
    std::vector<uint32_t> idx(6 * stacks * slices);
    if (axr_gen_uv_sphere(stacks, slices, 1.0f, reinterpret_cast<float*>(verts.data()), idx.data()) != AXR_OK) return 2;
    std::vector<uint8_t> checker(256 * 256 * 4);
    axr_gen_checker(256, 256, 8, checker.data());
    try {
        AR::Context ctx(0);
        AR::Texture tex(ctx, checker.data(), 256, 256, AR::Texture::Bilinear);
        AR::Material mat;
        mat.diffuseTexture = &tex;
        AR::Mesh mesh(ctx, verts, idx, {{0, 0, idx.size() / 3}}, {mat});
        AR::Framebuffer fb(ctx, W, H);
        AR::Camera cam({0, 0, 3}, {0, 0, 0}, 60.0f, (float)W / H);
        cam.setViewport(0, 0, W, H);
        cam.update(0.0f);                // the frame loop's update builds the view
        AR::GpuTiledPipeline pipe(&cam, &fb);
        AR::GouraudShader shader;
        shader.lightDirection = {0, -1, 0};
        pipe.setShader(&shader);
        for (int f = 0; f < frames; ++f) {
            fb.clearColor(AR::Color{0, 0, 0, 255});
            fb.clearDepth();
            pipe.drawMesh(AR::identity(), mesh);
        }
        std::vector<uint8_t> bgra(W * H * 4);
        std::vector<float> depth(W * H);
        fb.read(bgra.data(), depth.data());
        size_t covered;
        const uint64_t h = fnv(bgra, depth, &covered);
        std::printf("covered %zu hash %016llx\n", covered, (unsigned long long)h);
    } catch (const AR::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
