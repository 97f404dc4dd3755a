// ORACLE - test infrastructure only (see omath.hpp).
//
// The restated CPU tiled pipeline.
//   ClippedVertex, MAX_CLIPPED_VERTS   reference: include/pipeline.hpp:13-17
//   clipTriangle / clipTriangleSinglePlane / interpolateVertices / distFunc
//                                      reference: src/pipeline.cpp:176-370
//   Triangle (setup, backface)         reference: include/tiled_pipeline.hpp:14-26,
//                                                 src/tiled_pipeline.cpp:15-119
//   TiledPipeline::drawMesh and phases reference: src/tiled_pipeline.cpp:143-425
//   AVX2 tile raster                   reference: src/tiled_pipeline.cpp:427-594
//   mergeTileResults                   reference: src/tiled_pipeline.cpp:1125-1181
//   ThreadPool                         reference: include/thread_pool.hpp:9-36,
//                                                 src/thread_pool.cpp:5-31
// Deliberate differences from the reference, none of which changes a pixel:
//   * the triangle store grows (the reference's fixed 16 MiB arena throws
//     past ~36k-72k triangles);
//   * per-thread clip outputs are cleared for every material group (the
//     reference appends earlier groups again; the copies lose every tie);
//   * each raster worker uses its own copy of the shader (the reference writes
//     the shared shader's material from every worker);
//   * the merge stops at the tile's right edge instead of touching 8 lanes.
// It also records the winning primitive key per pixel (submission rank * 8 +
// clip sub-triangle), a debugging channel the reference does not have.
#pragma once
#include <array>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <future>
#include <mutex>
#include <queue>
#include <thread>
#include <vector>
#include "oshaders.hpp"

namespace orc {

struct ClippedVertex {
    Vertex vertex;
    vec4 clipPos;
};
constexpr int MAX_CLIPPED_VERTS = 24;
constexpr int TILE_SIZE = 16;
constexpr int BATCH_SIZE = 8;

float clampW(float w);
float distFunc(const vec4& v, int plane);
ClippedVertex interpolateVertices(const ClippedVertex& a, const ClippedVertex& b, float t);
int clipTriangleSinglePlane(int plane, ClippedVertex& v0, ClippedVertex& v1, ClippedVertex& v2, ClippedVertex& v3);
void clipTriangle(size_t& count, std::array<ClippedVertex, MAX_CLIPPED_VERTS>& tris);

struct Triangle {
    vec2 screenPos[3];
    vec3 ndcZ;
    std::array<Vertex, 3> vertices;
    float minX, minY, maxX, maxY;
    const Material* material;
    uint32_t primitiveId;   // face index in the mesh
    uint32_t key;           // submission rank * 8 + sub-triangle
    Triangle() = default;
    Triangle(const ClippedVertex& a, const ClippedVertex& b, const ClippedVertex& c, int W, int H, const Material* m);
    static bool isBackface(const ClippedVertex& v0, const ClippedVertex& v1, const ClippedVertex& v2, int W, int H);
};

class ThreadPool {
public:
    explicit ThreadPool(size_t n);
    ~ThreadPool();
    std::future<void> enqueue(std::function<void()> f);
    size_t size() const { return m_Workers.size(); }
private:
    std::vector<std::thread> m_Workers;
    std::queue<std::function<void()>> m_Tasks;
    std::mutex m_Mutex;
    std::condition_variable m_Cv;
    std::atomic<bool> m_Stop{false};
};

struct PhaseTimes {
    // seconds, named after the reference's profiling zones
    double clipping = 0, initTiles = 0, binning = 0, tilePrep = 0, raster = 0, merging = 0;
};

class TiledPipeline {
public:
    TiledPipeline(size_t threads, Framebuffer* fb);
    void setShader(IShader* s) { m_Shader = s; }
    void setFramebuffer(Framebuffer* fb) { m_Framebuffer = fb; }
    void setViewProj(const mat4& vp, const mat4& viewport, const vec3& camPos) {
        m_ViewProj = vp; m_Viewport = viewport; m_CamPos = camPos; m_HasCamera = true;
    }
    void drawMesh(const mat4& model, const Mesh& mesh);
    const PhaseTimes& times() const { return m_Times; }
    size_t lastTriangleCount() const { return m_Triangles.size(); }

private:
    struct Tile { int startX, startY, endX, endY; std::vector<const Triangle*> triangles; };
    struct TileBuffers {
        float depth[TILE_SIZE * TILE_SIZE];
        uint8_t color[TILE_SIZE * TILE_SIZE * 4];
        uint32_t ids[TILE_SIZE * TILE_SIZE];
        void clear();
    };
    struct TileResult { int startX, startY, endX, endY; TileBuffers buf; };

    void clipGroup(const Mesh& mesh, const MaterialGroup& g, size_t rankBase, const mat4& mvp);
    void initializeTiles();
    void binTrianglesToTiles();
    void processTile(size_t i, TileBuffers& buf, IShader& shader);
    void rasterizeTriangleInTile(const Triangle& tri, const Tile& tile, TileBuffers& buf,
                                 IShader& shader, const VSTransformedTriangle& vs);
    void mergeTileResults();

    ThreadPool m_Pool;
    Framebuffer* m_Framebuffer;
    IShader* m_Shader = nullptr;
    mat4 m_ViewProj, m_Viewport;
    vec3 m_CamPos;
    bool m_HasCamera = false;
    std::vector<Triangle> m_Triangles;
    std::vector<std::vector<Triangle>> m_ThreadClipped;
    std::vector<Tile> m_Tiles, m_RelevantTiles;
    std::vector<TileResult> m_TileResults;
    std::atomic<size_t> m_NextBatch{0};
    PhaseTimes m_Times;
};

}  // namespace orc
