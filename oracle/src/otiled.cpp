// ORACLE - test infrastructure only (see omath.hpp).  The restated CPU tiled
// pipeline; the header lists the reference lines each part follows.
#include "otiled.hpp"
#include <immintrin.h>
#include <chrono>
#include <cmath>
#include <stdexcept>

namespace orc {

namespace {
using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// (uint8_t)(std::clamp(c, 0, 1) * 255) as the reference's x86 build computes it
// (tiled_pipeline.cpp:577-580); a NaN channel converts through INT_MIN to 0.
inline uint8_t to_u8(float c) { return (uint8_t)(uint32_t)cvt_trunc_i32(sclamp(c, 0.0f, 1.0f) * 255.0f); }
}  // namespace

// ---- clipping (reference src/pipeline.cpp:176-370) ------------------------

float clampW(float w) {  // reference tiled_pipeline.cpp:15-23
    const float tiny = 1e-6f;
    if (std::fabs(w) < tiny) w = (w < 0.0f) ? -tiny : tiny;
    return w;
}

float distFunc(const vec4& v, int plane) {  // reference pipeline.cpp:274-285
    switch (plane) {
    case 0: return v.x + v.w;
    case 1: return v.w - v.x;
    case 2: return v.y + v.w;
    case 3: return v.w - v.y;
    case 4: return v.z + v.w;
    case 5: return v.w - v.z;
    default: return 0.0f;
    }
}

ClippedVertex interpolateVertices(const ClippedVertex& a, const ClippedVertex& b, float t) {  // pipeline.cpp:243-272
    ClippedVertex o;
    o.vertex.position = mix(a.vertex.position, b.vertex.position, t);
    o.vertex.normal = normalize(mix(a.vertex.normal, b.vertex.normal, t));
    o.vertex.tangent = normalize(mix(a.vertex.tangent, b.vertex.tangent, t));
    o.vertex.bitangent = normalize(mix(a.vertex.bitangent, b.vertex.bitangent, t));
    o.clipPos = mix(a.clipPos, b.clipPos, t);
    float w0 = a.clipPos.w, w1 = b.clipPos.w;
    vec2 uvw = mix(a.vertex.uv * w0, b.vertex.uv * w1, t);
    float iw = mix(w0, w1, t);
    o.vertex.uv = uvw / iw;
    return o;
}

static inline float edgeT(float din, float dout) {
    float denom = din - dout;
    return (std::fabs(denom) < 1e-7f) ? 0.5f : (din / denom);
}

int clipTriangleSinglePlane(int plane, ClippedVertex& v0, ClippedVertex& v1, ClippedVertex& v2, ClippedVertex& v3) {
    float d0 = distFunc(v0.clipPos, plane), d1 = distFunc(v1.clipPos, plane), d2 = distFunc(v2.clipPos, plane);
    if (d0 < 0.f && d1 < 0.f && d2 < 0.f) return 0;
    if (d0 >= 0.f && d1 >= 0.f && d2 >= 0.f) { v3 = v0; return 3; }
    const bool in0 = d0 >= 0.f, in1 = d1 >= 0.f, in2 = d2 >= 0.f;
    // Cycle the corners (keeping the winding) until corner 0 is inside.
    if (in1 && !in0) {
        std::swap(v0, v1); std::swap(d0, d1);
        std::swap(v1, v2); std::swap(d1, d2);
    } else if (in2 && !in1) {
        std::swap(v2, v1); std::swap(d2, d1);
        std::swap(v1, v0); std::swap(d1, d0);
    }
    v3 = interpolateVertices(v0, v2, edgeT(d0, d2));
    if (d1 >= 0.f) {
        v2 = interpolateVertices(v1, v2, edgeT(d1, d2));
        return 4;
    }
    v1 = interpolateVertices(v0, v1, edgeT(d0, d1));
    v2 = v3;
    return 3;
}

void clipTriangle(size_t& count, std::array<ClippedVertex, MAX_CLIPPED_VERTS>& tris) {
    std::array<ClippedVertex, MAX_CLIPPED_VERTS> next{};
    int n = 0;
    for (int plane = 0; plane < 6; ++plane) {
        n = 0;
        for (int i = 0; i < (int)count; i += 3) {
            ClippedVertex v3;
            int r = clipTriangleSinglePlane(plane, tris[i], tris[i + 1], tris[i + 2], v3);
            if (r == 3) {
                if (n + 3 <= MAX_CLIPPED_VERTS) {
                    next[n] = tris[i]; next[n + 1] = tris[i + 1]; next[n + 2] = tris[i + 2]; n += 3;
                }
            } else if (r == 4) {
                if (n + 6 <= MAX_CLIPPED_VERTS) {
                    next[n] = tris[i]; next[n + 1] = tris[i + 1]; next[n + 2] = tris[i + 2];
                    next[n + 3] = tris[i]; next[n + 4] = tris[i + 2]; next[n + 5] = v3;
                    n += 6;
                }
            }
        }
        count = (size_t)n;
        tris.swap(next);
        if (count == 0) break;
    }
}

// ---- triangle setup (reference tiled_pipeline.cpp:53-119) -----------------

static inline vec2 toScreen(const vec4& c, float invW, int W, int H) {
    return vec2((((c.x * invW) + 1.0f) * 0.5f) * (float)W, (((c.y * invW) + 1.0f) * 0.5f) * (float)H);
}

Triangle::Triangle(const ClippedVertex& a, const ClippedVertex& b, const ClippedVertex& c, int W, int H, const Material* m)
    : material(m), primitiveId(0), key(0) {
    const ClippedVertex* v[3] = {&a, &b, &c};
    for (int i = 0; i < 3; ++i) {
        vertices[i] = v[i]->vertex;
        float invW = 1.0f / clampW(v[i]->clipPos.w);
        comp(ndcZ)[i] = v[i]->clipPos.z * invW;
        screenPos[i] = toScreen(v[i]->clipPos, invW, W, H);
    }
    auto mn = [](float p, float q, float r) { float s = p; if (q < s) s = q; if (r < s) s = r; return s; };
    auto mx = [](float p, float q, float r) { float s = p; if (s < q) s = q; if (s < r) s = r; return s; };
    minX = mn(screenPos[0].x, screenPos[1].x, screenPos[2].x);
    minY = mn(screenPos[0].y, screenPos[1].y, screenPos[2].y);
    maxX = mx(screenPos[0].x, screenPos[1].x, screenPos[2].x);
    maxY = mx(screenPos[0].y, screenPos[1].y, screenPos[2].y);
}

bool Triangle::isBackface(const ClippedVertex& v0, const ClippedVertex& v1, const ClippedVertex& v2, int W, int H) {
    vec2 p0 = toScreen(v0.clipPos, 1.0f / clampW(v0.clipPos.w), W, H);
    vec2 p1 = toScreen(v1.clipPos, 1.0f / clampW(v1.clipPos.w), W, H);
    vec2 p2 = toScreen(v2.clipPos, 1.0f / clampW(v2.clipPos.w), W, H);
    float dx1 = p1.x - p0.x, dy1 = p1.y - p0.y;
    float dx2 = p2.x - p0.x, dy2 = p2.y - p0.y;
    return (dx1 * dy2 - dx2 * dy1) < 0;
}

// ---- thread pool (reference thread_pool.hpp:9-36, thread_pool.cpp:5-31) ---

ThreadPool::ThreadPool(size_t n) {
    for (size_t i = 0; i < n; ++i) {
        m_Workers.emplace_back([this] {
            for (;;) {
                std::function<void()> task;
                {
                    std::unique_lock<std::mutex> lk(m_Mutex);
                    m_Cv.wait(lk, [this] { return m_Stop.load() || !m_Tasks.empty(); });
                    if (m_Stop.load() && m_Tasks.empty()) return;
                    task = std::move(m_Tasks.front());
                    m_Tasks.pop();
                }
                task();
            }
        });
    }
}
ThreadPool::~ThreadPool() {
    m_Stop.store(true);
    m_Cv.notify_all();
    for (auto& w : m_Workers) w.join();
}
std::future<void> ThreadPool::enqueue(std::function<void()> f) {
    auto task = std::make_shared<std::packaged_task<void()>>(std::move(f));
    std::future<void> fut = task->get_future();
    {
        std::lock_guard<std::mutex> lk(m_Mutex);
        if (m_Stop.load()) throw std::runtime_error("enqueue on stopped ThreadPool");
        m_Tasks.emplace([task] { (*task)(); });
    }
    m_Cv.notify_one();
    return fut;
}

// ---- tiled pipeline (reference tiled_pipeline.cpp:143-425) ----------------

TiledPipeline::TiledPipeline(size_t threads, Framebuffer* fb)
    : m_Pool(threads ? threads : 1), m_Framebuffer(fb), m_ThreadClipped(threads ? threads : 1) {}

void TiledPipeline::TileBuffers::clear() {
    for (int i = 0; i < TILE_SIZE * TILE_SIZE; ++i) { depth[i] = std::numeric_limits<float>::infinity(); ids[i] = 0xFFFFFFFFu; }
    std::memset(color, 0, sizeof(color));
}

void TiledPipeline::clipGroup(const Mesh& mesh, const MaterialGroup& g, size_t rankBase, const mat4& mvp) {
    const Material* material = &mesh.materials.at(g.material);
    const size_t N = m_Pool.size();
    const int perThread = (int)((g.faceCount + N - 1) / N);
    const int W = m_Framebuffer->width(), H = m_Framebuffer->height();
    std::vector<std::future<void>> futs;
    for (size_t t = 0; t < N; ++t) {
        m_ThreadClipped[t].clear();
        futs.push_back(m_Pool.enqueue([&, t] {
            int start = (int)g.startIndex + (int)t * perThread;
            int end = std::min(start + perThread, (int)(g.startIndex + g.faceCount));
            std::vector<Triangle>& out = m_ThreadClipped[t];
            std::array<ClippedVertex, MAX_CLIPPED_VERTS> cv{};
            for (int f = start; f < end; ++f) {
                const uint32_t* ix = &mesh.indices[(size_t)f * 3];
                for (int k = 0; k < 3; ++k) {
                    const Vertex& v = mesh.vertices[ix[k]];
                    cv[k].vertex = v;
                    cv[k].clipPos = mvp * vec4(v.position, 1.0f);
                }
                size_t count = 3;
                clipTriangle(count, cv);
                const uint32_t rank = (uint32_t)(rankBase + (size_t)(f - (int)g.startIndex));
                for (size_t j = 0; j < count; j += 3) {
                    if (Triangle::isBackface(cv[j], cv[j + 1], cv[j + 2], W, H)) continue;
                    Triangle tri(cv[j], cv[j + 1], cv[j + 2], W, H, material);
                    tri.primitiveId = (uint32_t)f;
                    tri.key = rank * 8u + (uint32_t)(j / 3);
                    out.push_back(tri);
                }
            }
        }));
    }
    for (auto& f : futs) f.get();
    for (auto& v : m_ThreadClipped) m_Triangles.insert(m_Triangles.end(), v.begin(), v.end());
}

void TiledPipeline::initializeTiles() {
    const int W = m_Framebuffer->width(), H = m_Framebuffer->height();
    const int nx = (W + TILE_SIZE - 1) / TILE_SIZE, ny = (H + TILE_SIZE - 1) / TILE_SIZE;
    m_Tiles.clear();
    m_Tiles.reserve((size_t)nx * ny);
    for (int ty = 0; ty < ny; ++ty)
        for (int tx = 0; tx < nx; ++tx)
            m_Tiles.push_back(Tile{tx * TILE_SIZE, ty * TILE_SIZE, std::min((tx + 1) * TILE_SIZE, W),
                                   std::min((ty + 1) * TILE_SIZE, H), {}});
}

void TiledPipeline::binTrianglesToTiles() {  // reference tiled_pipeline.cpp:356-393
    const int W = m_Framebuffer->width();
    const int nx = (W + TILE_SIZE - 1) / TILE_SIZE;
    const int ny = (int)m_Tiles.size() / nx;
    for (const Triangle& t : m_Triangles) {
        int tx0 = std::max(0, cvt_trunc_i32(t.minX) / TILE_SIZE);
        int ty0 = std::max(0, cvt_trunc_i32(t.minY) / TILE_SIZE);
        int tx1 = std::min(nx - 1, cvt_trunc_i32(t.maxX) / TILE_SIZE);
        int ty1 = std::min(ny - 1, cvt_trunc_i32(t.maxY) / TILE_SIZE);
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) {
                Tile& tile = m_Tiles[(size_t)ty * nx + tx];
                bool miss = t.maxX < (float)tile.startX || t.minX > (float)tile.endX ||
                            t.maxY < (float)tile.startY || t.minY > (float)tile.endY;
                if (!miss) tile.triangles.push_back(&t);
            }
    }
}

void TiledPipeline::rasterizeTriangleInTile(const Triangle& tri, const Tile& tile, TileBuffers& buf,
                                            IShader& shader, const VSTransformedTriangle& vs) {
    // reference tiled_pipeline.cpp:427-594 (the AVX2 path that drawMesh uses)
    const int sx = std::max(tile.startX, cvt_trunc_i32(std::floor(tri.minX)));
    const int sy = std::max(tile.startY, cvt_trunc_i32(std::floor(tri.minY)));
    const int ex = std::min(tile.endX, cvt_trunc_i32(std::ceil(tri.maxX)));
    const int ey = std::min(tile.endY, cvt_trunc_i32(std::ceil(tri.maxY)));
    if (sx >= ex || sy >= ey) return;

    const float x0 = tri.screenPos[0].x, y0 = tri.screenPos[0].y;
    const float x1 = tri.screenPos[1].x, y1 = tri.screenPos[1].y;
    const float x2 = tri.screenPos[2].x, y2 = tri.screenPos[2].y;
    float c0 = x1 * y2 - x2 * y1;
    float c1 = x2 * y0 - x0 * y2;
    float c2 = x0 * y1 - x1 * y0;
    float area = (c0 + c1) + c2;
    if (area >= 0 && (double)area < 1.0E-12) return;
    float a0 = y1 - y2, b0 = x2 - x1;
    float a1 = y2 - y0, b1 = x0 - x2;
    float a2 = y0 - y1, b2 = x1 - x0;
    if (area < 0) {
        a0 = -a0; b0 = -b0; c0 = -c0;
        a1 = -a1; b1 = -b1; c1 = -c1;
        a2 = -a2; b2 = -b2; c2 = -c2;
        area = -area;
    }
    const float invArea = 1.0f / area;
    const __m256 va0 = _mm256_set1_ps(a0), va1 = _mm256_set1_ps(a1), va2 = _mm256_set1_ps(a2);
    const __m256 lane = _mm256_setr_ps(0.f, 1.f, 2.f, 3.f, 4.f, 5.f, 6.f, 7.f);
    const __m256 zero = _mm256_setzero_ps();
    const float sxc = (float)sx + 0.5f;

    for (int py = sy; py < ey; ++py) {
        const float pyc = (float)py + 0.5f;
        float r0 = (a0 * sxc + b0 * pyc) + c0;
        float r1 = (a1 * sxc + b1 * pyc) + c1;
        float r2 = (a2 * sxc + b2 * pyc) + c2;
        for (int px = sx; px < ex; px += 8) {
            __m256 e0 = _mm256_add_ps(_mm256_set1_ps(r0), _mm256_mul_ps(va0, lane));
            __m256 e1 = _mm256_add_ps(_mm256_set1_ps(r1), _mm256_mul_ps(va1, lane));
            __m256 e2 = _mm256_add_ps(_mm256_set1_ps(r2), _mm256_mul_ps(va2, lane));
            __m256 m = _mm256_and_ps(_mm256_cmp_ps(e0, zero, _CMP_GE_OQ),
                                     _mm256_and_ps(_mm256_cmp_ps(e1, zero, _CMP_GE_OQ), _mm256_cmp_ps(e2, zero, _CMP_GE_OQ)));
            int bits = _mm256_movemask_ps(m);
            if (bits) {
                alignas(32) float w0[8], w1[8], w2[8];
                _mm256_store_ps(w0, e0); _mm256_store_ps(w1, e1); _mm256_store_ps(w2, e2);
                for (int i = 0; i < 8; ++i) {
                    if (!(bits & (1 << i))) continue;
                    const int cx = px + i;
                    if (cx >= ex) break;
                    float alpha = w0[i] * invArea, beta = w1[i] * invArea, gamma = w2[i] * invArea;
                    float z = (tri.ndcZ.x * alpha + tri.ndcZ.y * beta) + tri.ndcZ.z * gamma;
                    const int idx = (py - tile.startY) * TILE_SIZE + (cx - tile.startX);
                    if (z < buf.depth[idx]) {
                        vec4 col;
                        vec3 bar(alpha, beta, gamma);
                        if (!shader.fragment(bar, col, vs)) {
                            buf.depth[idx] = z;
                            uint8_t* c = &buf.color[idx * 4];
                            c[0] = to_u8(col.x); c[1] = to_u8(col.y); c[2] = to_u8(col.z); c[3] = to_u8(col.w);
                            buf.ids[idx] = tri.key;
                        }
                    }
                }
            }
            r0 = r0 + a0 * 8.0f;
            r1 = r1 + a1 * 8.0f;
            r2 = r2 + a2 * 8.0f;
        }
    }
}

void TiledPipeline::processTile(size_t i, TileBuffers& buf, IShader& shader) {
    const Tile& tile = m_RelevantTiles[i];
    if (tile.triangles.empty()) return;
    buf.clear();
    VSTransformedTriangle vs;
    for (const Triangle* t : tile.triangles) {
        shader.material = t->material;
        for (int k = 0; k < 3; ++k) vs.vertices[k] = shader.vertex(t->vertices[k], k);
        vs.primitiveId = t->primitiveId;
        rasterizeTriangleInTile(*t, tile, buf, shader, vs);
    }
    TileResult& r = m_TileResults[i];
    r.startX = tile.startX; r.startY = tile.startY; r.endX = tile.endX; r.endY = tile.endY;
    r.buf = buf;
}

void TiledPipeline::mergeTileResults() {  // reference tiled_pipeline.cpp:1125-1181
    float* depth = m_Framebuffer->depth();
    uint8_t* color = m_Framebuffer->color();
    uint32_t* ids = m_Framebuffer->ids();
    const int W = m_Framebuffer->width();
    for (const TileResult& r : m_TileResults) {
        const int tw = r.endX - r.startX, th = r.endY - r.startY;
        for (int y = 0; y < th; ++y) {
            const size_t row = (size_t)(r.startY + y) * W + r.startX;
            for (int x = 0; x < tw; x += 8) {
                const int lanes = std::min(8, tw - x);
                int bits;
                if (lanes == 8) {
                    __m256 td = _mm256_loadu_ps(&r.buf.depth[y * TILE_SIZE + x]);
                    __m256 md = _mm256_loadu_ps(&depth[row + x]);
                    __m256 lt = _mm256_cmp_ps(td, md, _CMP_LT_OS);
                    bits = _mm256_movemask_ps(lt);
                    if (bits) _mm256_storeu_ps(&depth[row + x], _mm256_blendv_ps(md, td, lt));
                } else {
                    bits = 0;
                    for (int b = 0; b < lanes; ++b) {
                        float td = r.buf.depth[y * TILE_SIZE + x + b];
                        if (td < depth[row + x + b]) { depth[row + x + b] = td; bits |= 1 << b; }
                    }
                }
                for (int b = 0; b < lanes; ++b) {
                    if (!(bits & (1 << b))) continue;
                    const uint8_t* tc = &r.buf.color[(y * TILE_SIZE + x + b) * 4];
                    uint8_t* fc = &color[(row + x + b) * 4];
                    fc[0] = tc[2]; fc[1] = tc[1]; fc[2] = tc[0]; fc[3] = tc[3];
                    ids[row + x + b] = r.buf.ids[y * TILE_SIZE + x + b];
                }
            }
        }
    }
}

void TiledPipeline::drawMesh(const mat4& model, const Mesh& mesh) {
    if (!m_Shader || !m_Framebuffer || !m_HasCamera) return;
    const mat4 mvp = m_ViewProj * model;
    m_Shader->model = model;
    m_Shader->viewProj = m_ViewProj;
    m_Shader->mvp = mvp;
    m_Shader->viewportMat = m_Viewport;
    m_Shader->cameraPosition = m_CamPos;
    m_Shader->faceColors = mesh.faceColors.empty() ? nullptr : mesh.faceColors.data();

    auto t0 = Clock::now();
    m_Triangles.clear();
    size_t rankBase = 0;
    for (const MaterialGroup& g : mesh.groups) {
        clipGroup(mesh, g, rankBase, mvp);
        rankBase += g.faceCount;
    }
    auto t1 = Clock::now();
    initializeTiles();
    auto t2 = Clock::now();
    binTrianglesToTiles();
    auto t3 = Clock::now();
    m_RelevantTiles.clear();
    for (Tile& t : m_Tiles)
        if (!t.triangles.empty()) m_RelevantTiles.push_back(std::move(t));
    m_TileResults.assign(m_RelevantTiles.size(), TileResult{});
    const size_t batches = (m_RelevantTiles.size() + BATCH_SIZE - 1) / BATCH_SIZE;
    m_NextBatch.store(0);
    auto t4 = Clock::now();
    std::vector<std::future<void>> futs;
    for (size_t w = 0; w < m_Pool.size(); ++w) {
        futs.push_back(m_Pool.enqueue([this, batches] {
            std::unique_ptr<IShader> shader = m_Shader->clone();
            TileBuffers buf;
            for (;;) {
                size_t b = m_NextBatch.fetch_add(1, std::memory_order_relaxed);
                if (b >= batches) break;
                size_t s = b * BATCH_SIZE, e = std::min(s + BATCH_SIZE, m_RelevantTiles.size());
                for (size_t i = s; i < e; ++i) processTile(i, buf, *shader);
            }
        }));
    }
    for (auto& f : futs) f.get();
    auto t5 = Clock::now();
    mergeTileResults();
    auto t6 = Clock::now();
    m_TileResults.clear();
    m_RelevantTiles.clear();
    m_Times.clipping += secs(t0, t1); m_Times.initTiles += secs(t1, t2); m_Times.binning += secs(t2, t3);
    m_Times.tilePrep += secs(t3, t4); m_Times.raster += secs(t4, t5); m_Times.merging += secs(t5, t6);
}

}  // namespace orc
