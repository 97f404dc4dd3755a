// scene.cpp - synthetic workloads (host only): meshes and textures.  The
// reference loads meshes and textures from files (src/mesh.cpp:300-415, src/texture.cpp:21-35); the
// workloads here are procedural and seeded, so every generator uses its own
// integer RNG (splitmix64 seeding a PCG32 stream) and never a std::
// distribution, whose output differs between standard libraries.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>
#include "../../include/axiomr.h"
#include "axr_math.hpp"

using namespace axr;

namespace {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// PCG32 (O'Neill 2014), XSH-RR output.
struct Pcg32 {
    uint64_t state, inc;
    explicit Pcg32(uint64_t seed, uint64_t stream = 0) {
        uint64_t s = seed;
        state = splitmix64(s);
        inc = (splitmix64(s) ^ stream) | 1ull;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    // uniform in [0, 1) with 24 random bits
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    double uniform_d() { return (double)(((uint64_t)next() << 21) ^ (next() >> 11)) * (1.0 / 9007199254740992.0); }
    // standard normal by Box-Muller
    double normal() {
        double u1 = uniform_d(), u2 = uniform_d();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

void put_vertex(float* v, vec3 p, vec2 uv, vec3 n, vec3 t, vec3 b) {
    v[0] = p.x; v[1] = p.y; v[2] = p.z;
    v[3] = uv.x; v[4] = uv.y;
    v[5] = n.x; v[6] = n.y; v[7] = n.z;
    v[8] = t.x; v[9] = t.y; v[10] = t.z;
    v[11] = b.x; v[12] = b.y; v[13] = b.z;
}

// Lattice value noise: hashed values at integer points, trilinear blend with
// a smoothstep fade; fbm4 sums 4 octaves (lacunarity 2, gain 0.5) -> [-1, 1].
struct ValueNoise {
    uint64_t seed;
    double lattice(int64_t x, int64_t y, int64_t z) const {
        uint64_t h = seed ^ ((uint64_t)x * 0x8DA6B343ull) ^ ((uint64_t)y * 0xD8163841ull) ^ ((uint64_t)z * 0xCB1AB31Full);
        h = splitmix64(h);
        return (double)(h >> 11) * (1.0 / 9007199254740992.0);
    }
    double noise(double x, double y, double z) const {
        double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
        double tx = x - fx, ty = y - fy, tz = z - fz;
        auto fade = [](double t) { return t * t * (3.0 - 2.0 * t); };
        tx = fade(tx); ty = fade(ty); tz = fade(tz);
        double c[2][2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d);
        auto lerp = [](double p, double q, double t) { return p + (q - p) * t; };
        double x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
        double x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
        return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
    }
    double fbm4(double x, double y, double z) const {
        double sum = 0, amp = 1, freq = 2.0, norm = 0;
        for (int o = 0; o < 4; ++o) {
            sum += amp * (2.0 * noise(x * freq, y * freq, z * freq) - 1.0);
            norm += amp;
            amp *= 0.5;
            freq *= 2.0;
        }
        return sum / norm;
    }
};

struct D3 { double x, y, z; };
D3 dsub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 dcross(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 dnorm(D3 a) { double l = std::sqrt(ddot(a, a)); return l > 0 ? D3{a.x / l, a.y / l, a.z / l} : D3{0, 0, 1}; }

}  // namespace

extern "C" {

int axr_gen_uv_sphere(int stacks, int slices, float radius, float* verts56, uint32_t* idx) {
    if (stacks < 1 || slices < 3 || !verts56 || !idx) return AXR_ERR_INVALID;
    const double PI = 3.14159265358979323846;
    for (int i = 0; i <= stacks; ++i) {
        double th = PI * i / stacks;
        for (int j = 0; j <= slices; ++j) {
            double ph = 2.0 * PI * j / slices;
            D3 n{std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph)};
            D3 t{-std::sin(ph), 0.0, std::cos(ph)};
            D3 b = dcross(n, t);
            float* v = verts56 + 14 * (size_t)(i * (slices + 1) + j);
            put_vertex(v, make3((float)(radius * n.x), (float)(radius * n.y), (float)(radius * n.z)),
                       make2((float)j / slices, 1.0f - (float)i / stacks), make3((float)n.x, (float)n.y, (float)n.z),
                       make3((float)t.x, (float)t.y, (float)t.z), make3((float)b.x, (float)b.y, (float)b.z));
        }
    }
    // two counter-clockwise (seen from outside) triangles per quad
    size_t f = 0;
    for (int i = 0; i < stacks; ++i)
        for (int j = 0; j < slices; ++j) {
            uint32_t v00 = i * (slices + 1) + j, v01 = v00 + 1, v10 = v00 + slices + 1, v11 = v10 + 1;
            uint32_t* q = idx + 6 * f++;
            q[0] = v00; q[1] = v01; q[2] = v10;
            q[3] = v01; q[4] = v11; q[5] = v10;
        }
    return AXR_OK;
}

int axr_gen_icosphere(int n, float amp, uint64_t seed, float* verts56, uint32_t* idx) {
    if (n < 1 || n > 4096 || !verts56 || !idx) return AXR_ERR_INVALID;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    D3 c[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : c) p = dnorm(p);
    const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    // A lattice point of face (A,B,C) is i*A + j*B + k*C with i+j+k = n; its
    // key lists the (corner, weight) pairs with nonzero weight, sorted by
    // corner, so points on shared edges and corners get one vertex.
    std::unordered_map<uint64_t, uint32_t> ids;
    ids.reserve((size_t)12 * n * n);
    std::vector<D3> dirs;
    dirs.reserve((size_t)10 * n * n + 2);
    auto vid = [&](const int* f, int i, int j, int k) -> uint32_t {
        int cw[3][2] = {{f[0], i}, {f[1], j}, {f[2], k}};
        uint64_t parts[3];
        int np = 0;
        for (auto& p : cw) if (p[1]) parts[np++] = ((uint64_t)p[0] << 13) | (uint64_t)p[1];
        for (int a = 0; a < np; ++a)
            for (int b = a + 1; b < np; ++b)
                if (parts[b] < parts[a]) std::swap(parts[a], parts[b]);
        uint64_t key = 0;
        for (int a = 0; a < np; ++a) key = (key << 17) | parts[a];
        key |= (uint64_t)np << 60;
        auto it = ids.find(key);
        if (it != ids.end()) return it->second;
        D3 p{(i * c[f[0]].x + j * c[f[1]].x + k * c[f[2]].x) / n, (i * c[f[0]].y + j * c[f[1]].y + k * c[f[2]].y) / n,
             (i * c[f[0]].z + j * c[f[1]].z + k * c[f[2]].z) / n};
        uint32_t id = (uint32_t)dirs.size();
        dirs.push_back(dnorm(p));
        ids.emplace(key, id);
        return id;
    };
    std::vector<uint32_t> tri;
    tri.reserve((size_t)60 * n * n);
    for (auto& f : F) {
        // rows r = 0..n-1 from corner A toward edge BC
        for (int r = 0; r < n; ++r)
            for (int s = 0; s <= r; ++s) {
                // point (r, s) has weights i = n - r, j = r - s, k = s
                uint32_t a = vid(f, n - r, r - s, s);
                uint32_t b = vid(f, n - r - 1, r + 1 - s, s);
                uint32_t d = vid(f, n - r - 1, r - s, s + 1);
                tri.insert(tri.end(), {a, b, d});
                if (s < r) {
                    uint32_t e = vid(f, n - r, r - s - 1, s + 1);
                    tri.insert(tri.end(), {a, d, e});
                }
            }
    }
    const size_t nv = dirs.size(), nf = tri.size() / 3;
    if (nv != (size_t)10 * n * n + 2 || nf != (size_t)20 * n * n) return AXR_ERR_INVALID;
    ValueNoise vn{seed};
    std::vector<D3> P(nv);
    for (size_t i = 0; i < nv; ++i) {
        double r = 1.0 + amp * vn.fbm4(dirs[i].x, dirs[i].y, dirs[i].z);
        P[i] = D3{dirs[i].x * r, dirs[i].y * r, dirs[i].z * r};
    }
    // outward (counter-clockwise from outside) winding and area-weighted normals
    std::vector<D3> N(nv, D3{0, 0, 0});
    for (size_t f = 0; f < nf; ++f) {
        uint32_t* q = &tri[3 * f];
        D3 cr = dcross(dsub(P[q[1]], P[q[0]]), dsub(P[q[2]], P[q[0]]));
        D3 cen{P[q[0]].x + P[q[1]].x + P[q[2]].x, P[q[0]].y + P[q[1]].y + P[q[2]].y, P[q[0]].z + P[q[1]].z + P[q[2]].z};
        if (ddot(cr, cen) < 0) { std::swap(q[1], q[2]); cr = D3{-cr.x, -cr.y, -cr.z}; }
        for (int k = 0; k < 3; ++k) { N[q[k]].x += cr.x; N[q[k]].y += cr.y; N[q[k]].z += cr.z; }
        idx[3 * f] = q[0]; idx[3 * f + 1] = q[1]; idx[3 * f + 2] = q[2];
    }
    const double PI = 3.14159265358979323846;
    for (size_t i = 0; i < nv; ++i) {
        D3 nn = dnorm(N[i]);
        D3 d = dirs[i];
        double u = 0.5 + std::atan2(d.z, d.x) / (2.0 * PI);
        double v = 0.5 + std::asin(std::fmax(-1.0, std::fmin(1.0, d.y))) / PI;
        D3 tg = dcross(D3{0, 1, 0}, nn);
        if (ddot(tg, tg) < 1e-12) tg = D3{1, 0, 0};
        tg = dnorm(tg);
        D3 bt = dcross(nn, tg);
        put_vertex(verts56 + 14 * i, make3((float)P[i].x, (float)P[i].y, (float)P[i].z), make2((float)u, (float)v),
                   make3((float)nn.x, (float)nn.y, (float)nn.z), make3((float)tg.x, (float)tg.y, (float)tg.z),
                   make3((float)bt.x, (float)bt.y, (float)bt.z));
    }
    return AXR_OK;
}

int axr_gen_soup(uint64_t ntri, int W, int H, float fov_deg, float median_area, float sigma, float dup_frac,
                 uint64_t seed, float* verts56, uint32_t* idx, uint32_t* colors) {
    if (!ntri || W <= 0 || H <= 0 || !verts56 || !idx || !colors || ntri > 0x0FFFFFFFull) return AXR_ERR_INVALID;
    Pcg32 rng(seed, 7);
    const double tan_half = std::tan((double)fov_deg * 3.14159265358979323846 / 360.0);
    const double aspect = (double)W / H;
    const double PI = 3.14159265358979323846;
    for (uint64_t t = 0; t < ntri; ++t) {
        float* v = verts56 + 42 * t;
        const uint32_t col = rng.next();
        if (t > 0 && rng.uniform_d() < dup_frac) {
            // exact copy of an earlier triangle's corners: forces exact depth ties
            uint64_t src = (uint64_t)(rng.uniform_d() * (double)t);
            std::memcpy(v, verts56 + 42 * src, 42 * sizeof(float));
        } else {
            double cx = rng.uniform_d() * W, cy = rng.uniform_d() * H;
            double area = median_area * std::exp(sigma * rng.normal());
            double side = std::sqrt(4.0 * area / std::sqrt(3.0));
            double rad = side / std::sqrt(3.0);
            double th = rng.uniform_d() * 2.0 * PI;
            double d = 0.1 + rng.uniform_d() * 99.9;   // view depth in [0.1, 100]
            for (int k = 0; k < 3; ++k) {
                double a = th + k * (2.0 * PI / 3.0);   // counter-clockwise on screen
                double sx = cx + rad * std::cos(a), sy = cy + rad * std::sin(a);
                double ndx = 2.0 * sx / W - 1.0, ndy = 2.0 * sy / H - 1.0;
                double x = ndx * d * tan_half * aspect, y = ndy * d * tan_half;
                put_vertex(v + 14 * k, make3((float)x, (float)y, (float)(-d)), make2(0.0f, 0.0f),
                           make3(0.0f, 0.0f, 1.0f), make3(1.0f, 0.0f, 0.0f), make3(0.0f, 1.0f, 0.0f));
            }
        }
        idx[3 * t] = (uint32_t)(3 * t);
        idx[3 * t + 1] = (uint32_t)(3 * t + 1);
        idx[3 * t + 2] = (uint32_t)(3 * t + 2);
        colors[t] = col;
    }
    return AXR_OK;
}

int axr_gen_checker(int w, int h, int cell, uint8_t* rgba) {
    if (w <= 0 || h <= 0 || cell <= 0 || !rgba) return AXR_ERR_INVALID;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            bool on = (((x / cell) + (y / cell)) & 1) != 0;
            uint8_t* p = rgba + 4 * ((size_t)y * w + x);
            p[0] = on ? 230 : 40; p[1] = on ? 200 : 60; p[2] = on ? 160 : 90; p[3] = 255;
        }
    return AXR_OK;
}

int axr_gen_noise_texture(int w, int h, uint64_t seed, uint8_t* rgba) {
    if (w <= 0 || h <= 0 || !rgba) return AXR_ERR_INVALID;
    Pcg32 rng(seed, 11);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        uint32_t r = rng.next();
        rgba[4 * i] = (uint8_t)r; rgba[4 * i + 1] = (uint8_t)(r >> 8); rgba[4 * i + 2] = (uint8_t)(r >> 16);
        rgba[4 * i + 3] = 255;
    }
    return AXR_OK;
}

int axr_gen_normal_map(int w, int h, float min_z, uint64_t seed, uint8_t* rgba) {
    if (w <= 0 || h <= 0 || !rgba || !(min_z >= 0.0f && min_z <= 1.0f)) return AXR_ERR_INVALID;
    Pcg32 rng(seed, 13);
    const double PI = 3.14159265358979323846;
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        double z = min_z + (1.0 - min_z) * rng.uniform_d();
        double ph = 2.0 * PI * rng.uniform_d();
        double s = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        double n[3] = {s * std::cos(ph), s * std::sin(ph), z};
        for (int k = 0; k < 3; ++k) {
            double e = std::floor((n[k] * 0.5 + 0.5) * 255.0 + 0.5);
            rgba[4 * i + k] = (uint8_t)std::fmin(255.0, std::fmax(0.0, e));
        }
        rgba[4 * i + 3] = 255;
    }
    return AXR_OK;
}

}  // extern "C"
