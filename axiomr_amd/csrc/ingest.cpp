// ingest.cpp - Wavefront OBJ/MTL ingestion into the 56-byte vertex, index
// triple and material-group layout axr_mesh_create takes.  Host C++; the
// behaviour follows the reference loader (src/mesh.cpp:222-415):
//   - v / vt / vn / usemtl / f lines; 1-based indices; a corner whose
//     position index is missing or out of range is dropped; a missing or
//     out-of-range uv or normal index gives (0,0) / (0,0,0)
//   - corners are deduplicated on (position, uv, normal) with float equality
//     (so +0 and -0 merge and NaN never does), in first-seen order
//   - polygons are fanned from their first corner
//   - a face before the first usemtl belongs to no group (it is not drawn)
//   - materials come from <dir>/<basename>.mtl (not from mtllib), keyed by
//     name (a later newmtl of the same name replaces the earlier one)
//   - tangents and bitangents: per face, in face order, accumulated per
//     vertex, then orthonormalised against the normal (mesh.cpp:222-303)
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/axiomr.h"
#include "axr_math.hpp"

#pragma clang fp contract(off)
#pragma GCC optimize("no-fast-math")

using namespace axr;

struct axr_obj {
    std::vector<float> verts;          // 14 floats per vertex
    std::vector<uint32_t> idx;         // 3 per face
    struct Group {
        std::string material;
        uint64_t first, count;
    };
    std::vector<Group> groups;
    struct Mat {
        std::string name;
        float ka[3] = {0, 0, 0}, kd[3] = {0, 0, 0}, ks[3] = {0, 0, 0};
        float ns = 0, ni = 0, d = 0;
        int illum = 0;
        std::string maps[5];           // diffuse, metallic, bump, roughness, ao (full paths)
    };
    std::vector<Mat> mats;             // sorted by name
    bool mtl_found = false;
    std::string mtl_path;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

struct V14 {
    vec3 p;
    vec2 uv;
    vec3 n;
};

bool operator==(const V14& a, const V14& b) {
    return a.p.x == b.p.x && a.p.y == b.p.y && a.p.z == b.p.z && a.uv.x == b.uv.x && a.uv.y == b.uv.y &&
           a.n.x == b.n.x && a.n.y == b.n.y && a.n.z == b.n.z;
}

// Consistent with the float equality above: +0 and -0 hash alike.
struct V14Hash {
    size_t operator()(const V14& v) const {
        const float f[8] = {v.p.x, v.p.y, v.p.z, v.uv.x, v.uv.y, v.n.x, v.n.y, v.n.z};
        uint64_t h = 1469598103934665603ull;
        for (float x : f) {
            uint32_t u = 0;
            if (x != 0.0f) std::memcpy(&u, &x, 4);
            h = (h ^ u) * 1099511628211ull;
        }
        return (size_t)h;
    }
};

std::string directory_of(const std::string& path) {
    const size_t s = path.find_last_of("/\\");
    return s == std::string::npos ? std::string() : path.substr(0, s);
}

std::string basename_of(const std::string& path) {
    const size_t s = path.find_last_of("/\\");
    const size_t d = path.find_last_of('.');
    const size_t start = s == std::string::npos ? 0 : s + 1;
    if (d == std::string::npos) return path.substr(start);
    return path.substr(start, d - start);
}

// Leading integer of a face-index field as std::stoi reads it; false when
// there is none (the reference then throws).
bool parse_index(const std::string& s, int& out) {
    const char* b = s.c_str();
    char* e = nullptr;
    errno = 0;
    const long v = std::strtol(b, &e, 10);
    if (e == b || errno == ERANGE || v < INT32_MIN || v > INT32_MAX) return false;
    out = (int)v;
    return true;
}

std::string rest_trimmed(std::istringstream& ls) {
    std::string s;
    std::getline(ls, s);
    const size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return std::string();
    const size_t z = s.find_last_not_of(" \t\r");
    return s.substr(a, z - a + 1);
}

void read3(std::istringstream& ls, float* v) {
    for (int k = 0; k < 3; ++k) {
        v[k] = 0.0f;
        ls >> v[k];
    }
}

// Material count as the reference reads it (src/mesh.cpp:145-178): a
// "# Material Count: N" comment wins, else the number of newmtl lines.
uint32_t mtl_count(const std::string& path) {
    std::ifstream f(path);
    std::string line;
    uint32_t n = 0;
    bool found = false;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string t;
        ls >> t;
        if (t.empty()) continue;
        if (t == "#") {
            ls >> t;
            if (t == "Material") {
                ls >> t;
                if (t == "Count:") {
                    int c = 0;
                    ls >> c;
                    n = (uint32_t)c;
                    found = true;
                }
            }
        } else if (t == "newmtl" && !found) {
            ++n;
        }
    }
    return n;
}

// reference Mesh::loadMaterial / parseMaterialData (src/mesh.cpp:57-215).
// Material lines count only once the running material index has left -1,
// which (as in the reference) never happens if the count comment says 0.
void load_mtl(axr_obj& o, const std::string& obj_path) {
    o.mtl_path = directory_of(obj_path) + "/" + basename_of(obj_path) + ".mtl";
    std::ifstream f(o.mtl_path);
    if (!f.is_open()) return;
    o.mtl_found = true;
    const uint32_t count = mtl_count(o.mtl_path);
    int mat_index = -1;
    const std::string dir = directory_of(o.mtl_path);
    std::map<std::string, axr_obj::Mat> by_name;
    std::string line, cur;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::istringstream ls(line);
        std::string type;
        ls >> type;
        if (type == "newmtl") {
            if (mat_index < (int)(count - 1u)) ++mat_index;
            cur.clear();
            ls >> cur;
            by_name[cur] = axr_obj::Mat{};
            by_name[cur].name = cur;
            continue;
        }
        if (mat_index == -1 || by_name.find(cur) == by_name.end()) continue;
        axr_obj::Mat& m = by_name[cur];
        if (type == "Ka") read3(ls, m.ka);
        else if (type == "Kd") read3(ls, m.kd);
        else if (type == "Ks") read3(ls, m.ks);
        else if (type == "Ns") ls >> m.ns;
        else if (type == "Ni") ls >> m.ni;
        else if (type == "d") ls >> m.d;
        else if (type == "illum") ls >> m.illum;
        else {
            int slot = -1;
            if (type == "map_Kd") slot = 0;
            else if (type == "map_Ks" || type == "refl") slot = 1;
            else if (type == "map_Bump" || type == "bump" || type == "norm") slot = 2;
            else if (type == "map_Ns") slot = 3;
            else if (type == "map_A0") slot = 4;
            if (slot >= 0) m.maps[slot] = dir + "/" + rest_trimmed(ls);
        }
    }
    for (auto& kv : by_name) o.mats.push_back(kv.second);
}

vec3 v3(const float* f) { return make3(f[0], f[1], f[2]); }

// reference Mesh::calculateTangentBitangent (src/mesh.cpp:222-303).  The
// reference's "length() < 1e-8f" fallbacks compare glm's component count (3)
// and never run; they are left out.
void tangents(axr_obj& o) {
    const size_t nv = o.verts.size() / 14;
    std::vector<vec3> tsum(nv, make3(0, 0, 0)), bsum(nv, make3(0, 0, 0));
    const size_t nf = o.idx.size() / 3;
    for (size_t f = 0; f < nf; ++f) {
        const uint32_t* ix = &o.idx[3 * f];
        const float* a = &o.verts[14 * (size_t)ix[0]];
        const float* b = &o.verts[14 * (size_t)ix[1]];
        const float* c = &o.verts[14 * (size_t)ix[2]];
        const vec3 e1 = v3(b) - v3(a), e2 = v3(c) - v3(a);
        const vec2 d1 = make2(b[3], b[4]) - make2(a[3], a[4]);
        const vec2 d2 = make2(c[3], c[4]) - make2(a[3], a[4]);
        const vec3 n0 = v3(a + 5);
        const float den = d1.x * d2.y - d2.x * d1.y;
        vec3 t, bt;
        if (std::fabs(den) < 1e-8f) {
            t = (std::fabs(n0.x) > 0.8f) ? make3(0.0f, 1.0f, 0.0f) : make3(1.0f, 0.0f, 0.0f);
            bt = cross(n0, t);
        } else {
            const float fi = 1.0f / den;
            t = fi * (d2.y * e1 - d1.y * e2);
            bt = fi * (-d2.x * e1 + d1.x * e2);
            if (dot(cross(n0, t), bt) < 0.0f) bt = -bt;
        }
        for (int k = 0; k < 3; ++k) {
            tsum[ix[k]] = tsum[ix[k]] + t;
            bsum[ix[k]] = bsum[ix[k]] + bt;
        }
    }
    std::vector<bool> seen(nv, false);
    for (size_t f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k) seen[o.idx[3 * f + k]] = true;
    for (size_t i = 0; i < nv; ++i) {
        float* v = &o.verts[14 * i];
        const vec3 n = v3(v + 5);
        vec3 t = seen[i] ? tsum[i] : make3(1.0f, 0.0f, 0.0f);
        t = t - n * dot(n, t);
        t = normalize(t);
        vec3 b = cross(n, t);
        const float hand = dot(bsum[i], b) < 0.0f ? -1.0f : 1.0f;
        b = b * hand;
        v[8] = t.x; v[9] = t.y; v[10] = t.z;
        v[11] = b.x; v[12] = b.y; v[13] = b.z;
    }
}

int parse_obj(axr_obj& o, const std::string& path) {
    std::ifstream file(path, std::ios::binary);
    if (!file.is_open()) return fail(AXR_ERR_INVALID, "cannot open " + path);
    std::stringstream buf;
    buf << file.rdbuf();
    std::istringstream stream(buf.str());
    std::vector<vec3> pos, nrm;
    std::vector<vec2> uvs;
    std::unordered_map<V14, uint32_t, V14Hash> unique;
    std::string line;
    uint64_t nfaces = 0;
    size_t lineno = 0;
    while (std::getline(stream, line)) {
        ++lineno;
        if (line.empty()) continue;
        std::istringstream ls(line);
        std::string type;
        ls >> type;
        if (type == "v") {
            float f[3];
            read3(ls, f);
            pos.push_back(v3(f));
        } else if (type == "vt") {
            float u = 0.0f, v = 0.0f;
            ls >> u >> v;
            uvs.push_back(make2(u, v));
        } else if (type == "vn") {
            float f[3];
            read3(ls, f);
            nrm.push_back(v3(f));
        } else if (type == "usemtl") {
            std::string name;
            ls >> name;
            if (!o.groups.empty()) o.groups.back().count = nfaces - o.groups.back().first;
            o.groups.push_back({name, nfaces, 0});
        } else if (type == "f") {
            std::vector<uint32_t> corners;
            std::string tok;
            while (ls >> tok) {
                std::vector<std::string> parts;
                std::stringstream ss(tok);
                std::string sub;
                while (std::getline(ss, sub, '/')) parts.push_back(sub);
                int vi = -1, ti = -1, ni = -1, x;
                if (!parts.empty() && !parts[0].empty()) {
                    if (!parse_index(parts[0], x)) return fail(AXR_ERR_INVALID, "bad face index at line " + std::to_string(lineno));
                    vi = x - 1;
                }
                if (parts.size() > 1 && !parts[1].empty()) {
                    if (!parse_index(parts[1], x)) return fail(AXR_ERR_INVALID, "bad uv index at line " + std::to_string(lineno));
                    ti = x - 1;
                }
                if (parts.size() > 2 && !parts[2].empty()) {
                    if (!parse_index(parts[2], x)) return fail(AXR_ERR_INVALID, "bad normal index at line " + std::to_string(lineno));
                    ni = x - 1;
                }
                if (vi < 0 || vi >= (int)pos.size()) continue;
                V14 key;
                key.p = pos[vi];
                key.uv = (ti >= 0 && ti < (int)uvs.size()) ? uvs[ti] : make2(0.0f, 0.0f);
                key.n = (ni >= 0 && ni < (int)nrm.size()) ? nrm[ni] : make3(0.0f, 0.0f, 0.0f);
                auto it = unique.find(key);
                uint32_t id;
                if (it == unique.end()) {
                    id = (uint32_t)(o.verts.size() / 14);
                    const float f[14] = {key.p.x, key.p.y, key.p.z, key.uv.x, key.uv.y, key.n.x, key.n.y, key.n.z,
                                         0, 0, 0, 0, 0, 0};
                    o.verts.insert(o.verts.end(), f, f + 14);
                    if (key == key) unique.emplace(key, id);   // a NaN corner never matches, as the reference's map
                } else {
                    id = it->second;
                }
                corners.push_back(id);
            }
            for (size_t i = 1; i + 1 < corners.size(); ++i) {
                o.idx.push_back(corners[0]);
                o.idx.push_back(corners[i]);
                o.idx.push_back(corners[i + 1]);
                ++nfaces;
            }
        }
        if (!o.groups.empty()) o.groups.back().count = nfaces - o.groups.back().first;
    }
    return AXR_OK;
}

}  // namespace

extern "C" {

const char* axr_obj_last_error(void) { return g_err.c_str(); }

int axr_obj_load(const char* path, axr_obj** out) {
    if (!path || !out) return fail(AXR_ERR_INVALID, "null argument");
    *out = nullptr;
    axr_obj* o = new (std::nothrow) axr_obj();
    if (!o) return AXR_ERR_OOM;
    int r;
    try {
        r = parse_obj(*o, path);
        if (r == AXR_OK) {
            load_mtl(*o, path);
            tangents(*o);
        }
    } catch (const std::bad_alloc&) {
        r = fail(AXR_ERR_OOM, "out of memory reading " + std::string(path));
    }
    if (r != AXR_OK) {
        delete o;
        return r;
    }
    *out = o;
    return AXR_OK;
}

int axr_obj_destroy(axr_obj* o) {
    delete o;
    return AXR_OK;
}

int axr_obj_counts(const axr_obj* o, uint64_t* nverts, uint64_t* nfaces, uint32_t* ngroups, uint32_t* nmaterials) {
    if (!o) return fail(AXR_ERR_INVALID, "null model");
    if (nverts) *nverts = o->verts.size() / 14;
    if (nfaces) *nfaces = o->idx.size() / 3;
    if (ngroups) *ngroups = (uint32_t)o->groups.size();
    if (nmaterials) *nmaterials = (uint32_t)o->mats.size();
    return AXR_OK;
}

int axr_obj_arrays(const axr_obj* o, float* verts56, uint32_t* idx) {
    if (!o) return fail(AXR_ERR_INVALID, "null model");
    if (verts56 && !o->verts.empty()) std::memcpy(verts56, o->verts.data(), o->verts.size() * 4);
    if (idx && !o->idx.empty()) std::memcpy(idx, o->idx.data(), o->idx.size() * 4);
    return AXR_OK;
}

int axr_obj_get_group(const axr_obj* o, uint32_t i, axr_obj_group* out) {
    if (!o || !out || i >= o->groups.size()) return fail(AXR_ERR_INVALID, "bad group index");
    const axr_obj::Group& g = o->groups[i];
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->material, g.material.c_str(), sizeof(out->material) - 1);
    out->first_face = g.first;
    out->face_count = g.count;
    out->material_index = UINT32_MAX;
    for (size_t m = 0; m < o->mats.size(); ++m)
        if (o->mats[m].name == g.material) out->material_index = (uint32_t)m;
    return AXR_OK;
}

int axr_obj_get_material(const axr_obj* o, uint32_t i, axr_obj_material* out) {
    if (!o || !out || i >= o->mats.size()) return fail(AXR_ERR_INVALID, "bad material index");
    const axr_obj::Mat& m = o->mats[i];
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, m.name.c_str(), sizeof(out->name) - 1);
    for (int k = 0; k < 3; ++k) {
        out->ambient[k] = m.ka[k];
        out->diffuse[k] = m.kd[k];
        out->specular[k] = m.ks[k];
    }
    out->specular_exponent = m.ns;
    out->refractive_index = m.ni;
    out->dissolve = m.d;
    out->illumination_model = m.illum;
    char* dst[5] = {out->diffuse_map, out->metallic_map, out->bump_map, out->roughness_map, out->ao_map};
    for (int k = 0; k < 5; ++k) std::strncpy(dst[k], m.maps[k].c_str(), sizeof(out->diffuse_map) - 1);
    return AXR_OK;
}

}  // extern "C"
