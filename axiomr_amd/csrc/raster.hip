// raster.hip - the gfx950 kernels of the tiled render path.
//
// Stage map (reference: src/tiled_pipeline.cpp:143-322 drawMesh):
//   k_prep        zero the per-frame counters and bin counters
//   k_setup       one lane per face: transform, frustum classification, cull,
//                 triangle setup, bin counting           (tpl.cpp:185-240, 53-119)
//   k_clip        faces that cross a frustum plane: the full 6-plane clipper
//                 (pipeline.cpp:176-370), then cull/setup/count as above
//   k_scan_*      exclusive prefix sum of the bin counts (the frame arena layout)
//   k_scatter     write every (bin, record) pair          (tpl.cpp:368-393)
//   k_raster      one workgroup per bin: coverage and early-Z into LDS keys,
//                 then per-pixel resolve: depth test against the framebuffer,
//                 shading of the winner, BGRA/depth store  (tpl.cpp:395-594,
//                 1125-1181)
//
// Numerics: every coverage and depth value is computed with the reference's
// FP32 formulas in its evaluation order (no contraction), anchored at the
// reference's 16-pixel tiles and 8-pixel steps whatever the bin size here.
// The per-pixel winner is the minimum of the 64-bit key
// (ordered depth, submission key), which equals the reference's first-wins
// strict-less depth test over triangles in submission order.
#include <cstdlib>
#include <rocprofiler-sdk-roctx/roctx.h>
#include "axr_types.hpp"
#include "shade.hpp"
#include "internal.hpp"
#include <hip/hip_ext.h>

#pragma clang fp contract(off)

namespace axr {

constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_ITEMS;

// ---------------------------------------------------------------------------
// helpers

template <bool BATCH>
__device__ __forceinline__ const ViewDesc& view_at(const KArgs& ka, uint32_t v) {
    if constexpr (BATCH) return ka.table[v];
    else return ka.one;
}

// Largest i in [0, n) with a[i] <= x (a ascending, a[0] <= x).
__device__ __forceinline__ uint32_t find_le(const uint32_t* a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// Largest i in [0, n) with a[i] <= x (a ascending, a[0] <= x), searched by
// the whole wave: 64 probes per round, so about log64(n) dependent loads.
// Every lane of the wave must call it with the same arguments.
__device__ __forceinline__ uint32_t wave_find_le(const uint32_t* a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    const uint32_t lane = __lane_id();
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t p = lo + lane * step;
        const bool le = p < hi && a[p] <= x;
        const unsigned long long m = __ballot(le);      // a prefix of the lanes (a is ascending)
        const uint32_t last = 63u - (uint32_t)__clzll((long long)m);
        lo = lo + last * step;
        hi = lo + step < hi ? lo + step : hi;
    }
    return lo;
}

// View of face slot g for every lane of a wave (all 64 lanes call it):
// one 64-way search for the wave's first face, then only lanes past that
// view's last face search on their own (a wave rarely spans two views).
__device__ __forceinline__ uint32_t wave_view_of(const uint32_t* bases, uint32_t nviews, uint32_t g, uint32_t total) {
    uint32_t g0 = __shfl(g, 0);
    g0 = g0 < total ? g0 : total - 1;
    const uint32_t v0 = wave_find_le(bases, nviews, g0);
    if (g >= total) return v0;
    return g < bases[v0 + 1] ? v0 : find_le(bases, nviews, g);
}

__device__ __forceinline__ DevGroup group_of(const ViewDesc& V, uint32_t rank) {
    if (V.ngroups == 1) return V.g0;
    uint32_t lo = 0, hi = V.ngroups;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (gld(V.groups, mid).rank_base <= rank) lo = mid; else hi = mid;
    }
    return gld(V.groups, lo);
}

__device__ __forceinline__ vec3 load_pos(const float* p, uint32_t i) {
    return gld(reinterpret_cast<const vec3*>(p), i);
}

// One face as the setup stage sees it: submission rank s within its view,
// its group, mesh face index, vertex indices and clip-space corners
// (reference tiled_pipeline.cpp:185-212: gl_Position = mvp * position).
struct FaceGeo {
    DevGroup G;
    uint32_t face, i0, i1, i2;
    vec4 c0, c1, c2;
};

__device__ __forceinline__ void face_geo(const ViewDesc& V, uint32_t s, FaceGeo& f) {
    f.G = group_of(V, s);
    f.face = f.G.first + (s - f.G.rank_base);
    const uint3 ix = gld(reinterpret_cast<const uint3*>(V.idx), f.face);
    f.i0 = ix.x;
    f.i1 = ix.y;
    f.i2 = ix.z;
    f.c0 = V.mvp * make4(load_pos(V.pos, f.i0), 1.0f);
    f.c1 = V.mvp * make4(load_pos(V.pos, f.i1), 1.0f);
    f.c2 = V.mvp * make4(load_pos(V.pos, f.i2), 1.0f);
}

// Every 32-bit word of v as the wave's first active lane holds it: for a
// value the whole wave holds in common, so the compiler keeps it in scalar
// registers (and addresses through it with a scalar base).
template <class X>
__device__ __forceinline__ X wave_uniform(const X& v) {
    static_assert(sizeof(X) % 4 == 0, "whole words");
    X r;
    uint32_t w[sizeof(X) / 4];
    __builtin_memcpy(w, &v, sizeof(X));
    for (int i = 0; i < (int)(sizeof(X) / 4); ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    __builtin_memcpy(&r, w, sizeof(X));
    return r;
}

// Wave-aggregated append: returns this lane's slot for lanes with pred set.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    uint32_t base = 0;
    int leader = __ffsll((long long)mask) - 1;
    if (pred && (int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader < 0 ? 0 : leader);
    return base + (uint32_t)__popcll(mask & __lanemask_lt());
}

// A frame's sticky error bit.  The pinned host word is set as well (a
// system-scope store, seen by the host once the kernel has completed), so
// the host learns of an error without copying the counters back after every
// wait; errors are rare, so the store over the host link is too.
__device__ __forceinline__ void raise_error(const Arena& ar, uint32_t bit) {
    atomicOr(&ar.counters[CNT_ERROR], bit);
    if (ar.host_err) __hip_atomic_store(ar.host_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void set_error(const Arena& ar, uint32_t bit, uint32_t need_slot, uint32_t need) {
    raise_error(ar, bit);
    atomicMax(&ar.counters[need_slot], need);
}

// reference tiled_pipeline.cpp:15-23
__device__ __forceinline__ float clampW(float w) {
    const float tiny = 1e-6f;
    if (fabsf(w) < tiny) w = (w < 0.0f) ? -tiny : tiny;
    return w;
}

// reference pipeline.cpp:274-285
__device__ __forceinline__ float plane_dist(const vec4& v, int plane) {
    switch (plane) {
    case 0: return v.x + v.w;
    case 1: return v.w - v.x;
    case 2: return v.y + v.w;
    case 3: return v.w - v.y;
    case 4: return v.z + v.w;
    default: return v.w - v.z;
    }
}

enum { CLS_INSIDE = 0, CLS_DROP = 1, CLS_CLIP = 2 };
// The clipper's result for the cases it decides without interpolating: a
// triangle inside every plane passes through unchanged, and one entirely
// outside the first plane it is not entirely inside is dropped.  Every other
// triangle goes through the full clipper.
__device__ __forceinline__ int classify(const vec4& a, const vec4& b, const vec4& c) {
    for (int p = 0; p < 6; ++p) {
        float d0 = plane_dist(a, p), d1 = plane_dist(b, p), d2 = plane_dist(c, p);
        if (d0 >= 0.f && d1 >= 0.f && d2 >= 0.f) continue;
        if (d0 < 0.f && d1 < 0.f && d2 < 0.f) return CLS_DROP;
        return CLS_CLIP;
    }
    return CLS_INSIDE;
}

// Triangle setup: screen position (y up) and ndc z of one corner
// (reference tiled_pipeline.cpp:59-65).
__device__ __forceinline__ void to_screen(const vec4& c, int W, int H, float& sx, float& sy, float& z) {
    float invW = rcp_rn(clampW(c.w));
    z = c.z * invW;
    sx = (((c.x * invW) + 1.0f) * 0.5f) * (float)W;
    sy = (((c.y * invW) + 1.0f) * 0.5f) * (float)H;
}

// reference tiled_pipeline.cpp:89-119
__device__ __forceinline__ bool is_backface(const TriRec& r) {
    float dx1 = r.a.z - r.a.x, dy1 = r.a.w - r.a.y;
    float dx2 = r.b.x - r.a.x, dy2 = r.b.y - r.a.y;
    return (dx1 * dy2 - dx2 * dy1) < 0;
}

__device__ __forceinline__ TriRec make_rec(const vec4& c0, const vec4& c1, const vec4& c2, int W, int H) {
    TriRec r;
    float z0, z1, z2;
    to_screen(c0, W, H, r.a.x, r.a.y, z0);
    to_screen(c1, W, H, r.a.z, r.a.w, z1);
    to_screen(c2, W, H, r.b.x, r.b.y, z2);
    r.b.z = z0; r.b.w = z1;
    r.c.x = __float_as_uint(z2);
    return r;
}

// Edge functions of one triangle as the reference raster prepares them
// (tiled_pipeline.cpp:436-487), plus its pixel bounds.
struct RasterTri {
    float a0, b0, c0, a1, b1, c1, a2, b2, c2;
    float inv_area;
    float z0, z1, z2;
    int fx0, fy0, cx1, cy1;   // (int)floor(min), (int)ceil(max)
};

__device__ __forceinline__ bool raster_setup(const TriRec& r, RasterTri& t) {
    const float x0 = r.a.x, y0 = r.a.y, x1 = r.a.z, y1 = r.a.w, x2 = r.b.x, y2 = r.b.y;
    t.fx0 = cvt_i32(floorf(smin3(x0, x1, x2)));
    t.fy0 = cvt_i32(floorf(smin3(y0, y1, y2)));
    t.cx1 = cvt_i32(ceilf(smax3(x0, x1, x2)));
    t.cy1 = cvt_i32(ceilf(smax3(y0, y1, y2)));
    float c0 = x1 * y2 - x2 * y1;
    float c1 = x2 * y0 - x0 * y2;
    float c2 = x0 * y1 - x1 * y0;
    float area = (c0 + c1) + c2;
    if (area >= 0 && (double)area < 1.0e-12) return false;
    float a0 = y1 - y2, b0 = x2 - x1;
    float a1 = y2 - y0, b1 = x0 - x2;
    float a2 = y0 - y1, b2 = x1 - x0;
    if (area < 0) {
        a0 = -a0; b0 = -b0; c0 = -c0;
        a1 = -a1; b1 = -b1; c1 = -c1;
        a2 = -a2; b2 = -b2; c2 = -c2;
        area = -area;
    }
    t.a0 = a0; t.b0 = b0; t.c0 = c0;
    t.a1 = a1; t.b1 = b1; t.c1 = c1;
    t.a2 = a2; t.b2 = b2; t.c2 = c2;
    t.inv_area = rcp_rn(area);
    t.z0 = r.b.z; t.z1 = r.b.w; t.z2 = __uint_as_float(r.c.x);
    return true;
}

// Pixel rectangle [x0,x1) x [y0,y1) that the reference raster visits for
// this triangle over the whole screen.
struct Rect { int x0, y0, x1, y1; };
__device__ __forceinline__ Rect screen_rect(const RasterTri& t, int W, int H) {
    return Rect{imax(0, t.fx0), imax(0, t.fy0), imin(W, t.cx1), imin(H, t.cy1)};
}

// Coverage and depth of pixel (px, py), reference tiled_pipeline.cpp:502-562:
// the row value is anchored at sx = max(16*(px/16), floor(minX)); pixels 8..15
// past the anchor take one +8*a step; then + a*lane.
__device__ __forceinline__ bool cover(const RasterTri& t, int px, int py, float& al, float& be, float& ga, float& z) {
    const int sx = imax(px & ~15, t.fx0);
    const float sxc = (float)sx + 0.5f;
    const float pyc = (float)py + 0.5f;
    float r0 = (t.a0 * sxc + t.b0 * pyc) + t.c0;
    float r1 = (t.a1 * sxc + t.b1 * pyc) + t.c1;
    float r2 = (t.a2 * sxc + t.b2 * pyc) + t.c2;
    const int d = px - sx;
    if (d >= 8) {
        r0 = r0 + t.a0 * 8.0f;
        r1 = r1 + t.a1 * 8.0f;
        r2 = r2 + t.a2 * 8.0f;
    }
    const float lane = (float)(d & 7);
    const float e0 = r0 + t.a0 * lane;
    const float e1 = r1 + t.a1 * lane;
    const float e2 = r2 + t.a2 * lane;
    if (!(e0 >= 0.f && e1 >= 0.f && e2 >= 0.f)) return false;
    al = e0 * t.inv_area;
    be = e1 * t.inv_area;
    ga = e2 * t.inv_area;
    z = (t.z0 * al + t.z1 * be) + t.z2 * ga;
    return true;
}

// 64-bit depth key: order-preserving image of z (with -0 folded onto +0, the
// reference's strict-less test treats them as equal) above the submission key.
__device__ __forceinline__ unsigned long long depth_key(float z, uint32_t key) {
    uint32_t u = __float_as_uint(z);
    if ((u << 1) == 0) u = 0;
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((unsigned long long)u << 32) | key;
}
constexpr unsigned long long EMPTY_KEY = ~0ull;

// ---------------------------------------------------------------------------
// k_prep: zero per-frame counters (the sticky error word and the "need"
// high-water marks are left alone) and the bin counters/cursors.

// The arena cleans itself during a frame (k_raster zeroes the bin counts it
// consumed and the per-frame counters), so k_prep runs only after the arena
// is (re)allocated or after an overflow.
__global__ void k_prep(Arena ar, uint32_t ntiles) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < CNT_ERROR || tid == CNT_BIG_ITEMS || tid == CNT_BIG_RECS) ar.counters[tid] = 0;
    for (uint32_t i = tid; i < ntiles; i += gridDim.x * blockDim.x) {
        ar.tile_count[i] = 0;
    }
    for (uint32_t i = tid; i < ar.cap_vkeys; i += gridDim.x * blockDim.x) ar.vkeys[i] = ~0ull;
}

// ---------------------------------------------------------------------------
// binning helpers

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = (int)__lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// Bin range of a pixel rectangle for bins of T x T pixels.
template <int T>
struct BinRange {
    int tx0, ty0, nx, ny;
    __device__ __forceinline__ BinRange(const Rect& rc, bool on) {
        if (on) {
            tx0 = rc.x0 / T; ty0 = rc.y0 / T;
            nx = (rc.x1 - 1) / T - tx0 + 1; ny = (rc.y1 - 1) / T - ty0 + 1;
        } else {
            tx0 = ty0 = nx = ny = 0;
        }
    }
    __device__ __forceinline__ uint32_t count() const { return (uint32_t)(nx * ny); }
    __device__ __forceinline__ uint32_t bin(uint32_t k, const ViewDesc& V) const {
        return V.tile_base + (uint32_t)((ty0 + (int)(k / (uint32_t)nx)) * V.tiles_x + tx0 + (int)(k % (uint32_t)nx));
    }
};

// Emit the (bin, slot, ref) pairs of one record per lane, for a whole wave
// (all 64 lanes must call it).  The wave writes into its own region of the
// pair array (no shared counter); lanes that hit the same bin in the same
// round share one atomic on that bin's counter, which also hands every pair
// its final slot inside the bin, so the scatter pass needs no atomics.
template <int T>
__device__ __forceinline__ void emit_pairs_wave(const Arena& ar, const ViewDesc& V, bool on, const Rect& rc,
                                                uint32_t ref, uint32_t wave) {
    const BinRange<T> br(rc, on);
    const uint32_t n = br.count();
    const uint32_t incl = wave_incl_scan(n);
    const uint32_t total = __shfl(incl, 63);
    uint32_t base = wave * ar.wave_cap;
    bool fits = true;
    if (total > ar.wave_cap) {
        uint32_t o = 0;
        if (__lane_id() == 0) o = atomicAdd(&ar.counters[CNT_OVF_PAIRS], total);
        o = __shfl(o, 0);
        fits = o + total <= ar.cap_ovf;
        if (!fits && __lane_id() == 0) set_error(ar, ERR_PAIRS, CNT_NEED_PAIRS, o + total);
        base = ar.cap_waves * ar.wave_cap + o;
    }
    if (__lane_id() == 0) ar.wave_pairs[wave] = total > ar.wave_cap ? 0u : total;
    if (total == 0) return;
    const uint32_t p0 = base + incl - n;
    const uint32_t lane = __lane_id();
    for (uint32_t k = 0; __ballot(k < n); ++k) {
        const bool act = k < n;
        const uint32_t bin = act ? br.bin(k, V) : 0u;
        // group the lanes by bin (register work only), then every group's
        // first lane takes the group's slots with one atomic, all groups'
        // atomics in flight together
        unsigned long long pending = __ballot(act), mine = 0ull;
        int leader_of_mine = 0;
        while (pending) {
            const int leader = __ffsll((long long)pending) - 1;
            const uint32_t lb = __builtin_amdgcn_readlane(bin, leader);
            const unsigned long long m = __ballot(act && bin == lb) & pending;
            if ((m >> lane) & 1ull) {
                mine = m;
                leader_of_mine = leader;
            }
            pending &= ~m;
        }
        uint32_t slot = 0;
        if (act && leader_of_mine == (int)lane) slot = atomicAdd(&ar.tile_count[bin], (uint32_t)__popcll(mine));
        slot = __shfl(slot, leader_of_mine);
        if (act && fits) ar.pairs[p0 + k] = make_uint4(bin, slot + (uint32_t)__popcll(mine & __lanemask_lt()), ref, 0u);
    }
}

// Same for one lane on its own (divergent code such as the clipper).
// The pairs go to the overflow pool.
template <int T>
__device__ __forceinline__ void emit_pairs_lane(const Arena& ar, const ViewDesc& V, const Rect& rc, uint32_t ref) {
    const BinRange<T> br(rc, true);
    const uint32_t n = br.count();
    const uint32_t o = atomicAdd(&ar.counters[CNT_OVF_PAIRS], n);
    const bool fits = o + n <= ar.cap_ovf;
    if (!fits) set_error(ar, ERR_PAIRS, CNT_NEED_PAIRS, o + n);
    const uint32_t p0 = ar.cap_waves * ar.wave_cap + o;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t bin = br.bin(k, V);
        const uint32_t slot = atomicAdd(&ar.tile_count[bin], 1u);
        if (fits) ar.pairs[p0 + k] = make_uint4(bin, slot, ref, 0u);
    }
}

// ---------------------------------------------------------------------------
// visibility mode: helpers

// Pixels of a large triangle's screen rectangle per work item: one wave,
// four pixels per lane.
constexpr uint32_t BIG_CHUNK = 256;

__device__ __forceinline__ unsigned long long* big_counter(const Arena& ar) {
    return reinterpret_cast<unsigned long long*>(ar.counters + CNT_BIG_ITEMS);
}

// Mark the bins of rectangle rc as holding visibility keys.
template <int T>
__device__ __forceinline__ void flag_bins(const Arena& ar, const ViewDesc& V, const Rect& rc) {
    const int tx1 = (rc.x1 - 1) / T, ty1 = (rc.y1 - 1) / T;
    for (int ty = rc.y0 / T; ty <= ty1; ++ty)
        for (int tx = rc.x0 / T; tx <= tx1; ++tx) ar.tile_count[V.tile_base + (uint32_t)(ty * V.tiles_x + tx)] = 1u;
}

// One lane rasterizes a whole (small) triangle: every covered pixel of rc
// takes the minimum of its 64-bit key.
template <int T>
__device__ __forceinline__ void vis_raster_rect(const Arena& ar, const ViewDesc& V, const RasterTri& rt, const Rect& rc,
                                                uint32_t key) {
    unsigned long long* vk = ar.vkeys + V.pix_base;
    bool any = false;
    for (int py = rc.y0; py < rc.y1; ++py)
        for (int px = rc.x0; px < rc.x1; ++px) {
            float al, be, ga, z;
            if (cover(rt, px, py, al, be, ga, z) && z < INFINITY) {
                atomicMin(&vk[(size_t)py * (size_t)V.W + (size_t)px], depth_key(z, key));
                any = true;
            }
        }
    if (any) flag_bins<T>(ar, V, rc);
}

__device__ __forceinline__ uint32_t big_chunks(uint32_t area) { return (area + BIG_CHUNK - 1) / BIG_CHUNK; }

__device__ __forceinline__ void big_store(const Arena& ar, const TriRec& r, uint32_t ri, uint32_t i0, uint32_t nch) {
    if (ri >= ar.cap_big_recs || i0 + nch > ar.cap_big_items || i0 + nch < i0) {
        raise_error(ar, ERR_BIG);
        atomicMax(&ar.counters[CNT_NEED_BIG_RECS], ri + 1);
        atomicMax(&ar.counters[CNT_NEED_BIG_ITEMS], i0 + nch);
        return;
    }
    ar.big_rec[ri] = r;
    ar.big_start[ri] = i0;
}

// Queue large triangles (one per flagged lane, whole wave calls): one
// record each plus ceil(area / BIG_CHUNK) work items.  Records and items
// come from one 64-bit counter, so item starts grow with the record index.
__device__ __forceinline__ void big_append_wave(const Arena& ar, bool big, const TriRec& r, uint32_t area) {
    const unsigned long long mask = __ballot(big);
    if (mask == 0ull) return;
    const uint32_t nch = big ? big_chunks(area) : 0u;
    const uint32_t incl = wave_incl_scan(nch);
    const uint32_t total = __shfl(incl, 63);
    unsigned long long old = 0;
    if (__lane_id() == 0) old = atomicAdd(big_counter(ar), ((unsigned long long)__popcll(mask) << 32) | total);
    const uint32_t lo = __shfl((uint32_t)old, 0), hi = __shfl((uint32_t)(old >> 32), 0);
    if (!big) return;
    big_store(ar, r, hi + (uint32_t)__popcll(mask & __lanemask_lt()), lo + incl - nch, nch);
}

// Same for one lane on its own (the clipper).
__device__ __forceinline__ void big_append_lane(const Arena& ar, const TriRec& r, uint32_t area) {
    const uint32_t nch = big_chunks(area);
    const unsigned long long old = atomicAdd(big_counter(ar), (1ull << 32) | nch);
    big_store(ar, r, (uint32_t)(old >> 32), (uint32_t)old, nch);
}

// ---------------------------------------------------------------------------
// k_setup: one lane per face (submission order across views).

// Face slot g of view V (all 64 lanes of the wave call it).
// A triangle record stored with the non-temporal hint: the batch's records
// far exceed the caches and are read back only after the scatter.
__device__ __forceinline__ void store_rec_nt(TriRec* p, const TriRec& r) {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    v4u* q = reinterpret_cast<v4u*>(p);
    gst_nt(q, 0, v4u{__float_as_uint(r.a.x), __float_as_uint(r.a.y), __float_as_uint(r.a.z), __float_as_uint(r.a.w)});
    gst_nt(q, 1, v4u{__float_as_uint(r.b.x), __float_as_uint(r.b.y), __float_as_uint(r.b.z), __float_as_uint(r.b.w)});
    gst_nt(q, 2, v4u{r.c.x, r.c.y, r.c.z, r.c.w});
}

template <int T, bool INSIDE>
__device__ __forceinline__ void setup_face(const KArgs& ka, const Arena& ar, uint32_t g, const ViewDesc& V) {
    bool active = g < ka.total_faces;
    bool want_clip = false, binnable = false;
    Rect rc{0, 0, 0, 0};
    if (active) {
        const uint32_t s = g - V.face_base;
        FaceGeo f;
        face_geo(V, s, f);
        int cls = CLS_INSIDE;
        if constexpr (!INSIDE) cls = classify(f.c0, f.c1, f.c2);
        want_clip = cls == CLS_CLIP;
        if (cls == CLS_INSIDE) {
            TriRec r = make_rec(f.c0, f.c1, f.c2, V.W, V.H);
            if (!is_backface(r)) {
                RasterTri rt;
                if (raster_setup(r, rt)) {
                    rc = screen_rect(rt, V.W, V.H);
                    if (rc.x0 < rc.x1 && rc.y0 < rc.y1) {
                        r.c.y = s * 8u;
                        r.c.z = f.face;
                        r.c.w = f.G.material & REC_MATERIAL_MASK;
                        store_rec_nt(ar.rec + g, r);
                        binnable = true;
                    }
                }
            }
        }
    }
    const uint32_t cslot = wave_append(&ar.counters[CNT_CLIP_FACES], want_clip);
    if (want_clip) {
        if (cslot < ar.cap_clip_faces) ar.clip_faces[cslot] = g;
        else set_error(ar, ERR_CLIP_FACES, CNT_NEED_CLIP_FACES, cslot + 1);
        if (ka.skip_clip) raise_error(ar, ERR_SKIPPED);
    }
    emit_pairs_wave<T>(ar, V, binnable, rc, g, g >> 6);
}

// INSIDE: the host's bound put every view's mesh inside every frustum plane
// (ka.skip_clip), so the classification is left out, as in k_vis.  A batch
// wave whose faces lie in one view leaves it out for a view the bound puts
// inside (ViewDesc::inside), a branch on a value the wave holds in common.
template <int T, bool BATCH, bool INSIDE = false>
__global__ __launch_bounds__(256) void k_setup(KArgs ka, Arena ar) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (ar.counters[CNT_ERROR]) return;
    if constexpr (BATCH) {
        const uint32_t vi = wave_view_of(ka.face_bases, ka.nviews, g, ka.total_faces);
        const uint32_t v0 = __builtin_amdgcn_readfirstlane(vi);
        if (__ballot(vi != v0) == 0ull) {
            // the whole wave in one view (the usual case): the view's fields
            // in scalar registers, loaded once for the wave
            const ViewDesc V = wave_uniform(ka.table[v0]);
            if (INSIDE || V.inside) setup_face<T, true>(ka, ar, g, V);
            else setup_face<T, false>(ka, ar, g, V);
        } else {
            setup_face<T, INSIDE>(ka, ar, g, ka.table[vi]);
        }
    } else {
        setup_face<T, INSIDE>(ka, ar, g, ka.one);
    }
}

// ---------------------------------------------------------------------------
// k_vis: visibility mode's setup.  One lane per face: the same transform,
// classification, cull and setup as k_setup; a triangle whose screen
// rectangle has at most small_footprint pixels is rasterized by its lane
// straight into the key buffer, a larger one is queued as work items,
// faces that cross a plane go to the clipper.

// Faces per lane (their loads issued together: indices, then positions); a
// 256-thread block covers 256 * VIS_FACES_PER_LANE consecutive faces.  Two
// per lane measured slower than one on MI355X (fewer waves for small
// draws, no gain at 1M faces).
constexpr uint32_t VIS_FACES_PER_LANE = 1;

// INSIDE: the host's bound put the mesh's bounding sphere inside every
// frustum plane (ka.skip_clip), so every face classifies as inside and the
// classification is left out (a separate instantiation: a runtime test of the
// flag is compiled to a select, which still computes it).
template <int T, bool BATCH, bool INSIDE = false>
__global__ __launch_bounds__(256) void k_vis(KArgs ka, Arena ar) {
    constexpr uint32_t F = VIS_FACES_PER_LANE;
    if (ar.counters[CNT_ERROR]) return;
    uint32_t g[F], vi[F], s[F], face[F], i0[F], i1[F], i2[F];
    DevGroup G[F];
    bool act[F];
#pragma unroll
    for (uint32_t k = 0; k < F; ++k) {
        g[k] = blockIdx.x * (256u * F) + k * 256u + threadIdx.x;
        act[k] = g[k] < ka.total_faces;
        vi[k] = 0;
        if constexpr (BATCH) vi[k] = wave_view_of(ka.face_bases, ka.nviews, g[k], ka.total_faces);
        const ViewDesc& V = view_at<BATCH>(ka, vi[k]);
        s[k] = g[k] - V.face_base;
        if (act[k]) {
            G[k] = group_of(V, s[k]);
            face[k] = G[k].first + (s[k] - G[k].rank_base);
            const uint3 ix = gld(reinterpret_cast<const uint3*>(V.idx), face[k]);
            i0[k] = ix.x;
            i1[k] = ix.y;
            i2[k] = ix.z;
        }
    }
    vec3 p0[F], p1[F], p2[F];
#pragma unroll
    for (uint32_t k = 0; k < F; ++k) {
        if (act[k]) {
            const ViewDesc& V = view_at<BATCH>(ka, vi[k]);
            p0[k] = load_pos(V.pos, i0[k]);
            p1[k] = load_pos(V.pos, i1[k]);
            p2[k] = load_pos(V.pos, i2[k]);
        }
    }
    if constexpr (INSIDE) {
        // faces clearly facing away from the eye are culled before the
        // transform (the pre-test of KArgs::cull_sign; a face within its
        // margin goes on to the exact screen-space test)
        if (ka.cull_sign != 0.0f) {
            const vec3 e = make3(ka.cull_eye[0], ka.cull_eye[1], ka.cull_eye[2]);
            const float le = fabsf(e.x) + fabsf(e.y) + fabsf(e.z);
#pragma unroll
            for (uint32_t k = 0; k < F; ++k) {
                if (!act[k]) continue;
                const vec3 a = p0[k] - e, b = p1[k] - p0[k], c = p2[k] - p0[k];
                const float d = ka.cull_sign * dot(a, cross(b, c));
                const float la = fabsf(a.x) + fabsf(a.y) + fabsf(a.z) + le;
                const float lb = fabsf(b.x) + fabsf(b.y) + fabsf(b.z), lc = fabsf(c.x) + fabsf(c.y) + fabsf(c.z);
                if (d < -fmaxf(ka.cull_tau, 1e-3f * la * lb * lc)) act[k] = false;
            }
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < F; ++k) {
        bool want_clip = false, big = false;
        TriRec r;
        uint32_t area = 0;
        if (act[k]) {
            const ViewDesc& V = view_at<BATCH>(ka, vi[k]);
            const vec4 c0 = V.mvp * make4(p0[k], 1.0f);
            const vec4 c1 = V.mvp * make4(p1[k], 1.0f);
            const vec4 c2 = V.mvp * make4(p2[k], 1.0f);
            int cls = CLS_INSIDE;
            if constexpr (!INSIDE) cls = classify(c0, c1, c2);
            want_clip = cls == CLS_CLIP;
            if (cls == CLS_INSIDE) {
                r = make_rec(c0, c1, c2, V.W, V.H);
                RasterTri rt;
                if (!is_backface(r) && raster_setup(r, rt)) {
                    const Rect rc = screen_rect(rt, V.W, V.H);
                    if (rc.x0 < rc.x1 && rc.y0 < rc.y1) {
                        area = (uint32_t)((rc.x1 - rc.x0) * (rc.y1 - rc.y0));
                        if (area <= (uint32_t)ka.small_footprint) {
                            vis_raster_rect<T>(ar, V, rt, rc, s[k] * 8u);
                        } else {
                            big = true;
                            r.c.y = s[k] * 8u;
                            r.c.z = vi[k];
                            r.c.w = G[k].material & REC_MATERIAL_MASK;
                        }
                    }
                }
            }
        }
        const uint32_t cslot = wave_append(&ar.counters[CNT_CLIP_FACES], want_clip);
        if (want_clip) {
            if (cslot < ar.cap_clip_faces) ar.clip_faces[cslot] = g[k];
            else set_error(ar, ERR_CLIP_FACES, CNT_NEED_CLIP_FACES, cslot + 1);
        }
        if ((want_clip && ka.skip_clip) || (big && ka.skip_big)) raise_error(ar, ERR_SKIPPED);
        big_append_wave(ar, big, r, area);
    }
}

// ---------------------------------------------------------------------------
// k_clip: the reference's clipper on the faces k_setup queued.  A clipped
// vertex is the 14 vertex floats followed by the clip position.

struct CV { float f[18]; };

__device__ __forceinline__ vec4 cv_clip(const CV& v) { return make4(v.f[14], v.f[15], v.f[16], v.f[17]); }

__device__ void cv_interp(const CV& a, const CV& b, float t, CV& o) {  // pipeline.cpp:243-272
    const float s = 1.0f - t;
    for (int i = 0; i < 3; ++i) o.f[i] = a.f[i] * s + b.f[i] * t;   // position
    for (int k = 5; k < 14; k += 3) {                                 // normal, tangent, bitangent
        vec3 m = make3(a.f[k] * s + b.f[k] * t, a.f[k + 1] * s + b.f[k + 1] * t, a.f[k + 2] * s + b.f[k + 2] * t);
        m = normalize(m);
        o.f[k] = m.x; o.f[k + 1] = m.y; o.f[k + 2] = m.z;
    }
    for (int i = 14; i < 18; ++i) o.f[i] = a.f[i] * s + b.f[i] * t;  // clip position
    const float w0 = a.f[17], w1 = b.f[17];
    const float u = (a.f[3] * w0) * s + (b.f[3] * w1) * t;
    const float v = (a.f[4] * w0) * s + (b.f[4] * w1) * t;
    const float iw = w0 * s + w1 * t;
    o.f[3] = u / iw;
    o.f[4] = v / iw;
}

__device__ __forceinline__ float edge_t(float din, float dout) {
    float denom = din - dout;
    return (fabsf(denom) < 1e-7f) ? 0.5f : (din / denom);
}

__device__ __forceinline__ void cv_swap(CV& a, CV& b) { CV t = a; a = b; b = t; }

// reference pipeline.cpp:302-370
__device__ int clip_one_plane(int plane, CV& v0, CV& v1, CV& v2, CV& v3) {
    float d0 = plane_dist(cv_clip(v0), plane), d1 = plane_dist(cv_clip(v1), plane), d2 = plane_dist(cv_clip(v2), plane);
    if (d0 < 0.f && d1 < 0.f && d2 < 0.f) return 0;
    if (d0 >= 0.f && d1 >= 0.f && d2 >= 0.f) { v3 = v0; return 3; }
    const bool in0 = d0 >= 0.f, in1 = d1 >= 0.f, in2 = d2 >= 0.f;
    if (in1 && !in0) {
        cv_swap(v0, v1); float t = d0; d0 = d1; d1 = t;
        cv_swap(v1, v2); t = d1; d1 = d2; d2 = t;
    } else if (in2 && !in1) {
        cv_swap(v2, v1); float t = d2; d2 = d1; d1 = t;
        cv_swap(v1, v0); t = d1; d1 = d0; d0 = t;
    }
    cv_interp(v0, v2, edge_t(d0, d2), v3);
    if (d1 >= 0.f) {
        CV n2;
        cv_interp(v1, v2, edge_t(d1, d2), n2);
        v2 = n2;
        return 4;
    }
    CV n1;
    cv_interp(v0, v1, edge_t(d0, d1), n1);
    v1 = n1;
    v2 = v3;
    return 3;
}

constexpr int MAX_CLIPPED_VERTS = 24;

// reference pipeline.cpp:176-228; returns the vertex count (multiple of 3)
__device__ int clip_triangle(CV* tris, CV* next) {
    int count = 3;
    for (int plane = 0; plane < 6; ++plane) {
        int n = 0;
        for (int i = 0; i < count; i += 3) {
            CV v3;
            int r = clip_one_plane(plane, tris[i], tris[i + 1], tris[i + 2], v3);
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
        count = n;
        for (int i = 0; i < n; ++i) tris[i] = next[i];
        if (count == 0) break;
    }
    return count;
}

template <int T, bool BATCH, bool VIS>
__global__ __launch_bounds__(64) void k_clip(KArgs ka, Arena ar) {
    if (ar.counters[CNT_ERROR]) return;
    const uint32_t n = imin((int)ar.counters[CNT_CLIP_FACES], (int)ar.cap_clip_faces);
    CV tris[MAX_CLIPPED_VERTS], next[MAX_CLIPPED_VERTS];
    for (uint32_t q = blockIdx.x * 64u + threadIdx.x; q < n; q += gridDim.x * 64u) {
        const uint32_t g = ar.clip_faces[q];
        uint32_t vi = 0;
        if constexpr (BATCH) vi = find_le(ka.face_bases, ka.nviews, g);
        const ViewDesc& V = view_at<BATCH>(ka, vi);
        const uint32_t s = g - V.face_base;
        const DevGroup G = group_of(V, s);
        const uint32_t face = G.first + (s - G.rank_base);
        for (int k = 0; k < 3; ++k) {
            const uint32_t vtx = gld(V.idx, 3 * (size_t)face + k);
            const Vertex56 src = gld(V.verts, vtx);
            __builtin_memcpy(tris[k].f, &src, sizeof(Vertex56));
            vec4 c = V.mvp * make4(load_pos(V.pos, vtx), 1.0f);
            tris[k].f[14] = c.x; tris[k].f[15] = c.y; tris[k].f[16] = c.z; tris[k].f[17] = c.w;
        }
        const int nv = clip_triangle(tris, next);
        const uint32_t ntri = (uint32_t)(nv / 3);
        if (ntri == 0) continue;
        const uint32_t base = atomicAdd(&ar.counters[CNT_CLIP_RECS], ntri);
        if (base + ntri > ar.cap_clip_recs) {
            set_error(ar, ERR_CLIP_RECS, CNT_NEED_CLIP_RECS, base + ntri);
            continue;
        }
        for (uint32_t j = 0; j < ntri; ++j) {
            TriRec r = make_rec(cv_clip(tris[3 * j]), cv_clip(tris[3 * j + 1]), cv_clip(tris[3 * j + 2]), V.W, V.H);
            if (is_backface(r)) continue;
            RasterTri rt;
            if (!raster_setup(r, rt)) continue;
            Rect rc = screen_rect(rt, V.W, V.H);
            if (!(rc.x0 < rc.x1 && rc.y0 < rc.y1)) continue;
            const uint32_t slot = base + j;
            r.c.y = s * 8u + j;
            r.c.z = vi;
            r.c.w = G.material & REC_MATERIAL_MASK;
            ar.clip_rec[slot] = r;
            for (int k = 0; k < 3; ++k) {
                float* dst = reinterpret_cast<float*>(ar.clip_verts + 3 * (size_t)slot + k);
                for (int i = 0; i < 14; ++i) dst[i] = tris[3 * j + k].f[i];
            }
            if constexpr (VIS) {
                const uint32_t area = (uint32_t)((rc.x1 - rc.x0) * (rc.y1 - rc.y0));
                if (area <= (uint32_t)ka.small_footprint) vis_raster_rect<T>(ar, V, rt, rc, r.c.y);
                else big_append_lane(ar, r, area);
            } else {
                emit_pairs_lane<T>(ar, V, rc, REF_CLIP | slot);
            }
        }
        if constexpr (VIS) {
            ar.clip_base[g] = base;
        } else {
            TriRec m;
            m.a = make_float4(0, 0, 0, 0);
            m.b = make_float4(0, 0, 0, 0);
            m.c = make_uint4(0u, s * 8u, base, REC_CLIPPED | (G.material & REC_MATERIAL_MASK));
            ar.rec[g] = m;
        }
    }
}

// ---------------------------------------------------------------------------
// prefix sum of the bin counts.  Each bin contributes the 64-bit value
// (count > 0) << 32 | count, so one scan gives both the bin's offset in the
// bins array (low word) and its position in the list of non-empty bins
// (high word), which the raster walks instead of every bin.

typedef unsigned long long u64;

__device__ __forceinline__ u64 shfl_up64(u64 v, int o) {
    uint32_t lo = __shfl_up((uint32_t)v, o), hi = __shfl_up((uint32_t)(v >> 32), o);
    return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 wave_incl_scan64(u64 v) {
    const int lane = (int)__lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        u64 t = shfl_up64(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan across the 1024 threads of a block; returns the prefix and
// the block total.
__device__ __forceinline__ u64 block_excl_scan64(u64 v, u64* lds, u64& total) {
    const int lane = (int)__lane_id(), wave = threadIdx.x >> 6;
    u64 incl = wave_incl_scan64(v);
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        u64 w = lane < (SCAN_THREADS / 64) ? lds[lane] : 0ull;
        u64 wi = wave_incl_scan64(w);
        if (lane < (SCAN_THREADS / 64)) lds[16 + lane] = wi - w;
        if (lane == (SCAN_THREADS / 64) - 1) lds[32] = wi;
    }
    __syncthreads();
    total = lds[32];
    u64 r = lds[16 + wave] + incl - v;
    __syncthreads();
    return r;
}

__device__ __forceinline__ u64 bin_value(uint32_t count) { return ((u64)(count != 0) << 32) | count; }

// Write the offsets of SCAN_ITEMS consecutive bins starting at `base` and
// list the non-empty ones.
__device__ __forceinline__ void write_offsets(const Arena& ar, uint32_t base, uint32_t ntiles, const uint32_t* c, u64 ex) {
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < ntiles) {
            ar.tile_off[base + k] = (uint32_t)ex;
            if (c[k]) ar.active[ex >> 32] = base + k;
        }
        ex += bin_value(c[k]);
    }
}

// The per-frame counters the scan is the last to need: the clip queue and
// the clip records have been consumed by now.
__device__ __forceinline__ void reset_consumed_counters(const Arena& ar) {
    ar.counters[CNT_CLIP_FACES] = 0;
    ar.counters[CNT_CLIP_RECS] = 0;
}

__device__ __forceinline__ void finish_scan(const Arena& ar, u64 total) {
    const uint32_t pairs = (uint32_t)total;
    ar.counters[CNT_ACTIVE] = (uint32_t)(total >> 32);
    reset_consumed_counters(ar);
    if (pairs > ar.cap_bins) set_error(ar, ERR_PAIRS, CNT_NEED_PAIRS, pairs);
}

// One block scans all bin counts (used when there are at most
// SMALL_SCAN_TILES bins).
constexpr uint32_t SMALL_SCAN_TILES = 65536;
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_single(Arena ar, uint32_t ntiles) {
    __shared__ u64 lds[64];
    __shared__ u64 carry;
    if (ar.counters[CNT_ERROR]) return;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t start = 0; start < ntiles; start += SCAN_BLOCK) {
        const uint32_t base = start + threadIdx.x * SCAN_ITEMS;
        uint32_t c[SCAN_ITEMS];
        u64 s = 0;
        for (int k = 0; k < SCAN_ITEMS; ++k) { c[k] = (base + k < ntiles) ? ar.tile_count[base + k] : 0u; s += bin_value(c[k]); }
        u64 total;
        const u64 ex = block_excl_scan64(s, lds, total) + carry;
        write_offsets(ar, base, ntiles, c, ex);
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) finish_scan(ar, carry);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(Arena ar, uint32_t ntiles) {
    __shared__ u64 lds[64];
    if (ar.counters[CNT_ERROR]) return;
    const uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    u64 s = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k) if (base + k < ntiles) s += bin_value(ar.tile_count[base + k]);
    u64 total;
    block_excl_scan64(s, lds, total);
    if (threadIdx.x == 0) ar.scan_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(Arena ar, uint32_t nblocks) {
    __shared__ u64 lds[64];
    __shared__ u64 carry;
    if (ar.counters[CNT_ERROR]) return;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t start = 0; start < nblocks; start += SCAN_THREADS) {
        const uint32_t i = start + threadIdx.x;
        u64 v = i < nblocks ? ar.scan_sums[i] : 0ull;
        u64 total;
        u64 ex = block_excl_scan64(v, lds, total);
        const u64 c = carry;
        if (i < nblocks) ar.scan_sums[i] = c + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) finish_scan(ar, carry);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(Arena ar, uint32_t ntiles) {
    __shared__ u64 lds[64];
    if (ar.counters[CNT_ERROR]) return;
    const uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    uint32_t c[SCAN_ITEMS];
    u64 s = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k) { c[k] = (base + k < ntiles) ? ar.tile_count[base + k] : 0u; s += bin_value(c[k]); }
    u64 total;
    const u64 ex = block_excl_scan64(s, lds, total) + ar.scan_sums[blockIdx.x];
    write_offsets(ar, base, ntiles, c, ex);
}

// ---------------------------------------------------------------------------
// k_scatter: every pair already knows its slot inside its bin; place it.
// First the per-wave regions, then the overflow pool.

__global__ __launch_bounds__(256) void k_scatter(Arena ar, uint32_t nwaves) {
    if (ar.counters[CNT_ERROR]) return;
    const uint32_t stride = gridDim.x * 256u;
    const uint32_t cap = ar.wave_cap;
    // one wave per setup wave's region: its pair count is read once and
    // only the pairs it holds are visited.  A wave takes SCATTER_REGIONS
    // regions at a time with their loads in flight together (each region's
    // chain is count -> pair -> bin offset -> store; a region rarely holds
    // more than 64 pairs)
    constexpr uint32_t SCATTER_REGIONS = 4;
    const uint32_t lane = __lane_id();
    const uint32_t nw = gridDim.x * 4u;
    for (uint32_t w0 = (blockIdx.x * 256u + threadIdx.x) >> 6; w0 < nwaves; w0 += nw * SCATTER_REGIONS) {
        uint32_t n[SCATTER_REGIONS], off[SCATTER_REGIONS];
        uint4 p[SCATTER_REGIONS];
#pragma unroll
        for (uint32_t k = 0; k < SCATTER_REGIONS; ++k) {
            const uint32_t w = w0 + k * nw;
            n[k] = w < nwaves ? ar.wave_pairs[w] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < SCATTER_REGIONS; ++k)
            if (lane < n[k]) p[k] = ar.pairs[(size_t)(w0 + k * nw) * cap + lane];
#pragma unroll
        for (uint32_t k = 0; k < SCATTER_REGIONS; ++k)
            if (lane < n[k]) off[k] = ar.tile_off[p[k].x];
#pragma unroll
        for (uint32_t k = 0; k < SCATTER_REGIONS; ++k) {
            if (lane < n[k]) ar.bins[off[k] + p[k].y] = p[k].z;
            const uint4* reg = ar.pairs + (size_t)(w0 + k * nw) * cap;
            for (uint32_t i = lane + 64u; i < n[k]; i += 64u) {
                const uint4 q = reg[i];
                ar.bins[ar.tile_off[q.x] + q.y] = q.z;
            }
        }
    }
    const uint32_t novf = imin((int)ar.counters[CNT_OVF_PAIRS], (int)ar.cap_ovf);
    const uint4* ovf = ar.pairs + (size_t)ar.cap_waves * ar.wave_cap;
    for (uint32_t q = blockIdx.x * 256u + threadIdx.x; q < novf; q += stride) {
        const uint4 p = ovf[q];
        ar.bins[ar.tile_off[p.x] + p.y] = p.z;
    }
}

// ---------------------------------------------------------------------------
// k_raster: one workgroup per bin.


__device__ __forceinline__ const TriRec& rec_of(const Arena& ar, uint32_t ref) {
    return (ref & REF_CLIP) ? ar.clip_rec[ref & ~REF_CLIP] : ar.rec[ref];
}

// Vertex stage of the winner's three corners, then its fragment stage
// (reference tiled_pipeline.cpp:185-212 vertex, 564-575 fragment).  Clipped
// pieces take their corners from the clip records' vertices.
// um: the material every face of the view uses, held in scalar registers
// (a single-group mesh), or null: the record's material is loaded.
template <class S>
__device__ __forceinline__ vec4 shade_at(const ViewDesc& V, const Arena& ar, uint32_t material, uint32_t prim,
                                         bool clipped, uint32_t slot, uint32_t i0, uint32_t i1, uint32_t i2,
                                         vec3 bar, const DevMaterial* um = nullptr) {
    ShadeCtx c;
    c.u = &V.u;
    DevMaterial mat;
    if (um) mat = *um;
    else mat = gld(V.materials, material & REC_MATERIAL_MASK);
    c.m = &mat;
    c.face_colors = V.face_colors;
    c.prim = prim;
    typename S::Var var[3];
    if constexpr (sizeof(typename S::Var) > 1) {
        if (clipped) {
            for (int k = 0; k < 3; ++k) var[k] = S::vs(ar.clip_verts[3 * (size_t)slot + k], V.u);
        } else if (V.vary) {
            const typename S::Var* vy = static_cast<const typename S::Var*>(V.vary);
            var[0] = gld(vy, i0);
            var[1] = gld(vy, i1);
            var[2] = gld(vy, i2);
        } else {
            var[0] = S::vs(gld(V.verts, i0), V.u);
            var[1] = S::vs(gld(V.verts, i1), V.u);
            var[2] = S::vs(gld(V.verts, i2), V.u);
        }
    }
    return S::fs(bar, var, c);
}

template <class S>
__device__ __forceinline__ vec4 shade_winner(const ViewDesc& V, const Arena& ar, const TriRec& r, uint32_t slot,
                                            bool clipped, uint32_t rank, vec3 bar, const DevMaterial* um = nullptr) {
    uint32_t prim, i0 = 0, i1 = 0, i2 = 0;
    if (clipped) {
        const DevGroup G = group_of(V, rank);
        prim = G.first + (rank - G.rank_base);
    } else {
        prim = r.c.z;
        if constexpr (sizeof(typename S::Var) > 1) {
            const uint3 ix = gld(reinterpret_cast<const uint3*>(V.idx), prim);
            i0 = ix.x;
            i1 = ix.y;
            i2 = ix.z;
        }
    }
    return shade_at<S>(V, ar, r.c.w, prim, clipped, slot, i0, i1, i2, bar, um);
}

// known: the depth plane is known to hold known_depth here (a clear written
// just before the draw); else it is loaded.
template <class S>
__device__ __forceinline__ void resolve_pixel(const ViewDesc& V, const Arena& ar, unsigned long long k, int px, int py,
                                              bool known, float known_depth, const DevMaterial* um) {
    const uint32_t key = (uint32_t)k;
    const uint32_t rank = key >> 3, j = key & 7u;
    TriRec r = ar.rec[V.face_base + rank];
    bool clipped = false;
    uint32_t slot = 0;
    if (r.c.w & REC_CLIPPED) {
        clipped = true;
        slot = r.c.z + j;
        r = ar.clip_rec[slot];
    }
    RasterTri rt;
    raster_setup(r, rt);
    float al, be, ga, z;
    cover(rt, px, py, al, be, ga, z);
    const size_t pix = (size_t)py * (size_t)V.W + (size_t)px;
    if (!(z < (known ? known_depth : gld(V.depth, pix)))) return;   // reference mergeTileResults: strict less
    const vec4 col = shade_winner<S>(V, ar, r, slot, clipped, rank, make3(al, be, ga), um);
    gst_nt(V.depth, pix, z);
    gst_nt(V.color, pix, pack_bgra(col));
    if (V.ids) gst_nt(V.ids, pix, key);
}

template <int T, int NT, class S>
// The bin's covered pixels are listed first (plist, count in *ncov, which
// starts at zero), so that every wave shades covered pixels only: a bin on
// a silhouette takes fewer shading rounds than one pass over its rows.
__device__ __forceinline__ void resolve_tile(const ViewDesc& V, const Arena& ar, const unsigned long long* keys,
                                             int ox, int oy, bool known, float known_depth, uint16_t* plist,
                                             uint32_t* ncov) {
    static_assert((T * T) % 64 == 0 && NT % 64 == 0, "every wave makes the same number of passes");
    for (int p = threadIdx.x; p < T * T; p += NT) {
        const bool cov = keys[p] != EMPTY_KEY;
        const uint32_t slot = wave_append(ncov, cov);
        if (cov) plist[slot] = (uint16_t)p;
    }
    __syncthreads();
    const uint32_t n = *ncov;
    if (V.ngroups == 1) {
        // one material for the whole view: its fields (texture pointers
        // among them) in scalar registers, loaded once
        const DevMaterial um = wave_uniform(gld(V.materials, V.g0.material & REC_MATERIAL_MASK));
        for (uint32_t i = threadIdx.x; i < n; i += NT) {
            const int p = plist[i];
            resolve_pixel<S>(V, ar, keys[p], ox + (p % T), oy + (p / T), known, known_depth, &um);
        }
    } else {
        for (uint32_t i = threadIdx.x; i < n; i += NT) {
            const int p = plist[i];
            resolve_pixel<S>(V, ar, keys[p], ox + (p % T), oy + (p / T), known, known_depth, nullptr);
        }
    }
}

// Shader functor of a shader kind.
template <int SK> struct ShaderOf;
template <> struct ShaderOf<SHADER_FLAT> { using type = FlatShader; };
template <> struct ShaderOf<SHADER_PHONG> { using type = PhongShader; };
template <> struct ShaderOf<SHADER_PBR> { using type = PBRShader; };
template <> struct ShaderOf<SHADER_GOURAUD> { using type = GouraudShader; };
template <> struct ShaderOf<SHADER_PHONG_POINT> { using type = PhongPointShader; };
template <> struct ShaderOf<SHADER_FLAT_COLOR> { using type = FlatColorShader; };

template <int T, int NT, int SK>
__device__ __forceinline__ void resolve_tile_kind(const ViewDesc& V, const Arena& ar, const unsigned long long* keys,
                                                  int ox, int oy, bool kn, float kd, uint16_t* pl, uint32_t* nc) {
    if constexpr (SK >= 0) {
        resolve_tile<T, NT, typename ShaderOf<SK>::type>(V, ar, keys, ox, oy, kn, kd, pl, nc);
    } else {
        switch (V.u.kind) {
        case SHADER_FLAT: resolve_tile<T, NT, FlatShader>(V, ar, keys, ox, oy, kn, kd, pl, nc); break;
        case SHADER_PHONG: resolve_tile<T, NT, PhongShader>(V, ar, keys, ox, oy, kn, kd, pl, nc); break;
        case SHADER_PBR: resolve_tile<T, NT, PBRShader>(V, ar, keys, ox, oy, kn, kd, pl, nc); break;
        case SHADER_GOURAUD: resolve_tile<T, NT, GouraudShader>(V, ar, keys, ox, oy, kn, kd, pl, nc); break;
        case SHADER_PHONG_POINT: resolve_tile<T, NT, PhongPointShader>(V, ar, keys, ox, oy, kn, kd, pl, nc); break;
        default: resolve_tile<T, NT, FlatColorShader>(V, ar, keys, ox, oy, kn, kd, pl, nc); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Coverage of a bin's small triangles.  One lane per triangle walks its
// pixel rectangle (bbox within the bin), which leaves most lanes idle while
// the wave waits for its largest rectangle: on the 256-view batch the
// rectangles average 11.5 pixels but a wave's largest is far bigger, and
// 42% of the lane slots did work.  So a lane walks at most WALK_CAP pixels
// of its own rectangle; the rest of every longer rectangle is cut into
// pieces of WALK_CAP pixels that the wave's lanes then take in turn.

constexpr int WALK_CAP = 8;

// Pixels [q0, q1) of the rectangle [x0, x0 + w) x [y0, ...) in row-major
// order; every covered pixel takes the minimum of its key.
template <int T>
__device__ __forceinline__ void walk_span(const RasterTri& rt, int x0, int y0, int w, int q0, int q1, uint32_t key,
                                          unsigned long long* keys, int ox, int oy, float inv_w) {
    // row of q0 = floor(q0 / w): q0 < 2^16 and w <= 32, so (q0 + 0.5) / w is
    // at least 1/64 from an integer and its rounded product within 2^-7 of it
    const int dy = q0 == 0 ? 0 : (int)(((float)q0 + 0.5f) * inv_w);
    int px = x0 + (q0 - dy * w), py = y0 + dy;
    const int x1 = x0 + w;
    for (int q = q0; q < q1; ++q) {
        float al, be, ga, z;
        if (cover(rt, px, py, al, be, ga, z) && z < INFINITY)
            atomicMin(&keys[(py - oy) * T + (px - ox)], depth_key(z, key));
        if (++px == x1) {
            px = x0;
            ++py;
        }
    }
}

// A wave's triangles whose rectangles outlast WALK_CAP pixels, for the
// pieces: per lane what a walker needs of its triangle, and the first piece
// of each lane in the wave's numbering.  Two layouts:
//   WalkPieces  the edge setup itself, rectangle and key (17 words a lane);
//   WalkRefs    the bin entry's record reference and the rectangle (3 words):
//               the walker loads the record again and repeats the edge setup
//               (the same operations on the same bits), for a raster kernel
//               whose occupancy its LDS would otherwise bound.
struct WalkPieces {
    float4 e0[64];        // a0, b0, c0, a1
    float4 e1[64];        // b1, c1, a2, b2
    float4 e2[64];        // c2, inv_area, z0, z1
    float4 e3[64];        // z2, bits(fx0), bits(rect), bits(key)
    uint32_t start[64];
};
struct WalkRefs {
    uint32_t ref[64];
    uint32_t rect[64];
    uint32_t start[64];
};
// rect: x0 | y0 << 8 | w << 16 | h << 24, in bin coordinates

__device__ __forceinline__ void put_piece(WalkPieces* wp, uint32_t lane, const RasterTri& rt, uint32_t, uint32_t rect,
                                          uint32_t key) {
    wp->e0[lane] = make_float4(rt.a0, rt.b0, rt.c0, rt.a1);
    wp->e1[lane] = make_float4(rt.b1, rt.c1, rt.a2, rt.b2);
    wp->e2[lane] = make_float4(rt.c2, rt.inv_area, rt.z0, rt.z1);
    wp->e3[lane] = make_float4(rt.z2, __int_as_float(rt.fx0), __uint_as_float(rect), __uint_as_float(key));
}
__device__ __forceinline__ void get_piece(const WalkPieces* wp, const Arena&, uint32_t o, RasterTri& t, uint32_t& rect,
                                          uint32_t& key) {
    const float4 f0 = wp->e0[o], f1 = wp->e1[o], f2 = wp->e2[o], f3 = wp->e3[o];
    t.a0 = f0.x; t.b0 = f0.y; t.c0 = f0.z; t.a1 = f0.w;
    t.b1 = f1.x; t.c1 = f1.y; t.a2 = f1.z; t.b2 = f1.w;
    t.c2 = f2.x; t.inv_area = f2.y; t.z0 = f2.z; t.z1 = f2.w;
    t.z2 = f3.x; t.fx0 = __float_as_int(f3.y);
    rect = __float_as_uint(f3.z);
    key = __float_as_uint(f3.w);
}
__device__ __forceinline__ void put_piece(WalkRefs* wp, uint32_t lane, const RasterTri&, uint32_t ref, uint32_t rect,
                                          uint32_t) {
    wp->ref[lane] = ref;
    wp->rect[lane] = rect;
}
__device__ __forceinline__ void get_piece(const WalkRefs* wp, const Arena& ar, uint32_t o, RasterTri& t, uint32_t& rect,
                                          uint32_t& key) {
    const TriRec r = rec_of(ar, wp->ref[o]);
    raster_setup(r, t);
    rect = wp->rect[o];
    key = r.c.y;
}

// All 64 lanes call it with their own triangle (w = h = 0 for none); ref is
// its bin entry's record reference.
template <int T, class P>
__device__ __forceinline__ void walk_wave(P* wp, const Arena& ar, const RasterTri& rt, uint32_t ref, int x0, int y0,
                                          int w, int h, uint32_t key, unsigned long long* keys, int ox, int oy) {
    const int fp = w * h;
    const float inv_w = w ? 1.0f / (float)w : 0.0f;
    walk_span<T>(rt, x0 + ox, y0 + oy, w, 0, imin(fp, WALK_CAP), key, keys, ox, oy, inv_w);
    const uint32_t pieces = fp > WALK_CAP ? (uint32_t)(fp - 1) / WALK_CAP : 0u;
    const uint32_t incl = wave_incl_scan(pieces);
    const uint32_t total = __shfl(incl, 63);
    if (total == 0) return;
    const uint32_t lane = __lane_id();
    wp->start[lane] = incl - pieces;
    if (pieces)
        put_piece(wp, lane, rt, ref, (uint32_t)x0 | ((uint32_t)y0 << 8) | ((uint32_t)w << 16) | ((uint32_t)h << 24),
                  key);
    // the wave's LDS writes before its reads of other lanes' entries
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t j = lane; j < total; j += 64u) {
        // owner: the last lane whose pieces start at or before j (a lane
        // without pieces shares its start with the next lane that has some)
        uint32_t o = 0;
        for (uint32_t step = 32; step; step >>= 1)
            if (wp->start[o + step] <= j) o += step;
        RasterTri t;
        uint32_t rect, pkey;
        get_piece(wp, ar, o, t, rect, pkey);
        const int rx = (int)(rect & 255u), ry = (int)((rect >> 8) & 255u);
        const int rw = (int)((rect >> 16) & 255u), rh = (int)(rect >> 24);
        const int q0 = WALK_CAP * (int)(j - wp->start[o] + 1);
        walk_span<T>(t, rx + ox, ry + oy, rw, q0, imin(q0 + WALK_CAP, rw * rh), pkey, keys, ox, oy, 1.0f / (float)rw);
    }
}

// The point-light raster kernel takes the compact layout, which with its
// launch bound (raster_min_waves) lets five waves share a SIMD instead of
// four; the other kinds keep the walkers' loads in LDS.
template <int SK> struct PiecesOf { using type = WalkPieces; };
template <> struct PiecesOf<SHADER_PHONG_POINT> { using type = WalkRefs; };

template <bool BATCH>
__device__ __forceinline__ const ViewDesc& bin_view(const KArgs& ka, const ViewDesc* sview) {
    if constexpr (BATCH) return *sview;
    else return ka.one;
}

// sview (batches): the bin's view is copied into LDS once per bin, so that
// every per-pixel read of its uniforms and pointers is an LDS read.
template <int T, int NT, bool BATCH, int SK>
__device__ __forceinline__ void raster_bin(const KArgs& ka, const Arena& ar, uint32_t t, unsigned long long* keys,
                                           uint32_t* bigq, uint32_t& nbig, uint16_t* plist, ViewDesc* sview,
                                           typename PiecesOf<SK>::type* pieces) {
    if constexpr (BATCH) {
        const uint32_t vi = ka.tile_views[t];
        const uint32_t* src = reinterpret_cast<const uint32_t*>(ka.table + vi);
        for (uint32_t i = threadIdx.x; i < sizeof(ViewDesc) / 4; i += NT)
            reinterpret_cast<uint32_t*>(sview)[i] = gld(src, i);
    }
    const uint32_t n = ar.tile_count[t];
    const uint32_t off = ar.tile_off[t];
    for (int p = threadIdx.x; p < T * T; p += NT) keys[p] = EMPTY_KEY;
    if (threadIdx.x == 0) nbig = 0;
    __syncthreads();
    const ViewDesc& V = bin_view<BATCH>(ka, sview);
    const uint32_t local = t - V.tile_base;
    const int ox = (int)(local % (uint32_t)V.tiles_x) * T, oy = (int)(local / (uint32_t)V.tiles_x) * T;
    const int ex = imin(ox + T, V.W), ey = imin(oy + T, V.H);
    // every wave has read n: the count is consumed, leave it zero for the next frame
    if (threadIdx.x == 0) ar.tile_count[t] = 0;

    typename PiecesOf<SK>::type* wp = pieces + (threadIdx.x >> 6);
    for (uint32_t b0 = 0; b0 < n; b0 += NT) {
        const uint32_t i = b0 + threadIdx.x;
        RasterTri rt;
        int x0 = 0, y0 = 0, w = 0, h = 0;
        uint32_t key = 0, ref = 0;
        if (i < n) {
            ref = ar.bins[off + i];
            const TriRec r = rec_of(ar, ref);
            if (raster_setup(r, rt)) {
                x0 = imax(ox, rt.fx0);
                y0 = imax(oy, rt.fy0);
                w = imin(ex, rt.cx1) - x0;
                h = imin(ey, rt.cy1) - y0;
                if (w <= 0 || h <= 0) {
                    w = h = 0;
                } else if (w * h > ka.small_footprint) {
                    bigq[atomicAdd(&nbig, 1u)] = ref;
                    w = h = 0;
                }
                key = r.c.y;
            }
        }
        walk_wave<T>(wp, ar, rt, ref, x0 - ox, y0 - oy, w, h, key, keys, ox, oy);
        __syncthreads();
        const uint32_t nb = nbig;
        for (uint32_t q = 0; q < nb; ++q) {
            const TriRec r = rec_of(ar, bigq[q]);
            RasterTri rt;
            raster_setup(r, rt);
            const int x0 = imax(ox, rt.fx0), x1 = imin(ex, rt.cx1);
            const int y0 = imax(oy, rt.fy0), y1 = imin(ey, rt.cy1);
            const int w = x1 - x0, np = w * (y1 - y0);
            const uint32_t key = r.c.y;
            const float inv_w = 1.0f / (float)w;
            for (int p = threadIdx.x; p < np; p += NT) {
                // p / w as in walk_span (p < 2^16, w <= 32)
                const int dy = (int)(((float)p + 0.5f) * inv_w);
                const int px = x0 + (p - dy * w), py = y0 + dy;
                float al, be, ga, z;
                if (cover(rt, px, py, al, be, ga, z) && z < INFINITY)
                    atomicMin(&keys[(py - oy) * T + (px - ox)], depth_key(z, key));
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) nbig = 0;
        __syncthreads();
    }

    // nbig is zero again here: it counts the covered pixels
    resolve_tile_kind<T, NT, SK>(V, ar, keys, ox, oy, ka.depth_is_clear != 0, ka.clear_depth, plist, &nbig);
}

// One workgroup per entry of the list of non-empty bins (the grid covers
// every bin; blocks past the list exit after one load).  Workgroups are
// dispatched to the 8 XCDs round-robin, so block ids are permuted to give
// each XCD a contiguous run of the list: neighbouring bins share triangles,
// vertices and texels in that XCD's L2.  SK >= 0: built for one shader kind
// (see k_resolve).
constexpr uint32_t NUM_XCD = 8;

// Register bounds per shader kind (waves per SIMD the registers allow; the
// LDS of a 128-thread block at 32 px bins, 19.9 KB with the full walk
// pieces, allows four).  At most 80 VGPRs for the kinds that fit without
// spilling.  The point-light kind at five waves (at most 96 VGPRs, 8
// spilled) with the compact walk pieces (12.7 KB of LDS): the 256-view
// batch's k_raster 1.224 -> 1.192 ms; the compact pieces alone cost
// 0.04 ms and slowed the flat-colour soup by 11 %, so the other kinds keep
// the full ones.  Phong and PBR take the registers they need (held to 80,
// the point-light kind had spilled 22 VGPRs and run 1.68 against 1.38 ms).
constexpr int raster_min_waves(int sk) {
    return (sk == SHADER_FLAT || sk == SHADER_GOURAUD || sk == SHADER_FLAT_COLOR) ? 6
           : sk == SHADER_PHONG_POINT                                            ? 5
                                                                                 : 1;
}

template <int T, int NT, bool BATCH, int SK>
__global__ __launch_bounds__(NT, raster_min_waves(SK)) void k_raster(KArgs ka, Arena ar) {
    __shared__ unsigned long long keys[T * T];
    __shared__ uint32_t bigq[NT];
    __shared__ uint16_t plist[T * T];
    __shared__ uint32_t nbig;
    __shared__ ViewDesc sview[1];
    __shared__ typename PiecesOf<SK>::type pieces[NT / 64];
    if (ar.counters[CNT_ERROR]) return;
    // the pair overflow count was last needed by the scatter
    if (blockIdx.x == 0 && threadIdx.x == 0) ar.counters[CNT_OVF_PAIRS] = 0;
    const uint32_t nactive = ar.counters[CNT_ACTIVE];
    const uint32_t chunk = (nactive + NUM_XCD - 1) / NUM_XCD;
    const uint32_t j = blockIdx.x / NUM_XCD;
    const uint32_t a = (blockIdx.x % NUM_XCD) * chunk + j;
    if (j >= chunk || a >= nactive) return;
    raster_bin<T, NT, BATCH, SK>(ka, ar, ar.active[a], keys, bigq, nbig, plist, sview, pieces);
}

// ---------------------------------------------------------------------------
// k_big: visibility mode, the work items of large triangles.  One wave per
// item: BIG_CHUNK consecutive pixels (row-major) of the triangle's screen
// rectangle, four per lane.

template <int T, bool BATCH>
__global__ __launch_bounds__(256) void k_big(KArgs ka, Arena ar) {
    if (ar.counters[CNT_ERROR]) return;
    const unsigned long long c = *big_counter(ar);
    const uint32_t nitems = imin((int)(uint32_t)c, (int)ar.cap_big_items);
    const uint32_t nrecs = imin((int)(uint32_t)(c >> 32), (int)ar.cap_big_recs);
    const uint32_t lane = __lane_id();
    const uint32_t nw = gridDim.x * 4u;
    for (uint32_t it = (blockIdx.x * 256u + threadIdx.x) >> 6; it < nitems; it += nw) {
        const uint32_t ri = wave_find_le(ar.big_start, nrecs, it);
        const TriRec r = ar.big_rec[ri];
        uint32_t vi = 0;
        if constexpr (BATCH) vi = r.c.z;
        const ViewDesc& V = view_at<BATCH>(ka, vi);
        RasterTri rt;
        raster_setup(r, rt);
        const Rect rc = screen_rect(rt, V.W, V.H);
        const uint32_t w = (uint32_t)(rc.x1 - rc.x0);
        const uint32_t np = w * (uint32_t)(rc.y1 - rc.y0);
        const uint32_t p0 = (it - ar.big_start[ri]) * BIG_CHUNK;
        const uint32_t p1 = p0 + BIG_CHUNK < np ? p0 + BIG_CHUNK : np;
        unsigned long long* vk = ar.vkeys + V.pix_base;
        const uint32_t key = r.c.y;
        for (uint32_t p = p0 + lane; p < p1; p += 64u) {
            const int px = rc.x0 + (int)(p % w), py = rc.y0 + (int)(p / w);
            float al, be, ga, z;
            if (cover(rt, px, py, al, be, ga, z) && z < INFINITY) {
                atomicMin(&vk[(size_t)py * (size_t)V.W + (size_t)px], depth_key(z, key));
                ar.tile_count[V.tile_base + (uint32_t)((py / T) * V.tiles_x + px / T)] = 1u;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_resolve: visibility mode's last stage.  One wave per bin; bins
// without keys are skipped at once.  Every pixel with a key: recompute the winner's
// setup from its face (or read its clipped piece), depth-test against the
// framebuffer (strict less, reference mergeTileResults), shade, store, and
// leave the key empty for the next frame.

// Shades the winner of key k at (px, py) if it passes the depth test
// against dfb (the framebuffer's depth, or the depth of a clear the draw
// absorbs): returns whether it did, with its depth and packed colour.
// VARY: the view may carry the batch's vertex-stage outputs (V.vary); false
// for plain draws, whose views never do.
template <class S, bool VARY = true>
__device__ __forceinline__ bool shade_vis_pixel(const ViewDesc& V, const Arena& ar, unsigned long long k, int px,
                                                int py, float dfb, float& z, uint32_t& bgra, bool no_clip = false) {
    const uint32_t key = (uint32_t)k;
    const uint32_t s = key >> 3, j = key & 7u;
    const DevGroup G = group_of(V, s);
    const uint32_t face = G.first + (s - G.rank_base);
    const uint3 ix = gld(reinterpret_cast<const uint3*>(V.idx), face);
    const uint32_t i0 = ix.x, i1 = ix.y, i2 = ix.z;
    // the shaders' vertex stage reads the whole vertex; its first three
    // floats are the position (the packed copy setup read holds the same bits)
    constexpr bool kVerts = sizeof(typename S::Var) > 1;
    Vertex56 v0, v1, v2;
    vec3 p0, p1, p2;
    if (kVerts && !(VARY && V.vary)) {
        v0 = gld(V.verts, i0);
        v1 = gld(V.verts, i1);
        v2 = gld(V.verts, i2);
        p0 = make3(v0.px, v0.py, v0.pz);
        p1 = make3(v1.px, v1.py, v1.pz);
        p2 = make3(v2.px, v2.py, v2.pz);
    } else {
        p0 = load_pos(V.pos, i0);
        p1 = load_pos(V.pos, i1);
        p2 = load_pos(V.pos, i2);
    }
    const vec4 c0 = V.mvp * make4(p0, 1.0f), c1 = V.mvp * make4(p1, 1.0f), c2 = V.mvp * make4(p2, 1.0f);
    TriRec r;
    bool clipped = false;
    uint32_t slot = 0;
    // no_clip: the draw's faces all lie inside the frustum (the host's bound)
    if (!no_clip && classify(c0, c1, c2) == CLS_CLIP) {
        clipped = true;
        slot = ar.clip_base[V.face_base + s] + j;   // written by k_clip
        r = ar.clip_rec[slot];
    } else {
        r = make_rec(c0, c1, c2, V.W, V.H);
    }
    RasterTri rt;
    raster_setup(r, rt);
    float al, be, ga;
    cover(rt, px, py, al, be, ga, z);
    if (!(z < dfb)) return false;   // reference mergeTileResults: strict less
    ShadeCtx c;
    c.u = &V.u;
    const DevMaterial mat = gld(V.materials, G.material & REC_MATERIAL_MASK);
    c.m = &mat;
    c.face_colors = V.face_colors;
    c.prim = face;
    typename S::Var var[3];
    if constexpr (kVerts) {
        if (clipped) {
            for (int q = 0; q < 3; ++q) var[q] = S::vs(ar.clip_verts[3 * (size_t)slot + q], V.u);
        } else if (VARY && V.vary) {
            const typename S::Var* vy = static_cast<const typename S::Var*>(V.vary);
            var[0] = gld(vy, i0);
            var[1] = gld(vy, i1);
            var[2] = gld(vy, i2);
        } else {
            var[0] = S::vs(v0, V.u);
            var[1] = S::vs(v1, V.u);
            var[2] = S::vs(v2, V.u);
        }
    }
    bgra = pack_bgra(S::fs(make3(al, be, ga), var, c));
    return true;
}

// Returns whether the winner passed the depth test against dfb and was
// written.
// VARY, INSIDE: as for shade_vis_pixel and k_shade.
template <class S, bool VARY, bool INSIDE>
__device__ __forceinline__ bool resolve_vis_pixel(const ViewDesc& V, const Arena& ar, unsigned long long k, int px,
                                                  int py, size_t pix, float dfb) {
    float z;
    uint32_t bgra;
    if (!shade_vis_pixel<S, VARY>(V, ar, k, px, py, dfb, z, bgra, INSIDE)) return false;
    gst(V.depth, pix, z);
    gst(V.color, pix, bgra);
    if (V.ids) gst(V.ids, pix, (uint32_t)k);
    return true;
}

// A clear that a draw absorbs (its values; on = false: none).
struct ClearSpec {
    bool on;
    uint32_t bgra;
    float depth;
};

__device__ __forceinline__ void clear_pixel(const ViewDesc& V, size_t pix, const ClearSpec& cs) {
    gst(V.color, pix, cs.bgra);
    gst(V.depth, pix, cs.depth);
    if (V.ids) gst(V.ids, pix, 0xFFFFFFFFu);
}

// Pixels p0, p0 + step, ... of the bin at (ox, oy), row-major.
template <int T, class S, bool VARY, bool INSIDE>
__device__ __forceinline__ void resolve_vis_tile(const ViewDesc& V, const Arena& ar, int ox, int oy,
                                                 const ClearSpec& cs, int p0, int step) {
    unsigned long long* vk = ar.vkeys + V.pix_base;
    for (int p = p0; p < T * T; p += step) {
        const int px = ox + (p % T), py = oy + (p / T);
        if (px >= V.W || py >= V.H) continue;
        const size_t pix = (size_t)py * (size_t)V.W + (size_t)px;
        const unsigned long long k = vk[pix];
        bool drawn = false;
        if (k != EMPTY_KEY) {
            vk[pix] = EMPTY_KEY;
            drawn = resolve_vis_pixel<S, VARY, INSIDE>(V, ar, k, px, py, pix, cs.on ? cs.depth : gld(V.depth, pix));
        }
        if (!drawn && cs.on) clear_pixel(V, pix, cs);
    }
}

// Pixels p0, p0 + step, ... of a bin without keys, when the draw absorbs a
// clear.
template <int T>
__device__ __forceinline__ void clear_vis_tile(const ViewDesc& V, int ox, int oy, const ClearSpec& cs, int p0,
                                               int step) {
    for (int p = p0; p < T * T; p += step) {
        const int px = ox + (p % T), py = oy + (p / T);
        if (px >= V.W || py >= V.H) continue;
        clear_pixel(V, (size_t)py * (size_t)V.W + (size_t)px, cs);
    }
}

template <int T, int SK, bool VARY, bool INSIDE>
__device__ __forceinline__ void resolve_vis_tile_kind(const ViewDesc& V, const Arena& ar, int ox, int oy,
                                                      const ClearSpec& cs, int p0, int step) {
    if constexpr (SK >= 0) {
        resolve_vis_tile<T, typename ShaderOf<SK>::type, VARY, INSIDE>(V, ar, ox, oy, cs, p0, step);
    } else {
        switch (V.u.kind) {
        case SHADER_FLAT: resolve_vis_tile<T, FlatShader, VARY, INSIDE>(V, ar, ox, oy, cs, p0, step); break;
        case SHADER_PHONG: resolve_vis_tile<T, PhongShader, VARY, INSIDE>(V, ar, ox, oy, cs, p0, step); break;
        case SHADER_PBR: resolve_vis_tile<T, PBRShader, VARY, INSIDE>(V, ar, ox, oy, cs, p0, step); break;
        case SHADER_GOURAUD: resolve_vis_tile<T, GouraudShader, VARY, INSIDE>(V, ar, ox, oy, cs, p0, step); break;
        case SHADER_PHONG_POINT: resolve_vis_tile<T, PhongPointShader, VARY, INSIDE>(V, ar, ox, oy, cs, p0, step); break;
        default: resolve_vis_tile<T, FlatColorShader, VARY, INSIDE>(V, ar, ox, oy, cs, p0, step); break;
        }
    }
}

// One wave per bin (four per workgroup, grid-stride over the bins): small
// bins cost no workgroup dispatch of their own and need no barrier.  SK >= 0
// builds the kernel for that shader kind alone (every view uses it), so its
// register budget is that shader's and not the largest of all six; SK = -1
// picks the shader per bin.
//
// ka.clear: the framebuffers' clear (a plain draw's pending clear, a batch's
// or a server batch's clear, the same values for every view) is written here
// instead of by a pass of its own: every bin is visited, a pixel no winner
// is written to takes the clear values, and winners are depth-tested
// against the clear depth without reading the framebuffer.  If the frame
// failed (an arena overflow), only the clear is written, as a separate
// clear pass would have.
// INSIDE (plain draws only): every face lies inside the frustum
// (ka.skip_clip), so the winners' classification is left out.  Only a
// batch's views carry vertex-stage outputs (VARY = BATCH).
template <int T, bool BATCH, int SK, bool INSIDE = false>
__global__ __launch_bounds__(256) void k_resolve(KArgs ka, Arena ar) {
    const bool failed = ar.counters[CNT_ERROR] != 0u;
    const ClearSpec cs{ka.clear != 0, ka.clear_bgra, ka.clear_depth};
    if (failed && !cs.on) return;
    if (!failed && blockIdx.x == 0 && threadIdx.x == 0) {
        // the per-frame counters of this mode were last needed by k_big
        ar.counters[CNT_CLIP_FACES] = 0;
        ar.counters[CNT_CLIP_RECS] = 0;
        *big_counter(ar) = 0ull;
    }
    if constexpr (T == 8) {
        // one wave per bin
        const uint32_t nw = gridDim.x * 4u;
        for (uint32_t t = (blockIdx.x * 256u + threadIdx.x) >> 6; t < ka.total_tiles; t += nw) {
            // the same word for the whole wave; the branch waits for the
            // load, so every lane has read the flag before lane 0 clears it
            const bool keys = !failed && ar.tile_count[t] != 0u;
            if (!keys && !cs.on) continue;
            if (keys && __lane_id() == 0) ar.tile_count[t] = 0u;
            uint32_t vi = 0;
            if constexpr (BATCH) vi = ka.tile_views[t];
            const ViewDesc& V = view_at<BATCH>(ka, vi);
            const uint32_t local = t - V.tile_base;
            const int ox = (int)(local % (uint32_t)V.tiles_x) * T, oy = (int)(local / (uint32_t)V.tiles_x) * T;
            if (keys) resolve_vis_tile_kind<T, SK, BATCH, INSIDE>(V, ar, ox, oy, cs, (int)__lane_id(), 64);
            else clear_vis_tile<T>(V, ox, oy, cs, (int)__lane_id(), 64);
        }
    } else {
        // one workgroup (four waves, a pixel or more per thread) per bin:
        // every wave reads the flag before the barrier, thread 0 clears it
        // after
        for (uint32_t t = blockIdx.x; t < ka.total_tiles; t += gridDim.x) {
            const bool keys = !failed && ar.tile_count[t] != 0u;
            __syncthreads();
            if (keys && threadIdx.x == 0) ar.tile_count[t] = 0u;
            if (!keys && !cs.on) continue;
            uint32_t vi = 0;
            if constexpr (BATCH) vi = ka.tile_views[t];
            const ViewDesc& V = view_at<BATCH>(ka, vi);
            const uint32_t local = t - V.tile_base;
            const int ox = (int)(local % (uint32_t)V.tiles_x) * T, oy = (int)(local / (uint32_t)V.tiles_x) * T;
            if (keys) resolve_vis_tile_kind<T, SK, BATCH, INSIDE>(V, ar, ox, oy, cs, (int)threadIdx.x, 256);
            else clear_vis_tile<T>(V, ox, oy, cs, (int)threadIdx.x, 256);
        }
    }
}

// ---------------------------------------------------------------------------
// The resolve of a plain draw that absorbs a clear, in two passes, so that
// the shading need not wait for an earlier draw into the same framebuffer:
// winners are tested against the clear depth, never against the planes.
// k_shade shades every winner into its own key slot (depth bits << 32 |
// packed colour; a winner behind the clear empties the slot); k_finish, the
// only pass that touches the framebuffer, writes every pixel from its slot
// or the clear and leaves the slots and bin flags empty.  For a framebuffer
// with an id plane (IDS) the slot keeps the winner's key in its low word
// and the colour goes to the arena's colour slots (vcol).

// INSIDE: as for k_vis, every face lies inside the frustum (ka.skip_clip),
// so the winners' classification is left out.  That instantiation without an
// id plane asks for seven waves per SIMD (at most 72 VGPRs, which the Phong
// kinds fit without spilling and PBR would not): the 1080p frame's k_shade
// 17.2 -> 16.3 us against six waves; eight spilled and ran 19.2 us.
template <int T, int SK, bool IDS, bool INSIDE = false>
__global__ __launch_bounds__(256, (INSIDE && !IDS && SK != SHADER_PBR) ? 7 : 1) void k_shade(KArgs ka, Arena ar) {
    using S = typename ShaderOf<SK>::type;
    if (ar.counters[CNT_ERROR]) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the per-frame counters of this mode were last needed by k_big
        ar.counters[CNT_CLIP_FACES] = 0;
        ar.counters[CNT_CLIP_RECS] = 0;
        *big_counter(ar) = 0ull;
    }
    const ViewDesc& V = ka.one;
    unsigned long long* vk = ar.vkeys + V.pix_base;
    // one wave per 64 pixels of a bin (SUB waves per bin), one pixel per
    // lane: a bin edge of 16 or 32 gives the pass as many waves as one of 8
    constexpr uint32_t SUB = (uint32_t)(T * T + 63) / 64;
    const uint32_t nunits = ka.total_tiles * SUB;
    const int lane = (int)__lane_id();
    const uint32_t nw = gridDim.x * 4u;
    for (uint32_t u = (blockIdx.x * 256u + threadIdx.x) >> 6; u < nunits; u += nw) {
        const uint32_t t = u / SUB;
        if (ar.tile_count[t] == 0u) continue;
        const int p = (int)(u % SUB) * 64 + lane;   // row-major within the bin
        const uint32_t local = t - V.tile_base;
        const int px = (int)(local % (uint32_t)V.tiles_x) * T + (p % T);
        const int py = (int)(local / (uint32_t)V.tiles_x) * T + (p / T);
        if (p >= T * T || px >= V.W || py >= V.H) continue;
        const size_t pix = (size_t)py * (size_t)V.W + (size_t)px;
        const unsigned long long k = vk[pix];
        if (k == EMPTY_KEY) continue;
        float z;
        uint32_t bgra;
        const bool drawn = shade_vis_pixel<S, false>(V, ar, k, px, py, ka.clear_depth, z, bgra, INSIDE);
        vk[pix] = drawn ? ((unsigned long long)__float_as_uint(z) << 32) | (IDS ? (uint32_t)k : bgra) : EMPTY_KEY;
        if (IDS && drawn) ar.vcol[V.pix_base + pix] = bgra;
    }
}

// One wave per strip of 64 pixels by one bin row (64 / T bins side by
// side), one lane per column: each row of the strip is one 256-byte store
// per plane.  A slot is read only in bins that had keys.
template <int T, bool IDS>
__global__ __launch_bounds__(256) void k_finish(KArgs ka, Arena ar) {
    static_assert(64 % T == 0, "a strip holds whole bins");
    const bool failed = ar.counters[CNT_ERROR] != 0u;
    const ViewDesc& V = ka.one;
    const ClearSpec cs{true, ka.clear_bgra, ka.clear_depth};
    unsigned long long* vk = ar.vkeys + V.pix_base;
    const uint32_t lane = __lane_id();
    const uint32_t sx_n = ((uint32_t)V.W + 63u) / 64u;
    const uint32_t nstrips = sx_n * (uint32_t)V.tiles_y;
    const uint32_t nw = gridDim.x * 4u;
    for (uint32_t st = (blockIdx.x * 256u + threadIdx.x) >> 6; st < nstrips; st += nw) {
        const uint32_t sx = st % sx_n, by = st / sx_n;
        const int px = (int)(sx * 64u + lane);
        const bool in_x = px < V.W;
        const uint32_t t = V.tile_base + by * (uint32_t)V.tiles_x + (uint32_t)px / (uint32_t)T;
        const bool keys = !failed && in_x && ar.tile_count[t] != 0u;
        const int rows = imin(T, V.H - (int)by * T);
        const size_t pix0 = (size_t)by * T * (size_t)V.W + (size_t)px;
        constexpr int RB = T < 8 ? T : 8;   // rows whose slots are loaded together
        for (int r0 = 0; r0 < rows; r0 += RB) {
            // all slot loads of RB rows first: the stores below may alias
            // them for the compiler, which would otherwise wait for each
            unsigned long long sl[RB];
#pragma unroll
            for (int r = 0; r < RB; ++r)
                sl[r] = keys && r0 + r < rows ? vk[pix0 + (size_t)(r0 + r) * V.W] : EMPTY_KEY;
            if (!in_x) continue;
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                if (r0 + r >= rows) break;
                const size_t pix = pix0 + (size_t)(r0 + r) * V.W;
                if (sl[r] != EMPTY_KEY) {
                    vk[pix] = EMPTY_KEY;
                    V.color[pix] = IDS ? ar.vcol[V.pix_base + pix] : (uint32_t)sl[r];
                    V.depth[pix] = __uint_as_float((uint32_t)(sl[r] >> 32));
                    if (IDS) V.ids[pix] = (uint32_t)sl[r];
                } else {
                    clear_pixel(V, pix, cs);
                }
            }
        }
        // every lane of a bin read its flag before the loop's loads returned
        if (keys && (uint32_t)px % (uint32_t)T == 0u) ar.tile_count[t] = 0u;
    }
}

// k_finish for a width that is a multiple of 4 (every row then starts
// 16-byte aligned in the 256-byte aligned planes): one wave per strip of
// 64 pixels by one bin row, each lane four adjacent pixels of every fourth
// row, 16-byte loads and stores (each store instruction covers four rows
// of 256 contiguous bytes).
template <int T, bool IDS>
__global__ __launch_bounds__(256) void k_finish4(KArgs ka, Arena ar) {
    static_assert(64 % T == 0 && T % 4 == 0, "a strip holds whole bins, a lane's pixels one bin");
    constexpr int RL = T / 4;   // rows per lane
    const bool failed = ar.counters[CNT_ERROR] != 0u;
    const ViewDesc& V = ka.one;
    unsigned long long* vk = ar.vkeys + V.pix_base;
    const uint4 cv = make_uint4(ka.clear_bgra, ka.clear_bgra, ka.clear_bgra, ka.clear_bgra);
    const float4 dv = make_float4(ka.clear_depth, ka.clear_depth, ka.clear_depth, ka.clear_depth);
    const ulonglong2 ev2 = make_ulonglong2(EMPTY_KEY, EMPTY_KEY);
    const uint32_t lane = __lane_id();
    const uint32_t cg = lane & 15u, rg = lane >> 4;
    const uint32_t sx_n = ((uint32_t)V.W + 63u) / 64u;
    const uint32_t nstrips = sx_n * (uint32_t)V.tiles_y;
    const uint32_t nw = gridDim.x * 4u;
    for (uint32_t st = (blockIdx.x * 256u + threadIdx.x) >> 6; st < nstrips; st += nw) {
        const uint32_t sx = st % sx_n, by = st / sx_n;
        const int px = (int)(sx * 64u + 4u * cg);
        if (px >= V.W) continue;   // W % 4 == 0: the lane's four pixels are all inside or all outside
        const uint32_t t = V.tile_base + by * (uint32_t)V.tiles_x + (uint32_t)px / (uint32_t)T;
        const bool keys = !failed && ar.tile_count[t] != 0u;
        const int y0 = (int)by * T;
        // all slot loads first: the stores below may alias them for the
        // compiler, which would otherwise wait for each
        ulonglong2 a[RL], b[RL];
#pragma unroll
        for (int k = 0; k < RL; ++k) {
            a[k] = b[k] = ev2;
            const int py = y0 + (int)rg + 4 * k;
            if (keys && py < V.H) {
                const ulonglong2* q = reinterpret_cast<const ulonglong2*>(vk + (size_t)py * V.W + (size_t)px);
                a[k] = q[0];
                b[k] = q[1];
            }
        }
#pragma unroll
        for (int k = 0; k < RL; ++k) {
            const int py = y0 + (int)rg + 4 * k;
            if (py >= V.H) break;
            const size_t pix = (size_t)py * V.W + (size_t)px;
            const unsigned long long s0 = a[k].x, s1 = a[k].y, s2 = b[k].x, s3 = b[k].y;
            uint4 c = cv;
            float4 d = dv;
            uint4 id = make_uint4(~0u, ~0u, ~0u, ~0u);
            if ((s0 & s1 & s2 & s3) != EMPTY_KEY) {
                if (s0 != EMPTY_KEY) { c.x = (uint32_t)s0; d.x = __uint_as_float((uint32_t)(s0 >> 32)); }
                if (s1 != EMPTY_KEY) { c.y = (uint32_t)s1; d.y = __uint_as_float((uint32_t)(s1 >> 32)); }
                if (s2 != EMPTY_KEY) { c.z = (uint32_t)s2; d.z = __uint_as_float((uint32_t)(s2 >> 32)); }
                if (s3 != EMPTY_KEY) { c.w = (uint32_t)s3; d.w = __uint_as_float((uint32_t)(s3 >> 32)); }
                if constexpr (IDS) {
                    // the slots hold the winners' keys; their colours are in vcol
                    id = c;
                    const uint4 vc = *reinterpret_cast<const uint4*>(ar.vcol + V.pix_base + pix);
                    c = make_uint4(s0 != EMPTY_KEY ? vc.x : cv.x, s1 != EMPTY_KEY ? vc.y : cv.y,
                                   s2 != EMPTY_KEY ? vc.z : cv.z, s3 != EMPTY_KEY ? vc.w : cv.w);
                    id = make_uint4(s0 != EMPTY_KEY ? id.x : ~0u, s1 != EMPTY_KEY ? id.y : ~0u,
                                    s2 != EMPTY_KEY ? id.z : ~0u, s3 != EMPTY_KEY ? id.w : ~0u);
                }
                ulonglong2* q = reinterpret_cast<ulonglong2*>(vk + pix);
                q[0] = ev2;
                q[1] = ev2;
            }
            typedef unsigned int v4u __attribute__((ext_vector_type(4)));
            gst_nt(reinterpret_cast<v4u*>(V.color + pix), 0, v4u{c.x, c.y, c.z, c.w});
            gst_nt(reinterpret_cast<v4u*>(V.depth + pix), 0,
                   v4u{__float_as_uint(d.x), __float_as_uint(d.y), __float_as_uint(d.z), __float_as_uint(d.w)});
            if constexpr (IDS) gst_nt(reinterpret_cast<v4u*>(V.ids + pix), 0, v4u{id.x, id.y, id.z, id.w});
        }
        // every lane of a bin read its flag before the loop's loads returned
        if (keys && rg == 0u && (uint32_t)px % (uint32_t)T == 0u) ar.tile_count[t] = 0u;
    }
}

// ---------------------------------------------------------------------------
// framebuffer clear (reference framebuffer.cpp:26-42)

// Planes start 256-byte aligned (hipMalloc, or 256-byte aligned offsets in
// a server block), so four pixels go per 16-byte store; the tail is scalar.
__device__ __forceinline__ void clear_planes(uint32_t* color, float* depth, uint32_t* ids, size_t n, uint32_t bgra,
                                             float d, size_t first, size_t stride) {
    const size_t n4 = n / 4;
    // non-temporal 16-byte stores (a cache hint only): a clear streams whole
    // planes, which later passes read back at most once
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    v4u* c4 = reinterpret_cast<v4u*>(color);
    v4u* d4 = reinterpret_cast<v4u*>(depth);
    v4u* i4 = reinterpret_cast<v4u*>(ids);
    const uint32_t du = __float_as_uint(d);
    const v4u cv = {bgra, bgra, bgra, bgra}, dv = {du, du, du, du}, iv = {~0u, ~0u, ~0u, ~0u};
    for (size_t i = first; i < n4; i += stride) {
        __builtin_nontemporal_store(cv, c4 + i);
        __builtin_nontemporal_store(dv, d4 + i);
        if (ids) __builtin_nontemporal_store(iv, i4 + i);
    }
    for (size_t i = 4 * n4 + first; i < n; i += stride) {
        color[i] = bgra;
        depth[i] = d;
        if (ids) ids[i] = 0xFFFFFFFFu;
    }
}

__global__ void k_clear(uint32_t* color, float* depth, uint32_t* ids, size_t n, uint32_t bgra, float d) {
    clear_planes(color, depth, ids, n, bgra, d, blockIdx.x * (size_t)blockDim.x + threadIdx.x,
                 (size_t)gridDim.x * blockDim.x);
}

// Clear of every view of a batch: blockIdx.y picks the view (views
// [view0, view0 + gridDim.y)), the x blocks stride over its pixels.
__global__ void k_clear_views(const ViewDesc* views, uint32_t view0, uint32_t bgra, float d) {
    const ViewDesc& V = views[view0 + blockIdx.y];
    clear_planes(V.color, V.depth, V.ids, (size_t)V.W * (size_t)V.H, bgra, d, blockIdx.x * (size_t)blockDim.x + threadIdx.x,
                 (size_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// k_vary: the vertex stage of every vertex of every job (a mesh shared by
// views whose vertex stage reads the same uniforms), one lane per vertex;
// the resolve then reads each winner's three outputs instead of running
// the vertex stage per pixel.  The same function on the same inputs, so
// the varyings are bit for bit those the resolve would compute.

template <class S>
__device__ __forceinline__ void vary_one(const VaryJob& j, uint32_t v, const ViewDesc& V) {
    static_cast<typename S::Var*>(j.out)[v] = S::vs(j.verts[v], V.u);
}

__global__ __launch_bounds__(256) void k_vary(const VaryJob* jobs, const uint32_t* firsts, uint32_t njobs,
                                              uint32_t total, const ViewDesc* views) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= total) return;
    const VaryJob& j = jobs[find_le(firsts, njobs, g)];
    const uint32_t v = g - j.first;
    const ViewDesc& V = views[j.view];
    switch (j.kind) {
    case SHADER_FLAT: vary_one<FlatShader>(j, v, V); break;
    case SHADER_PHONG: vary_one<PhongShader>(j, v, V); break;
    case SHADER_PBR: vary_one<PBRShader>(j, v, V); break;
    case SHADER_GOURAUD: vary_one<GouraudShader>(j, v, V); break;
    case SHADER_PHONG_POINT: vary_one<PhongPointShader>(j, v, V); break;
    default: break;   // FLAT_COLOR: no vertex stage
    }
}

// Bytes of the vertex stage's output per vertex for a shader kind (0: the
// shader has none).
uint32_t vary_bytes(int kind) {
    switch (kind) {
    case SHADER_FLAT: return sizeof(FlatShader::Var);
    case SHADER_PHONG: return sizeof(PhongShader::Var);
    case SHADER_PBR: return sizeof(PBRShader::Var);
    case SHADER_GOURAUD: return sizeof(GouraudShader::Var);
    case SHADER_PHONG_POINT: return sizeof(PhongPointShader::Var);
    default: return 0;
    }
}

// ---------------------------------------------------------------------------
// host-side launch wrappers

// Expects a clean arena (bin counts and per-frame counters zero) and leaves
// it clean.
constexpr int RESOLVE_MAX_BLOCKS = 16384;

// Tool ranges named after the reference's phases (its Tracy zones,
// tiled_pipeline.cpp:173, 264, 268, 273, 283, 312) around each stage's
// launches, for a profiler's marker trace; on when AXR_ROCTX=1 at library
// load.  The stages of a raster mode map onto the phases as in
// kBinnedPhases / kVisPhases (the vertex pass of a batch belongs to the
// reference's raster phase, where its fragment shading runs).
static bool phase_ranges_on() {
    static const bool on = [] {
        const char* v = getenv("AXR_ROCTX");
        return v && v[0] == '1';
    }();
    return on;
}
static const char* const kBinnedPhases[kStages] = {"ClippingAllTriangles", "ClippingAllTriangles", "TilePreparation",
                                                   "BinAllTriangles", "RasterizingTiles", "RasterizingTiles"};
static const char* const kVisPhases[kStages] = {"ClippingAllTriangles", "ClippingAllTriangles", "RasterizingTiles",
                                                "RasterizingTiles", "MergingAllTiles", "RasterizingTiles"};
static thread_local const char* const* t_phases = kBinnedPhases;

// One kernel of stage `stage`.  With kernel timing (ev), the first kernel
// of a stage carries the stage's start event and its last kernel the stop
// event, attached to the dispatches (hipExtLaunchKernel).
template <class... KA, class... A>
static void stage_launch(void (*kern)(KA...), dim3 grid, dim3 block, hipStream_t st, StageEvents* ev, int stage,
                         bool first, bool last, A... args) {
    const bool ranges = phase_ranges_on();
    if (ranges && first) roctxRangePush(t_phases[stage]);
    struct Pop {
        bool on;
        ~Pop() { if (on) roctxRangePop(); }
    } pop{ranges && last};
    if (ev) {
        hipExtLaunchKernelGGL(kern, grid, block, 0, st, first ? ev->start[stage] : nullptr,
                              last ? ev->stop[stage] : nullptr, 0, args...);
        ev->used |= 1u << stage;
    } else {
        hipLaunchKernelGGL(kern, grid, block, 0, st, args...);
    }
}
template <class... KA, class... A>
static void stage_launch1(void (*kern)(KA...), dim3 grid, dim3 block, hipStream_t st, StageEvents* ev, int stage,
                          A... args) {
    stage_launch(kern, grid, block, st, ev, stage, true, true, args...);
}

static hipError_t fb_wait(hipStream_t st, const FbWait& w) {
    if (w.sig) return hipStreamWaitValue64(st, w.sig, w.value, hipStreamWaitValueGte);
    if (w.ev) return hipStreamWaitEvent(st, w.ev, 0);
    return hipSuccess;
}

// The shader-specialised kernel of ka.shader (-1: the switching one).
template <template <int> class K>
static auto by_shader(int shader) -> decltype(K<-1>::fn) {
    switch (shader) {
    case SHADER_FLAT: return K<SHADER_FLAT>::fn;
    case SHADER_PHONG: return K<SHADER_PHONG>::fn;
    case SHADER_PBR: return K<SHADER_PBR>::fn;
    case SHADER_GOURAUD: return K<SHADER_GOURAUD>::fn;
    case SHADER_PHONG_POINT: return K<SHADER_PHONG_POINT>::fn;
    case SHADER_FLAT_COLOR: return K<SHADER_FLAT_COLOR>::fn;
    default: return K<-1>::fn;
    }
}
template <int T, bool IDS, bool INSIDE> struct ShadeK {
    template <int SK> struct K {
        static constexpr auto fn = &k_shade<T, (SK < 0 ? SHADER_FLAT_COLOR : SK), IDS, INSIDE>;
    };
};
template <int T, bool BATCH, bool INSIDE> struct ResolveK {
    template <int SK> struct K { static constexpr auto fn = &k_resolve<T, BATCH, SK, INSIDE>; };
};
template <int T, int NT, bool BATCH> struct RasterK {
    template <int SK> struct K { static constexpr auto fn = &k_raster<T, NT, BATCH, SK>; };
};

// k_shade, then (after the wait) k_finish4 or k_finish: stages 3 and 4.
template <int T, bool IDS>
static hipError_t launch_two_pass(const KArgs& ka, const Arena& ar, hipStream_t st, StageEvents* ev,
                                  const FbWait& wait) {
    const uint64_t units = (uint64_t)ka.total_tiles * (uint64_t)((T * T + 63) / 64);
    const int sb = (int)std::min<uint64_t>(std::max<uint64_t>(1, (units + 3) / 4), RESOLVE_MAX_BLOCKS);
    const auto shade = ka.skip_clip ? by_shader<ShadeK<T, IDS, true>::template K>(ka.shader)
                                    : by_shader<ShadeK<T, IDS, false>::template K>(ka.shader);
    stage_launch1(shade, dim3(sb), dim3(256), st, ev, 3, ka, ar);
    hipError_t e = fb_wait(st, wait);
    if (e != hipSuccess) return e;
    const uint32_t strips = (((uint32_t)ka.one.W + 63u) / 64u) * (uint32_t)ka.one.tiles_y;
    const int fblocks = imax(1, imin((int)((strips + 3) / 4), RESOLVE_MAX_BLOCKS));
    if (ka.one.W % 4 == 0) stage_launch1(k_finish4<T, IDS>, dim3(fblocks), dim3(256), st, ev, 4, ka, ar);
    else stage_launch1(k_finish<T, IDS>, dim3(fblocks), dim3(256), st, ev, 4, ka, ar);
    return hipGetLastError();
}

// Visibility mode: stage 0 k_vis, 1 k_clip, 2 k_big, then either 4
// k_resolve or (two-pass) 3 k_shade and 4 k_finish.
template <int T, bool BATCH>
static hipError_t launch_vis_t(const KArgs& ka, const Arena& ar, hipStream_t st, StageEvents* ev, const FbWait& wait) {
    t_phases = kVisPhases;
    const uint32_t vis_blocks = (ka.total_faces + 256u * VIS_FACES_PER_LANE - 1) / (256u * VIS_FACES_PER_LANE);
    if (vis_blocks) {
        if (!BATCH && ka.skip_clip) stage_launch1(k_vis<T, false, true>, dim3(vis_blocks), dim3(256), st, ev, 0, ka, ar);
        else stage_launch1(k_vis<T, BATCH>, dim3(vis_blocks), dim3(256), st, ev, 0, ka, ar);
    }
    // passes the host showed to have no work are not launched
    if (BATCH || !ka.skip_clip) stage_launch1(k_clip<T, BATCH, true>, dim3(256), dim3(64), st, ev, 1, ka, ar);
    if (BATCH || !ka.skip_big) stage_launch1(k_big<T, BATCH>, dim3(1024), dim3(256), st, ev, 2, ka, ar);
    if (!BATCH && ka.two_pass && ka.clear && ka.shader >= 0 && ka.shader < SHADER_COUNT) {
        // shade at once, write the framebuffer after the earlier draws
        return ka.one.ids ? launch_two_pass<T, true>(ka, ar, st, ev, wait)
                          : launch_two_pass<T, false>(ka, ar, st, ev, wait);
    }
    // the framebuffer pass waits for what other streams were asked to do first
    hipError_t e = fb_wait(st, wait);
    if (e != hipSuccess) return e;
    // T = 8: four bins per workgroup; larger bins: one bin per workgroup
    const uint32_t per_block = T == 8 ? 4u : 1u;
    const int rb = imax(1, imin((int)((ka.total_tiles + per_block - 1) / per_block), RESOLVE_MAX_BLOCKS));
    const auto resolve = (!BATCH && ka.skip_clip) ? by_shader<ResolveK<T, false, true>::template K>(ka.shader)
                                                  : by_shader<ResolveK<T, BATCH, false>::template K>(ka.shader);
    stage_launch1(resolve, dim3(rb), dim3(256), st, ev, 4, ka, ar);
    return hipGetLastError();
}

// Binned mode: stage 0 k_setup, 1 k_clip, 2 the scan, 3 k_scatter, 4 k_raster.
template <int T, int NT, bool BATCH>
static hipError_t launch_frame_t(const KArgs& ka, const Arena& ar, hipStream_t st, StageEvents* ev,
                                 const FbWait& wait) {
    if (ka.mode == RASTER_VISIBILITY) return launch_vis_t<T, BATCH>(ka, ar, st, ev, wait);
    t_phases = kBinnedPhases;
    const uint32_t ntiles = ka.total_tiles;
    const uint32_t setup_blocks = (ka.total_faces + 255) / 256;
    if (setup_blocks) {
        if (ka.skip_clip) stage_launch1(k_setup<T, BATCH, true>, dim3(setup_blocks), dim3(256), st, ev, 0, ka, ar);
        else stage_launch1(k_setup<T, BATCH>, dim3(setup_blocks), dim3(256), st, ev, 0, ka, ar);
    }
    if (!ka.skip_clip) stage_launch1(k_clip<T, BATCH, false>, dim3(256), dim3(64), st, ev, 1, ka, ar);
    if (ntiles <= SMALL_SCAN_TILES) {
        stage_launch1(k_scan_single, dim3(1), dim3(SCAN_THREADS), st, ev, 2, ar, ntiles);
    } else {
        const uint32_t scan_blocks = (ntiles + SCAN_BLOCK - 1) / SCAN_BLOCK;
        stage_launch(k_scan_reduce, dim3(scan_blocks), dim3(SCAN_THREADS), st, ev, 2, true, false, ar, ntiles);
        stage_launch(k_scan_top, dim3(1), dim3(SCAN_THREADS), st, ev, 2, false, false, ar, scan_blocks);
        stage_launch(k_scan_down, dim3(scan_blocks), dim3(SCAN_THREADS), st, ev, 2, false, true, ar, ntiles);
    }
    const uint32_t nwaves = setup_blocks * 4;
    const uint32_t sblocks = imax(64, imin(8192, (int)((nwaves + 3) / 4)));
    stage_launch1(k_scatter, dim3(sblocks), dim3(256), st, ev, 3, ar, nwaves);
    const uint32_t rblocks = ((imax(1, (int)ntiles) + NUM_XCD - 1) / NUM_XCD) * NUM_XCD;
    hipError_t e = fb_wait(st, wait);
    if (e != hipSuccess) return e;
    stage_launch1(by_shader<RasterK<T, NT, BATCH>::template K>(ka.shader), dim3(rblocks), dim3(NT), st, ev, 4, ka, ar);
    return hipGetLastError();
}

template <int T, bool BATCH>
static hipError_t launch_frame_nt(const KArgs& ka, const Arena& ar, hipStream_t st, StageEvents* ev, const FbWait& w) {
    switch (ka.raster_threads) {
    case 64: return launch_frame_t<T, 64, BATCH>(ka, ar, st, ev, w);
    case 128: return launch_frame_t<T, 128, BATCH>(ka, ar, st, ev, w);
    default: return launch_frame_t<T, 256, BATCH>(ka, ar, st, ev, w);
    }
}

template <bool BATCH>
static hipError_t launch_frame_b(const KArgs& ka, const Arena& ar, hipStream_t st, StageEvents* ev, const FbWait& w) {
    switch (ka.tile_size) {
    case 8: return launch_frame_nt<8, BATCH>(ka, ar, st, ev, w);
    case 32: return launch_frame_nt<32, BATCH>(ka, ar, st, ev, w);
    default: return launch_frame_nt<16, BATCH>(ka, ar, st, ev, w);
    }
}

hipError_t launch_frame(const KArgs& ka, const Arena& ar, hipStream_t st, StageEvents* ev, const FbWait& wait) {
    // stage 5: a batch's vertex stage once per vertex of each shared mesh,
    // before the resolve reads it
    if (ka.table && ka.vary_total)
        stage_launch1(k_vary, dim3((ka.vary_total + 255) / 256), dim3(256), st, ev, 5, ka.vary_jobs, ka.vary_firsts,
                      ka.vary_njobs, ka.vary_total, ka.table);
    return ka.table ? launch_frame_b<true>(ka, ar, st, ev, wait) : launch_frame_b<false>(ka, ar, st, ev, wait);
}

hipError_t launch_clear_views(const ViewDesc* views, const uint32_t* pix_bases, uint32_t nviews, uint32_t total,
                              uint32_t bgra, float d, hipStream_t st) {
    (void)pix_bases;
    // about 8 Ki pixels of the average view per block, at most 64 blocks a view
    const uint32_t avg = nviews ? total / nviews : 0;
    const int bx = imax(1, imin(64, (int)((avg + 8191) / 8192)));
    for (uint32_t v0 = 0; v0 < nviews; v0 += 65535u) {
        const uint32_t nv = nviews - v0 < 65535u ? nviews - v0 : 65535u;
        k_clear_views<<<dim3(bx, nv), 256, 0, st>>>(views, v0, bgra, d);
    }
    return hipGetLastError();
}

hipError_t launch_prep(const Arena& ar, uint32_t ntiles, hipStream_t st) {
    k_prep<<<imax(1, imin(1024, (int)((ntiles + 255) / 256))), 256, 0, st>>>(ar, ntiles);
    return hipGetLastError();
}

hipError_t launch_clear(uint32_t* color, float* depth, uint32_t* ids, size_t n, uint32_t bgra, float d, hipStream_t st) {
    int blocks = (int)imin(2048, (int)((n / 4 + 255) / 256));
    k_clear<<<imax(1, blocks), 256, 0, st>>>(color, depth, ids, n, bgra, d);
    return hipGetLastError();
}

}  // namespace axr
