// axr_math.hpp - the small vector/matrix library of the render path, usable
// from host and device code.
//
// Every operation spells out one fixed evaluation order, and all HIP and host
// sources of this library are compiled with FP contraction disabled, so the
// same expression gives the same bits on the CPU and on gfx950.  The orders
// follow the GLM conventions that the reference renderer's arithmetic goes
// through (reference: include/math.hpp:9-14 pulls in GLM; src/tiled_pipeline.cpp
// :149,210-212 and src/pipeline.cpp:252-269 use it):
//   mat4 * vec4 = (c0*x + c1*y) + (c2*z + c3*w)
//   mat4 * mat4 : column i = ((A0*b.x + A1*b.y) + A2*b.z) + A3*b.w
//   mat3 * vec3 : row r    = (m0r*x + m1r*y) + m2r*z
//   dot3        = (x*x' + y*y') + z*z'
//   normalize   = v * (1 / sqrt(dot(v, v)))
//   mix(x,y,a)  = x*(1-a) + y*a
//   reflect     = I - (N * dot(N, I)) * 2
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AXR_HD __host__ __device__ __forceinline__
#else
#define AXR_HD inline
#endif

#pragma clang fp contract(off)

namespace axr {

// p[i], loaded through the global address space.  The kernels read many
// pointers from memory (a batch's view table, a material, a texture
// descriptor); to the compiler those are generic pointers, and a load
// through one is a flat load, which waits on the LDS and scalar counters as
// well as the vector-memory one.  They always point to global memory.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ T gld(const T* p, size_t i) {
    T v;
    __builtin_memcpy(&v, (const __attribute__((address_space(1))) T*)p + i, sizeof(T));
    return v;
}
template <class T>
__device__ __forceinline__ void gst(T* p, size_t i, const T& v) {
    __builtin_memcpy((__attribute__((address_space(1))) T*)p + i, &v, sizeof(T));
}
// gst with the non-temporal hint (a cache hint only) for a scalar or native
// vector T: framebuffer writes that the frame's later passes do not read.
template <class T>
__device__ __forceinline__ void gst_nt(T* p, size_t i, T v) {
    __builtin_nontemporal_store(v, (__attribute__((address_space(1))) T*)p + i);
}
#else
template <class T>
AXR_HD T gld(const T* p, size_t i) { return p[i]; }
template <class T>
AXR_HD void gst(T* p, size_t i, const T& v) { p[i] = v; }
template <class T>
AXR_HD void gst_nt(T* p, size_t i, T v) { p[i] = v; }
#endif

struct vec2 { float x, y; };
struct vec3 { float x, y, z; };
struct vec4 { float x, y, z, w; };
// Column-major like the reference's matrices: c[i] is column i.
struct mat3 { vec3 c[3]; };
struct mat4 { vec4 c[4]; };

AXR_HD vec2 make2(float x, float y) { return vec2{x, y}; }
AXR_HD vec3 make3(float x, float y, float z) { return vec3{x, y, z}; }
AXR_HD vec4 make4(float x, float y, float z, float w) { return vec4{x, y, z, w}; }
AXR_HD vec4 make4(vec3 v, float w) { return vec4{v.x, v.y, v.z, w}; }
AXR_HD vec3 xyz(vec4 v) { return vec3{v.x, v.y, v.z}; }

// std::min / std::max / std::clamp semantics (first argument wins on
// unordered comparisons), which is what the reference's std:: calls do.
AXR_HD float smin(float a, float b) { return (b < a) ? b : a; }
AXR_HD float smax(float a, float b) { return (a < b) ? b : a; }
AXR_HD float sclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

// 1/x, correctly rounded (the same bits as the IEEE division 1.0f / x).  On
// the device the hardware reciprocal refined by one FMA Newton step gives
// those bits for every x whose biased exponent lies in [2, 252] (checked on
// gfx950 for all 4,211,081,216 such floats: tests/native/rcp_check.hip);
// the rest take the division.  Four such reciprocals per triangle setup
// cost a quarter of the instructions of the division sequence.
#if defined(__HIP_DEVICE_COMPILE__)
AXR_HD float rcp_rn(float x) {
    const uint32_t ex = (__float_as_uint(x) >> 23) & 0xFFu;
    if (ex - 2u <= 250u) {
        const float r = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
    }
    return 1.0f / x;
}
#else
AXR_HD float rcp_rn(float x) { return 1.0f / x; }
#endif
AXR_HD int imin(int a, int b) { return (b < a) ? b : a; }
AXR_HD int imax(int a, int b) { return (a < b) ? b : a; }
// std::min({a,b,c}) / std::max({a,b,c}): scan left to right, keep the first
// extreme value.
AXR_HD float smin3(float a, float b, float c) { float r = a; if (b < r) r = b; if (c < r) r = c; return r; }
AXR_HD float smax3(float a, float b, float c) { float r = a; if (r < b) r = b; if (r < c) r = c; return r; }

// vec2
AXR_HD vec2 operator+(vec2 a, vec2 b) { return vec2{a.x + b.x, a.y + b.y}; }
AXR_HD vec2 operator-(vec2 a, vec2 b) { return vec2{a.x - b.x, a.y - b.y}; }
AXR_HD vec2 operator*(vec2 a, float s) { return vec2{a.x * s, a.y * s}; }
AXR_HD vec2 operator*(float s, vec2 a) { return vec2{s * a.x, s * a.y}; }
AXR_HD vec2 operator/(vec2 a, float s) { return vec2{a.x / s, a.y / s}; }

// vec3
AXR_HD vec3 operator+(vec3 a, vec3 b) { return vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AXR_HD vec3 operator-(vec3 a, vec3 b) { return vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AXR_HD vec3 operator*(vec3 a, vec3 b) { return vec3{a.x * b.x, a.y * b.y, a.z * b.z}; }
AXR_HD vec3 operator/(vec3 a, vec3 b) { return vec3{a.x / b.x, a.y / b.y, a.z / b.z}; }
AXR_HD vec3 operator*(vec3 a, float s) { return vec3{a.x * s, a.y * s, a.z * s}; }
AXR_HD vec3 operator*(float s, vec3 a) { return vec3{s * a.x, s * a.y, s * a.z}; }
AXR_HD vec3 operator/(vec3 a, float s) { return vec3{a.x / s, a.y / s, a.z / s}; }
AXR_HD vec3 operator+(vec3 a, float s) { return vec3{a.x + s, a.y + s, a.z + s}; }
AXR_HD vec3 operator-(vec3 a, float s) { return vec3{a.x - s, a.y - s, a.z - s}; }
AXR_HD vec3 operator-(vec3 a) { return vec3{-a.x, -a.y, -a.z}; }

// vec4
AXR_HD vec4 operator+(vec4 a, vec4 b) { return vec4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
AXR_HD vec4 operator-(vec4 a, vec4 b) { return vec4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
AXR_HD vec4 operator*(vec4 a, float s) { return vec4{a.x * s, a.y * s, a.z * s, a.w * s}; }
AXR_HD vec4 operator*(float s, vec4 a) { return vec4{s * a.x, s * a.y, s * a.z, s * a.w}; }

AXR_HD float dot(vec3 a, vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
AXR_HD vec3 cross(vec3 a, vec3 b) {
    return vec3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
AXR_HD float length(vec3 v) { return sqrtf(dot(v, v)); }
AXR_HD vec3 normalize(vec3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
AXR_HD vec3 reflect(vec3 i, vec3 n) { return i - (n * dot(n, i)) * 2.0f; }
AXR_HD float mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
AXR_HD vec2 mix(vec2 x, vec2 y, float a) { return x * (1.0f - a) + y * a; }
AXR_HD vec3 mix(vec3 x, vec3 y, float a) { return x * (1.0f - a) + y * a; }
AXR_HD vec4 mix(vec4 x, vec4 y, float a) { return x * (1.0f - a) + y * a; }

AXR_HD vec4 operator*(const mat4& m, vec4 v) {
    vec4 add0 = m.c[0] * v.x + m.c[1] * v.y;
    vec4 add1 = m.c[2] * v.z + m.c[3] * v.w;
    return add0 + add1;
}
AXR_HD mat4 operator*(const mat4& a, const mat4& b) {
    mat4 r;
    for (int i = 0; i < 4; ++i)
        r.c[i] = ((a.c[0] * b.c[i].x + a.c[1] * b.c[i].y) + a.c[2] * b.c[i].z) + a.c[3] * b.c[i].w;
    return r;
}
AXR_HD vec3 operator*(const mat3& m, vec3 v) {
    return vec3{(m.c[0].x * v.x + m.c[1].x * v.y) + m.c[2].x * v.z,
                (m.c[0].y * v.x + m.c[1].y * v.y) + m.c[2].y * v.z,
                (m.c[0].z * v.x + m.c[1].z * v.y) + m.c[2].z * v.z};
}
AXR_HD mat3 operator*(float s, const mat3& m) { return mat3{{s * m.c[0], s * m.c[1], s * m.c[2]}}; }
AXR_HD mat3 operator+(const mat3& a, const mat3& b) { return mat3{{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }

AXR_HD mat4 identity4() {
    mat4 m;
    m.c[0] = vec4{1, 0, 0, 0}; m.c[1] = vec4{0, 1, 0, 0};
    m.c[2] = vec4{0, 0, 1, 0}; m.c[3] = vec4{0, 0, 0, 1};
    return m;
}
AXR_HD mat3 upper3(const mat4& m) { return mat3{{xyz(m.c[0]), xyz(m.c[1]), xyz(m.c[2])}}; }
AXR_HD mat4 transpose(const mat4& m) {
    mat4 t;
    t.c[0] = vec4{m.c[0].x, m.c[1].x, m.c[2].x, m.c[3].x};
    t.c[1] = vec4{m.c[0].y, m.c[1].y, m.c[2].y, m.c[3].y};
    t.c[2] = vec4{m.c[0].z, m.c[1].z, m.c[2].z, m.c[3].z};
    t.c[3] = vec4{m.c[0].w, m.c[1].w, m.c[2].w, m.c[3].w};
    return t;
}
AXR_HD float at(const mat4& m, int col, int row) {
    const float* p = &m.c[col].x; return p[row];
}

// General 4x4 inverse by cofactor expansion along 2x2 sub-determinants.
// Used once per draw on the host (normal matrix of the flat shader).
inline mat4 inverse(const mat4& m) {
    float a[16];
    for (int c = 0; c < 4; ++c) for (int r = 0; r < 4; ++r) a[c * 4 + r] = at(m, c, r);
    // a[c*4+r] = element (row r, column c)
    auto e = [&](int r, int c) { return a[c * 4 + r]; };
    float s0 = e(0,0) * e(1,1) - e(1,0) * e(0,1);
    float s1 = e(0,0) * e(1,2) - e(1,0) * e(0,2);
    float s2 = e(0,0) * e(1,3) - e(1,0) * e(0,3);
    float s3 = e(0,1) * e(1,2) - e(1,1) * e(0,2);
    float s4 = e(0,1) * e(1,3) - e(1,1) * e(0,3);
    float s5 = e(0,2) * e(1,3) - e(1,2) * e(0,3);
    float c5 = e(2,2) * e(3,3) - e(3,2) * e(2,3);
    float c4 = e(2,1) * e(3,3) - e(3,1) * e(2,3);
    float c3 = e(2,1) * e(3,2) - e(3,1) * e(2,2);
    float c2 = e(2,0) * e(3,3) - e(3,0) * e(2,3);
    float c1 = e(2,0) * e(3,2) - e(3,0) * e(2,2);
    float c0 = e(2,0) * e(3,1) - e(3,0) * e(2,1);
    float det = ((((s0 * c5 - s1 * c4) + s2 * c3) + s3 * c2) - s4 * c1) + s5 * c0;
    float inv = 1.0f / det;
    float r[4][4];  // r[row][col]
    r[0][0] = ( e(1,1) * c5 - e(1,2) * c4 + e(1,3) * c3) * inv;
    r[0][1] = (-e(0,1) * c5 + e(0,2) * c4 - e(0,3) * c3) * inv;
    r[0][2] = ( e(3,1) * s5 - e(3,2) * s4 + e(3,3) * s3) * inv;
    r[0][3] = (-e(2,1) * s5 + e(2,2) * s4 - e(2,3) * s3) * inv;
    r[1][0] = (-e(1,0) * c5 + e(1,2) * c2 - e(1,3) * c1) * inv;
    r[1][1] = ( e(0,0) * c5 - e(0,2) * c2 + e(0,3) * c1) * inv;
    r[1][2] = (-e(3,0) * s5 + e(3,2) * s2 - e(3,3) * s1) * inv;
    r[1][3] = ( e(2,0) * s5 - e(2,2) * s2 + e(2,3) * s1) * inv;
    r[2][0] = ( e(1,0) * c4 - e(1,1) * c2 + e(1,3) * c0) * inv;
    r[2][1] = (-e(0,0) * c4 + e(0,1) * c2 - e(0,3) * c0) * inv;
    r[2][2] = ( e(3,0) * s4 - e(3,1) * s2 + e(3,3) * s0) * inv;
    r[2][3] = (-e(2,0) * s4 + e(2,1) * s2 - e(2,3) * s0) * inv;
    r[3][0] = (-e(1,0) * c3 + e(1,1) * c1 - e(1,2) * c0) * inv;
    r[3][1] = ( e(0,0) * c3 - e(0,1) * c1 + e(0,2) * c0) * inv;
    r[3][2] = (-e(3,0) * s3 + e(3,1) * s1 - e(3,2) * s0) * inv;
    r[3][3] = ( e(2,0) * s3 - e(2,1) * s1 + e(2,2) * s0) * inv;
    mat4 out;
    for (int c = 0; c < 4; ++c) out.c[c] = vec4{r[0][c], r[1][c], r[2][c], r[3][c]};
    return out;
}

// Right-handed perspective with clip-space z in [-w, w]
// (reference: src/camera.cpp:77 calls glm::perspective with these conventions).
inline mat4 perspective(float fovy, float aspect, float znear, float zfar) {
    float t = tanf(fovy / 2.0f);
    mat4 m;
    m.c[0] = vec4{1.0f / (aspect * t), 0, 0, 0};
    m.c[1] = vec4{0, 1.0f / t, 0, 0};
    m.c[2] = vec4{0, 0, -(zfar + znear) / (zfar - znear), -1.0f};
    m.c[3] = vec4{0, 0, -(2.0f * zfar * znear) / (zfar - znear), 0};
    return m;
}
inline mat4 translate(const mat4& m, vec3 v) {
    mat4 r = m;
    r.c[3] = ((m.c[0] * v.x + m.c[1] * v.y) + m.c[2] * v.z) + m.c[3];
    return r;
}
// Rotation of `angle` radians about `axis` applied after m (m * R).
inline mat4 rotate(const mat4& m, float angle, vec3 axis) {
    float c = cosf(angle), s = sinf(angle);
    vec3 ax = normalize(axis);
    vec3 tmp = ax * (1.0f - c);
    float r00 = c + tmp.x * ax.x, r01 = tmp.x * ax.y + s * ax.z, r02 = tmp.x * ax.z - s * ax.y;
    float r10 = tmp.y * ax.x - s * ax.z, r11 = c + tmp.y * ax.y, r12 = tmp.y * ax.z + s * ax.x;
    float r20 = tmp.z * ax.x + s * ax.y, r21 = tmp.z * ax.y - s * ax.x, r22 = c + tmp.z * ax.z;
    mat4 r;
    r.c[0] = (m.c[0] * r00 + m.c[1] * r01) + m.c[2] * r02;
    r.c[1] = (m.c[0] * r10 + m.c[1] * r11) + m.c[2] * r12;
    r.c[2] = (m.c[0] * r20 + m.c[1] * r21) + m.c[2] * r22;
    r.c[3] = m.c[3];
    return r;
}

// Quaternion (w, x, y, z) for the camera's orientation (host only).  The
// operation orders follow GLM 0.9.9 (gtc/quaternion.inl, detail/type_quat.inl),
// the reference's un-vendored math library.
struct quat { float w, x, y, z; };
inline quat angle_axis(float angle, vec3 axis) {
    float s = sinf(angle * 0.5f);
    return quat{cosf(angle * 0.5f), axis.x * s, axis.y * s, axis.z * s};
}
inline float qdot(quat a, quat b) { return (a.w * b.w + a.x * b.x) + (a.y * b.y + a.z * b.z); }
inline quat qnormalize(quat q) {
    float len = sqrtf(qdot(q, q));
    if (len <= 0.0f) return quat{1, 0, 0, 0};
    float inv = 1.0f / len;
    return quat{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}
inline quat qinverse(quat q) {
    float d = qdot(q, q);
    return quat{q.w / d, -q.x / d, -q.y / d, -q.z / d};
}
// Hamilton product p*q (rotate by q, then by p), terms summed left to right.
inline quat qmul(quat p, quat q) {
    return quat{p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z,
                p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y,
                p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z,
                p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x};
}
inline vec3 qrotate(quat q, vec3 v) {
    vec3 u{q.x, q.y, q.z};
    vec3 uv = cross(u, v);
    vec3 uuv = cross(u, uv);
    return v + ((uv * q.w) + uuv) * 2.0f;
}
// Shortest-arc rotation taking unit vector a onto unit vector b (S. Melax,
// "The Shortest Arc Quaternion", Game Programming Gems 1, 2000), with GLM's
// handling of the parallel and antiparallel cases (epsilon = FLT_EPSILON).
inline quat rotation_between(vec3 a, vec3 b) {
    const float eps = 1.1920928955078125e-7f;
    float c = dot(a, b);
    if (c >= 1.0f - eps) return quat{1, 0, 0, 0};
    if (c < -1.0f + eps) {
        vec3 axis = cross(make3(0.0f, 0.0f, 1.0f), a);
        if (dot(axis, axis) < eps) axis = cross(make3(1.0f, 0.0f, 0.0f), a);
        return angle_axis(3.14159265358979323846f, normalize(axis));
    }
    vec3 axis = cross(a, b);
    float s = sqrtf((1.0f + c) * 2.0f);
    float invs = 1.0f / s;
    return quat{s * 0.5f, axis.x * invs, axis.y * invs, axis.z * invs};
}
inline mat4 mat4_cast(quat q) {
    float qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
    float qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
    float qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    mat4 r = identity4();
    r.c[0].x = 1.0f - 2.0f * (qyy + qzz); r.c[0].y = 2.0f * (qxy + qwz); r.c[0].z = 2.0f * (qxz - qwy);
    r.c[1].x = 2.0f * (qxy - qwz); r.c[1].y = 1.0f - 2.0f * (qxx + qzz); r.c[1].z = 2.0f * (qyz + qwx);
    r.c[2].x = 2.0f * (qxz + qwy); r.c[2].y = 2.0f * (qyz - qwx); r.c[2].z = 1.0f - 2.0f * (qxx + qyy);
    return r;
}

}  // namespace axr
