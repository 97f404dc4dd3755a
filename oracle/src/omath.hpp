// ORACLE - test infrastructure only.  This directory is a CPU restatement of
// the reference renderer's tiled pipeline.  Only the test suite, the smoke
// check and the CPU-baseline leg of bench.py may load it, and only as the
// checker; the product library never links or calls it.
//
// omath.hpp: the GLM subset the reference's hot path uses
// (reference: include/math.hpp:9-14 configures GLM; GLM itself is not vendored
// in the reference tree - external/CMakeLists.txt:13-14 expects it at
// external/glm - so its rounding is restated here from GLM's documented
// operator definitions, not pinned against GLM).
//
// Evaluation orders (one rounding per operation, no FMA contraction, built with
// -ffp-contract=off):
//   M*v (4x4) : (m0*x + m1*y) + (m2*z + m3*w)        per component
//   A*B (4x4) : ((A0*b.x + A1*b.y) + A2*b.z) + A3*b.w per column
//   M*v (3x3) : (m0*x + m1*y) + m2*z                  per component
//   dot3      : (x0*x1 + y0*y1) + z0*z1
//   normalize : v * (1/sqrt(dot(v,v)))
//   mix       : x*(1-a) + y*a
//   reflect   : I - (N*dot(N,I))*2
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

template <int N> struct vecn;
template <> struct vecn<2> { float x = 0, y = 0; vecn() = default; vecn(float a, float b) : x(a), y(b) {} };
template <> struct vecn<3> { float x = 0, y = 0, z = 0; vecn() = default; explicit vecn(float s) : x(s), y(s), z(s) {}
    vecn(float a, float b, float c) : x(a), y(b), z(c) {} };
template <> struct vecn<4> { float x = 0, y = 0, z = 0, w = 0; vecn() = default;
    vecn(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    vecn(const vecn<3>& v, float d) : x(v.x), y(v.y), z(v.z), w(d) {} };
using vec2 = vecn<2>;
using vec3 = vecn<3>;
using vec4 = vecn<4>;

inline float* comp(vec2& v) { return &v.x; }
inline float* comp(vec3& v) { return &v.x; }
inline float* comp(vec4& v) { return &v.x; }
inline const float* comp(const vec2& v) { return &v.x; }
inline const float* comp(const vec3& v) { return &v.x; }
inline const float* comp(const vec4& v) { return &v.x; }

// Component-wise binary operations, one rounding each.
template <int N, class F> inline vecn<N> zipv(const vecn<N>& a, const vecn<N>& b, F f) {
    vecn<N> r; for (int i = 0; i < N; ++i) comp(r)[i] = f(comp(a)[i], comp(b)[i]); return r;
}
template <int N, class F> inline vecn<N> mapv(const vecn<N>& a, F f) {
    vecn<N> r; for (int i = 0; i < N; ++i) comp(r)[i] = f(comp(a)[i]); return r;
}
template <int N> inline vecn<N> operator+(const vecn<N>& a, const vecn<N>& b) { return zipv(a, b, [](float p, float q) { return p + q; }); }
template <int N> inline vecn<N> operator-(const vecn<N>& a, const vecn<N>& b) { return zipv(a, b, [](float p, float q) { return p - q; }); }
template <int N> inline vecn<N> operator*(const vecn<N>& a, const vecn<N>& b) { return zipv(a, b, [](float p, float q) { return p * q; }); }
template <int N> inline vecn<N> operator*(const vecn<N>& a, float s) { return mapv(a, [s](float p) { return p * s; }); }
template <int N> inline vecn<N> operator*(float s, const vecn<N>& a) { return mapv(a, [s](float p) { return s * p; }); }
template <int N> inline vecn<N> operator/(const vecn<N>& a, float s) { return mapv(a, [s](float p) { return p / s; }); }
template <int N> inline vecn<N> operator/(const vecn<N>& a, const vecn<N>& b) { return zipv(a, b, [](float p, float q) { return p / q; }); }
template <int N> inline vecn<N> operator-(const vecn<N>& a) { return mapv(a, [](float p) { return -p; }); }
template <int N> inline vecn<N> operator-(const vecn<N>& a, float s) { return mapv(a, [s](float p) { return p - s; }); }
template <int N> inline vecn<N> operator+(const vecn<N>& a, float s) { return mapv(a, [s](float p) { return p + s; }); }

inline float dot(const vec3& a, const vec3& b) { float r = a.x * b.x; r = r + a.y * b.y; return r + a.z * b.z; }
inline vec3 cross(const vec3& a, const vec3& b) {
    return vec3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline vec3 normalize(const vec3& v) { float inv = 1.0f / std::sqrt(dot(v, v)); return v * inv; }
inline vec3 reflect(const vec3& i, const vec3& n) { return i - (n * dot(n, i)) * 2.0f; }
template <class T> inline T mix(const T& x, const T& y, float a) { return x * (1.0f - a) + y * a; }
inline float mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
template <int N> inline vecn<N> vpow(const vecn<N>& a, const vecn<N>& e) { return zipv(a, e, [](float p, float q) { return std::pow(p, q); }); }

struct mat3 {
    vec3 col[3];
    vec3& operator[](int i) { return col[i]; }
    const vec3& operator[](int i) const { return col[i]; }
};
struct mat4 {
    vec4 col[4];
    vec4& operator[](int i) { return col[i]; }
    const vec4& operator[](int i) const { return col[i]; }
    static mat4 identity() {
        mat4 m; m[0] = vec4(1, 0, 0, 0); m[1] = vec4(0, 1, 0, 0); m[2] = vec4(0, 0, 1, 0); m[3] = vec4(0, 0, 0, 1); return m;
    }
    static mat4 from_cols(const float* f) {  // 16 floats, column-major
        mat4 m; for (int c = 0; c < 4; ++c) m[c] = vec4(f[4 * c], f[4 * c + 1], f[4 * c + 2], f[4 * c + 3]); return m;
    }
};

inline vec4 operator*(const mat4& m, const vec4& v) {
    vec4 lo = m[0] * v.x + m[1] * v.y;
    vec4 hi = m[2] * v.z + m[3] * v.w;
    return lo + hi;
}
inline mat4 operator*(const mat4& a, const mat4& b) {
    mat4 r;
    for (int c = 0; c < 4; ++c) {
        vec4 acc = a[0] * b[c].x + a[1] * b[c].y;
        acc = acc + a[2] * b[c].z;
        r[c] = acc + a[3] * b[c].w;
    }
    return r;
}
inline vec3 operator*(const mat3& m, const vec3& v) {
    vec3 r;
    for (int i = 0; i < 3; ++i) {
        float s = comp(m[0])[i] * v.x + comp(m[1])[i] * v.y;
        comp(r)[i] = s + comp(m[2])[i] * v.z;
    }
    return r;
}
inline mat3 operator*(float s, const mat3& m) { mat3 r; for (int c = 0; c < 3; ++c) r[c] = s * m[c]; return r; }
inline mat3 operator+(const mat3& a, const mat3& b) { mat3 r; for (int c = 0; c < 3; ++c) r[c] = a[c] + b[c]; return r; }
inline mat3 to_mat3(const mat4& m) { mat3 r; for (int c = 0; c < 3; ++c) r[c] = vec3(m[c].x, m[c].y, m[c].z); return r; }
inline mat4 transpose(const mat4& m) {
    mat4 t; for (int c = 0; c < 4; ++c) for (int r = 0; r < 4; ++r) comp(t[r])[c] = comp(m[c])[r]; return t;
}
inline vec3 xyz(const vec4& v) { return vec3(v.x, v.y, v.z); }

// 4x4 inverse through the adjugate, 2x2 minors of rows (0,1) and (2,3).
inline mat4 inverse(const mat4& m) {
    auto e = [&](int row, int col) { return comp(m[col])[row]; };
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
    float k = 1.0f / det;
    float rr[4][4];
    rr[0][0] = ( e(1,1) * c5 - e(1,2) * c4 + e(1,3) * c3) * k;
    rr[0][1] = (-e(0,1) * c5 + e(0,2) * c4 - e(0,3) * c3) * k;
    rr[0][2] = ( e(3,1) * s5 - e(3,2) * s4 + e(3,3) * s3) * k;
    rr[0][3] = (-e(2,1) * s5 + e(2,2) * s4 - e(2,3) * s3) * k;
    rr[1][0] = (-e(1,0) * c5 + e(1,2) * c2 - e(1,3) * c1) * k;
    rr[1][1] = ( e(0,0) * c5 - e(0,2) * c2 + e(0,3) * c1) * k;
    rr[1][2] = (-e(3,0) * s5 + e(3,2) * s2 - e(3,3) * s1) * k;
    rr[1][3] = ( e(2,0) * s5 - e(2,2) * s2 + e(2,3) * s1) * k;
    rr[2][0] = ( e(1,0) * c4 - e(1,1) * c2 + e(1,3) * c0) * k;
    rr[2][1] = (-e(0,0) * c4 + e(0,1) * c2 - e(0,3) * c0) * k;
    rr[2][2] = ( e(3,0) * s4 - e(3,1) * s2 + e(3,3) * s0) * k;
    rr[2][3] = (-e(2,0) * s4 + e(2,1) * s2 - e(2,3) * s0) * k;
    rr[3][0] = (-e(1,0) * c3 + e(1,1) * c1 - e(1,2) * c0) * k;
    rr[3][1] = ( e(0,0) * c3 - e(0,1) * c1 + e(0,2) * c0) * k;
    rr[3][2] = (-e(3,0) * s3 + e(3,1) * s1 - e(3,2) * s0) * k;
    rr[3][3] = ( e(2,0) * s3 - e(2,1) * s1 + e(2,2) * s0) * k;
    mat4 out; for (int c = 0; c < 4; ++c) out[c] = vec4(rr[0][c], rr[1][c], rr[2][c], rr[3][c]);
    return out;
}

// std::clamp / std::max / std::min with the standard library's comparison
// order (an unordered comparison keeps the first argument).
inline float sclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
inline float smax(float a, float b) { return (a < b) ? b : a; }
inline float smin(float a, float b) { return (b < a) ? b : a; }

// float -> int the way x86 cvttss2si does it (truncation; NaN and values
// outside int range give INT_MIN).  The reference's (int) casts compile to
// this instruction on its x86 target.
inline int cvt_trunc_i32(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int)f;
}

}  // namespace orc
