// Oracle common math: a strict-IEEE restatement of the WoodenMath routines the
// reference's ray-tracing path calls.  TEST INFRASTRUCTURE ONLY: nothing in the
// product library may include or link anything under oracle/.
//
// WoodenMath is not vendored with the reference (pch.h:22-26), so every routine
// below restates the textbook (pbrt-v3) definition the reference names, with a
// fixed evaluation order:
//   * no FMA contraction (built with -ffp-contract=off);
//   * dot3 = (x*x' + y*y') + z*z';  cross as pbrt-v3;
//   * normalize(v) = v / sqrt(dot3(v, v)), component-wise IEEE division;
//   * every transcendental is correctly rounded: evaluated in FP64 and rounded
//     once to FP32 (CR_* below);
//   * affine transform of a point: x' = ((m00*x + m01*y) + m02*z) + m03.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <algorithm>

namespace orc {

constexpr double PI_REF = 3.14;                       // pch.h:30 (a double literal)
constexpr double PI_TRUE = 3.14159265358979323846;    // WoodenMath's own pi (textbook)
constexpr float PI_TRUE_F = 3.14159265358979323846f;

inline float cr_sin(float x) { return (float)std::sin((double)x); }
inline float cr_cos(float x) { return (float)std::cos((double)x); }
inline float cr_tan(float x) { return (float)std::tan((double)x); }
inline float cr_atan(float x) { return (float)std::atan((double)x); }
inline float cr_atan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline float cr_acos(float x) { return (float)std::acos((double)x); }
inline float cr_log(float x) { return (float)std::log((double)x); }
inline float cr_exp(float x) { return (float)std::exp((double)x); }
inline float cr_pow(float x, float y) { return (float)std::pow((double)x, (double)y); }
inline float cr_log2(float x) { return (float)std::log2((double)x); }

struct V3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
inline float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross3(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length2(V3 a) { return dot3(a, a); }
inline V3 normalize3(V3 a) { float l = std::sqrt(length2(a)); return a / l; }
inline bool is_zero3(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
inline V3 face_forward(V3 n, V3 v) { return dot3(n, v) < 0.0f ? -n : n; }

// std::min / std::max semantics: first argument kept on ties.
inline float fmin_s(float a, float b) { return (b < a) ? b : a; }
inline float fmax_s(float a, float b) { return (a < b) ? b : a; }

struct Bounds {
    float mn[3], mx[3];
};
inline Bounds bounds_empty() {
    Bounds b;
    for (int i = 0; i < 3; i++) {
        b.mn[i] = std::numeric_limits<float>::max();
        b.mx[i] = std::numeric_limits<float>::lowest();
    }
    return b;
}
inline Bounds bounds_union(const Bounds& a, const Bounds& b) {
    Bounds r;
    for (int i = 0; i < 3; i++) {
        r.mn[i] = fmin_s(a.mn[i], b.mn[i]);
        r.mx[i] = fmax_s(a.mx[i], b.mx[i]);
    }
    return r;
}
inline Bounds bounds_union_pt(const Bounds& a, V3 p) {
    Bounds r;
    for (int i = 0; i < 3; i++) {
        r.mn[i] = fmin_s(a.mn[i], p[i]);
        r.mx[i] = fmax_s(a.mx[i], p[i]);
    }
    return r;
}
// pbrt-v3 Bounds3::SurfaceArea.
inline float bounds_area(const Bounds& b) {
    float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    return 2 * ((dx * dy + dx * dz) + dy * dz);
}
// pbrt-v3 Bounds3::MaximumExtent.
inline int bounds_max_extent(const Bounds& b) {
    float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    if (dx > dy && dx > dz) return 0;
    if (dy > dz) return 1;
    return 2;
}
// pbrt-v3 Bounds3::Offset ("getLerpFactors", SLBVH.cpp:201).
inline V3 bounds_offset(const Bounds& b, V3 p) {
    V3 o = v3(p.x - b.mn[0], p.y - b.mn[1], p.z - b.mn[2]);
    if (b.mx[0] > b.mn[0]) o.x /= b.mx[0] - b.mn[0];
    if (b.mx[1] > b.mn[1]) o.y /= b.mx[1] - b.mn[1];
    if (b.mx[2] > b.mn[2]) o.z /= b.mx[2] - b.mn[2];
    return o;
}

// Row-major 3x4 affine matrix.
struct Affine { float m[3][4]; };
inline V3 xf_point(const Affine& a, V3 p) {
    return v3(((a.m[0][0] * p.x + a.m[0][1] * p.y) + a.m[0][2] * p.z) + a.m[0][3],
              ((a.m[1][0] * p.x + a.m[1][1] * p.y) + a.m[1][2] * p.z) + a.m[1][3],
              ((a.m[2][0] * p.x + a.m[2][1] * p.y) + a.m[2][2] * p.z) + a.m[2][3]);
}
inline V3 xf_vector(const Affine& a, V3 v) {
    return v3((a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z,
              (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
              (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z);
}
// Normals transform by the inverse transpose: n'_i = sum_j inv[j][i] n_j.
inline V3 xf_normal(const Affine& inv, V3 n) {
    return v3((inv.m[0][0] * n.x + inv.m[1][0] * n.y) + inv.m[2][0] * n.z,
              (inv.m[0][1] * n.x + inv.m[1][1] * n.y) + inv.m[2][1] * n.z,
              (inv.m[0][2] * n.x + inv.m[1][2] * n.y) + inv.m[2][2] * n.z);
}
// pbrt-v3 Transform::operator()(Bounds3): union of the 8 transformed corners.
inline Bounds xf_bounds(const Affine& a, const Bounds& b) {
    Bounds r = bounds_empty();
    for (int c = 0; c < 8; c++) {
        V3 p = v3((c & 1) ? b.mx[0] : b.mn[0], (c & 2) ? b.mx[1] : b.mn[1],
                  (c & 4) ? b.mx[2] : b.mn[2]);
        r = bounds_union_pt(r, xf_point(a, p));
    }
    return r;
}

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

}  // namespace orc
