// Device-side numerics shared by the HIP kernels.
//
// The kernels are compiled with -ffp-contract=off and IEEE division/sqrt, so
// every float expression below rounds exactly as written.  Conventions (the
// same ones the CPU restatement of the reference uses):
//   * dot3 = (x*x' + y*y') + z*z';  cross as pbrt-v3;
//   * normalize(v) = v / sqrt(dot3(v, v)) component-wise;
//   * transcendentals are correctly rounded: evaluated in FP64, rounded once;
//   * min/max follow std::min/std::max (first argument kept on ties);
//   * affine point transform x' = ((m00*x + m01*y) + m02*z) + m03.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define WPBR_D __device__ __forceinline__
#define WPBR_HD __host__ __device__ __forceinline__

namespace wpbr {

constexpr double PI_REF = 3.14;  // the reference's `#define PI 3.14` (pch.h:30), a double
constexpr float PI_TRUE_F = 3.14159265358979323846f;
constexpr float INF = __builtin_huge_valf();
constexpr float FLT_MAX_ = 3.402823466e+38f;
constexpr float FLT_LOWEST_ = -3.402823466e+38f;

WPBR_D float cr_sin(float x) { return (float)sin((double)x); }
WPBR_D float cr_cos(float x) { return (float)cos((double)x); }
WPBR_D float cr_atan(float x) { return (float)atan((double)x); }
WPBR_D float cr_atan2(float y, float x) { return (float)atan2((double)y, (double)x); }
WPBR_D float cr_acos(float x) { return (float)acos((double)x); }
WPBR_D float cr_log(float x) { return (float)log((double)x); }
WPBR_D float cr_pow(float x, float y) { return (float)pow((double)x, (double)y); }
WPBR_D float cr_log2(float x) { return (float)log2((double)x); }

struct f3 { float x, y, z; };
WPBR_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
WPBR_HD f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
WPBR_HD f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
WPBR_HD f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
WPBR_HD f3 mul3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
WPBR_HD f3 div3(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
WPBR_HD float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
WPBR_HD f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
WPBR_HD float len2(f3 a) { return dot3(a, a); }
WPBR_D f3 normalize3(f3 a) { float l = sqrtf(len2(a)); return div3(a, l); }
WPBR_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

WPBR_HD float smin(float a, float b) { return (b < a) ? b : a; }
WPBR_HD float smax(float a, float b) { return (a < b) ? b : a; }

struct bbox { float mn[3], mx[3]; };
WPBR_HD bbox box_empty() {
    bbox b;
    b.mn[0] = b.mn[1] = b.mn[2] = FLT_MAX_;
    b.mx[0] = b.mx[1] = b.mx[2] = FLT_LOWEST_;
    return b;
}
WPBR_HD bbox box_union(const bbox& a, const bbox& b) {
    bbox r;
#pragma unroll
    for (int i = 0; i < 3; i++) { r.mn[i] = smin(a.mn[i], b.mn[i]); r.mx[i] = smax(a.mx[i], b.mx[i]); }
    return r;
}
WPBR_HD bbox box_union_pt(const bbox& a, f3 p) {
    bbox r;
    r.mn[0] = smin(a.mn[0], p.x); r.mx[0] = smax(a.mx[0], p.x);
    r.mn[1] = smin(a.mn[1], p.y); r.mx[1] = smax(a.mx[1], p.y);
    r.mn[2] = smin(a.mn[2], p.z); r.mx[2] = smax(a.mx[2], p.z);
    return r;
}
WPBR_HD float box_area(const bbox& b) {
    float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    return 2 * ((dx * dy + dx * dz) + dy * dz);
}
WPBR_HD int box_max_extent(const bbox& b) {
    float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    if (dx > dy && dx > dz) return 0;
    if (dy > dz) return 1;
    return 2;
}
WPBR_D bbox load_box(const float* p) {
    bbox b;
    b.mn[0] = p[0]; b.mn[1] = p[1]; b.mn[2] = p[2];
    b.mx[0] = p[3]; b.mx[1] = p[4]; b.mx[2] = p[5];
    return b;
}
WPBR_D void store_box(float* p, const bbox& b) {
    p[0] = b.mn[0]; p[1] = b.mn[1]; p[2] = b.mn[2];
    p[3] = b.mx[0]; p[4] = b.mx[1]; p[5] = b.mx[2];
}

struct affine { float m[3][4]; };
WPBR_D affine load_affine(const float* p) {
    affine a;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) a.m[r][c] = p[4 * r + c];
    return a;
}
WPBR_HD f3 xf_point(const affine& a, f3 p) {
    return mk3(((a.m[0][0] * p.x + a.m[0][1] * p.y) + a.m[0][2] * p.z) + a.m[0][3],
               ((a.m[1][0] * p.x + a.m[1][1] * p.y) + a.m[1][2] * p.z) + a.m[1][3],
               ((a.m[2][0] * p.x + a.m[2][1] * p.y) + a.m[2][2] * p.z) + a.m[2][3]);
}
WPBR_HD f3 xf_vector(const affine& a, f3 v) {
    return mk3((a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z,
               (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
               (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z);
}
WPBR_HD f3 xf_normal(const affine& inv, f3 n) {
    return mk3((inv.m[0][0] * n.x + inv.m[1][0] * n.y) + inv.m[2][0] * n.z,
               (inv.m[0][1] * n.x + inv.m[1][1] * n.y) + inv.m[2][1] * n.z,
               (inv.m[0][2] * n.x + inv.m[1][2] * n.y) + inv.m[2][2] * n.z);
}

// Float <-> order-preserving uint (for atomic min/max reductions whose result
// is insensitive to the sign of zero).
WPBR_D uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
WPBR_D float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Wave64 helpers.
WPBR_D uint64_t ballot64(bool p) { return __ballot(p); }
WPBR_D uint32_t lane_id() { return __lane_id(); }
WPBR_D uint32_t popc64(uint64_t m) { return __popcll(m); }
WPBR_D uint64_t lanemask_lt() { uint32_t l = __lane_id(); return l ? (~0ull >> (64 - l)) : 0ull; }

// Wave-wide smin / smax of a float through DPP: row_shr 1, 2, 4, 8 leave each
// row's result in its lane 15, row_bcast 15 and 31 carry rows 0..2 into lane
// 63, which every lane reads.  Lanes whose DPP source lies outside the row
// take the identity (+-inf).  All 64 lanes must be active.
template <bool MAX>
WPBR_D float wave_reduce_dpp(float v) {
    const int id = __float_as_int(MAX ? -INF : INF);
    auto op = [](float a, float b) { return MAX ? smax(a, b) : smin(a, b); };
#define WPBR_DPP_STEP(ctrl, rows) \
    v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), ctrl, rows, 0xF, false)))
    WPBR_DPP_STEP(0x111, 0xF);
    WPBR_DPP_STEP(0x112, 0xF);
    WPBR_DPP_STEP(0x114, 0xF);
    WPBR_DPP_STEP(0x118, 0xF);
    WPBR_DPP_STEP(0x142, 0xA);
    WPBR_DPP_STEP(0x143, 0xC);
#undef WPBR_DPP_STEP
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
WPBR_D float wave_min_dpp(float v) { return wave_reduce_dpp<false>(v); }
WPBR_D float wave_max_dpp(float v) { return wave_reduce_dpp<true>(v); }

// Lane value of the lane `lane ^ 1` / `lane ^ 2` (within quads).
WPBR_D float quad_xor1(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false)); }
WPBR_D float quad_xor2(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false)); }
WPBR_D uint32_t quad_xor1(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false); }
WPBR_D uint32_t quad_xor2(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false); }

}  // namespace wpbr
