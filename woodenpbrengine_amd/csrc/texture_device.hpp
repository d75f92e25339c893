// Anisotropic texture filtering on the device: the MIP pyramid's texels and the
// EWA lookup of STexture2DSamplerAnisotropic (CSTextureProcess.h:134-251).
#pragma once
#include "wpbr_device.hpp"

namespace wpbr {

constexpr int TEX_MAX_LEVELS = 16;
constexpr int TEX_FILTER_SIZE = 64;         // CFilterTableGaussing::size (MEngine.cpp:92-94)
constexpr float TEX_MAX_ANISOTROPY = 32.0f; // CMaterialMetal.cpp:28

// An RGB texture and its MIP levels: level l is lw[l] x lh[l] texels (level 0
// the image, each next level max(1, half) per axis), row-major float4 texels (w
// unused) from texels + off[l].
struct TexView {
    const float4* texels;
    const float* filter;  // TEX_FILTER_SIZE Gaussian weights
    uint32_t n_levels;
    uint32_t off[TEX_MAX_LEVELS];
    uint32_t lw[TEX_MAX_LEVELS], lh[TEX_MAX_LEVELS];
};

WPBR_D float3 tex_texel(const TexView& T, int x, int y, uint32_t l) {
    const float4 v = T.texels[T.off[l] + (size_t)y * T.lw[l] + x];
    return make_float3(v.x, v.y, v.z);
}

WPBR_D float3 f3s(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
WPBR_D float3 f3add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }

// STexture2DSampler::triangle (CSTextureProcess.h:57-81) at level 0: texel
// indices clamped below as the reference does and also above (st beyond 1 would
// read past the level).
WPBR_D float3 tex_bilinear(const TexView& T, float s_, float t_) {
    const int rw = (int)T.lw[0], rh = (int)T.lh[0];
    const float s = s_ * (float)rw - 0.5f, t = t_ * (float)rh - 0.5f;
    int s0 = (int)floorf(s), t0 = (int)floorf(t);
    s0 = s0 < 0 ? 0 : s0;
    t0 = t0 < 0 ? 0 : t0;
    const float ds = s - (float)s0, dt = t - (float)t0;
    int s0n = s0 + 1 < rw - 1 ? s0 + 1 : rw - 1;
    int t0n = t0 + 1 < rh - 1 ? t0 + 1 : rh - 1;
    s0 = s0 < rw - 1 ? s0 : rw - 1;
    t0 = t0 < rh - 1 ? t0 : rh - 1;
    float3 v = f3s(f3s(tex_texel(T, s0, t0, 0), 1 - ds), 1 - dt);
    v = f3add(v, f3s(f3s(tex_texel(T, s0, t0n, 0), 1 - ds), dt));
    v = f3add(v, f3s(f3s(tex_texel(T, s0n, t0, 0), ds), 1 - dt));
    v = f3add(v, f3s(f3s(tex_texel(T, s0n, t0n, 0), ds), dt));
    return v;
}

// STexture2DSamplerAnisotropic::ewa (CSTextureProcess.h:182-251) on level l.
WPBR_D float3 tex_ewa(const TexView& T, float s_, float t_, float dsx, float dtx, float dsy, float dty,
                      uint32_t l) {
    // st and both derivatives scaled by the level's size per axis
    const float rw = (float)T.lw[l], rh = (float)T.lh[l];
    const float s = s_ * rw - 0.5f, t = t_ * rh - 0.5f;
    dsx *= rw; dtx *= rh; dsy *= rw; dty *= rh;
    float A = (dtx * dtx + dty * dty) + 1;
    float B = -2.0f * (dsx * dtx + dsy * dty);
    float C = (dsx * dsx + dsy * dsy) + 1;
    const float invF = 1 / (A * C - (B * B) * 0.25f);
    A *= invF; B *= invF; C *= invF;
    const float det = (-B) * B + (4.0f * A) * C;
    const float invDet = 1 / det;
    const float us = sqrtf(C * det) * (2.0f * invDet), vs = sqrtf(A * det) * (2.0f * invDet);
    const float hiw = rw - 1.0f, hih = rh - 1.0f;
    const float st0s = fminf(fmaxf(s - us, 0.0f), hiw), st0t = fminf(fmaxf(t - vs, 0.0f), hih);
    const float st1s = fminf(fmaxf(s + us, 0.0f), hiw), st1t = fminf(fmaxf(t + vs, 0.0f), hih);
    const int s0 = (int)ceilf(st0s), t0 = (int)ceilf(st0t);
    const int s1 = (int)floorf(st1s), t1 = (int)floorf(st1t);
    float3 sum = make_float3(0.f, 0.f, 0.f);
    float sumW = 0.0f;
    for (int it = t0; it <= t1; ++it) {
        const float tt = (float)it - t;
        for (int is = s0; is <= s1; ++is) {
            const float ss = (float)is - s;
            const float r2 = ((A * ss) * ss + (B * ss) * tt) + (C * tt) * tt;
            if (r2 < 1) {
                int idx = (int)(r2 * (float)TEX_FILTER_SIZE);
                idx = idx < TEX_FILTER_SIZE - 1 ? idx : TEX_FILTER_SIZE - 1;
                const float w = T.filter[idx];
                sum = f3add(sum, f3s(tex_texel(T, is, it, l), w));
                sumW += w;
            }
        }
    }
    return make_float3(sum.x / sumW, sum.y / sumW, sum.z / sumW);
}

// STexture2DSamplerAnisotropic::eval (CSTextureProcess.h:136-180).  The
// reference's swap test compares dstdx with itself, so the major axis is always
// dstdx; the lerp between levels is (1 - d) a + d b.  The anisotropy clamp
// lengthens the minor axis for the level choice only: ewa() reads the caller's
// derivatives again (:183-185), so the filter ellipse keeps them unscaled.
WPBR_D float3 tex_lookup(const TexView& T, float s, float t, float dsx, float dtx, float dsy, float dty) {
    const float major = sqrtf(dsx * dsx + dtx * dtx);
    float minor = sqrtf(dsy * dsy + dty * dty);
    if (minor * TEX_MAX_ANISOTROPY < major && minor > 0) {
        const float scale = major / (minor * TEX_MAX_ANISOTROPY);
        minor *= scale;
    }
    if (minor == 0) return tex_bilinear(T, s, t);
    const float lod = fmaxf(0.0f, (float)(T.n_levels - 1) + cr_log2(fmaxf(minor, 0.0f)));
    const int ilod = (int)floorf(lod);
    if ((uint32_t)ilod >= T.n_levels - 1) return tex_ewa(T, s, t, dsx, dtx, dsy, dty, T.n_levels - 1);
    const float d = lod - (float)ilod;
    const float3 a = tex_ewa(T, s, t, dsx, dtx, dsy, dty, (uint32_t)ilod);
    const float3 b = tex_ewa(T, s, t, dsx, dtx, dsy, dty, (uint32_t)ilod + 1);
    return f3add(f3s(a, 1 - d), f3s(b, d));
}

}  // namespace wpbr
