// Texture filtering restated on the CPU (test infrastructure only): the MIP
// pyramid of SLoadTexture::generateMIPS (CSTextureProcess.h:301-345), the
// Gaussian table of JobFilterTableGaussing (CSTextureProcess.cpp:16-32,
// size 64 at MEngine.cpp:92-94) and STexture2DSamplerAnisotropic
// (CSTextureProcess.h:134-251) with maxAnisotropy 32 (CMaterialMetal.cpp:28).
// Unpinned by the reference (WoodenMath is absent): wml::lerp is taken as
// (1 - t) a + t b and a vector divided by a scalar as a divide per component.
#include "oracle.h"

namespace orc {

TexRGB operator+(const TexRGB& a, const TexRGB& b) { return {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}}; }
TexRGB operator*(const TexRGB& a, float s) { return {{a.v[0] * s, a.v[1] * s, a.v[2] * s}}; }

// generateMIPS (CSTextureProcess.h:349-371): (uint16_t)(log2(max side) - 1)
// levels, each max(1, half) of the previous one per axis.
Texture generate_mips(uint32_t w0, uint32_t h0, const float* rgb) {
    uint16_t nLevels = (uint16_t)(std::log2((double)std::max(w0, h0)) - 1);
    Texture tex;
    tex.w.resize(nLevels);
    tex.h.resize(nLevels);
    tex.mips.resize(nLevels);
    tex.w[0] = w0;
    tex.h[0] = h0;
    tex.mips[0].resize((size_t)w0 * h0);
    for (uint32_t y = 0; y < h0; y++)
        for (uint32_t x = 0; x < w0; x++) {
            const float* p = rgb + 3 * ((size_t)x + (size_t)y * w0);
            tex.texel(x, y, 0) = {{p[0], p[1], p[2]}};
        }
    for (uint16_t i = 1; i < nLevels; i++) {
        const uint32_t rx = std::max(1u, tex.w[i - 1] / 2), ry = std::max(1u, tex.h[i - 1] / 2);
        const uint32_t px = tex.w[i - 1], py = tex.h[i - 1];
        tex.w[i] = rx;
        tex.h[i] = ry;
        tex.mips[i].resize((size_t)rx * ry);
        for (uint32_t y = 0; y < ry - 1; y++) {
            for (uint32_t x = 0; x < rx - 1; x++)
                tex.texel(x, y, i) = (tex.texel(2 * x, 2 * y, i - 1) + tex.texel(2 * x + 1, 2 * y, i - 1) +
                                      tex.texel(2 * x, 2 * y + 1, i - 1) + tex.texel(2 * x + 1, 2 * y + 1, i - 1)) *
                                     0.25f;
            tex.texel(rx - 1, y, i) = (tex.texel(px - 1, 2 * y, i - 1) + tex.texel(px - 1, 2 * y + 1, i - 1)) * 0.5f;
        }
        for (uint32_t x = 0; x < rx - 1; x++)
            tex.texel(x, ry - 1, i) = (tex.texel(2 * x, py - 1, i - 1) + tex.texel(2 * x + 1, py - 1, i - 1)) * 0.5f;
        tex.texel(rx - 1, ry - 1, i) = tex.texel(px - 1, py - 1, i - 1);
    }
    return tex;
}

std::vector<float> texture_filter_table() {
    const uint32_t size = 64;
    std::vector<float> w(size);
    for (int i = 0; i < (int)size; ++i) {
        float alpha = 2;
        float r2 = float(i) / float(size - 1);
        w[i] = cr_exp(-alpha * r2) - cr_exp(-alpha);
    }
    return w;
}

// STexture2DSampler::triangle at `mip`, indices clamped below (as written) and
// above (reads past a level are undefined in the reference).
namespace {

TexRGB triangle(const float* st, const Texture& tex, uint32_t mip) {
    uint32_t nLevels = (uint32_t)tex.mips.size();
    mip = std::min(mip, nLevels - 1);
    const int rw = (int)tex.w[mip], rh = (int)tex.h[mip];
    float s = st[0] * (float)rw - 0.5f;
    float t = st[1] * (float)rh - 0.5f;
    int s0 = (int)std::floor(s), t0 = (int)std::floor(t);
    s0 = std::max(s0, 0);
    t0 = std::max(t0, 0);
    float ds = s - s0, dt = t - t0;
    int s0n = std::min(s0 + 1, rw - 1);
    int t0n = std::min(t0 + 1, rh - 1);
    s0 = std::min(s0, rw - 1);
    t0 = std::min(t0, rh - 1);
    return tex.texel(s0, t0, mip) * (1 - ds) * (1 - dt) + tex.texel(s0, t0n, mip) * (1 - ds) * dt +
           tex.texel(s0n, t0, mip) * ds * (1 - dt) + tex.texel(s0n, t0n, mip) * ds * dt;
}

}  // namespace

// st and both derivatives scaled by the level's resolution per axis
// (DPoint2f times DPoint2i, CSTextureProcess.h:192-194).
TexRGB texture_ewa(const float* stp, const float* dx, const float* dy, const Texture& tex,
                   const std::vector<float>& filter, uint32_t mip) {
    const float r[2] = {(float)tex.w[mip], (float)tex.h[mip]};
    float st[2] = {stp[0] * r[0] - 0.5f, stp[1] * r[1] - 0.5f};
    float dst0[2] = {dx[0] * r[0], dx[1] * r[1]};
    float dst1[2] = {dy[0] * r[0], dy[1] * r[1]};
    float abc[3];
    abc[0] = dst0[1] * dst0[1] + dst1[1] * dst1[1] + 1;
    abc[1] = -2 * (dst0[0] * dst0[1] + dst1[0] * dst1[1]);
    abc[2] = dst0[0] * dst0[0] + dst1[0] * dst1[0] + 1;
    const float& A = abc[0];
    const float& B = abc[1];
    const float& C = abc[2];
    float invF = 1 / (A * C - B * B * 0.25f);
    for (float& v : abc) v *= invF;
    float det = -B * B + 4 * A * C;
    float invDet = 1 / det;
    float uvSqrt[2] = {C * det, A * det};
    for (float& v : uvSqrt) v = std::sqrt(v) * (2 * invDet);
    float st0[2], st1[2];
    for (int k = 0; k < 2; k++) {
        st0[k] = std::min(std::max(st[k] - uvSqrt[k], 0.0f), r[k] - 1.0f);
        st1[k] = std::min(std::max(st[k] + uvSqrt[k], 0.0f), r[k] - 1.0f);
    }
    int s0 = (int)std::ceil(st0[0]), t0 = (int)std::ceil(st0[1]);
    int s1 = (int)std::floor(st1[0]), t1 = (int)std::floor(st1[1]);
    TexRGB sum = {{0.f, 0.f, 0.f}};
    float sumWts = 0;
    const int size = (int)filter.size();
    for (int it = t0; it <= t1; ++it) {
        float tt = it - st[1];
        for (int is = s0; is <= s1; ++is) {
            float ss = is - st[0];
            float r2 = A * ss * ss + B * ss * tt + C * tt * tt;
            if (r2 < 1) {
                int index = std::min((int)(r2 * size), size - 1);
                float weight = filter[index];
                sum = sum + tex.texel(is, it, mip) * weight;
                sumWts += weight;
            }
        }
    }
    return {{sum.v[0] / sumWts, sum.v[1] / sumWts, sum.v[2] / sumWts}};
}

TexRGB texture_eval(const float* q, const Texture& tex, const std::vector<float>& filter) {
    float dst0[2] = {q[2], q[3]}, dst1[2] = {q[4], q[5]};
    const float maxAnisotropy = 32;
    // the reference's swap test compares dstdx with itself: never taken
    float majorLength = std::sqrt(dst0[0] * dst0[0] + dst0[1] * dst0[1]);
    float minorLength = std::sqrt(dst1[0] * dst1[0] + dst1[1] * dst1[1]);
    // The clamp lengthens the minor axis for the level choice only: ewa() reads
    // the caller's dstdx / dstdy again (CSTextureProcess.h:183-185), so the
    // filter ellipse keeps the unscaled derivatives.
    if (minorLength * maxAnisotropy < majorLength && minorLength > 0) {
        float scale = majorLength / (minorLength * maxAnisotropy);
        minorLength *= scale;
    }
    if (minorLength == 0) return triangle(q, tex, 0);
    uint32_t nLevels = (uint32_t)tex.mips.size();
    float lod = std::max(0.0f, (float)(nLevels - 1) + cr_log2(std::max(minorLength, 0.0f)));
    int ilod = (int)std::floor(lod);
    if ((uint32_t)ilod >= nLevels - 1) return texture_ewa(q, dst0, dst1, tex, filter, nLevels - 1);
    float d = lod - ilod;
    return texture_ewa(q, dst0, dst1, tex, filter, ilod) * (1 - d) +
           texture_ewa(q, dst0, dst1, tex, filter, ilod + 1) * d;
}

}  // namespace orc

extern "C" {

// Level sizes (w, h per level) to `sizes` and, when `out` is set, the levels'
// texels one after another; returns the level count.
uint32_t orc_texture_levels(uint32_t w, uint32_t h, const float* rgb, uint32_t* sizes, float* out) {
    orc::Texture t = orc::generate_mips(w, h, rgb);
    for (size_t l = 0; l < t.mips.size() && sizes; l++) {
        sizes[2 * l] = t.w[l];
        sizes[2 * l + 1] = t.h[l];
    }
    if (out) {
        size_t o = 0;
        for (auto& m : t.mips)
            for (auto& v : m)
                for (int c = 0; c < 3; c++) out[o++] = v.v[c];
    }
    return (uint32_t)t.mips.size();
}

void orc_texture_filter(float* w64) {
    std::vector<float> w = orc::texture_filter_table();
    for (size_t i = 0; i < w.size(); i++) w64[i] = w[i];
}

void orc_texture_ewa(uint32_t w, uint32_t h, const float* rgb, uint32_t level, uint32_t n, const float* q,
                     float* out) {
    orc::Texture t = orc::generate_mips(w, h, rgb);
    std::vector<float> f = orc::texture_filter_table();
    for (uint32_t i = 0; i < n; i++) {
        const float* qi = q + 6 * (size_t)i;
        orc::TexRGB v = orc::texture_ewa(qi, qi + 2, qi + 4, t, f, level);
        for (int c = 0; c < 3; c++) out[3 * (size_t)i + c] = v.v[c];
    }
}

void orc_texture_lookup(uint32_t w, uint32_t h, const float* rgb, uint32_t n, const float* q, float* out) {
    orc::Texture t = orc::generate_mips(w, h, rgb);
    std::vector<float> f = orc::texture_filter_table();
    for (uint32_t i = 0; i < n; i++) {
        orc::TexRGB v = orc::texture_eval(q + 6 * (size_t)i, t, f);
        for (int c = 0; c < 3; c++) out[3 * (size_t)i + c] = v.v[c];
    }
}

}  // extern "C"
