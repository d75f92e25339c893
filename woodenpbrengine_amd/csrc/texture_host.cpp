// Textures: the MIP pyramid of SLoadTexture::generateMIPS (CSTextureProcess.h:
// 301-345), the Gaussian filter table of JobFilterTableGaussing
// (CSTextureProcess.cpp:16-32) and the C ABI of the anisotropic lookup.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "texture_host.hpp"

using namespace wpbr;

namespace wpbr {

wpbr_status tex_lookup_launch(const TexView& T, uint32_t n, const float* q, float* out, hipStream_t st);

bool tex_res_ok(uint32_t res) { return res >= 4 && res <= 16384 && (res & (res - 1)) == 0; }

// Level 0 = the texels; level i halves the previous one: 2x2 means inside, the
// last column and row from the previous level's last column / row only, the
// corner copied.  nLevels = log2(res) - 1, so the coarsest level is 4 x 4.
void tex_build_mips(uint32_t res, const float* rgb, std::vector<float4>& texels, uint32_t* off,
                    uint32_t& n_levels) {
    n_levels = (uint32_t)(uint16_t)(std::log2((double)res) - 1);
    size_t total = 0;
    for (uint32_t l = 0; l < n_levels; l++) {
        off[l] = (uint32_t)total;
        total += (size_t)(res >> l) * (res >> l);
    }
    texels.assign(total, make_float4(0.f, 0.f, 0.f, 0.f));
    for (size_t i = 0; i < (size_t)res * res; i++)
        texels[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.f);
    auto at = [&](uint32_t l, uint32_t x, uint32_t y) -> float4& {
        return texels[off[l] + (size_t)y * (res >> l) + x];
    };
    auto mean4 = [](float4 a, float4 b, float4 c, float4 d) {
        return make_float4(((a.x + b.x) + c.x + d.x) * 0.25f, ((a.y + b.y) + c.y + d.y) * 0.25f,
                           ((a.z + b.z) + c.z + d.z) * 0.25f, 0.f);
    };
    auto mean2 = [](float4 a, float4 b) {
        return make_float4((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f, 0.f);
    };
    for (uint32_t l = 1; l < n_levels; l++) {
        const uint32_t r = res >> l, rp = res >> (l - 1);
        for (uint32_t y = 0; y + 1 < r; y++) {
            for (uint32_t x = 0; x + 1 < r; x++)
                at(l, x, y) = mean4(at(l - 1, 2 * x, 2 * y), at(l - 1, 2 * x + 1, 2 * y), at(l - 1, 2 * x, 2 * y + 1),
                                    at(l - 1, 2 * x + 1, 2 * y + 1));
            at(l, r - 1, y) = mean2(at(l - 1, rp - 1, 2 * y), at(l - 1, rp - 1, 2 * y + 1));
        }
        for (uint32_t x = 0; x + 1 < r; x++)
            at(l, x, r - 1) = mean2(at(l - 1, 2 * x, rp - 1), at(l - 1, 2 * x + 1, rp - 1));
        at(l, r - 1, r - 1) = at(l - 1, rp - 1, rp - 1);
    }
}

// weights[i] = exp(-2 r2) - exp(-2), r2 = i / (size - 1), exp correctly rounded.
void tex_filter_table(float* w) {
    for (int i = 0; i < TEX_FILTER_SIZE; ++i) {
        const float alpha = 2;
        const float r2 = float(i) / float(TEX_FILTER_SIZE - 1);
        w[i] = (float)std::exp((double)(-alpha * r2)) - (float)std::exp((double)(-alpha));
    }
}

wpbr_status tex_upload(wpbr_ctx* ctx, uint32_t res, const float* rgb, DevTexture& t) {
    std::vector<float4> texels;
    TexView& v = t.view;
    std::memset(&v, 0, sizeof(v));
    v.res = res;
    tex_build_mips(res, rgb, texels, v.off, v.n_levels);
    float filt[TEX_FILTER_SIZE];
    tex_filter_table(filt);
    wpbr_status s = hip_check(ctx, hipMalloc((void**)&t.d_texels, texels.size() * sizeof(float4)), "hipMalloc texture");
    if (!s) s = hip_check(ctx, hipMalloc((void**)&t.d_filter, sizeof(filt)), "hipMalloc filter table");
    if (!s) s = hip_check(ctx, hipMemcpy(t.d_texels, texels.data(), texels.size() * sizeof(float4), hipMemcpyHostToDevice),
                          "texture upload");
    if (!s) s = hip_check(ctx, hipMemcpy(t.d_filter, filt, sizeof(filt), hipMemcpyHostToDevice), "filter upload");
    v.texels = t.d_texels;
    v.filter = t.d_filter;
    return s;
}

void tex_free(DevTexture& t) {
    if (t.d_texels) (void)hipFree(t.d_texels);
    if (t.d_filter) (void)hipFree(t.d_filter);
    t.d_texels = nullptr;
    t.d_filter = nullptr;
}

}  // namespace wpbr

struct wpbr_texture {
    wpbr_ctx* ctx = nullptr;
    DevTexture t{};
};

extern "C" {

wpbr_status wpbr_texture_create(wpbr_ctx* ctx, uint32_t res, const float* rgb, wpbr_texture** out) {
    if (!ctx || !rgb || !out) return WPBR_ERR_INVALID_ARG;
    if (!tex_res_ok(res)) return set_err(ctx, WPBR_ERR_INVALID_ARG, "texture size must be a power of two in [4, 16384]");
    (void)hipSetDevice(ctx->device);
    wpbr_texture* tx = new wpbr_texture();
    tx->ctx = ctx;
    wpbr_status s = tex_upload(ctx, res, rgb, tx->t);
    if (s) {
        tex_free(tx->t);
        delete tx;
        return s;
    }
    *out = tx;
    return WPBR_OK;
}

void wpbr_texture_destroy(wpbr_texture* tex) {
    if (!tex) return;
    (void)hipSetDevice(tex->ctx->device);
    tex_free(tex->t);
    delete tex;
}

wpbr_status wpbr_texture_level(wpbr_ctx* ctx, const wpbr_texture* tex, uint32_t level, float* rgb) {
    if (!ctx || !tex || !rgb) return WPBR_ERR_INVALID_ARG;
    const TexView& v = tex->t.view;
    if (level >= v.n_levels) return set_err(ctx, WPBR_ERR_INVALID_ARG, "no such MIP level");
    const size_t n = (size_t)(v.res >> level) * (v.res >> level);
    std::vector<float4> buf(n);
    (void)hipSetDevice(ctx->device);
    wpbr_status s = hip_check(ctx, hipMemcpy(buf.data(), v.texels + v.off[level], n * sizeof(float4),
                                             hipMemcpyDeviceToHost), "texture download");
    if (s) return s;
    for (size_t i = 0; i < n; i++) {
        rgb[3 * i] = buf[i].x;
        rgb[3 * i + 1] = buf[i].y;
        rgb[3 * i + 2] = buf[i].z;
    }
    return WPBR_OK;
}

uint32_t wpbr_texture_levels(const wpbr_texture* tex) { return tex ? tex->t.view.n_levels : 0; }

wpbr_status wpbr_texture_lookup(wpbr_ctx* ctx, const wpbr_texture* tex, uint32_t n, const float* d_query,
                                float* d_rgb) {
    if (!ctx || !tex || (n && (!d_query || !d_rgb))) return WPBR_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipGetLastError();
    hipEventRecord(ctx->ev0, ctx->stream);
    wpbr_status s = tex_lookup_launch(tex->t.view, n, d_query, d_rgb, ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("texture lookup launch failed: ") + hipGetErrorString(last_launch_error));
    s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "texture lookup");
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

}  // extern "C"
