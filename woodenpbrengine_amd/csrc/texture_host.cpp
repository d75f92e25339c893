// Textures: the MIP pyramid of SLoadTexture::generateMIPS (CSTextureProcess.h:
// 349-371, built on the device), the Gaussian filter table of JobFilterTableGaussing
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
wpbr_status tex_mips_launch(const TexView& T, hipStream_t st);

bool tex_size_ok(uint32_t w, uint32_t h) {
    return w >= 1 && h >= 1 && w <= 16384 && h <= 16384 && (w >= 4 || h >= 4);
}

// The pyramid's shape (generateMIPS, CSTextureProcess.h:349-371): log2 of the
// longer side minus 1 levels (truncated to an integer), each level max(1, half)
// of the previous one per axis; levels stored one after another.
size_t tex_layout(uint32_t w, uint32_t h, TexView& v) {
    v.n_levels = (uint32_t)(uint16_t)(std::log2((double)std::max(w, h)) - 1);
    size_t total = 0;
    for (uint32_t l = 0; l < v.n_levels; l++) {
        v.lw[l] = l ? std::max(1u, v.lw[l - 1] / 2) : w;
        v.lh[l] = l ? std::max(1u, v.lh[l - 1] / 2) : h;
        v.off[l] = (uint32_t)total;
        total += (size_t)v.lw[l] * v.lh[l];
    }
    return total;
}

// weights[i] = exp(-2 r2) - exp(-2), r2 = i / (size - 1), exp correctly rounded.
void tex_filter_table(float* w) {
    for (int i = 0; i < TEX_FILTER_SIZE; ++i) {
        const float alpha = 2;
        const float r2 = float(i) / float(TEX_FILTER_SIZE - 1);
        w[i] = (float)std::exp((double)(-alpha * r2)) - (float)std::exp((double)(-alpha));
    }
}

// Level 0 from the host's texels, the coarser levels built on the device.
wpbr_status tex_upload(wpbr_ctx* ctx, uint32_t w, uint32_t h, const float* rgb, DevTexture& t) {
    TexView& v = t.view;
    std::memset(&v, 0, sizeof(v));
    const size_t total = tex_layout(w, h, v);
    std::vector<float4> base((size_t)w * h);
    for (size_t i = 0; i < base.size(); i++)
        base[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.f);
    float filt[TEX_FILTER_SIZE];
    tex_filter_table(filt);
    wpbr_status s = hip_check(ctx, hipMalloc((void**)&t.d_texels, total * sizeof(float4)), "hipMalloc texture");
    if (!s) s = hip_check(ctx, hipMalloc((void**)&t.d_filter, sizeof(filt)), "hipMalloc filter table");
    if (!s) s = hip_check(ctx, hipMemcpyAsync(t.d_texels, base.data(), base.size() * sizeof(float4),
                                              hipMemcpyHostToDevice, ctx->stream), "texture upload");
    if (!s) s = hip_check(ctx, hipMemcpyAsync(t.d_filter, filt, sizeof(filt), hipMemcpyHostToDevice, ctx->stream),
                          "filter upload");
    v.texels = t.d_texels;
    v.filter = t.d_filter;
    if (!s) {
        (void)hipGetLastError();
        s = tex_mips_launch(v, ctx->stream);
        if (s) s = set_err(ctx, s, std::string("MIP build launch failed: ") + hipGetErrorString(last_launch_error));
    }
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "MIP build");
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

namespace wpbr {
const TexView* texture_view(const wpbr_texture* t) { return &t->t.view; }
}  // namespace wpbr

extern "C" {

wpbr_status wpbr_texture_create(wpbr_ctx* ctx, uint32_t width, uint32_t height, const float* rgb,
                                wpbr_texture** out) {
    if (!ctx || !rgb || !out) return WPBR_ERR_INVALID_ARG;
    if (!tex_size_ok(width, height))
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "texture sides must be in [1, 16384], the longer at least 4");
    (void)hipSetDevice(ctx->device);
    wpbr_texture* tx = new wpbr_texture();
    tx->ctx = ctx;
    wpbr_status s = tex_upload(ctx, width, height, rgb, tx->t);
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
    const size_t n = (size_t)v.lw[level] * v.lh[level];
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

wpbr_status wpbr_texture_size(const wpbr_texture* tex, uint32_t level, uint32_t* width, uint32_t* height) {
    if (!tex || !width || !height || level >= tex->t.view.n_levels) return WPBR_ERR_INVALID_ARG;
    *width = tex->t.view.lw[level];
    *height = tex->t.view.lh[level];
    return WPBR_OK;
}

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
