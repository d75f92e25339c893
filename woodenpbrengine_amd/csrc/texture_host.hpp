// Host side of the textures: MIP pyramid, filter table, device upload.
#pragma once
#include <vector>
#include "texture_device.hpp"
#include "wpbr_ctx.hpp"
#include "wpbr_internal.hpp"

namespace wpbr {

struct DevTexture {
    TexView view;
    float4* d_texels = nullptr;
    float* d_filter = nullptr;
};

bool tex_size_ok(uint32_t w, uint32_t h);
size_t tex_layout(uint32_t w, uint32_t h, TexView& v);
void tex_filter_table(float* w);
wpbr_status tex_upload(wpbr_ctx* ctx, uint32_t w, uint32_t h, const float* rgb, DevTexture& t);
void tex_free(DevTexture& t);

}  // namespace wpbr
