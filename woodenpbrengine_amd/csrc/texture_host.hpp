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

bool tex_res_ok(uint32_t res);
void tex_build_mips(uint32_t res, const float* rgb, std::vector<float4>& texels, uint32_t* off, uint32_t& n_levels);
void tex_filter_table(float* w);
wpbr_status tex_upload(wpbr_ctx* ctx, uint32_t res, const float* rgb, DevTexture& t);
void tex_free(DevTexture& t);

}  // namespace wpbr
