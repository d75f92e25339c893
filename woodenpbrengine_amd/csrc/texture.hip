// The MIP pyramid's levels (wpbr_texture_create) and texture lookups over a
// batch of queries (wpbr_texture_lookup: one thread per query, the EWA filter
// of texture_device.hpp).
#include <hip/hip_runtime.h>
#include "texture_device.hpp"
#include "wpbr_internal.hpp"

namespace wpbr {

__global__ __launch_bounds__(256) void k_tex_lookup(TexView T, uint32_t n, const float* __restrict__ q,
                                                    float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* a = q + 6 * (size_t)i;
    const float3 v = tex_lookup(T, a[0], a[1], a[2], a[3], a[4], a[5]);
    out[3 * (size_t)i] = v.x;
    out[3 * (size_t)i + 1] = v.y;
    out[3 * (size_t)i + 2] = v.z;
}

// MIP level l from level l - 1 (generateMIPS, CSTextureProcess.h:359-371): the
// mean of a 2x2 block inside; the last column averages the previous level's
// last column over two rows, the last row its last row over two columns; the
// last texel copies the previous level's last texel.  One thread per texel.
WPBR_D float4 mip_mean4(float4 a, float4 b, float4 c, float4 d) {
    return make_float4(((a.x + b.x) + c.x + d.x) * 0.25f, ((a.y + b.y) + c.y + d.y) * 0.25f,
                       ((a.z + b.z) + c.z + d.z) * 0.25f, 0.f);
}
WPBR_D float4 mip_mean2(float4 a, float4 b) {
    return make_float4((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f, 0.f);
}
__global__ __launch_bounds__(256) void k_mip_level(TexView T, uint32_t l) {
    const uint32_t w = T.lw[l], h = T.lh[l], pw = T.lw[l - 1], ph = T.lh[l - 1];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    const float4* src = T.texels + T.off[l - 1];
    auto at = [&](uint32_t xx, uint32_t yy) { return src[(size_t)yy * pw + xx]; };
    float4 v;
    if (x + 1 < w && y + 1 < h) v = mip_mean4(at(2 * x, 2 * y), at(2 * x + 1, 2 * y), at(2 * x, 2 * y + 1),
                                              at(2 * x + 1, 2 * y + 1));
    else if (y + 1 < h) v = mip_mean2(at(pw - 1, 2 * y), at(pw - 1, 2 * y + 1));
    else if (x + 1 < w) v = mip_mean2(at(2 * x, ph - 1), at(2 * x + 1, ph - 1));
    else v = at(pw - 1, ph - 1);
    const_cast<float4*>(T.texels)[T.off[l] + (size_t)y * w + x] = v;
}

wpbr_status tex_mips_launch(const TexView& T, hipStream_t st) {
    for (uint32_t l = 1; l < T.n_levels; l++) {
        const uint32_t n = T.lw[l] * T.lh[l];
        hipLaunchKernelGGL(k_mip_level, dim3((n + 255) / 256), dim3(256), 0, st, T, l);
    }
    return launch_status();
}

wpbr_status tex_lookup_launch(const TexView& T, uint32_t n, const float* q, float* out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_tex_lookup, dim3((n + 255) / 256), dim3(256), 0, st, T, n, q, out);
    return launch_status();
}

}  // namespace wpbr
