// Texture lookups over a batch of queries (wpbr_texture_lookup): one thread per
// query, the EWA filter of texture_device.hpp.
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

wpbr_status tex_lookup_launch(const TexView& T, uint32_t n, const float* q, float* out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_tex_lookup, dim3((n + 255) / 256), dim3(256), 0, st, T, n, q, out);
    return launch_status();
}

}  // namespace wpbr
