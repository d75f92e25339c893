// Scene ingestion on the device: the world-space mesh bake of
// STriangleMesh::create (CSTriangleMesh.cpp:11-39) fused with the gather into
// the triangle vertex array the geometry takes (what CSTriangle reads through
// CTriangleMesh::iVertices, CSTriangleMesh.h:24-29), and the 8-bit texel decode
// of SLoadTexture::load (readImagePNG's division by 255, CSTextureProcess.h:
// 462-496, then gammaCorrectInv, :322-329 and :517-522).
//
// All three kernels are HBM-bound streams: 12 B read + 12 B written per vertex
// attribute, 12 B of indices + 36 B of gathered positions read + 36 B written
// per triangle, 1 B read + 4 B written per texel channel.
#include <hip/hip_runtime.h>
#include "wpbr_device.hpp"
#include "wpbr_internal.hpp"

namespace wpbr {

struct BakeArgs {
    affine world, inv;
    uint32_t n_vertices, n_triangles;
    const float* pos;       // n_vertices x 3, object space
    const float* nrm;       // n_vertices x 3 or null
    const float* tan;       // n_vertices x 3 or null
    const uint32_t* idx;    // n_triangles x 3
    float* pos_w;           // outputs, each may be null
    float* nrm_w;
    float* tan_w;
    float* tri;             // n_triangles x 9
    uint32_t* bad;          // count of indices >= n_vertices
};

WPBR_D f3 ld3(const float* p, size_t i) { return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
WPBR_D void st3(float* p, size_t i, f3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// CSTriangleMesh.cpp:21-34: positions as points, normals by the transpose of
// the inverse, tangents as vectors (DTransform applied to DPoint3f / DNormal3f /
// DVector3f).
__global__ __launch_bounds__(256) void k_bake_vertices(BakeArgs a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_vertices) return;
    if (a.pos_w) st3(a.pos_w, i, xf_point(a.world, ld3(a.pos, i)));
    if (a.nrm && a.nrm_w) st3(a.nrm_w, i, xf_normal(a.inv, ld3(a.nrm, i)));
    if (a.tan && a.tan_w) st3(a.tan_w, i, xf_vector(a.world, ld3(a.tan, i)));
}

// Corner k of triangle i is the world position of vertex idx[3 i + k]; the
// transform is evaluated again at the corner (same arithmetic, same bits as the
// baked vertex), so this kernel does not wait for k_bake_vertices.  An index
// out of range is counted and its corner written as zeros.  A workgroup's 256
// triangles move through LDS: their indices are read and their 9 floats
// written as contiguous dwords (strides 3 and 9 in LDS are free of bank
// conflicts), one triangle per lane in between.
__global__ __launch_bounds__(256) void k_bake_triangles(BakeArgs a) {
    __shared__ uint32_t s_idx[256 * 3];
    __shared__ float s_out[256 * 9];
    const size_t base = (size_t)blockIdx.x * 256;
    const uint32_t nb = (uint32_t)min((size_t)256, (size_t)a.n_triangles - base);
    for (uint32_t k = threadIdx.x; k < 3 * nb; k += 256) s_idx[k] = a.idx[3 * base + k];
    __syncthreads();
    const uint32_t t = threadIdx.x;
    if (t < nb) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint32_t v = s_idx[3 * t + k];
            f3 p = mk3(0.0f, 0.0f, 0.0f);
            if (v < a.n_vertices)
                p = xf_point(a.world, ld3(a.pos, v));
            else
                atomicAdd(a.bad, 1u);
            s_out[9 * t + 3 * k] = p.x;
            s_out[9 * t + 3 * k + 1] = p.y;
            s_out[9 * t + 3 * k + 2] = p.z;
        }
    }
    __syncthreads();
    float* o = a.tri + 9 * base;
    for (uint32_t k = threadIdx.x; k < 9 * nb; k += 256) o[k] = s_out[k];
}

// gammaCorrectInv(i / 255.0f) (CSTextureProcess.h:517-522): value / 12.92 up to
// 0.04045, else ((value + 0.055) / 1.055)^2.4f, the power correctly rounded
// (evaluated in FP64 and rounded once).
WPBR_HD float srgb8_to_linear(uint32_t i) {
    const float v = (float)i / 255.0f;
    if (v <= 0.04045f) return v * 1.f / 12.92f;
    const float b = (v + 0.055f) * 1.f / 1.055f;
    return (float)pow((double)b, (double)2.4f);
}

// One channel per thread; each workgroup first builds the 256-entry table in
// LDS (one power per lane), then strides over the texels.
__global__ __launch_bounds__(256) void k_srgb8_linear(size_t n, const uint8_t* __restrict__ in,
                                                      float* __restrict__ out) {
    __shared__ float lut[256];
    lut[threadIdx.x] = srgb8_to_linear(threadIdx.x);
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = lut[in[i]];
}

wpbr_status bake_launch(const float* xform, uint32_t n_vertices, const float* pos, const float* nrm,
                        const float* tan, uint32_t n_triangles, const uint32_t* idx, float* pos_w,
                        float* nrm_w, float* tan_w, float* tri, uint32_t* bad, hipStream_t st) {
    BakeArgs a;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) {
            a.world.m[r][c] = xform[4 * r + c];
            a.inv.m[r][c] = xform[12 + 4 * r + c];
        }
    a.n_vertices = n_vertices;
    a.n_triangles = n_triangles;
    a.pos = pos;
    a.nrm = nrm;
    a.tan = tan;
    a.idx = idx;
    a.pos_w = pos_w;
    a.nrm_w = nrm_w;
    a.tan_w = tan_w;
    a.tri = tri;
    a.bad = bad;
    if (n_vertices && (pos_w || (nrm && nrm_w) || (tan && tan_w)))
        hipLaunchKernelGGL(k_bake_vertices, dim3((n_vertices + 255) / 256), dim3(256), 0, st, a);
    if (n_triangles && tri)
        hipLaunchKernelGGL(k_bake_triangles, dim3((n_triangles + 255) / 256), dim3(256), 0, st, a);
    return launch_status();
}

wpbr_status srgb8_launch(size_t n, const uint8_t* in, float* out, int num_cus, hipStream_t st) {
    if (!n) return WPBR_OK;
    const size_t want = (n + 255) / 256;
    const size_t cap = (size_t)num_cus * 8;
    const uint32_t grid = (uint32_t)(want < cap ? want : cap);
    hipLaunchKernelGGL(k_srgb8_linear, dim3(grid), dim3(256), 0, st, n, in, out);
    return launch_status();
}

}  // namespace wpbr
