// Scene ingestion restated on the CPU (test infrastructure only): the world-space
// bake of STriangleMesh::create (CSTriangleMesh.cpp:11-39) followed by the
// triangle vertex gather CSTriangle performs through CTriangleMesh::iVertices
// (CSTriangleMesh.h:24-29, CSTriangle.cpp:116-142), and the texel decode of
// SLoadTexture::load (readImagePNG's /255.0f, CSTextureProcess.h:480-490, then
// gammaCorrectInv, :322-329 and :517-522).
// Unpinned by the reference (WoodenMath is absent): DTransform applied to a
// DNormal3f is taken as the transpose of the inverse (pbrt-v3's convention),
// and std::pow(float, 2.4f) as the correctly rounded power.
#include "oracle.h"

extern "C" {

// CSTriangleMesh.cpp:21-34, then corner k of triangle i = world position of
// idx[3 i + k].  xform: world 3x4 then inverse 3x4.
void orc_mesh_bake(const float* xform, uint32_t nv, const float* pos, const float* nrm, const float* tan,
                   uint32_t nt, const uint32_t* idx, float* pos_w, float* nrm_w, float* tan_w, float* tri) {
    orc::Affine world, inv;
    std::memcpy(&world, xform, 48);
    std::memcpy(&inv, xform + 12, 48);
    std::vector<orc::V3> P(nv);
    for (uint32_t i = 0; i < nv; i++) {
        P[i] = orc::xf_point(world, orc::v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
        std::memcpy(pos_w + 3 * (size_t)i, &P[i], 12);
        if (nrm) {
            const orc::V3 n = orc::xf_normal(inv, orc::v3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
            std::memcpy(nrm_w + 3 * (size_t)i, &n, 12);
        }
        if (tan) {
            const orc::V3 t = orc::xf_vector(world, orc::v3(tan[3 * i], tan[3 * i + 1], tan[3 * i + 2]));
            std::memcpy(tan_w + 3 * (size_t)i, &t, 12);
        }
    }
    for (uint32_t i = 0; i < nt; i++)
        for (int k = 0; k < 3; k++) std::memcpy(tri + 9 * (size_t)i + 3 * k, &P[idx[3 * (size_t)i + k]], 12);
}

// gammaCorrectInv(v / 255.0f) for each byte.
void orc_srgb8_to_linear(size_t n, const uint8_t* in, float* out) {
    for (size_t i = 0; i < n; i++) {
        const float v = (float)in[i] / 255.0f;
        out[i] = v <= 0.04045f ? v * 1.f / 12.92f : orc::cr_pow((v + 0.055f) * 1.f / 1.055f, 2.4f);
    }
}

}  // extern "C"
