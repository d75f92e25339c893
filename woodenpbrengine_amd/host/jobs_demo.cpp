// Runs the reference's job sequences through wpbr_jobs.hpp on one scene:
// MEngine::buildLBVH then MEngine::runCollisionSystem over a list of ray casts.
//
// usage: wpbr_jobs_demo IN OUT
// IN  (little-endian): u32 n_spheres, u32 n_triangles, u32 n_rays, u32 options,
//     f32 sphere_xform[24 n_spheres], f32 radius[n_spheres],
//     f32 tri_verts[9 n_triangles], f32 rays[6 n_rays] (origin xyz, dir xyz),
//     then if options & 1: u8 bSurfInteraction[n_rays],
//     if options & 2: f32 ray differentials[12 n_rays] (difXRay origin, dir,
//     difYRay origin, dir), if options & 4: the triangle attributes u8
//     flags[n_triangles], f32 uv[6 n_triangles], normal[9 n_triangles],
//     tangent[9 n_triangles]
// OUT: u32 n_nodes, u32 n_shapes, then per node u32 offset, u32 nShapes, u32 axis,
//     f32 bounds[6 n_nodes], u32 shapesEntities[n_shapes], then per ray
//     u32 hCollision, f32 time, f32 p[3], f32 wo[3], then per ray the 49 floats
//     of its interaction in the WPBR_SI_* order (the CSurfaceInteraction of a
//     bSurfInteraction hit, else p, wo, time and zeros)
#include <cstdio>
#include <vector>

#include "wpbr_jobs.hpp"

template <class T>
static bool read_vec(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    uint32_t hdr[4];
    std::vector<float> xf, rad, tris, rays, diff, uv, nrm, tan;
    std::vector<uint8_t> surf, flags;
    bool ok = fread(hdr, 4, 4, in) == 4 && read_vec(in, xf, 24 * (size_t)hdr[0]) &&
              read_vec(in, rad, hdr[0]) && read_vec(in, tris, 9 * (size_t)hdr[1]) &&
              read_vec(in, rays, 6 * (size_t)hdr[2]);
    const uint32_t opt = ok ? hdr[3] : 0;
    if (ok && (opt & 1)) ok = read_vec(in, surf, hdr[2]);
    if (ok && (opt & 2)) ok = read_vec(in, diff, 12 * (size_t)hdr[2]);
    if (ok && (opt & 4))
        ok = read_vec(in, flags, hdr[1]) && read_vec(in, uv, 6 * (size_t)hdr[1]) &&
             read_vec(in, nrm, 9 * (size_t)hdr[1]) && read_vec(in, tan, 9 * (size_t)hdr[1]);
    fclose(in);
    if (!ok) {
        fprintf(stderr, "short input\n");
        return 2;
    }
    try {
        wpbr_host::MEngine engine(0);
        engine.setShapes(xf, rad, tris);
        if (opt & 4) engine.setTriangleAttributes(flags, uv, nrm, tan);
        engine.buildLBVH();
        engine.rayCasts.resize(hdr[2]);
        for (uint32_t i = 0; i < hdr[2]; i++) {
            const float* r = &rays[6 * (size_t)i];
            engine.rayCasts[i] = {{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}, 0.0f}, (opt & 1) ? surf[i] != 0 : false};
        }
        if (opt & 2) {
            engine.rayDifferentials.resize(hdr[2]);
            for (uint32_t i = 0; i < hdr[2]; i++) {
                const float* d = &diff[12 * (size_t)i];
                wpbr_host::CRayDifferential& rd = engine.rayDifferentials[i];
                static_cast<wpbr_host::DRayf&>(rd) = engine.rayCasts[i].ray;
                rd.difXRay = {{d[0], d[1], d[2]}, {d[3], d[4], d[5]}, 0.0f};
                rd.difYRay = {{d[6], d[7], d[8]}, {d[9], d[10], d[11]}, 0.0f};
            }
        }
        engine.runCollisionSystem();
        {
            // the render's step after the collision (MEngine.cpp:397)
            wpbr_host::JobComputeDifferentialsForSurfInter j;
            j.run(&engine);
        }
        wpbr_host::CSceneTreeLBVH t = engine.tree();
        FILE* out = fopen(argv[2], "wb");
        if (!out) return 2;
        uint32_t counts[2] = {(uint32_t)t.nodes.size(), (uint32_t)t.shapesEntities.size()};
        fwrite(counts, 4, 2, out);
        for (const wpbr_host::LinearBVHNode& nd : t.nodes) {
            uint32_t w[3] = {nd.offset, nd.nShapes, nd.axis};
            fwrite(w, 4, 3, out);
        }
        fwrite(t.nodesBounds.data(), sizeof(wpbr_host::DBounds3f), t.nodesBounds.size(), out);
        fwrite(t.shapesEntities.data(), 4, t.shapesEntities.size(), out);
        for (const wpbr_host::CInteraction& it : engine.interactions) {
            float rec[8];
            uint32_t h = it.hCollision;
            std::memcpy(&rec[0], &h, 4);
            rec[1] = it.time;
            rec[2] = it.p.x; rec[3] = it.p.y; rec[4] = it.p.z;
            rec[5] = it.wo.x; rec[6] = it.wo.y; rec[7] = it.wo.z;
            fwrite(rec, 4, 8, out);
        }
        for (uint32_t i = 0; i < hdr[2]; i++) {
            const wpbr_host::CInteraction& it = engine.interactions[i];
            const wpbr_host::CSurfaceInteraction& s = engine.surfaceInteractions[i];
            float r[WPBR_SI_FLOATS] = {};
            auto put = [&](int f, const wpbr_host::DVector3f& v) { r[f] = v.x; r[f + 1] = v.y; r[f + 2] = v.z; };
            if (it.hCollision != wpbr_host::INVALID_HANDLE && !engine.rayCasts[i].bSurfInteraction) {
                put(WPBR_SI_P, it.p);
                put(WPBR_SI_WO, it.wo);
                r[WPBR_SI_TIME] = it.time;
            } else if (it.hCollision != wpbr_host::INVALID_HANDLE) {
                put(WPBR_SI_P, s.p);
                put(WPBR_SI_WO, s.wo);
                r[WPBR_SI_TIME] = s.time;
                put(WPBR_SI_N, s.n);
                r[WPBR_SI_UV] = s.uv.x;
                r[WPBR_SI_UV + 1] = s.uv.y;
                put(WPBR_SI_DPDU, s.dpdu);
                put(WPBR_SI_DPDV, s.dpdv);
                put(WPBR_SI_DNDU, s.dndu);
                put(WPBR_SI_DNDV, s.dndv);
                put(WPBR_SI_SHADING_N, s.shading.n);
                put(WPBR_SI_SHADING_DPDU, s.shading.dpdu);
                put(WPBR_SI_SHADING_DPDV, s.shading.dpdv);
                put(WPBR_SI_SHADING_DNDU, s.shading.dndu);
                put(WPBR_SI_SHADING_DNDV, s.shading.dndv);
                put(WPBR_SI_DPDX, s.dpdx);
                put(WPBR_SI_DPDY, s.dpdy);
                r[WPBR_SI_DUDX] = s.dudx;
                r[WPBR_SI_DVDX] = s.dvdx;
                r[WPBR_SI_DUDY] = s.dudy;
                r[WPBR_SI_DVDY] = s.dvdy;
            }
            fwrite(r, 4, WPBR_SI_FLOATS, out);
        }
        fclose(out);
    } catch (const wpbr_host::Error& e) {
        fprintf(stderr, "wpbr error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
