// Runs the reference's job sequences through wpbr_jobs.hpp on one scene:
// MEngine::buildLBVH then MEngine::runCollisionSystem over a list of ray casts.
//
// usage: wpbr_jobs_demo IN OUT
// IN  (little-endian): u32 n_spheres, u32 n_triangles, u32 n_rays,
//     f32 sphere_xform[24 n_spheres], f32 radius[n_spheres],
//     f32 tri_verts[9 n_triangles], f32 rays[6 n_rays] (origin xyz, dir xyz)
// OUT: u32 n_nodes, u32 n_shapes, then per node u32 offset, u32 nShapes, u32 axis,
//     f32 bounds[6 n_nodes], u32 shapesEntities[n_shapes], then per ray
//     u32 hCollision, f32 time, f32 p[3], f32 wo[3]
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
    uint32_t hdr[3];
    std::vector<float> xf, rad, tris, rays;
    bool ok = fread(hdr, 4, 3, in) == 3 && read_vec(in, xf, 24 * (size_t)hdr[0]) &&
              read_vec(in, rad, hdr[0]) && read_vec(in, tris, 9 * (size_t)hdr[1]) &&
              read_vec(in, rays, 6 * (size_t)hdr[2]);
    fclose(in);
    if (!ok) {
        fprintf(stderr, "short input\n");
        return 2;
    }
    try {
        wpbr_host::MEngine engine(0);
        engine.setShapes(xf, rad, tris);
        engine.buildLBVH();
        engine.rayCasts.resize(hdr[2]);
        for (uint32_t i = 0; i < hdr[2]; i++) {
            const float* r = &rays[6 * (size_t)i];
            engine.rayCasts[i] = {{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}, 0.0f}, false};
        }
        engine.runCollisionSystem();
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
        fclose(out);
    } catch (const wpbr_host::Error& e) {
        fprintf(stderr, "wpbr error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
