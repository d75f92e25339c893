// Runs the reference's render job sequence (MEngine::render, MEngine.cpp:380-448)
// through wpbr_jobs.hpp on one scene, and the BSDF Compute / Sample jobs over a
// list of requests.
//
// usage: wpbr_render_demo SCENE OUT [BSDF_IN BSDF_OUT]
// SCENE (little-endian): the fields of wpbr_scene_desc in declaration order,
//     each uint32_t field as 4 bytes and each array as u64 byte count + bytes
//     (count 0: null); no textures (n_textures 0).
// OUT: the film (width*height*4 f32), then the 8-bit film (width*height*3 u8);
//     the PNG of JobOutputFilm is written to OUT.png.
// BSDF_IN: u32 n, f32 frame[9 n] (rows ss, ts, ns; SoA), f32 wo[3 n], f32 wi[3 n],
//     f32 u[2 n], u32 material[n].
// BSDF_OUT: Compute value[n][64], pdf[n], then Sample value[n][64], pdf[n], wi[n][3].
#include <cstdio>
#include <vector>

#include "wpbr_jobs.hpp"

namespace {

struct Reader {
    FILE* f;
    std::vector<std::vector<uint8_t>> store;
    bool ok = true;
    void u32(uint32_t& v) { ok = ok && fread(&v, 4, 1, f) == 1; }
    template <class T>
    void arr(const T*& p) {
        uint64_t nb = 0;
        ok = ok && fread(&nb, 8, 1, f) == 1;
        p = nullptr;
        if (!ok || nb == 0) return;
        store.emplace_back(nb);
        ok = fread(store.back().data(), 1, nb, f) == nb;
        p = reinterpret_cast<const T*>(store.back().data());
    }
};

bool read_desc(Reader& r, wpbr_scene_desc& d) {
    d = wpbr_scene_desc{};
    r.u32(d.n_spheres);
    r.arr(d.sphere_xform);
    r.arr(d.sphere_radius);
    r.u32(d.n_triangles);
    r.arr(d.tri_verts);
    r.arr(d.tri_flags);
    r.arr(d.tri_uv);
    r.arr(d.tri_normal);
    r.arr(d.tri_tangent);
    r.arr(d.shape_material);
    r.u32(d.n_materials);
    r.arr(d.materials);
    r.u32(d.n_sphere_lights);
    r.arr(d.light_shape);
    r.arr(d.light_emit);
    r.u32(d.has_env);
    r.arr(d.env_emit);
    r.arr(d.env_rgb);
    r.u32(d.env_w);
    r.u32(d.env_h);
    r.arr(d.env_xform);
    r.arr(d.raster_camera);
    r.arr(d.camera_world);
    r.u32(d.width);
    r.u32(d.height);
    r.u32(d.spp);
    r.u32(d.seed);
    r.u32(d.sampler);
    // textures are device objects: a serialised scene has none
    const uint8_t* unused = nullptr;
    r.u32(d.n_textures);
    r.arr(unused);
    r.arr(d.material_textures);
    return r.ok && d.n_textures == 0;
}

template <class T>
bool read_vec(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3 && argc != 5) {
        fprintf(stderr, "usage: %s SCENE OUT [BSDF_IN BSDF_OUT]\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    Reader rd{in, {}};
    wpbr_scene_desc desc;
    const bool ok = read_desc(rd, desc);
    fclose(in);
    if (!ok) {
        fprintf(stderr, "short scene\n");
        return 2;
    }
    try {
        wpbr_host::MEngine engine(0);
        engine.setScene(desc);
        engine.outputFilename = std::string(argv[2]) + ".png";
        engine.render();
        FILE* out = fopen(argv[2], "wb");
        if (!out) return 2;
        fwrite(engine.film.data(), 4, engine.film.size(), out);
        fwrite(engine.filmRGB8.data(), 1, engine.filmRGB8.size(), out);
        fclose(out);
        if (argc == 5) {
            FILE* bi = fopen(argv[3], "rb");
            if (!bi) return 2;
            uint32_t n = 0;
            std::vector<float> frame, wo, wi, u;
            std::vector<uint32_t> mat;
            bool bok = fread(&n, 4, 1, bi) == 1 && read_vec(bi, frame, 9 * (size_t)n) &&
                       read_vec(bi, wo, 3 * (size_t)n) && read_vec(bi, wi, 3 * (size_t)n) &&
                       read_vec(bi, u, 2 * (size_t)n) && read_vec(bi, mat, n);
            fclose(bi);
            if (!bok) {
                fprintf(stderr, "short BSDF requests\n");
                return 2;
            }
            engine.bsdfRequests.resize(n);
            auto v3 = [&](const std::vector<float>& a, int row, uint32_t i) {
                return wpbr_host::DVector3f{a[(size_t)(3 * row) * n + i], a[(size_t)(3 * row + 1) * n + i],
                                            a[(size_t)(3 * row + 2) * n + i]};
            };
            for (uint32_t i = 0; i < n; i++) {
                wpbr_host::CBSDFRequest& r = engine.bsdfRequests[i];
                r.transform = {v3(frame, 0, i), v3(frame, 1, i), v3(frame, 2, i)};
                r.wo = v3(wo, 0, i);
                r.wi = v3(wi, 0, i);
                r.u = {u[i], u[n + i]};
                r.material = mat[i];
            }
            FILE* bo = fopen(argv[4], "wb");
            if (!bo) return 2;
            engine.computeBSDFs();
            fwrite(engine.sampledBSDFValues.data(), sizeof(wpbr_host::CSampledBSDFValue), n, bo);
            fwrite(engine.sampledBSDFPDFs.data(), 4, n, bo);
            engine.sampleBSDFs();
            fwrite(engine.sampledBSDFValues.data(), sizeof(wpbr_host::CSampledBSDFValue), n, bo);
            fwrite(engine.sampledBSDFPDFs.data(), 4, n, bo);
            fwrite(engine.sampledWIs.data(), sizeof(wpbr_host::CSampledWI), n, bo);
            fclose(bo);
        }
    } catch (const wpbr_host::Error& e) {
        fprintf(stderr, "wpbr error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
