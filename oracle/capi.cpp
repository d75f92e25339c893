// Oracle C entry points for the pytest checkers and bench.py's CPU baseline.
// TEST INFRASTRUCTURE ONLY.
#include "oracle.h"
#include <thread>
#include <atomic>
#include <chrono>

using namespace orc;

namespace {

template <class F>
void parallel_for(uint32_t n, int nthreads, F f) {
    if (nthreads <= 1 || n < 1024) {
        for (uint32_t i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<uint32_t> next(0);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        th.emplace_back([&]() {
            while (true) {
                uint32_t b = next.fetch_add(256);
                if (b >= n) break;
                uint32_t e = std::min(n, b + 256);
                for (uint32_t i = b; i < e; i++) f(i);
            }
        });
    }
    for (auto& x : th) x.join();
}

Scene make_scene(uint32_t ns, const float* sx, const float* sr, uint32_t nt, const float* tv) {
    Scene s;
    s.n_spheres = ns;
    s.sphere_xform = sx;
    s.sphere_radius = sr;
    s.n_triangles = nt;
    s.tri_verts = tv;
    return s;
}

}  // namespace

extern "C" {

// a1: bounds + centroids for spheres (first) then triangles.
void orc_shape_bounds(uint32_t ns, const float* sx, const float* sr, uint32_t nt, const float* tv,
                      float* bounds, float* centroid) {
    sphere_bounds(ns, sx, sr, bounds, centroid);
    triangle_bounds(nt, tv, bounds + 6 * (size_t)ns, centroid + 3 * (size_t)ns);
}

void orc_centroid_bounds(uint32_t n, const float* centroid, float* out6) {
    Bounds b = centroid_bounds(n, centroid);
    std::memcpy(out6, &b, 24);
}

// order: LBVH_ORDER_REFERENCE (0) or LBVH_ORDER_FULL (1), see oracle.h.
void* orc_lbvh_build(uint32_t n, const float* bounds, const float* centroid,
                     const uint32_t* entities, int order) {
    LbvhResult* r = new LbvhResult();
    lbvh_build(n, bounds, centroid, entities, *r, order);
    return r;
}

void orc_lbvh_free(void* h) { delete (LbvhResult*)h; }

// counts: n_nodes, n_sah_nodes, n_treelets
void orc_lbvh_info(void* h, uint32_t* counts) {
    LbvhResult* r = (LbvhResult*)h;
    counts[0] = r->n_nodes;
    counts[1] = r->n_sah_nodes;
    counts[2] = r->n_treelets;
}

// field: 0 morton, 1 sorted_code, 2 sorted_index, 3 shape_entities, 4 node_offset,
// 5 node_nshapes, 6 node_axis (u8), 7 node_bounds (f32 x6), 8 treelet_start,
// 9 centroid bounds (f32 x6)
int orc_lbvh_copy(void* h, int field, void* dst) {
    LbvhResult* r = (LbvhResult*)h;
    switch (field) {
        case 0: std::memcpy(dst, r->morton.data(), 4 * r->morton.size()); return 0;
        case 1: std::memcpy(dst, r->sorted_code.data(), 4 * r->sorted_code.size()); return 0;
        case 2: std::memcpy(dst, r->sorted_index.data(), 4 * r->sorted_index.size()); return 0;
        case 3: std::memcpy(dst, r->shape_entities.data(), 4 * r->shape_entities.size()); return 0;
        case 4: std::memcpy(dst, r->node_offset.data(), 4 * (size_t)r->n_nodes); return 0;
        case 5: std::memcpy(dst, r->node_nshapes.data(), 4 * (size_t)r->n_nodes); return 0;
        case 6: std::memcpy(dst, r->node_axis.data(), (size_t)r->n_nodes); return 0;
        case 7: std::memcpy(dst, r->node_bounds.data(), 24 * (size_t)r->n_nodes); return 0;
        case 8: std::memcpy(dst, r->treelet_start.data(), 4 * r->treelet_start.size()); return 0;
        case 9: std::memcpy(dst, &r->centroid_bounds, 24); return 0;
    }
    return -1;
}

// Closest hit for n rays (SoA: rays[0..6][n] = ox, oy, oz, dx, dy, dz).
// Writes the hit shape index (0xFFFFFFFF = miss) and t (+inf on miss).
// Returns the number of candidate primitive tests; *overflow set on stack overflow.
uint64_t orc_collide(uint32_t ns, const float* sx, const float* sr, uint32_t nt, const float* tv,
                     void* h, uint32_t n, const float* rays, uint32_t* hitShape, float* hitT,
                     int nthreads, int* overflow) {
    LbvhResult* r = (LbvhResult*)h;
    Scene sc = make_scene(ns, sx, sr, nt, tv);
    std::atomic<uint64_t> cand(0);
    std::atomic<int> ovf(0);
    parallel_for(n, nthreads, [&](uint32_t i) {
        V3 o = v3(rays[i], rays[(size_t)n + i], rays[2 * (size_t)n + i]);
        V3 d = v3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]);
        float t;
        uint64_t c = 0;
        int so = 0;
        int64_t slot = trace_ray(sc, *r, o, d, &t, &c, &so);
        hitShape[i] = slot < 0 ? 0xFFFFFFFFu : r->sorted_index[(size_t)slot];
        hitT[i] = t;
        cand += c;
        if (so) ovf = 1;
    });
    if (overflow) *overflow = ovf.load();
    return cand.load();
}

// Brute force: minimum t over every primitive in shape order (no BVH).
void orc_brute_force(uint32_t ns, const float* sx, const float* sr, uint32_t nt, const float* tv,
                     uint32_t n, const float* rays, uint32_t* hitShape, float* hitT, int nthreads) {
    Scene sc = make_scene(ns, sx, sr, nt, tv);
    parallel_for(n, nthreads, [&](uint32_t i) {
        V3 o = v3(rays[i], rays[(size_t)n + i], rays[2 * (size_t)n + i]);
        V3 d = v3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]);
        float best = std::numeric_limits<float>::infinity();
        uint32_t id = 0xFFFFFFFFu;
        for (uint32_t s = 0; s < sc.n_spheres + sc.n_triangles; s++) {
            float t = s < sc.n_spheres
                          ? sphere_intersect(sc.sphere_xform + 24 * (size_t)s, sc.sphere_radius[s],
                                             o, d, nullptr, nullptr)
                          : tri_intersect(sc.tri_verts + 9 * (size_t)(s - sc.n_spheres), o, d,
                                          nullptr);
            if (t == 0.0f) continue;
            if (best > t) { best = t; id = s; }
        }
        hitShape[i] = id;
        hitT[i] = best;
    });
}

float orc_tri_intersect(const float* v, const float* o, const float* d, float* bary) {
    return tri_intersect(v, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), bary);
}

float orc_sphere_intersect(const float* xform, float radius, const float* o, const float* d) {
    return sphere_intersect(xform, radius, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), nullptr,
                            nullptr);
}

int orc_box_hit(const float* b6, const float* o, const float* d, float tMax, float* tNear) {
    V3 oo = v3(o[0], o[1], o[2]);
    V3 inv = v3(1.0f / d[0], 1.0f / d[1], 1.0f / d[2]);
    int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
    return box_hit(b6, b6 + 3, oo, inv, neg, tMax, tNear) ? 1 : 0;
}

// libstdc++ std::partition reference used to pin the oracle's two-pointer form.
void orc_partition_check(uint32_t n, const uint8_t* buckets, int best, uint32_t* order_oracle,
                         uint32_t* order_std, uint32_t* mids) {
    std::vector<std::pair<uint8_t, uint32_t>> a(n), b(n);
    for (uint32_t i = 0; i < n; i++) a[i] = b[i] = {buckets[i], i};
    auto pred = [best](const std::pair<uint8_t, uint32_t>& x) { return x.first <= best; };
    auto ma = partition_bidir(a.data(), a.data() + n, pred);
    auto mb = std::partition(b.data(), b.data() + n, pred);
    for (uint32_t i = 0; i < n; i++) { order_oracle[i] = a[i].second; order_std[i] = b[i].second; }
    mids[0] = (uint32_t)(ma - a.data());
    mids[1] = (uint32_t)(mb - b.data());
}

}  // extern "C"

// ---- render -------------------------------------------------------------------
// Scene description; the layout equals wpbr_scene_desc in include/wpbr.h.
struct orc_scene_desc {
    uint32_t n_spheres;
    const float* sphere_xform;
    const float* sphere_radius;
    uint32_t n_triangles;
    const float* tri_verts;
    const uint8_t* tri_flags;
    const float* tri_uv;
    const float* tri_normal;
    const float* tri_tangent;
    const uint32_t* shape_material;
    uint32_t n_materials;
    const float* materials;  // 9 floats each: type(bits), alpha, eta, eta_rgb[3], k_rgb[3]
    uint32_t n_sphere_lights;
    const uint32_t* light_shape;
    const float* light_emit;
    uint32_t has_env;
    const float* env_emit;
    const float* env_rgb;
    uint32_t env_w, env_h;
    const float* env_xform;
    const float* raster_camera;
    const float* camera_world;
    uint32_t width, height, spp, seed;
    uint32_t sampler;
    // textured metals: textures (w, h, texels h x w x 3 each) and per material
    // [metallic, roughness] texture indices (-1 none)
    uint32_t n_textures;
    const uint32_t* tex_size;        // n_textures x 2: w, h
    const float* const* tex_rgb;     // n_textures pointers
    const int32_t* material_textures;  // n_materials x 2
};

extern "C" {

}  // extern "C"

struct SceneHold {
    RenderScene sc;
    std::vector<Material> mats;
    LbvhResult bvh;
};

static SceneHold* make_scene(const orc_scene_desc* d) {
    SceneHold* H = new SceneHold();
    RenderScene& sc = H->sc;
    sc.geom.n_spheres = d->n_spheres;
    sc.geom.sphere_xform = d->sphere_xform;
    sc.geom.sphere_radius = d->sphere_radius;
    sc.geom.n_triangles = d->n_triangles;
    sc.geom.tri_verts = d->tri_verts;
    sc.tri_flags = d->tri_flags;
    sc.tri_uv = d->tri_uv;
    sc.tri_normal = d->tri_normal;
    sc.tri_tangent = d->tri_tangent;
    sc.shape_material = d->shape_material;
    std::vector<Material>& mats = H->mats;
    mats.resize(d->n_materials);
    for (uint32_t i = 0; i < d->n_materials; i++) {
        const float* m = d->materials + 9 * (size_t)i;
        std::memcpy(&mats[i].type, &m[0], 4);
        mats[i].alpha = m[1];
        mats[i].eta = m[2];
        for (int c = 0; c < 3; c++) { mats[i].eta_rgb[c] = m[3 + c]; mats[i].k_rgb[c] = m[6 + c]; }
    }
    sc.n_materials = d->n_materials;
    sc.materials = mats.data();
    sc.n_sphere_lights = d->n_sphere_lights;
    sc.light_shape = d->light_shape;
    sc.light_emit = d->light_emit;
    sc.has_env = (int)d->has_env;
    sc.env_emit = d->env_emit;
    sc.env_rgb = d->env_rgb;
    sc.env_w = d->env_w;
    sc.env_h = d->env_h;
    sc.env_xform = d->env_xform;
    sc.raster_camera = d->raster_camera;
    sc.camera_world = d->camera_world;
    sc.width = d->width;
    sc.height = d->height;
    sc.spp = d->spp;
    sc.seed = d->seed;
    sc.smp = make_sampler(d->seed, d->sampler, 16u + d->spp, 16u + 4u * d->spp);
    for (uint32_t t = 0; t < d->n_textures; t++)
        sc.textures.push_back(generate_mips(d->tex_size[2 * t], d->tex_size[2 * t + 1], d->tex_rgb[t]));
    sc.material_textures = d->n_textures ? d->material_textures : nullptr;
    uint32_t n = d->n_spheres + d->n_triangles;
    std::vector<float> b(6 * (size_t)n), c(3 * (size_t)n);
    sphere_bounds(d->n_spheres, d->sphere_xform, d->sphere_radius, b.data(), c.data());
    triangle_bounds(d->n_triangles, d->tri_verts, b.data() + 6 * (size_t)d->n_spheres,
                    c.data() + 3 * (size_t)d->n_spheres);
    lbvh_build(n, b.data(), c.data(), nullptr, H->bvh);
    return H;
}


extern "C" {

// The BSDF Compute / Sample jobs over n requests of the scene's materials (render.cpp).
void orc_bsdf_eval(const orc_scene_desc* d, int sample, uint32_t n, const float* frame, const float* wo,
                   const float* wi, const float* u, const uint32_t* material, float* value, float* pdf,
                   float* wi_out) {
    SceneHold* H = make_scene(d);
    bsdf_eval(H->sc, sample, n, frame, wo, wi, u, material, value, pdf, wi_out);
    delete H;
}

// The scene's closest hits and interactions for n ray casts (interaction.cpp).
void orc_scene_collide(const orc_scene_desc* d, int nthreads, uint32_t n, const float* rays, const uint8_t* surf,
                       const float* diff, uint32_t* hit_shape, float* hit_t, float* si) {
    SceneHold* H = make_scene(d);
    scene_collide(H->sc, H->bvh, n, rays, surf, diff, hit_shape, hit_t, si, nthreads);
    delete H;
}

void* orc_render(const orc_scene_desc* d, int nthreads, const uint32_t* tiles, uint32_t n_tiles) {
    SceneHold* H = make_scene(d);
    RenderOut* out = new RenderOut();
    auto t0 = std::chrono::steady_clock::now();
    render(H->sc, H->bvh, *out, nthreads, tiles, n_tiles);
    auto t1 = std::chrono::steady_clock::now();
    // counters[7]: wall time of the render loop alone (ns), excluding the build
    out->counters[7] = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    delete H;
    return out;
}

// Finite-difference reference: record the base render's decisions, then
// evaluate the L2 image loss at other parameters (see diff.cpp).
void* orc_diff_record(const orc_scene_desc* d, int nthreads) {
    SceneHold* H = make_scene(d);
    DiffState* D = diff_record(H->sc, H->bvh, nthreads);
    delete H;
    return D;
}

double orc_diff_loss(void* h, const double* alpha, const double* etak, const double* env_rgb,
                     const float* target_rgb, float* rgb_out, int nthreads) {
    return diff_loss(*(DiffState*)h, alpha, etak, env_rgb, target_rgb, rgb_out, nthreads);
}

void orc_diff_free(void* h) { diff_free((DiffState*)h); }

// field: 0 sample_xyz, 1 sample_pfilm, 2 film_rgb, 3 film_xyz, 4 film_w, 5 counters (u64 x 8)
int orc_render_copy(void* h, int field, void* dst) {
    RenderOut* o = (RenderOut*)h;
    switch (field) {
        case 0: std::memcpy(dst, o->sample_xyz.data(), 4 * o->sample_xyz.size()); return 0;
        case 1: std::memcpy(dst, o->sample_pfilm.data(), 4 * o->sample_pfilm.size()); return 0;
        case 2: std::memcpy(dst, o->film_rgb.data(), 4 * o->film_rgb.size()); return 0;
        case 3: std::memcpy(dst, o->film_xyz.data(), 4 * o->film_xyz.size()); return 0;
        case 4: std::memcpy(dst, o->film_w.data(), 4 * o->film_w.size()); return 0;
        case 5: std::memcpy(dst, o->counters, 64); return 0;
    }
    return -1;
}

void orc_render_free(void* h) { delete (RenderOut*)h; }

// Spectrum tables and conversions (for the pinning tests).
void orc_spectrum_tables(float* out) {  // 3 x 64 XYZ, 7 x 64 refl, 7 x 64 illum
    const Tables& T = tables();
    std::memcpy(out, T.X, 256);
    std::memcpy(out + 64, T.Y, 256);
    std::memcpy(out + 128, T.Z, 256);
    std::memcpy(out + 192, T.refl, 7 * 256);
    std::memcpy(out + 192 + 448, T.illum, 7 * 256);
}

void orc_from_rgb(const float* rgb, int illuminant, float* out64) {
    Spec s = from_rgb(rgb, illuminant != 0);
    std::memcpy(out64, s.data(), 256);
}

void orc_to_xyz(const float* s64, float* xyz) {
    Spec s;
    std::memcpy(s.data(), s64, 256);
    to_xyz(s, xyz);
}

void orc_film_rgb8(uint32_t npix, const float* xyz, const float* w, uint8_t* out) {
    film_rgb8(npix, xyz, w, out);
}

void orc_shade_probe(int fn, const float* in, float* out) { shade_probe(fn, in, out); }

void orc_texture_points(const orc_scene_desc* d, uint32_t n, const float* pfilm, uint32_t* shape, float* q) {
    SceneHold* H = make_scene(d);
    texture_points(H->sc, H->bvh, n, pfilm, shape, q);
    delete H;
}

void orc_samples(uint32_t seed, uint32_t pixel, uint32_t n1, uint32_t n2, float* u1, float* u2) {
    const Sampler S = make_sampler(seed, 0, 1, 1);
    for (uint32_t i = 0; i < n1; i++) u1[i] = sample_1d(S, pixel, i);
    for (uint32_t i = 0; i < n2; i++) sample_2d(S, pixel, i, u2 + 2 * i);
}

// The sample arrays of pixels pixel0 .. pixel0 + npix - 1 (npix x n1, npix x n2 x 2):
// by_element = 0 builds them as the reference's jobs do (stratified_arrays),
// 1 reads them one position at a time as the render does (sample_1d / sample_2d).
void orc_sample_arrays(uint32_t seed, uint32_t mode, uint32_t pixel0, uint32_t npix, uint32_t n1,
                       uint32_t n2, int by_element, float* u1, float* u2) {
    const Sampler S = make_sampler(seed, mode, n1, n2);
    for (uint32_t p = 0; p < npix; p++) {
        float* a = u1 + (size_t)p * n1;
        float* b = u2 + 2 * (size_t)p * n2;
        if (!by_element && mode != 0) {
            stratified_arrays(S, pixel0 + p, a, b);
            continue;
        }
        for (uint32_t c = 0; c < n1; c++) a[c] = sample_1d(S, pixel0 + p, c);
        for (uint32_t c = 0; c < n2; c++) sample_2d(S, pixel0 + p, c, b + 2 * c);
    }
}

uint32_t orc_permute_index(uint32_t c, uint32_t n, uint32_t key) { return permute_index(c, n, key); }

}  // extern "C"
