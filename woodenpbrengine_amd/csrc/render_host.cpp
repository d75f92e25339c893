// Host side of the renderer: scene upload, load-time preprocessing and the
// per-pass launch loop of wpbr_render.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <vector>
#include "render_host.hpp"
#include "spectrum_tables.inc"

using namespace wpbr;

namespace wpbr {

constexpr double PI_REF_H = 3.14;  // pch.h:30

// ---- SampledSpectrum::init (CSpectrumSampled.h:46-89, 236-281) ---------------
float lerp_h(float a, float b, float t) { return (1 - t) * a + t * b; }

float average_samples(const float* lambda, const float* vals, int n, float l0, float l1) {
    if (l1 <= lambda[0]) return vals[0];
    if (l0 >= lambda[n - 1]) return vals[n - 1];
    if (n == 1) return vals[0];
    float sum = 0;
    if (l0 < lambda[0]) sum += vals[0] * (lambda[0] - l0);
    if (l1 > lambda[n - 1]) sum += vals[n - 1] * (l1 - lambda[n - 1]);
    int j = 0;
    while (l0 > lambda[j + 1]) j++;
    for (; j + 1 < n && l1 >= lambda[j]; j++) {
        float s0 = l0 > lambda[j] ? l0 : lambda[j];
        float s1 = l1 < lambda[j + 1] ? l1 : lambda[j + 1];
        float a = lerp_h(vals[j], vals[j + 1], (s0 - lambda[j]) / (lambda[j + 1] - lambda[j]));
        float b = lerp_h(vals[j], vals[j + 1], (s1 - lambda[j]) / (lambda[j + 1] - lambda[j]));
        float prod = (a + b) * (s1 - s0);
        sum = (float)((double)sum + (double)prod * 0.5);
    }
    return sum / (l1 - l0);
}

const HostTables& host_tables() {
    static HostTables t = [] {
        HostTables h;
        const float* refl[7] = {WPBR_RGBRefl2SpectWhite, WPBR_RGBRefl2SpectCyan,
                                WPBR_RGBRefl2SpectMagenta, WPBR_RGBRefl2SpectYellow,
                                WPBR_RGBRefl2SpectRed, WPBR_RGBRefl2SpectGreen,
                                WPBR_RGBRefl2SpectBlue};
        const float* illum[7] = {WPBR_RGBIllum2SpectWhite, WPBR_RGBIllum2SpectCyan,
                                 WPBR_RGBIllum2SpectMagenta, WPBR_RGBIllum2SpectYellow,
                                 WPBR_RGBIllum2SpectRed, WPBR_RGBIllum2SpectGreen,
                                 WPBR_RGBIllum2SpectBlue};
        for (int i = 0; i < NSPEC; i++) {
            float wl0 = lerp_h(400.0f, 700.0f, float(i) / float(NSPEC));
            float wl1 = lerp_h(400.0f, 700.0f, float(i + 1) / float(NSPEC));
            h.xyz[i] = average_samples(WPBR_CIE_lambda, WPBR_CIE_X, WPBR_N_CIE, wl0, wl1);
            h.xyz[NSPEC + i] = average_samples(WPBR_CIE_lambda, WPBR_CIE_Y, WPBR_N_CIE, wl0, wl1);
            h.xyz[2 * NSPEC + i] = average_samples(WPBR_CIE_lambda, WPBR_CIE_Z, WPBR_N_CIE, wl0, wl1);
            for (int b = 0; b < 7; b++) {
                h.refl[b * NSPEC + i] = average_samples(WPBR_RGB2SpectLambda, refl[b], WPBR_N_RGB2SPECT, wl0, wl1);
                h.illum[b * NSPEC + i] = average_samples(WPBR_RGB2SpectLambda, illum[b], WPBR_N_RGB2SPECT, wl0, wl1);
            }
        }
        return h;
    }();
    return t;
}

// fromRGB(rgb, Reflectance) (CSpectrumSampled.h:110-158).
void from_rgb_refl(const float* rgb, float* s) {
    const float* B = host_tables().refl;
    for (int i = 0; i < NSPEC; i++) s[i] = 0.0f;
    auto add = [&](int b, float k) { for (int i = 0; i < NSPEC; i++) s[i] += B[b * NSPEC + i] * k; };
    float r = rgb[0], g = rgb[1], b = rgb[2];
    if (r <= g && r <= b) {
        add(0, r);
        if (g <= b) { add(1, g - r); add(6, b - g); }
        else { add(1, b - r); add(5, g - b); }
    } else if (g <= r && g <= b) {
        add(0, g);
        if (r <= b) { add(2, r - g); add(6, b - r); }
        else { add(2, b - g); add(4, r - b); }
    } else {
        add(0, b);
        if (r <= g) { add(3, r - b); add(5, g - r); }
        else { add(3, g - b); add(4, r - g); }
    }
}

// Device materials and conductor spectra (eta[64], k[64] each) from the
// descriptor's 9-float records: conductor spectra via fromRGB(Reflectance); a
// textured conductor (CMaterialMetal.cpp:24-38) takes the constant spectrum of
// its eta (the reference's Spectrum(1.5f)) and, without a metallic texture,
// fromRGB of its k RGB.
void build_materials(const float* h, uint32_t nm, const int32_t* mat_tex, std::vector<DevMaterial>& mats,
                     std::vector<float>& specs) {
    mats.assign(nm, DevMaterial{});
    specs.clear();
    for (uint32_t m = 0; m < nm; m++) {
        const float* src = h + 9 * (size_t)m;
        std::memcpy(&mats[m].type, &src[0], 4);
        mats[m].alpha = src[1];
        mats[m].eta = src[2];
        mats[m].spec = 0;
        mats[m].tex_k = mat_tex ? mat_tex[2 * (size_t)m] : -1;
        mats[m].tex_r = mat_tex ? mat_tex[2 * (size_t)m + 1] : -1;
        if (mats[m].type == MAT_CONDUCTOR) {
            mats[m].spec = (uint32_t)(specs.size() / (2 * NSPEC));
            float e[NSPEC], k[NSPEC];
            if (mats[m].tex_k >= 0 || mats[m].tex_r >= 0)
                for (int i = 0; i < NSPEC; i++) e[i] = src[2];
            else
                from_rgb_refl(src + 3, e);
            from_rgb_refl(src + 6, k);
            specs.insert(specs.end(), e, e + NSPEC);
            specs.insert(specs.end(), k, k + NSPEC);
        }
    }
}

// Bilinear mip-0 lookup (CSTextureProcess.h:58-81), indices clamped to the texture.
void env_bilerp_h(const float* rgbTex, int W, int H, float s_, float t_, float* out) {
    float s = s_ * (float)W - 0.5f, t = t_ * (float)H - 0.5f;
    int s0 = (int)std::floor(s), t0 = (int)std::floor(t);
    s0 = s0 < 0 ? 0 : s0; t0 = t0 < 0 ? 0 : t0;
    s0 = s0 > W - 1 ? W - 1 : s0; t0 = t0 > H - 1 ? H - 1 : t0;
    float ds = s - (float)s0, dt = t - (float)t0;
    int s0n = s0 + 1 > W - 1 ? W - 1 : s0 + 1;
    int t0n = t0 + 1 > H - 1 ? H - 1 : t0 + 1;
    s0n = s0n < 0 ? 0 : s0n; t0n = t0n < 0 ? 0 : t0n;
    auto tex = [&](int x, int y, int c) { return rgbTex[((size_t)y * W + x) * 3 + c]; };
    for (int c = 0; c < 3; c++) {
        float v = ((tex(s0, t0, c) * (1 - ds)) * (1 - dt) + (tex(s0, t0n, c) * (1 - ds)) * dt) +
                  (tex(s0n, t0, c) * ds) * (1 - dt);
        out[c] = v + (tex(s0n, t0n, c) * ds) * dt;
    }
}

// JobLightInfiniteAreaPreprocessSampling (CLightInfiniteArea.cpp:44-70) followed by
// the pbrt-v3 Distribution2D constructor: returns the per-row functions (the
// reference's data array, whose index is x + y * H) and the marginal integral.
void env_distribution(const float* rgbTex, uint32_t W, uint32_t H, std::vector<float>& func,
                      std::vector<float>& rowInt, float& margInt) {
    func.assign((size_t)W * H, 0.0f);
    for (uint32_t y = 0; y < H; y++) {
        float yN = (float)y / (float)H;
        float sinTheta = (float)std::sin(PI_REF_H * (double)float(y + 0.5f) / (double)float(H));
        for (uint32_t x = 0; x < W; x++) {
            float xN = (float)x / (float)W;
            float rgb[3];
            env_bilerp_h(rgbTex, (int)W, (int)H, xN, yN, rgb);
            func[x + (size_t)y * H] = rgb[1];
            func[x + (size_t)y * H] *= sinTheta;
        }
    }
    rowInt.assign(H, 0.0f);
    for (uint32_t v = 0; v < H; v++) {
        float c = 0.0f;
        for (uint32_t i = 1; i < W + 1; i++) c = c + func[(size_t)v * W + i - 1] / (int)W;
        rowInt[v] = c;
    }
    float c = 0.0f;
    for (uint32_t i = 1; i < H + 1; i++) c = c + rowInt[i - 1] / (int)H;
    margInt = c;
}

}  // namespace wpbr

namespace wpbr {

void ctx_release_render(wpbr_ctx* ctx) {
    if (!ctx->render) return;
    wpbr_render_state* r = ctx->render;
    if (r->mem) (void)hipFree(r->mem);
    if (r->spill2) (void)hipFree(r->spill2);
    if (r->claim2) (void)hipFree(r->claim2);
    if (r->stream2) (void)hipStreamDestroy(r->stream2);
    for (auto& e : ctx->render->evs) (void)hipEventDestroy(e);
    delete ctx->render;
    ctx->render = nullptr;
}

size_t carve_wave(char* p, uint32_t n, WaveState& w) {
    size_t off = 0;
    auto take = [&](size_t count) -> void* {
        off = (off + 255) & ~size_t(255);
        void* q = p ? p + off : nullptr;
        off += count * 4;
        return q;
    };
    w.c1 = (uint32_t*)take(n);
    w.c2 = (uint32_t*)take(n);
    w.pfilm = (float*)take(2 * (size_t)n);
    w.rays = (float*)take(6 * (size_t)n);
    w.hit = (uint32_t*)take(n);
    w.hit_t = (float*)take(n);
    w.bary = (float*)take(3 * (size_t)n);
    w.flags = (uint32_t*)take(n);
    w.light = (uint32_t*)take(n);
    w.mat = (uint32_t*)take(n);
    w.l_cos = (float*)take(n);
    w.l_sc = (float*)take(n);
    w.l_ac = (float*)take(n);
    w.l_pdf = (float*)take(n);
    w.l_bpdf = (float*)take(n);
    w.l_zero = (uint32_t*)take(n);
    w.b_cos = (float*)take(n);
    w.b_sc = (float*)take(n);
    w.b_ac = (float*)take(n);
    w.b_pdf = (float*)take(n);
    w.b_zero = (uint32_t*)take(n);
    w.b_wi = (float*)take(3 * (size_t)n);
    w.le = (float*)take(3 * (size_t)n);
    w.tk = (float*)take(3 * (size_t)n);
    w.qrays = (float*)take(12 * (size_t)n);
    w.qhit = (uint32_t*)take(2 * (size_t)n);
    w.qt = (float*)take(2 * (size_t)n);
    w.qcount = (uint32_t*)take(64);
    w.l_ray = (uint32_t*)take(n);
    w.b_ray = (uint32_t*)take(n);
    w.qtarget = (uint32_t*)take(2 * (size_t)n);
    w.r_wo = (float*)take(3 * (size_t)n);
    w.r_frame = (float*)take(9 * (size_t)n);
    w.r_wl = (float*)take(3 * (size_t)n);
    w.r_pdfl = (float*)take(n);
    w.r_ub = (float*)take(2 * (size_t)n);
    return (off + 255) & ~size_t(255);
}

wpbr_status ensure_wave(wpbr_ctx* ctx, uint32_t n) {
    if (!ctx->render) ctx->render = new wpbr_render_state();
    wpbr_render_state* r = ctx->render;
    if (n <= r->cap) return WPBR_OK;
    if (r->mem) { hipFree(r->mem); r->mem = nullptr; r->cap = 0; }
    WaveState tmp;
    const size_t bytes = carve_wave(nullptr, n, tmp);
    wpbr_status s = hip_check(ctx, hipMalloc(&r->mem, 2 * bytes), "hipMalloc render state");
    if (s) return s;
    r->cap = n;
    carve_wave((char*)r->mem, n, r->ws);
    carve_wave((char*)r->mem + bytes, n, r->ws2);
    // one set of per-pixel sample-stream cursors: each pass continues them
    r->ws2.c1 = r->ws.c1;
    r->ws2.c2 = r->ws.c2;
    return WPBR_OK;
}

// The second stream of the render loop and its traversal scratch (a spill area
// like the context's, and claim counters, zeroed once; the kernels leave them
// zero).
wpbr_status ensure_stream2(wpbr_ctx* ctx) {
    wpbr_render_state* r = ctx->render;
    if (!r->stream2) {
        wpbr_status s = hip_check(ctx, hipStreamCreateWithFlags(&r->stream2, hipStreamNonBlocking), "stream");
        if (s) return s;
    }
    if (!r->claim2) {
        wpbr_status s = hip_check(ctx, hipMalloc(&r->claim2, 16), "hipMalloc");
        if (!s) s = hip_check(ctx, hipMemsetAsync(r->claim2, 0, 16, ctx->stream), "memset");
        if (s) return s;
    }
    if (r->spill2_words < ctx->spill_words) {
        if (r->spill2) (void)hipFree(r->spill2);
        r->spill2 = nullptr;
        r->spill2_words = 0;
        wpbr_status s = hip_check(ctx, hipMalloc(&r->spill2, ctx->spill_words * 4), "hipMalloc spill");
        if (s) return s;
        r->spill2_words = ctx->spill_words;
    }
    return WPBR_OK;
}

}  // namespace wpbr


namespace wpbr {
// Uploads a scene description (everything but the LBVH build).
wpbr_status scene_upload(wpbr_ctx* ctx, const wpbr_scene_desc* d, wpbr_scene** out) {
    if (!ctx || !d || !out) return WPBR_ERR_INVALID_ARG;
    *out = nullptr;
    if (d->width == 0 || d->height == 0 || d->spp == 0 || (d->width % 32) || (d->height % 32))
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "film must be a non-empty multiple of 32x32 tiles");
    if (d->sampler > WPBR_SAMPLER_STRATIFIED_N) return set_err(ctx, WPBR_ERR_INVALID_ARG, "unknown sampler mode");
    uint32_t nLights = d->n_sphere_lights + (d->has_env ? 1u : 0u);
    if (nLights == 0) return set_err(ctx, WPBR_ERR_INVALID_ARG, "scene has no light");
    uint64_t n = (uint64_t)d->n_spheres + d->n_triangles;
    if (n == 0) return set_err(ctx, WPBR_ERR_INVALID_ARG, "scene has no shapes");
    for (uint32_t l = 0; l < d->n_sphere_lights; l++)
        if (d->light_shape[l] >= d->n_spheres)
            return set_err(ctx, WPBR_ERR_INVALID_ARG, "light shape is not a sphere");
    for (uint64_t s = 0; s < n; s++)
        if (d->shape_material[s] != 0xFFFFFFFFu && d->shape_material[s] >= d->n_materials)
            return set_err(ctx, WPBR_ERR_INVALID_ARG, "material index out of range");
    if (d->n_textures) {
        if (!d->textures || !d->material_textures)
            return set_err(ctx, WPBR_ERR_INVALID_ARG, "textures: null texture or binding array");
        for (uint32_t t = 0; t < d->n_textures; t++)
            if (!d->textures[t]) return set_err(ctx, WPBR_ERR_INVALID_ARG, "textures: null texture");
        for (uint32_t m = 0; m < d->n_materials; m++) {
            uint32_t type;
            std::memcpy(&type, d->materials + 9 * (size_t)m, 4);
            for (int k = 0; k < 2; k++) {
                const int32_t ti = d->material_textures[2 * (size_t)m + k];
                if (ti < -1 || ti >= (int32_t)d->n_textures)
                    return set_err(ctx, WPBR_ERR_INVALID_ARG, "material " + std::to_string(m) + ": no such texture");
                if (ti >= 0 && type != MAT_CONDUCTOR)
                    return set_err(ctx, WPBR_ERR_INVALID_ARG,
                                   "material " + std::to_string(m) + ": only conductors bind textures");
            }
        }
    }
    hipSetDevice(ctx->device);
    wpbr_scene* sc = new wpbr_scene();
    sc->ctx = ctx;
    sc->spp = d->spp;
    sc->seed = d->seed;
    sc->smp = make_sampler(d->seed, d->sampler, 16u + d->spp, 16u + 4u * d->spp);
    wpbr_status st = WPBR_OK;
    auto up = [&](const void* src, size_t bytes) -> void* {
        if (st) return nullptr;
        void* p = nullptr;
        st = hip_check(ctx, hipMalloc(&p, bytes ? bytes : 4), "hipMalloc scene");
        if (st) return nullptr;
        sc->bufs.push_back(p);
        if (bytes && src) st = hip_check(ctx, hipMemcpy(p, src, bytes, hipMemcpyHostToDevice), "scene upload");
        return p;
    };
    const uint32_t ns = d->n_spheres, nt = d->n_triangles;
    SceneView& v = sc->sv;
    v.g.n_spheres = ns;
    v.g.sphere_xform = (const float*)up(d->sphere_xform, 96 * (size_t)ns);
    v.g.sphere_radius = (const float*)up(d->sphere_radius, 4 * (size_t)ns);
    v.g.n_triangles = nt;
    v.g.tri_verts = (const float*)up(d->tri_verts, 36 * (size_t)nt);
    std::vector<uint8_t> flags(nt, 0);
    if (d->tri_flags) std::memcpy(flags.data(), d->tri_flags, nt);
    v.tri_flags = (const uint8_t*)up(flags.data(), nt);
    v.tri_uv = (const float*)up(d->tri_uv, d->tri_uv ? 24 * (size_t)nt : 0);
    v.tri_normal = (const float*)up(d->tri_normal, d->tri_normal ? 36 * (size_t)nt : 0);
    v.tri_tangent = (const float*)up(d->tri_tangent, d->tri_tangent ? 36 * (size_t)nt : 0);
    v.shape_material = (const uint32_t*)up(d->shape_material, 4 * (size_t)n);
    // materials: conductor spectra via fromRGB(Reflectance); texture bindings
    sc->h_mat_tex.assign(2 * (size_t)d->n_materials, -1);
    std::vector<TexView> views(d->n_textures);
    if (d->n_textures) {
        std::memcpy(sc->h_mat_tex.data(), d->material_textures, 8 * (size_t)d->n_materials);
        for (uint32_t t = 0; t < d->n_textures; t++) views[t] = *texture_view(d->textures[t]);
        for (int32_t ti : sc->h_mat_tex) sc->textured = sc->textured || ti >= 0;
    }
    v.textures = (const TexView*)up(views.data(), sizeof(TexView) * views.size());
    std::vector<DevMaterial> mats;
    std::vector<float> specs;
    build_materials(d->materials, d->n_materials, sc->h_mat_tex.data(), mats, specs);
    v.materials = (const DevMaterial*)up(mats.data(), sizeof(DevMaterial) * mats.size());
    v.mat_spec = (const float*)up(specs.data(), 4 * specs.size());
    v.n_spec_floats = (uint32_t)specs.size();
    sc->d_materials = const_cast<DevMaterial*>(v.materials);
    sc->d_mat_spec = const_cast<float*>(v.mat_spec);
    sc->h_materials.assign(d->materials, d->materials + 9 * (size_t)d->n_materials);
    v.n_sphere_lights = d->n_sphere_lights;
    v.light_shape = (const uint32_t*)up(d->light_shape, 4 * (size_t)d->n_sphere_lights);
    v.light_emit = (const float*)up(d->light_emit, 4 * NSPEC * (size_t)d->n_sphere_lights);
    v.has_env = d->has_env;
    v.env_w = d->has_env ? d->env_w : 1;
    v.env_h = d->has_env ? d->env_h : 1;
    if (d->has_env) {
        v.env_emit = (const float*)up(d->env_emit, 4 * NSPEC);
        v.env_rgb = (const float*)up(d->env_rgb, 12 * (size_t)d->env_w * d->env_h);
        v.env_xform = (const float*)up(d->env_xform, 96);
        std::vector<float> func, rowInt;
        float margInt = 0;
        env_distribution(d->env_rgb, d->env_w, d->env_h, func, rowInt, margInt);
        v.env_cond_func = (const float*)up(func.data(), 4 * func.size());
        v.env_cond_int = (const float*)up(rowInt.data(), 4 * rowInt.size());
        v.env_marg_int = (const float*)up(&margInt, 4);
        sc->d_env_rgb = const_cast<float*>(v.env_rgb);
        sc->d_env_func = const_cast<float*>(v.env_cond_func);
        sc->d_env_rowint = const_cast<float*>(v.env_cond_int);
        sc->d_env_margint = const_cast<float*>(v.env_marg_int);
        sc->h_env.assign(d->env_rgb, d->env_rgb + 3 * (size_t)d->env_w * d->env_h);
    }
    const HostTables& T = host_tables();
    v.xyz = (const float*)up(T.xyz, sizeof(T.xyz));
    v.illum = (const float*)up(T.illum, sizeof(T.illum));
    v.refl = (const float*)up(T.refl, sizeof(T.refl));
    v.raster_camera = (const float*)up(d->raster_camera, 64);
    v.camera_world = (const float*)up(d->camera_world, 48);
    v.width = d->width;
    v.height = d->height;
    if (st) { wpbr_scene_destroy(sc); return st; }
    sc->geom.n_spheres = ns;
    sc->geom.d_sphere_xform = v.g.sphere_xform;
    sc->geom.d_sphere_radius = v.g.sphere_radius;
    sc->geom.n_triangles = nt;
    sc->geom.d_tri_verts = v.g.tri_verts;
    st = wpbr_lbvh_create(ctx, &sc->bvh);
    if (st) { wpbr_scene_destroy(sc); return st; }
    *out = sc;
    return WPBR_OK;
}
}  // namespace wpbr

extern "C" {

wpbr_status wpbr_scene_create(wpbr_ctx* ctx, const wpbr_scene_desc* d, wpbr_scene** out) {
    wpbr_status st = scene_upload(ctx, d, out);
    if (st) return st;
    st = wpbr_lbvh_build(ctx, &(*out)->geom, (*out)->bvh);
    if (st) { wpbr_scene_destroy(*out); *out = nullptr; }
    return st;
}

wpbr_status wpbr_scenes_create(wpbr_ctx* ctx, uint32_t n_scenes, const wpbr_scene_desc* descs, wpbr_scene** out) {
    if (!ctx || !descs || !out || n_scenes == 0) return WPBR_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < n_scenes; i++) out[i] = nullptr;
    std::vector<wpbr_geometry> geoms(n_scenes);
    std::vector<wpbr_lbvh*> bvhs(n_scenes);
    wpbr_status st = WPBR_OK;
    for (uint32_t i = 0; i < n_scenes && !st; i++) {
        st = scene_upload(ctx, &descs[i], &out[i]);
        if (st) { st = set_err(ctx, st, "scene " + std::to_string(i) + ": " + ctx->last_error); break; }
        geoms[i] = out[i]->geom;
        bvhs[i] = out[i]->bvh;
    }
    if (!st) st = wpbr_lbvh_build_segmented(ctx, n_scenes, geoms.data(), bvhs.data());
    if (st) {
        for (uint32_t i = 0; i < n_scenes; i++) { wpbr_scene_destroy(out[i]); out[i] = nullptr; }
    }
    return st;
}

void wpbr_scene_destroy(wpbr_scene* sc) {
    if (!sc) return;
    hipSetDevice(sc->ctx->device);
    hipStreamSynchronize(sc->ctx->stream);
    if (sc->bvh) wpbr_lbvh_destroy(sc->bvh);
    for (void* p : sc->bufs) hipFree(p);
    delete sc;
}

wpbr_status wpbr_scene_rebuild(wpbr_ctx* ctx, wpbr_scene* sc) {
    if (!ctx || !sc) return WPBR_ERR_INVALID_ARG;
    return wpbr_lbvh_build(ctx, &sc->geom, sc->bvh);
}

wpbr_status wpbr_render(wpbr_ctx* ctx, wpbr_scene* sc, float* d_film, float* d_sample_xyz) {
    if (!ctx || !sc || !d_film) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    (void)hipGetLastError();  // launch checks below must not see stale errors
    const uint32_t n = sc->sv.width * sc->sv.height;
    wpbr_status s = ensure_wave(ctx, n);
    if (!s) s = ctx_ensure_trace(ctx, 2 * n);
    if (!s) s = ensure_stream2(ctx);
    if (s) return s;
    wpbr_render_state* r = ctx->render;
    hipStream_t st = ctx->stream;
    s = hip_check(ctx, hipMemsetAsync(d_film, 0, (size_t)n * 16, st), "film clear");
    // qcount[0]: queue length of the current pass; qcount[2..3]: secondary rays traced
    if (!s) s = hip_check(ctx, hipMemsetAsync(r->ws.qcount, 0, 64, st), "queue counters");
    if (!s) s = hip_check(ctx, hipMemsetAsync(r->ws2.qcount, 0, 64, st), "queue counters");
    if (s) return s;
    RenderLaunch L;
    L.sv = sc->sv;
    L.n = n;
    L.spp = sc->spp;
    L.smp = sc->smp;
    L.film = d_film;
    L.sample_xyz = d_sample_xyz;
    L.record = 0;
    const wpbr_lbvh* b = sc->bvh;
    TraceArgs ta;
    ta.g = sc->sv.g;
    ta.trav = (const float4*)b->s.trav;
    ta.trav4 = (const float4*)b->s.trav4;
    ta.prims = (const float4*)b->s.prims;
    ta.root_box = b->s.node_bounds;
    ta.root_word = b->root_word;
    ta.spill_depth = ctx->spill_depth;
    ta.err = ctx->err_flag;
    ta.n_dev = nullptr;
    ta.stride = 0;
    const uint32_t dyn_blocks = ctx_trace_shape(ctx, ta);
    // Pass j runs on stream j % 2 with wavefront state j % 2 (the per-pixel
    // sample-stream cursors are shared): its camera waits for pass j - 1's shade1
    // (the last to advance the cursors), so pass j's camera, primary trace and
    // shading overlap pass j - 1's secondary trace and splat on the other
    // stream; its splat waits for pass j - 1's splat, so the film accumulates
    // the passes in order.
    hipStream_t sts[2] = {st, r->stream2};
    const WaveState* wss[2] = {&r->ws, &r->ws2};
    uint32_t* spills[2] = {ctx->spill, r->spill2};
    uint32_t* claims[2] = {ta.claim, r->claim2};
    // events: start, then per pass 6 (pass start | camera | trace | shade1 | trace | shade2)
    const size_t nev = 1 + 6 * (size_t)sc->spp;
    while (r->evs.size() < nev) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return set_err(ctx, WPBR_ERR_HIP, "hipEventCreate");
        r->evs.push_back(e);
    }
    hipEventRecord(r->evs[0], st);
    hipStreamWaitEvent(sts[1], r->evs[0], 0);
    for (uint32_t j = 0; j < sc->spp && !s; j++) {
        hipEvent_t* e = &r->evs[1 + 6 * (size_t)j];
        const int k = (int)(j & 1u);
        hipStream_t sj = sts[k];
        const WaveState& w = *wss[k];
        L.ws = w;
        L.spp_index = j;
        ta.spill = spills[k];
        ta.claim = claims[k];
        if (j) hipStreamWaitEvent(sj, r->evs[6 * (size_t)j - 2], 0);  // pass j - 1's shade1
        if (j && !ctx->render_overlap) hipStreamWaitEvent(sj, r->evs[6 * (size_t)j], 0);  // all of pass j - 1
        hipEventRecord(e[0], sj);
        s = render_pass_launch(L, 0, sj);
        hipEventRecord(e[1], sj);
        ta.n = n;
        ta.n_dev = nullptr;
        ta.stride = 0;
        ta.rays = w.rays;
        ta.hit_shape = w.hit;
        ta.hit_t = w.hit_t;
        ta.bary = w.bary;
        ta.target = nullptr;
        if (!s) s = trace_launch(ta, ctx->trace_block, ctx->trace_stack, sj, dyn_blocks);
        hipEventRecord(e[2], sj);
        if (!s) s = render_pass_launch(L, 1, sj);
        hipEventRecord(e[3], sj);
        // secondary rays: the queue holds *qcount rays with SoA stride 2n
        ta.n = 2 * n;
        ta.n_dev = w.qcount;
        ta.stride = 2 * n;
        ta.rays = w.qrays;
        ta.hit_shape = w.qhit;
        ta.hit_t = w.qt;
        ta.bary = nullptr;
        ta.target = w.qtarget;
        if (!s) s = trace_launch(ta, ctx->trace_block, ctx->trace_stack, sj, dyn_blocks);
        hipEventRecord(e[4], sj);
        if (j) hipStreamWaitEvent(sj, r->evs[6 * (size_t)j], 0);  // pass j - 1's splat
        if (!s) s = render_pass_launch(L, 2, sj);
        hipEventRecord(e[5], sj);
    }
    const hipEvent_t last = r->evs[6 * (size_t)sc->spp];
    if (sc->spp % 2 == 0) hipStreamWaitEvent(st, last, 0);
    if (s) {
        return set_err(ctx, s, std::string("render launch failed: ") +
                                   hipGetErrorString(last_launch_error));
    }
    unsigned long long q2[2] = {0, 0};
    s = hip_check(ctx, hipMemcpyAsync(&q2[0], r->ws.qcount + 2, 8, hipMemcpyDeviceToHost, st), "d2h");
    if (!s) s = hip_check(ctx, hipMemcpyAsync(&q2[1], r->ws2.qcount + 2, 8, hipMemcpyDeviceToHost, st), "d2h");
    if (!s) s = ctx_finish_trace(ctx);
    if (s) return s;
    const unsigned long long qtotal = q2[0] + q2[1];
    // per-pass stage times, each on its own stream (the two streams overlap)
    double trace_ms = 0, shade_ms = 0, cam_ms = 0, prim_ms = 0;
    for (uint32_t j = 0; j < sc->spp; j++) {
        hipEvent_t* e = &r->evs[1 + 6 * (size_t)j];
        float t0, t1, t2, t3, t4;
        hipEventElapsedTime(&t0, e[0], e[1]);
        hipEventElapsedTime(&t1, e[1], e[2]);
        hipEventElapsedTime(&t2, e[2], e[3]);
        hipEventElapsedTime(&t3, e[3], e[4]);
        hipEventElapsedTime(&t4, e[4], e[5]);
        cam_ms += t0;
        trace_ms += t1 + t3;
        prim_ms += t1;
        shade_ms += t2 + t4;
    }
    float tot = 0;
    hipEventElapsedTime(&tot, r->evs[0], last);
    ctx->last_ms = tot;
    r->stats[0] = tot;
    r->stats[1] = trace_ms;
    r->stats[2] = 2.0 * sc->spp;
    r->stats[3] = (double)n * sc->spp + (double)qtotal;
    r->stats[4] = shade_ms;
    r->stats[5] = cam_ms;
    r->stats[6] = (double)qtotal;
    r->stats[7] = prim_ms;
    return WPBR_OK;
}

wpbr_status wpbr_render_stage(wpbr_ctx* ctx, wpbr_scene* sc, float* d_film, float* d_sample_xyz, uint32_t pass,
                              int stage) {
    if (!ctx || !sc || !d_film || stage < WPBR_STAGE_BEGIN || stage > WPBR_STAGE_INTEGRATE)
        return WPBR_ERR_INVALID_ARG;
    if (stage != WPBR_STAGE_BEGIN && pass >= sc->spp)
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "render stage: pass beyond the scene's spp");
    hipSetDevice(ctx->device);
    (void)hipGetLastError();  // launch checks below must not see stale errors
    const uint32_t n = sc->sv.width * sc->sv.height;
    wpbr_status s = ensure_wave(ctx, n);
    if (!s) s = ctx_ensure_trace(ctx, 2 * n);
    if (s) return s;
    wpbr_render_state* r = ctx->render;
    hipStream_t st = ctx->stream;
    const WaveState& w = r->ws;
    hipEventRecord(ctx->ev0, st);
    if (stage == WPBR_STAGE_BEGIN) {
        s = hip_check(ctx, hipMemsetAsync(d_film, 0, (size_t)n * 16, st), "film clear");
        if (!s) s = hip_check(ctx, hipMemsetAsync(w.qcount, 0, 64, st), "queue counters");
    } else if (stage == WPBR_STAGE_PRIMARY_TRACE || stage == WPBR_STAGE_SHADOW_TRACE) {
        const wpbr_lbvh* b = sc->bvh;
        TraceArgs ta;
        ta.g = sc->sv.g;
        ta.trav = (const float4*)b->s.trav;
        ta.trav4 = (const float4*)b->s.trav4;
        ta.prims = (const float4*)b->s.prims;
        ta.root_box = b->s.node_bounds;
        ta.root_word = b->root_word;
        ta.spill_depth = ctx->spill_depth;
        ta.err = ctx->err_flag;
        const uint32_t dyn_blocks = ctx_trace_shape(ctx, ta);
        ta.spill = ctx->spill;
        if (stage == WPBR_STAGE_PRIMARY_TRACE) {
            ta.n = n;
            ta.n_dev = nullptr;
            ta.stride = 0;
            ta.rays = w.rays;
            ta.hit_shape = w.hit;
            ta.hit_t = w.hit_t;
            ta.bary = w.bary;
            ta.target = nullptr;
        } else {
            // the queue holds *qcount rays with SoA stride 2n
            ta.n = 2 * n;
            ta.n_dev = w.qcount;
            ta.stride = 2 * n;
            ta.rays = w.qrays;
            ta.hit_shape = w.qhit;
            ta.hit_t = w.qt;
            ta.bary = nullptr;
            ta.target = w.qtarget;
        }
        s = trace_launch(ta, ctx->trace_block, ctx->trace_stack, st, dyn_blocks);
    } else {
        RenderLaunch L;
        L.sv = sc->sv;
        L.n = n;
        L.spp = sc->spp;
        L.smp = sc->smp;
        L.film = d_film;
        L.sample_xyz = d_sample_xyz;
        L.record = 0;
        L.ws = w;
        L.spp_index = pass;
        s = render_pass_launch(L, stage == WPBR_STAGE_CAMERA ? 0 : stage == WPBR_STAGE_SCATTER ? 1 : 2, st);
    }
    hipEventRecord(ctx->ev1, st);
    if (s) return set_err(ctx, s, std::string("render stage launch failed: ") + hipGetErrorString(last_launch_error));
    s = ctx_finish_trace(ctx);
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

}  // extern "C"

// Closest hits of n ray casts over the tree b of geometry sv.g, then (d_si) the
// interaction of each from sv's per-triangle attributes (the only other fields
// of sv it reads).
static wpbr_status collide_interactions(wpbr_ctx* ctx, const SceneView& sv, const wpbr_lbvh* b, uint32_t n,
                                        const float* d_rays, const uint8_t* d_surf, const float* d_ray_diff,
                                        uint32_t* d_hit_shape, float* d_hit_t, float* d_si) {
    hipSetDevice(ctx->device);
    (void)hipGetLastError();  // launch checks below must not see stale errors
    if (!b->built) return set_err(ctx, WPBR_ERR_NOT_BUILT, "lbvh not built");
    wpbr_status s = ctx_ensure_trace(ctx, n);
    if (s || !n) return s;
    TraceArgs ta;
    ta.g = sv.g;
    ta.trav = (const float4*)b->s.trav;
    ta.trav4 = (const float4*)b->s.trav4;
    ta.prims = (const float4*)b->s.prims;
    ta.root_box = b->s.node_bounds;
    ta.root_word = b->root_word;
    ta.n = n;
    ta.n_dev = nullptr;
    ta.stride = 0;
    ta.rays = d_rays;
    ta.hit_shape = d_hit_shape;
    ta.hit_t = d_hit_t;
    // the closest triangle's barycentrics go to the rows the interaction kernel
    // fills last (it reads each ray's before writing them)
    ta.bary = d_si ? d_si + (size_t)WPBR_SI_DUDX * n : nullptr;
    ta.target = nullptr;
    ta.spill = ctx->spill;
    ta.spill_depth = ctx->spill_depth;
    ta.err = ctx->err_flag;
    const uint32_t dyn_blocks = ctx_trace_shape(ctx, ta);
    hipEventRecord(ctx->ev0, ctx->stream);
    s = trace_launch(ta, ctx->trace_block, ctx->trace_stack, ctx->stream, dyn_blocks);
    if (!s && d_si)
        s = scene_si_launch(sv, n, d_rays, d_surf, d_ray_diff, d_hit_shape, d_hit_t, d_si, ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("collide launch failed: ") + hipGetErrorString(last_launch_error));
    s = ctx_finish_trace(ctx);
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

extern "C" {

wpbr_status wpbr_scene_collide(wpbr_ctx* ctx, wpbr_scene* sc, uint32_t n, const float* d_rays,
                               const uint8_t* d_surf, const float* d_ray_diff, uint32_t* d_hit_shape,
                               float* d_hit_t, float* d_si) {
    if (!ctx || !sc || (n && (!d_rays || !d_hit_shape || !d_hit_t))) return WPBR_ERR_INVALID_ARG;
    return collide_interactions(ctx, sc->sv, sc->bvh, n, d_rays, d_surf, d_ray_diff, d_hit_shape, d_hit_t, d_si);
}

wpbr_status wpbr_collide_interactions(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_tri_attributes* attrs,
                                      const wpbr_lbvh* bvh, uint32_t n, const float* d_rays, const uint8_t* d_surf,
                                      const float* d_ray_diff, uint32_t* d_hit_shape, float* d_hit_t,
                                      float* d_si) {
    if (!ctx || !geom || !bvh || (n && (!d_rays || !d_hit_shape || !d_hit_t))) return WPBR_ERR_INVALID_ARG;
    if ((uint64_t)geom->n_spheres + geom->n_triangles != bvh->n)
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "geometry does not match the lbvh");
    if (attrs && attrs->d_tri_flags && geom->n_triangles &&
        (!attrs->d_tri_uv || !attrs->d_tri_normal || !attrs->d_tri_tangent))
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "triangle flags given without uv, normal and tangent arrays");
    SceneView sv;
    std::memset(&sv, 0, sizeof(sv));
    sv.g = to_view(geom);
    if (attrs) {
        sv.tri_flags = attrs->d_tri_flags;
        sv.tri_uv = attrs->d_tri_uv;
        sv.tri_normal = attrs->d_tri_normal;
        sv.tri_tangent = attrs->d_tri_tangent;
    }
    return collide_interactions(ctx, sv, bvh, n, d_rays, d_surf, d_ray_diff, d_hit_shape, d_hit_t, d_si);
}

}  // extern "C"

static wpbr_status bsdf_args(wpbr_ctx* ctx, wpbr_scene* sc, const wpbr_bsdf_requests* req, uint32_t n, bool sample,
                             float* d_value, float* d_pdf, float* d_wi, BsdfArgs& a) {
    if (!ctx || !sc || !req) return WPBR_ERR_INVALID_ARG;
    if (n > (1u << 30)) return set_err(ctx, WPBR_ERR_INVALID_ARG, "bsdf requests: more than 2^30");
    if (n && (!req->d_frame || !req->d_wo || !req->d_material || !d_value || !d_pdf ||
              (sample ? (!req->d_u || !d_wi) : !req->d_wi)))
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "bsdf requests: a required array is null");
    if (reinterpret_cast<uintptr_t>(d_value) % 16)
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "bsdf value array: not 16-byte aligned");
    a.frame = req->d_frame;
    a.wo = req->d_wo;
    a.wi = req->d_wi;
    a.u = req->d_u;
    a.material = req->d_material;
    a.n = n;
    a.n_materials = (uint32_t)(sc->h_materials.size() / 9);
    a.value = d_value;
    a.pdf = d_pdf;
    a.wi_out = d_wi;
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    return WPBR_OK;
}

static wpbr_status bsdf_run(wpbr_ctx* ctx, wpbr_scene* sc, int kind, const BsdfArgs& a, bool sample,
                            const uint32_t* d_queue, uint32_t q) {
    if (kind != WPBR_BSDF_ANY && (kind < 0 || kind > 2)) return set_err(ctx, WPBR_ERR_INVALID_ARG, "bsdf kind");
    if (kind != WPBR_BSDF_ANY && q && !d_queue) return set_err(ctx, WPBR_ERR_INVALID_ARG, "null queue");
    hipEventRecord(ctx->ev0, ctx->stream);
    wpbr_status s = bsdf_launch(sc->sv, a, kind == WPBR_BSDF_ANY ? -1 : kind, sample,
                                kind == WPBR_BSDF_ANY ? nullptr : d_queue, kind == WPBR_BSDF_ANY ? a.n : q,
                                ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("bsdf launch failed: ") + hipGetErrorString(last_launch_error));
    s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "bsdf");
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

extern "C" {

wpbr_status wpbr_bsdf_queues(wpbr_ctx* ctx, wpbr_scene* sc, const wpbr_bsdf_requests* req, uint32_t n,
                             uint32_t* d_queue, uint32_t counts[3]) {
    if (!ctx || !sc || !req || !counts || (n && (!req->d_material || !d_queue))) return WPBR_ERR_INVALID_ARG;
    BsdfArgs a{};
    a.material = req->d_material;
    a.n = n;
    a.n_materials = (uint32_t)(sc->h_materials.size() / 9);
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    counts[0] = counts[1] = counts[2] = 0;
    if (!n) return WPBR_OK;
    const size_t words = 3 * (((size_t)n + 255) / 256) + 4;
    if (words > ctx->bsdf_blk_words) {
        if (ctx->bsdf_blk) hipFree(ctx->bsdf_blk);
        ctx->bsdf_blk = nullptr;
        ctx->bsdf_blk_words = 0;
        wpbr_status s = hip_check(ctx, hipMalloc(&ctx->bsdf_blk, 4 * words), "hipMalloc bsdf queues");
        if (s) return s;
        ctx->bsdf_blk_words = words;
    }
    uint32_t* d_counts = ctx->bsdf_blk + words - 4;
    hipEventRecord(ctx->ev0, ctx->stream);
    wpbr_status s = bsdf_queues_launch(sc->sv, a, ctx->bsdf_blk, d_counts, d_queue, ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("bsdf queues launch failed: ") + hipGetErrorString(last_launch_error));
    s = hip_check(ctx, hipMemcpyAsync(counts, d_counts, 12, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "bsdf queues");
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

wpbr_status wpbr_bsdf_compute(wpbr_ctx* ctx, wpbr_scene* sc, int kind, const wpbr_bsdf_requests* req, uint32_t n,
                              const uint32_t* d_queue, uint32_t q, float* d_value, float* d_pdf) {
    BsdfArgs a{};
    wpbr_status s = bsdf_args(ctx, sc, req, n, false, d_value, d_pdf, nullptr, a);
    return s ? s : bsdf_run(ctx, sc, kind, a, false, d_queue, q);
}

wpbr_status wpbr_bsdf_sample(wpbr_ctx* ctx, wpbr_scene* sc, int kind, const wpbr_bsdf_requests* req, uint32_t n,
                             const uint32_t* d_queue, uint32_t q, float* d_wi, float* d_value, float* d_pdf) {
    BsdfArgs a{};
    wpbr_status s = bsdf_args(ctx, sc, req, n, true, d_value, d_pdf, d_wi, a);
    return s ? s : bsdf_run(ctx, sc, kind, a, true, d_queue, q);
}

wpbr_status wpbr_set_render_overlap(wpbr_ctx* ctx, int overlap) {
    if (!ctx || (overlap != 0 && overlap != 1)) return WPBR_ERR_INVALID_ARG;
    ctx->render_overlap = overlap;
    return WPBR_OK;
}

int wpbr_render_stats(wpbr_ctx* ctx, double* out, int cap) {
    if (!ctx || !out || !ctx->render) return 0;
    int k = cap < 8 ? cap : 8;
    for (int i = 0; i < k; i++) out[i] = ctx->render->stats[i];
    return k;
}

}  // extern "C"

extern "C" wpbr_status wpbr_autotune_render(wpbr_ctx* ctx, wpbr_scene* sc, float* d_film,
                                            wpbr_trace_config* best) {
    if (!ctx || !sc || !d_film) return WPBR_ERR_INVALID_ARG;
    return wpbr::autotune(ctx, wpbr::trace_candidates(true), [&](float* ms) {
        wpbr_status s = wpbr_render(ctx, sc, d_film, nullptr);
        *ms = (float)ctx->render->stats[0];
        return s;
    }, best);
}
