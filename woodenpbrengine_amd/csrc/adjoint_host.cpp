// Host side of the differentiable render: parameter updates, the adjoint's
// device tables, the image loss and the backward pass (path replay of every
// sample pass followed by the adjoint kernel).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "render_host.hpp"

namespace wpbr {

wpbr_status adjoint_launch(const AdjArgs& a, hipStream_t st);
uint32_t adjoint_blocks(uint32_t n);
wpbr_status adjoint_finish(const float* part, uint32_t rows, uint32_t n_mat, float* grad, float* dI,
                           const float* env_dI, uint32_t ntex, hipStream_t st);
wpbr_status film_loss_launch(const float4* film, const float* target, uint32_t npix, float* rgb, float* dL,
                             float* partial, hipStream_t st);
wpbr_status bsdf_grad_launch(const float* in, float* out, uint32_t n, hipStream_t st);
wpbr_status film_loss_sum_launch(const float* partial, uint32_t nb, double* out, hipStream_t st);

namespace {

int rgb_case_h(const float* rgb) {
    float r = rgb[0], g = rgb[1], b = rgb[2];
    if (r <= g && r <= b) return g <= b ? 0 : 1;
    if (g <= r && g <= b) return r <= b ? 2 : 3;
    return r <= g ? 4 : 5;
}

wpbr_status upload(wpbr_ctx* ctx, void* d, const void* h, size_t bytes) {
    if (!bytes) return WPBR_OK;
    return hip_check(ctx, hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream), "param upload");
}

// Device copies of the materials and conductor spectra from the host parameters.
wpbr_status push_materials(wpbr_ctx* ctx, wpbr_scene* sc) {
    const size_t nm = sc->h_materials.size() / 9;
    std::vector<DevMaterial> mats;
    std::vector<float> specs;
    build_materials(sc->h_materials.data(), (uint32_t)nm, sc->h_mat_tex.data(), mats, specs);
    wpbr_status s = upload(ctx, sc->d_materials, mats.data(), sizeof(DevMaterial) * nm);
    if (!s) s = upload(ctx, sc->d_mat_spec, specs.data(), 4 * specs.size());
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "param upload");
    return s;
}

wpbr_status push_env(wpbr_ctx* ctx, wpbr_scene* sc) {
    const uint32_t W = sc->sv.env_w, H = sc->sv.env_h;
    std::vector<float> func, rowInt;
    float margInt = 0;
    env_distribution(sc->h_env.data(), W, H, func, rowInt, margInt);
    wpbr_status s = upload(ctx, sc->d_env_rgb, sc->h_env.data(), 4 * sc->h_env.size());
    if (!s) s = upload(ctx, sc->d_env_func, func.data(), 4 * func.size());
    if (!s) s = upload(ctx, sc->d_env_rowint, rowInt.data(), 4 * rowInt.size());
    if (!s) s = upload(ctx, sc->d_env_margint, &margInt, 4);
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "param upload");
    return s;
}

// Device tables of the adjoint (AdjView): Smits reflectance bases, the
// materials' RGB eta/k and fromRGB cases, and for the env light's
// distribution the texels and weights behind every cell (the reference writes
// cell x + y * H from the lookup at (x / W, y / H), a later write replacing an
// earlier one) and the derivative of its integral by each texel's green.
wpbr_status build_adjoint_tables(wpbr_ctx* ctx, wpbr_scene* sc) {
    const size_t nm = sc->h_materials.size() / 9;
    std::vector<float> rgb(6 * std::max<size_t>(nm, 1), 0.0f);
    std::vector<int> cases(2 * std::max<size_t>(nm, 1), 0);
    for (size_t m = 0; m < nm; m++) {
        for (int c = 0; c < 6; c++) rgb[6 * m + c] = sc->h_materials[9 * m + 3 + c];
        cases[2 * m] = rgb_case_h(&rgb[6 * m]);
        cases[2 * m + 1] = rgb_case_h(&rgb[6 * m + 3]);
    }
    const uint32_t W = sc->sv.env_w, H = sc->sv.env_h;
    const size_t ncell = (size_t)W * H;
    std::vector<int> src(4 * ncell, -1);
    std::vector<float> srcw(4 * ncell, 0.0f), dI(ncell, 0.0f);
    if (sc->sv.has_env) {
        for (uint32_t y = 0; y < H; y++) {
            const float yN = (float)y / (float)H;
            const float sinTheta = (float)std::sin(3.14 * (double)float(y + 0.5f) / (double)float(H));
            const float t = yN * (float)H - 0.5f;
            int t0 = (int)std::floor(t);
            t0 = std::min(std::max(t0, 0), (int)H - 1);
            const int t0n = std::max(std::min(t0 + 1, (int)H - 1), 0);
            const float dt = t - (float)t0;
            for (uint32_t x = 0; x < W; x++) {
                const float xN = (float)x / (float)W;
                const float s = xN * (float)W - 0.5f;
                int s0 = (int)std::floor(s);
                s0 = std::min(std::max(s0, 0), (int)W - 1);
                const int s0n = std::max(std::min(s0 + 1, (int)W - 1), 0);
                const float ds = s - (float)s0;
                const size_t k = x + (size_t)y * H;
                if (k >= ncell) continue;
                const int xs[4] = {s0, s0, s0n, s0n}, ys[4] = {t0, t0n, t0, t0n};
                const float w[4] = {(1 - ds) * (1 - dt), (1 - ds) * dt, ds * (1 - dt), ds * dt};
                for (int q = 0; q < 4; q++) {
                    src[4 * k + q] = ys[q] * (int)W + xs[q];
                    srcw[4 * k + q] = w[q] * sinTheta;
                }
            }
        }
        for (size_t k = 0; k < ncell; k++)
            for (int q = 0; q < 4; q++)
                if (src[4 * k + q] >= 0) dI[src[4 * k + q]] += srcw[4 * k + q] / (float)ncell;
    }
    auto dev = [&](const void* h, size_t bytes) -> void* {
        void* p = nullptr;
        if (hip_check(ctx, hipMalloc(&p, bytes ? bytes : 4), "hipMalloc adjoint tables")) return nullptr;
        sc->bufs.push_back(p);
        if (bytes && upload(ctx, p, h, bytes)) return nullptr;
        return p;
    };
    AdjView& A = sc->adj;
    A.n_mat = (uint32_t)nm;
    A.refl = (const float*)dev(host_tables().refl, sizeof(host_tables().refl));
    A.mat_rgb = (const float*)dev(rgb.data(), 4 * rgb.size());
    A.mat_case = (const int*)dev(cases.data(), 4 * cases.size());
    A.env_src = (const int*)dev(src.data(), 4 * src.size());
    A.env_srcw = (const float*)dev(srcw.data(), 4 * srcw.size());
    A.env_dI = (const float*)dev(dI.data(), 4 * dI.size());
    if (!A.refl || !A.mat_rgb || !A.mat_case || !A.env_src || !A.env_srcw || !A.env_dI) return WPBR_ERR_OOM;
    wpbr_status s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "adjoint tables");
    if (!s) sc->adj_ready = true;
    return s;
}

}  // namespace
}  // namespace wpbr

using namespace wpbr;

extern "C" {

wpbr_status wpbr_scene_param_counts(wpbr_scene* sc, uint32_t counts[3]) {
    if (!sc || !counts) return WPBR_ERR_INVALID_ARG;
    counts[0] = (uint32_t)(sc->h_materials.size() / 9);
    counts[1] = 6 * counts[0];
    counts[2] = sc->sv.has_env ? 3 * sc->sv.env_w * sc->sv.env_h : 0;
    return WPBR_OK;
}

wpbr_status wpbr_scene_set_params(wpbr_ctx* ctx, wpbr_scene* sc, const float* alpha, const float* etak,
                                  const float* env_rgb) {
    if (!ctx || !sc) return WPBR_ERR_INVALID_ARG;
    if (env_rgb && !sc->sv.has_env) return set_err(ctx, WPBR_ERR_INVALID_ARG, "scene has no env map");
    hipSetDevice(ctx->device);
    const size_t nm = sc->h_materials.size() / 9;
    wpbr_status s = WPBR_OK;
    if (alpha || etak) {
        for (size_t m = 0; m < nm; m++) {
            if (alpha) sc->h_materials[9 * m + 1] = alpha[m];
            if (etak)
                for (int c = 0; c < 6; c++) sc->h_materials[9 * m + 3 + c] = etak[6 * m + c];
        }
        s = push_materials(ctx, sc);
        sc->adj_ready = false;
    }
    if (!s && env_rgb) {
        std::memcpy(sc->h_env.data(), env_rgb, 4 * sc->h_env.size());
        s = push_env(ctx, sc);
        sc->adj_ready = false;
    }
    return s;
}

wpbr_status wpbr_film_loss(wpbr_ctx* ctx, wpbr_scene* sc, const float* d_film, const float* d_target_rgb,
                           float* d_dL_drgb, float* d_rgb, double* loss) {
    if (!ctx || !sc || !d_film) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    const uint32_t npix = sc->sv.width * sc->sv.height;
    const uint32_t nb = (npix + 255) / 256;
    float* partial = nullptr;
    wpbr_status s = WPBR_OK;
    if (loss && d_target_rgb) s = hip_check(ctx, hipMallocAsync((void**)&partial, 4 * (size_t)nb, ctx->stream), "hipMalloc");
    if (!s) s = film_loss_launch((const float4*)d_film, d_target_rgb, npix, d_rgb, d_dL_drgb, partial, ctx->stream);
    if (!s && partial) {
        std::vector<float> h(nb);
        s = hip_check(ctx, hipMemcpyAsync(h.data(), partial, 4 * (size_t)nb, hipMemcpyDeviceToHost, ctx->stream), "d2h");
        if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "film loss");
        double acc = 0;
        for (float v : h) acc += v;
        *loss = acc;
    } else if (!s) {
        s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "film loss");
        if (loss) *loss = 0.0;
    }
    if (partial) hipFreeAsync(partial, ctx->stream);
    return s;
}

wpbr_status wpbr_film_loss_noalloc(wpbr_ctx* ctx, wpbr_scene* sc, const float* d_film, const float* d_target_rgb,
                                   float* d_dL_drgb, float* d_rgb, double* loss) {
    if (!ctx || !sc || !d_film) return WPBR_ERR_INVALID_ARG;
    const uint32_t npix = sc->sv.width * sc->sv.height;
    const uint32_t nb = (npix + 255) / 256;
    if (nb > LOSS_NOALLOC_PARTIALS)
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "wpbr_film_loss_noalloc: films of at most 16777216 pixels");
    if (!ctx->loss_scratch)
        return set_err(ctx, WPBR_ERR_HIP, "wpbr_film_loss_noalloc: scratch symbol lookup failed at wpbr_open");
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    // every partial the sum reads is written by this call's k_film_loss
    float* partial = (loss && d_target_rgb) ? ctx->loss_scratch : nullptr;
    double* total = reinterpret_cast<double*>(ctx->loss_scratch + LOSS_NOALLOC_PARTIALS);
    wpbr_status s = film_loss_launch((const float4*)d_film, d_target_rgb, npix, d_rgb, d_dL_drgb, partial, ctx->stream);
    if (!s && partial) s = film_loss_sum_launch(partial, nb, total, ctx->stream);
    if (s) return set_err(ctx, s, std::string("film loss launch failed: ") + hipGetErrorString(last_launch_error));
    double acc = 0.0;
    if (partial) s = hip_check(ctx, hipMemcpyAsync(&acc, total, 8, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "film loss");
    if (!s && loss) *loss = acc;
    return s;
}

wpbr_status wpbr_bsdf_derivatives(wpbr_ctx* ctx, uint32_t n, const float* d_in, float* d_out) {
    if (!ctx || (n && (!d_in || !d_out))) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    wpbr_status s = bsdf_grad_launch(d_in, d_out, n, ctx->stream);
    if (s) return set_err(ctx, s, std::string("bsdf derivatives launch failed: ") + hipGetErrorString(last_launch_error));
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "bsdf derivatives");
}

wpbr_status wpbr_render_backward(wpbr_ctx* ctx, wpbr_scene* sc, const float* d_film, const float* d_dL_drgb,
                                 float* d_grad) {
    if (!ctx || !sc || !d_film || !d_dL_drgb || !d_grad) return WPBR_ERR_INVALID_ARG;
    if (sc->textured)
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "render backward: textured materials are not differentiable");
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    const uint32_t n = sc->sv.width * sc->sv.height;
    wpbr_status s = ensure_wave(ctx, n);
    if (!s) s = ctx_ensure_trace(ctx, 2 * n);
    if (!s) s = ensure_stream2(ctx);
    if (!s && !sc->adj_ready) s = build_adjoint_tables(ctx, sc);
    if (s) return s;
    wpbr_render_state* r = ctx->render;
    hipStream_t st = ctx->stream;
    const uint32_t nm = sc->adj.n_mat;
    const uint32_t rows = adjoint_blocks(n), ncol = 7 * nm + 1;
    const uint32_t ntex = sc->sv.has_env ? sc->sv.env_w * sc->sv.env_h : 0;
    const size_t ngrad = 7 * (size_t)nm + 3 * (size_t)ntex;
    float* part = nullptr;
    float* dI = nullptr;
    s = hip_check(ctx, hipMallocAsync((void**)&part, 4 * ((size_t)rows * ncol + 1), st), "hipMalloc adjoint");
    if (s) return s;
    dI = part + (size_t)rows * ncol;
    s = hip_check(ctx, hipMemsetAsync(part, 0, 4 * ((size_t)rows * ncol + 1), st), "memset");
    if (!s) s = hip_check(ctx, hipMemsetAsync(d_grad, 0, 4 * ngrad, st), "memset");
    if (!s) s = hip_check(ctx, hipMemsetAsync(r->ws.qcount, 0, 64, st), "queue counters");
    if (!s) s = hip_check(ctx, hipMemsetAsync(r->ws2.qcount, 0, 64, st), "queue counters");
    RenderLaunch L;
    L.sv = sc->sv;
    L.n = n;
    L.spp = sc->spp;
    L.smp = sc->smp;
    L.film = nullptr;
    L.sample_xyz = nullptr;
    L.record = 1;
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
    // Passes alternate between the context's stream and stream2 with their own
    // wavefront state (as in wpbr_render): pass j's camera waits for pass j - 1's
    // shade1 (the sample-stream cursors), its k_adjoint for pass j - 1's (the
    // gradient sums accumulate the passes in order).
    hipStream_t sts[2] = {st, r->stream2};
    const WaveState* wss[2] = {&r->ws, &r->ws2};
    uint32_t* spills[2] = {ctx->spill, r->spill2};
    uint32_t* claims[2] = {ta.claim, r->claim2};
    while (r->evs.size() < 2 * (size_t)sc->spp) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) {
            hipFreeAsync(part, st);
            return set_err(ctx, WPBR_ERR_HIP, "hipEventCreate");
        }
        r->evs.push_back(e);
    }
    AdjArgs aa;
    aa.A = sc->adj;
    aa.film = (const float4*)d_film;
    aa.dL = d_dL_drgb;
    aa.part = part;
    aa.grad_env = d_grad + 7 * (size_t)nm;
    hipEventRecord(ctx->ev0, st);
    hipStreamWaitEvent(sts[1], ctx->ev0, 0);
    for (uint32_t j = 0; j < sc->spp && !s; j++) {
        const int k = (int)(j & 1u);
        hipStream_t sj = sts[k];
        const WaveState& w = *wss[k];
        L.ws = w;
        L.spp_index = j;
        ta.spill = spills[k];
        ta.claim = claims[k];
        // path replay: the forward's camera, primary trace, shade1 and secondary trace
        if (j) hipStreamWaitEvent(sj, r->evs[2 * (size_t)(j - 1)], 0);  // pass j - 1's shade1
        s = render_pass_launch(L, 0, sj);
        ta.n = n;
        ta.n_dev = nullptr;
        ta.stride = 0;
        ta.rays = w.rays;
        ta.hit_shape = w.hit;
        ta.hit_t = w.hit_t;
        ta.bary = w.bary;
        ta.target = nullptr;
        if (!s) s = trace_launch(ta, ctx->trace_block, ctx->trace_stack, sj, dyn_blocks);
        if (!s) s = render_pass_launch(L, 1, sj);
        hipEventRecord(r->evs[2 * (size_t)j], sj);
        ta.n = 2 * n;
        ta.n_dev = w.qcount;
        ta.stride = 2 * n;
        ta.rays = w.qrays;
        ta.hit_shape = w.qhit;
        ta.hit_t = w.qt;
        ta.bary = nullptr;
        ta.target = w.qtarget;
        if (!s) s = trace_launch(ta, ctx->trace_block, ctx->trace_stack, sj, dyn_blocks);
        if (j) hipStreamWaitEvent(sj, r->evs[2 * (size_t)j - 1], 0);  // pass j - 1's k_adjoint
        aa.L = L;
        if (!s) s = adjoint_launch(aa, sj);
        hipEventRecord(r->evs[2 * (size_t)j + 1], sj);
    }
    if (sc->spp % 2 == 0) hipStreamWaitEvent(st, r->evs[2 * (size_t)sc->spp - 1], 0);
    if (!s) s = adjoint_finish(part, rows, nm, d_grad, dI, sc->adj.env_dI, ntex, st);
    hipEventRecord(ctx->ev1, st);
    if (s) {
        hipFreeAsync(part, st);
        return set_err(ctx, s, std::string("backward launch failed: ") + hipGetErrorString(last_launch_error));
    }
    hipFreeAsync(part, st);
    s = ctx_finish_trace(ctx);
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

}  // extern "C"
