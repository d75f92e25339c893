// Host side of the batched many-scene render: every stage of a sample pass
// runs as one launch over the pixels (or tiles, or rays) of all scenes, and the
// whole pass loop is captured once in a HIP graph and replayed.
#include <hip/hip_runtime.h>
#include <cstring>
#include <string>
#include <vector>
#include "render_host.hpp"

using namespace wpbr;

struct wpbr_batch {
    wpbr_ctx* ctx = nullptr;
    std::vector<wpbr_scene*> scenes;
    std::vector<float*> films;
    // per scene two wave states (sharing the sample-stream cursors) and two
    // spill areas, inside `mem`: consecutive passes alternate between them and
    // between the context's stream and stream2 (as wpbr_render does)
    std::vector<WaveState> ws[2];
    std::vector<size_t> spill_off[2];  // per scene: first spill word
    hipStream_t stream2 = nullptr;
    std::vector<hipEvent_t> evs;     // per pass: shade1 done, splat done
    uint32_t spp = 0;
    uint64_t pixels = 0;
    void* mem = nullptr;             // wave states + spill
    void* tables = nullptr;          // launch records and block maps
    size_t tables_cap = 0;
    RenderLaunch* d_L[2] = {nullptr, nullptr};
    TraceArgs* d_Tp[2] = {nullptr, nullptr};  // primary rays
    TraceArgs* d_Tq[2] = {nullptr, nullptr};  // secondary-ray queues
    SegMap map_pix{}, map_tile{}, map_tp{}, map_tq{};
    uint32_t grid_pix = 0, grid_tile = 0, grid_tp = 0, grid_tq = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    wpbr_trace_config captured{};    // traversal configuration of the captured graph
    double stats[4] = {0, 0, 0, 0};
};

namespace {

constexpr uint32_t SPILL_DEPTH_MAX = 256 - 8;  // LDS stack of at least 8 entries

void release_graph(wpbr_batch* b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Block maps: per scene `count` blocks, recorded as (scene per block, first
// block per scene, blocks per scene) in a host image at `off`.
struct MapImage {
    std::vector<uint16_t> scene;
    std::vector<uint32_t> first, count;
    uint32_t total = 0;
    void add(uint32_t sc, uint32_t c) {
        first.push_back(total);
        count.push_back(c);
        scene.insert(scene.end(), c, (uint16_t)sc);
        total += c;
    }
};

// Launch records and block maps for the context's traversal configuration,
// uploaded to b->tables.
wpbr_status upload_tables(wpbr_ctx* ctx, wpbr_batch* b) {
    const uint32_t S = (uint32_t)b->scenes.size();
    const uint32_t block = (uint32_t)ctx->trace_block;
    const uint32_t depth = 256 - (uint32_t)ctx->trace_stack;
    std::vector<RenderLaunch> L(2 * S);
    std::vector<TraceArgs> Tp(2 * S), Tq(2 * S);
    MapImage pix, tile, tp, tq;
    for (uint32_t s2 = 0; s2 < 2 * S; s2++) {
        const uint32_t s = s2 % S, k = s2 / S;
        const wpbr_scene* sc = b->scenes[s];
        const uint32_t n = sc->sv.width * sc->sv.height;
        const WaveState& w = b->ws[k][s];
        RenderLaunch& l = L[s2];
        l.sv = sc->sv;
        l.ws = w;
        l.n = n;
        l.spp_index = 0;
        l.spp = sc->spp;
        l.smp = sc->smp;
        l.film = b->films[s];
        l.sample_xyz = nullptr;
        l.record = 0;
        const wpbr_lbvh* bvh = sc->bvh;
        TraceArgs a{};
        a.g = sc->sv.g;
        a.trav = (const float4*)bvh->s.trav;
        a.trav4 = (const float4*)bvh->s.trav4;
        a.width = (uint32_t)ctx->trace_width;
        a.prims = (const float4*)bvh->s.prims;
        a.root_box = bvh->s.node_bounds;
        a.root_word = bvh->root_word;
        a.parts = 1;
        a.spill = (uint32_t*)((char*)b->mem + 4 * b->spill_off[k][s]);
        a.spill_depth = depth;
        a.err = ctx->err_flag;
        Tp[s2] = a;
        Tp[s2].n = n;
        Tp[s2].n_dev = nullptr;
        Tp[s2].stride = 0;
        Tp[s2].rays = w.rays;
        Tp[s2].hit_shape = w.hit;
        Tp[s2].hit_t = w.hit_t;
        Tp[s2].bary = w.bary;
        Tp[s2].target = nullptr;
        Tq[s2] = a;
        Tq[s2].n = 2 * n;
        Tq[s2].n_dev = w.qcount;
        Tq[s2].stride = 2 * n;
        Tq[s2].rays = w.qrays;
        Tq[s2].hit_shape = w.qhit;
        Tq[s2].hit_t = w.qt;
        Tq[s2].bary = nullptr;
        Tq[s2].target = w.qtarget;
        if (k) continue;  // the block maps are the same for both states
        pix.add(s, (n + 255) / 256);
        tile.add(s, n / 1024);
        tp.add(s, (n + block - 1) / block);
        tq.add(s, (2 * n + block - 1) / block);
    }
    size_t bytes = 0;
    auto reserve = [&](size_t len) { size_t o = (bytes + 255) & ~size_t(255); bytes = o + len; return o; };
    const size_t oL = reserve(sizeof(RenderLaunch) * 2 * S), oTp = reserve(sizeof(TraceArgs) * 2 * S),
                 oTq = reserve(sizeof(TraceArgs) * 2 * S);
    MapImage* maps[4] = {&pix, &tile, &tp, &tq};
    size_t om[4][3];
    for (int k = 0; k < 4; k++) {
        om[k][0] = reserve(2 * (size_t)maps[k]->total);
        om[k][1] = reserve(4 * (size_t)S);
        om[k][2] = reserve(4 * (size_t)S);
    }
    if (bytes > b->tables_cap) {
        if (b->tables) hipFree(b->tables);
        b->tables = nullptr;
        b->tables_cap = 0;
        wpbr_status st = hip_check(ctx, hipMalloc(&b->tables, bytes), "hipMalloc batch tables");
        if (st) return st;
        b->tables_cap = bytes;
    }
    std::vector<char> img(bytes, 0);
    auto put = [&](size_t o, const void* p, size_t len) { if (len) std::memcpy(img.data() + o, p, len); };
    put(oL, L.data(), sizeof(RenderLaunch) * 2 * S);
    put(oTp, Tp.data(), sizeof(TraceArgs) * 2 * S);
    put(oTq, Tq.data(), sizeof(TraceArgs) * 2 * S);
    char* d = (char*)b->tables;
    SegMap* dm[4] = {&b->map_pix, &b->map_tile, &b->map_tp, &b->map_tq};
    uint32_t* grids[4] = {&b->grid_pix, &b->grid_tile, &b->grid_tp, &b->grid_tq};
    for (int k = 0; k < 4; k++) {
        put(om[k][0], maps[k]->scene.data(), 2 * (size_t)maps[k]->total);
        put(om[k][1], maps[k]->first.data(), 4 * (size_t)S);
        put(om[k][2], maps[k]->count.data(), 4 * (size_t)S);
        dm[k]->scene = (const uint16_t*)(d + om[k][0]);
        dm[k]->first = (const uint32_t*)(d + om[k][1]);
        dm[k]->count = (const uint32_t*)(d + om[k][2]);
        *grids[k] = maps[k]->total;
    }
    for (int k = 0; k < 2; k++) {
        b->d_L[k] = (RenderLaunch*)(d + oL) + k * S;
        b->d_Tp[k] = (TraceArgs*)(d + oTp) + k * S;
        b->d_Tq[k] = (TraceArgs*)(d + oTq) + k * S;
    }
    wpbr_status st = hip_check(ctx, hipMemcpyAsync(b->tables, img.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
                               "batch tables upload");
    if (!st) st = hip_check(ctx, hipStreamSynchronize(ctx->stream), "batch tables upload");
    return st;
}

// The pass loop: clear, then per sample pass camera, primary trace, shade1,
// secondary trace, radiance + splat.  Pass j runs on stream j % 2 with wave
// states j % 2; its camera waits for pass j - 1's shade1 (the sample-stream
// cursors), its splat for pass j - 1's splat (the films accumulate the passes
// in order).  Captured, the two streams become two branches of the graph.
wpbr_status launch_passes(wpbr_ctx* ctx, wpbr_batch* b, hipStream_t st) {
    const int block = ctx->trace_block, sd = ctx->trace_stack, width = ctx->trace_width;
    bool tables = true;
    for (const wpbr_scene* sc : b->scenes) tables = tables && splat_tables_fit_h(sc->sv);
    hipStream_t sts[2] = {st, b->stream2};
    hipEvent_t* ev = b->evs.data();  // [0]: start; pass j: [1 + 2j] shade1, [2 + 2j] splat
    wpbr_status s = render_batch_launch(b->d_L[0], b->map_pix, b->grid_pix, 3, 0, st);
    if (!s) s = render_batch_launch(b->d_L[1], b->map_pix, b->grid_pix, 3, 0, st);
    hipEventRecord(ev[0], st);
    hipStreamWaitEvent(sts[1], ev[0], 0);
    for (uint32_t j = 0; j < b->spp && !s; j++) {
        const int k = (int)(j & 1u);
        hipStream_t sj = sts[k];
        if (j) hipStreamWaitEvent(sj, ev[2 * j - 1], 0);  // pass j - 1's shade1
        s = render_batch_launch(b->d_L[k], b->map_pix, b->grid_pix, 0, j, sj);
        if (!s) s = trace_launch_seg(b->d_Tp[k], b->map_tp, b->grid_tp, block, sd, width, sj);
        if (!s) s = render_batch_launch(b->d_L[k], b->map_pix, b->grid_pix, 1, j, sj);
        hipEventRecord(ev[1 + 2 * j], sj);
        if (!s) s = trace_launch_seg(b->d_Tq[k], b->map_tq, b->grid_tq, block, sd, width, sj);
        if (j) hipStreamWaitEvent(sj, ev[2 * j], 0);  // pass j - 1's splat
        if (!s) s = render_batch_launch(b->d_L[k], b->map_tile, b->grid_tile, tables ? 4 : 2, j, sj);
        hipEventRecord(ev[2 + 2 * j], sj);
    }
    if (b->spp % 2 == 0) hipStreamWaitEvent(st, ev[2 * b->spp], 0);
    return s;
}

bool same_config(const wpbr_trace_config& a, const wpbr_ctx* ctx) {
    return a.block == ctx->trace_block && a.stack_depth == ctx->trace_stack && a.width == ctx->trace_width;
}

// Captures the pass loop for the context's traversal configuration.
wpbr_status capture(wpbr_ctx* ctx, wpbr_batch* b) {
    release_graph(b);
    wpbr_status s = upload_tables(ctx, b);
    if (s) return s;
    hipStream_t st = ctx->stream;
    s = hip_check(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
    if (s) return s;
    wpbr_status ls = launch_passes(ctx, b, st);
    hipGraph_t g = nullptr;
    s = hip_check(ctx, hipStreamEndCapture(st, &g), "end capture");
    if (ls && !s) { hipGraphDestroy(g); return set_err(ctx, ls, "batch launch capture failed"); }
    if (s) return s;
    b->graph = g;
    s = hip_check(ctx, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
    if (s) { release_graph(b); return s; }
    b->captured.block = ctx->trace_block;
    b->captured.stack_depth = ctx->trace_stack;
    b->captured.width = ctx->trace_width;
    return WPBR_OK;
}

}  // namespace

extern "C" {

wpbr_status wpbr_batch_create(wpbr_ctx* ctx, uint32_t n_scenes, wpbr_scene* const* scenes, float* const* d_films,
                              wpbr_batch** out) {
    if (!ctx || !scenes || !d_films || !out || n_scenes == 0) return WPBR_ERR_INVALID_ARG;
    *out = nullptr;
    if (n_scenes > 65535) return set_err(ctx, WPBR_ERR_INVALID_ARG, "at most 65535 scenes per batch");
    for (uint32_t s = 0; s < n_scenes; s++) {
        if (!scenes[s] || !d_films[s]) return set_err(ctx, WPBR_ERR_INVALID_ARG, "null scene or film");
        if (scenes[s]->spp != scenes[0]->spp)
            return set_err(ctx, WPBR_ERR_INVALID_ARG, "every scene of a batch has the same spp");
        if (!scenes[s]->bvh || !scenes[s]->bvh->built) return set_err(ctx, WPBR_ERR_NOT_BUILT, "scene not built");
    }
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    wpbr_status st = ctx_ensure_trace(ctx, 0);  // the error flag
    if (st) return st;
    wpbr_batch* b = new wpbr_batch();
    b->ctx = ctx;
    b->scenes.assign(scenes, scenes + n_scenes);
    b->films.assign(d_films, d_films + n_scenes);
    b->spp = scenes[0]->spp;
    // one allocation: every scene's wave state, then its spill slots (2n rays)
    std::vector<size_t> wave_off[2];
    size_t bytes = 0;
    for (int k = 0; k < 2; k++) {
        wave_off[k].resize(n_scenes);
        for (uint32_t s = 0; s < n_scenes; s++) {
            const uint32_t n = scenes[s]->sv.width * scenes[s]->sv.height;
            WaveState tmp;
            wave_off[k][s] = bytes;
            bytes += carve_wave(nullptr, n, tmp);
            if (!k) b->pixels += n;
        }
    }
    size_t spill_words = 0;
    for (int k = 0; k < 2; k++) {
        b->spill_off[k].resize(n_scenes);
        for (uint32_t s = 0; s < n_scenes; s++) {
            const uint32_t n = scenes[s]->sv.width * scenes[s]->sv.height;
            b->spill_off[k][s] = bytes / 4 + spill_words;
            spill_words += (size_t)SPILL_DEPTH_MAX * 2 * n;
        }
    }
    bytes += 4 * spill_words;
    st = hip_check(ctx, hipMalloc(&b->mem, bytes), "hipMalloc batch state");
    if (!st) st = hip_check(ctx, hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking), "stream");
    for (uint32_t e = 0; !st && e < 1 + 2 * b->spp; e++) {
        hipEvent_t ev;
        st = hip_check(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        if (!st) b->evs.push_back(ev);
    }
    if (st) { wpbr_batch_destroy(b); return st; }
    for (int k = 0; k < 2; k++) {
        b->ws[k].resize(n_scenes);
        for (uint32_t s = 0; s < n_scenes; s++) {
            carve_wave((char*)b->mem + wave_off[k][s], scenes[s]->sv.width * scenes[s]->sv.height, b->ws[k][s]);
            // one set of sample-stream cursors per scene
            b->ws[k][s].c1 = b->ws[0][s].c1;
            b->ws[k][s].c2 = b->ws[0][s].c2;
        }
    }
    *out = b;
    return WPBR_OK;
}

void wpbr_batch_destroy(wpbr_batch* b) {
    if (!b) return;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    release_graph(b);
    if (b->tables) (void)hipFree(b->tables);
    if (b->mem) (void)hipFree(b->mem);
    for (hipEvent_t e : b->evs) (void)hipEventDestroy(e);
    if (b->stream2) (void)hipStreamDestroy(b->stream2);
    delete b;
}

wpbr_status wpbr_batch_render(wpbr_ctx* ctx, wpbr_batch* b) {
    if (!ctx || !b || b->ctx != ctx) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    wpbr_status s = ctx_ensure_trace(ctx, 0);
    if (s) return s;
    if (!b->exec || !same_config(b->captured, ctx)) {
        s = capture(ctx, b);
        if (s) return s;
    }
    hipStream_t st = ctx->stream;
    hipEventRecord(ctx->ev0, st);
    s = hip_check(ctx, hipGraphLaunch(b->exec, st), "graph launch");
    hipEventRecord(ctx->ev1, st);
    if (s) return s;
    s = ctx_finish_trace(ctx);
    if (s) return s;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) ms = 0;
    ctx->last_ms = ms;
    b->stats[0] = ms;
    b->stats[1] = (double)b->pixels * b->spp;
    b->stats[2] = (double)b->scenes.size();
    b->stats[3] = 5.0 * b->spp + 1.0;
    return WPBR_OK;
}

int wpbr_batch_stats(wpbr_ctx* ctx, wpbr_batch* b, double* out, int cap) {
    if (!ctx || !b || !out) return 0;
    // [4]: secondary rays traced in the last render (sum over the scenes)
    unsigned long long total = 0;
    for (int k = 0; k < 2; k++)
        for (const WaveState& w : b->ws[k]) {
            unsigned long long q = 0;
            if (hipMemcpy(&q, w.qcount + 2, 8, hipMemcpyDeviceToHost) != hipSuccess) return 0;
            total += q;
        }
    double v[5] = {b->stats[0], b->stats[1], b->stats[2], b->stats[3], (double)total};
    int k = cap < 5 ? cap : 5;
    for (int i = 0; i < k; i++) out[i] = v[i];
    return k;
}

wpbr_status wpbr_autotune_batch(wpbr_ctx* ctx, wpbr_batch* b, wpbr_trace_config* best) {
    if (!ctx || !b) return WPBR_ERR_INVALID_ARG;
    std::vector<wpbr_trace_config> cands;
    for (const wpbr_trace_config& c : trace_candidates(true))
        if (c.mapping == 1) cands.push_back(c);  // the batched traversal is the static kernel
    return autotune(ctx, cands, [&](float* ms) {
        wpbr_status s = wpbr_batch_render(ctx, b);
        *ms = ctx->last_ms;
        return s;
    }, best);
}

}  // extern "C"
