// C ABI of libwpbr_hip.so: contexts, device memory, LBVH build and collision.
// Host code only; kernels live in the .hip translation units.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>
#include "wpbr_internal.hpp"
#include "wpbr_ctx.hpp"

using namespace wpbr;

namespace wpbr {

wpbr_status set_err(wpbr_ctx* ctx, wpbr_status s, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return s;
}

wpbr_status hip_check(wpbr_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return WPBR_OK;
    return set_err(ctx, e == hipErrorOutOfMemory ? WPBR_ERR_OOM : WPBR_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

// Scratch of wpbr_film_loss_noalloc: a file-scope device array (one copy per
// GPU the library runs on, held until the process ends) with the per-block
// loss partials of k_film_loss and, after them, the loss they sum to (a double,
// 8-byte aligned).  In host code `scratch` names no device memory: kernels and
// copies use the address hipGetSymbolAddress gives for the current device,
// looked up once per device (at wpbr_open) and kept here.
static __device__ __align__(256) float scratch[LOSS_NOALLOC_PARTIALS + 2];
static_assert(LOSS_NOALLOC_PARTIALS % 2 == 0, "the double after the partials is 8-byte aligned");
static std::mutex g_scratch_mu;
static float* g_scratch[LOSS_NOALLOC_DEVICES] = {};

// The device copy of `scratch` on `device` (which must be current); null if
// the lookup failed.
float* noalloc_scratch(int device) {
    if (device < 0 || device >= LOSS_NOALLOC_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch[device] = static_cast<float*>(p);
    }
    return g_scratch[device];
}

// Bump allocator over one device allocation.
struct Carver {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += count * sizeof(T);
        return p;
    }
};

static size_t lbvh_layout(LbvhScratch& s, char* base, uint32_t n) {
    Carver c{base};
    const size_t N = std::max<uint32_t>(n, 1);
    const size_t tiles = (N + LBVH_SORT_TILE - 1) / LBVH_SORT_TILE;
    // treelets number at most 4096 in the complete order and n in the reference order
    const uint32_t CT = std::max<uint32_t>(4096, (uint32_t)N);
    s.n = n;
    s.cap_t = CT;
    s.cap_r = ref_cap_r(CT);
    s.cap_c = ref_cap_c(CT);
    s.slot_bounds = c.take<float>(6 * N);
    s.centroid = c.take<float>(3 * N);
    s.cb_ord = c.take<uint32_t>(8 + 6 * 2048);  // also the reference order's 2 x 1024 root partials
    s.morton = c.take<uint32_t>(N);
    s.keys_a = c.take<uint32_t>(N);
    s.vals_a = c.take<uint32_t>(N);
    s.keys_b = c.take<uint32_t>(N);
    s.vals_b = c.take<uint32_t>(N);
    s.hist = c.take<uint32_t>(4 * 256);
    s.status_words = (int)(tiles * 256);
    s.status0 = c.take<uint32_t>(tiles * 256);
    s.status1 = c.take<uint32_t>(tiles * 256);
    s.tile_counter = c.take<uint32_t>(8);
    s.err = c.take<uint32_t>(4);
    s.bin_start = c.take<uint32_t>(4096);
    s.presence = c.take<uint32_t>(128);
    s.t_start = c.take<uint32_t>(CT);
    s.t_count = c.take<uint32_t>(CT);
    s.t_slot = c.take<uint32_t>(CT);
    s.t_nodes = c.take<uint32_t>(CT);
    s.t_final = c.take<uint32_t>(CT);
    s.t_box = c.take<float>(6 * (size_t)CT);
    s.r_start = c.take<uint32_t>(2 * N);
    s.r_n = c.take<uint32_t>(2 * N);
    s.r_c0 = c.take<uint32_t>(2 * N);
    s.r_meta = c.take<uint32_t>(2 * N);
    s.r_size = c.take<uint32_t>(2 * N);
    s.r_off = c.take<uint32_t>(2 * N);
    s.r_box = c.take<float>(12 * N);
    s.task_i0 = c.take<uint32_t>(SAH_TASKS);
    s.task_i1 = c.take<uint32_t>(SAH_TASKS);
    s.task_dep = c.take<uint32_t>(SAH_TASKS);
    s.task_rt = c.take<uint32_t>(SAH_TASKS);
    s.g_perm = c.take<uint32_t>(CT);
    s.g_la = c.take<uint32_t>(CT);
    s.g_lb = c.take<uint32_t>(CT);
    s.sub_i0 = c.take<uint32_t>(CT);
    s.sub_i1 = c.take<uint32_t>(CT);
    s.sub_dep = c.take<uint32_t>(CT);
    s.sub_rt = c.take<uint32_t>(CT);
    s.sub_n = c.take<uint32_t>(1);
    s.sub_ub = c.take<uint32_t>(CT);
    s.u_i0 = c.take<uint32_t>(CT);
    s.u_i1 = c.take<uint32_t>(CT);
    s.u_mid = c.take<uint32_t>(CT);
    s.u_dep = c.take<uint32_t>(CT);
    s.u_rt = c.take<uint32_t>(CT);
    s.u_axis = c.take<uint32_t>(CT);
    s.u_box = c.take<float>(6 * (size_t)CT);
    s.upper_count = c.take<uint32_t>(1);
    s.n_nodes = c.take<uint32_t>(1);
    s.n_treelets = c.take<uint32_t>(1);
    s.node_offset = c.take<uint32_t>(2 * N);
    s.node_nshapes = c.take<uint32_t>(2 * N);
    s.node_axis = c.take<uint8_t>(2 * N);
    s.node_bounds = c.take<float>(12 * N);
    s.trav = c.take<float4>(4 * 3 * N);
    s.trav4 = c.take<float4>(8 * 3 * N);
    s.prims = c.take<float4>(3 * N);
    s.big_list = c.take<uint32_t>(N / 64 + 1);
    s.big_count = c.take<uint32_t>(1);
    // reference order
    s.head_blk = c.take<uint32_t>((N + 255) / 256);
    s.bigt = c.take<uint32_t>(N / 30 + 1);
    s.bigt_n = c.take<uint32_t>(1);
    s.big_pre = c.take<uint32_t>(N / 30 + 2);
    s.gp_box = c.take<float>(6 * (size_t)CT);
    s.gp_bucket = c.take<uint8_t>(CT);
    s.pl_a = c.take<uint32_t>(CT);
    s.pl_b = c.take<uint32_t>(CT);
    s.lv_n = c.take<uint32_t>(4);
    s.lv_rec = c.take<uint32_t>(2 * (size_t)s.cap_r * REF_RG_WORDS);
    s.lv_stat = c.take<uint32_t>(2 * (size_t)s.cap_r * REF_ST_WORDS);
    s.ch_range = c.take<uint32_t>(2 * (size_t)s.cap_c);
    s.ch_cnt = c.take<uint32_t>(REF_CH_WORDS * (size_t)s.cap_c);
    s.fin_p = c.take<uint32_t>((size_t)CT + 1);
    s.fin_blk = c.take<uint32_t>((size_t)CT / REF_FIN_BLOCK + 1);
    return c.off + 256;
}

GeomView to_view(const wpbr_geometry* g) {
    GeomView v;
    v.n_spheres = g->n_spheres;
    v.sphere_xform = g->d_sphere_xform;
    v.sphere_radius = g->d_sphere_radius;
    v.n_triangles = g->n_triangles;
    v.tri_verts = g->d_tri_verts;
    return v;
}

// Checks a scene's geometry and sizes the LBVH's buffers for it (stage 0 of a build).
static wpbr_status lbvh_prepare(wpbr_ctx* ctx, const wpbr_geometry* geom, wpbr_lbvh* b) {
    uint64_t n64 = (uint64_t)geom->n_spheres + geom->n_triangles;
    if (n64 == 0 || n64 >= (1ull << 28)) return set_err(ctx, WPBR_ERR_INVALID_ARG, "shape count (1 .. 2^28-1)");
    if ((geom->n_spheres && (!geom->d_sphere_xform || !geom->d_sphere_radius)) ||
        (geom->n_triangles && !geom->d_tri_verts))
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "null geometry pointer");
    uint32_t n = (uint32_t)n64;
    // the scratch arrays, then a device copy of their descriptor (the
    // reference order's kernels read it)
    if (n > b->cap) {
        if (b->mem) { hipFree(b->mem); b->mem = nullptr; }
        LbvhScratch tmp;
        size_t bytes = ((lbvh_layout(tmp, nullptr, n) + 255) & ~size_t(255)) + sizeof(LbvhScratch);
        wpbr_status s = hip_check(ctx, hipMalloc(&b->mem, bytes), "hipMalloc lbvh");
        if (s) { b->cap = 0; return s; }
        b->cap = n;
    }
    const size_t desc_off = (lbvh_layout(b->s, (char*)b->mem, b->cap) + 255) & ~size_t(255);
    b->s.n = n;
    b->s.order = (uint32_t)b->order;
    b->d_self = reinterpret_cast<LbvhScratch*>((char*)b->mem + desc_off);
    wpbr_status su = hip_check(ctx, hipMemcpyAsync(b->d_self, &b->s, sizeof(LbvhScratch), hipMemcpyHostToDevice,
                                                   ctx->stream), "lbvh descriptor");
    if (!su) su = hip_check(ctx, hipStreamSynchronize(ctx->stream), "lbvh descriptor");
    if (su) return su;
    b->n = n;
    b->geom = to_view(geom);
    b->built = false;
    b->next_stage = 0;
    return WPBR_OK;
}

// The host-visible results of a finished build: node and treelet counts, the
// root's traversal word and the device error bits.
static wpbr_status lbvh_finish(wpbr_ctx* ctx, wpbr_lbvh* b, const uint32_t* info) {
    b->n_nodes = info[0];
    b->n_treelets = info[1];
    const uint32_t err = info[2], ns0 = info[3], off0 = info[4];
    b->root_is_leaf = ns0 > 0;
    // the root record, or a primitive range when the root is a leaf of <= 4 shapes
    b->root_word = (ns0 > 0 && ns0 <= 4) ? (0x80000000u | ((ns0 - 1) << 28) | off0) : 0u;
    b->built = err == 0;
    if (err & 1u) return set_err(ctx, WPBR_ERR_TIMEOUT, "radix sort look-back timed out");
    if (err) return set_err(ctx, WPBR_ERR_HIP, "upper SAH build: internal stack exhausted");
    return WPBR_OK;
}

}  // namespace wpbr

extern "C" {

wpbr_status wpbr_open(int device, wpbr_ctx** out) {
    if (!out) return WPBR_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return WPBR_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return WPBR_ERR_HIP;
    wpbr_ctx* c = new wpbr_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return WPBR_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    c->loss_scratch = noalloc_scratch(device);  // null: wpbr_film_loss_noalloc reports it
    {
        wpbr::LevelSync& ls = c->level_sync;
        if (hipHostMalloc((void**)&ls.host, 64, hipHostMallocDefault) != hipSuccess) ls.host = nullptr;
        if (ls.host && (hipEventCreateWithFlags(&ls.ev[0], hipEventDisableTiming) != hipSuccess ||
                        hipEventCreateWithFlags(&ls.ev[1], hipEventDisableTiming) != hipSuccess)) {
            if (ls.ev[0]) (void)hipEventDestroy(ls.ev[0]);
            (void)hipHostFree(ls.host);
            ls = wpbr::LevelSync{nullptr, {nullptr, nullptr}};
        }
    }
    *out = c;
    return WPBR_OK;
}

void wpbr_close(wpbr_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (ctx->spill) hipFree(ctx->spill);
    if (ctx->err_flag) hipFree(ctx->err_flag);
    if (ctx->batch_mem) hipFree(ctx->batch_mem);
    if (ctx->bsdf_blk) hipFree(ctx->bsdf_blk);
    ctx_release_render(ctx);
    if (ctx->level_sync.host) {
        hipEventDestroy(ctx->level_sync.ev[0]);
        hipEventDestroy(ctx->level_sync.ev[1]);
        hipHostFree(ctx->level_sync.host);
    }
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* wpbr_status_string(wpbr_status s) {
    switch (s) {
        case WPBR_OK: return "ok";
        case WPBR_ERR_INVALID_ARG: return "invalid argument";
        case WPBR_ERR_OOM: return "out of device memory";
        case WPBR_ERR_HIP: return "HIP runtime error";
        case WPBR_ERR_STACK_OVERFLOW: return "traversal stack overflow";
        case WPBR_ERR_TIMEOUT: return "device-side wait timed out";
        case WPBR_ERR_NOT_BUILT: return "acceleration structure not built";
    }
    return "unknown status";
}

const char* wpbr_last_error(wpbr_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

wpbr_status wpbr_malloc(wpbr_ctx* ctx, size_t bytes, void** d_ptr) {
    if (!ctx || !d_ptr) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    return hip_check(ctx, hipMalloc(d_ptr, bytes ? bytes : 1), "hipMalloc");
}

wpbr_status wpbr_free(wpbr_ctx* ctx, void* d_ptr) {
    if (!ctx) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    return hip_check(ctx, hipFree(d_ptr), "hipFree");
}

wpbr_status wpbr_upload(wpbr_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
    if (!ctx || (!d_dst && bytes) || (!src && bytes)) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    wpbr_status s = hip_check(ctx, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream),
                              "upload");
    if (s) return s;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "upload sync");
}

wpbr_status wpbr_download(wpbr_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
    if (!ctx || (!d_src && bytes) || (!dst && bytes)) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    wpbr_status s = hip_check(ctx, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream),
                              "download");
    if (s) return s;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "download sync");
}

wpbr_status wpbr_sync(wpbr_ctx* ctx) {
    if (!ctx) return WPBR_ERR_INVALID_ARG;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
}

float wpbr_last_time_ms(wpbr_ctx* ctx) { return ctx ? ctx->last_ms : 0.0f; }

// ---- a1 -------------------------------------------------------------------
wpbr_status wpbr_shape_bounds(wpbr_ctx* ctx, const wpbr_geometry* geom, float* d_bounds,
                              float* d_centroid) {
    if (!ctx || !geom || !d_bounds || !d_centroid) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    uint32_t* cb = nullptr;
    const size_t nshape = (size_t)geom->n_spheres + geom->n_triangles;
    (void)nshape;
    wpbr_status s = hip_check(ctx, hipMalloc(&cb, 4 * (8 + 6 * 512)), "hipMalloc");
    if (s) return s;
    s = shape_bounds_launch(to_view(geom), d_bounds, d_centroid, cb, ctx->stream);
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "shape bounds");
    hipFree(cb);
    return s;
}

// ---- LBVH -----------------------------------------------------------------
wpbr_status wpbr_lbvh_create(wpbr_ctx* ctx, wpbr_lbvh** out) {
    if (!ctx || !out) return WPBR_ERR_INVALID_ARG;
    wpbr_lbvh* b = new wpbr_lbvh();
    b->ctx = ctx;
    b->order = ctx->lbvh_order;
    hipSetDevice(ctx->device);
    for (int i = 0; i < wpbr_lbvh::NEV; i++) hipEventCreate(&b->ev[i]);
    *out = b;
    return WPBR_OK;
}

void wpbr_lbvh_destroy(wpbr_lbvh* b) {
    if (!b) return;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    if (b->mem) hipFree(b->mem);
    for (int i = 0; i < wpbr_lbvh::NEV; i++) hipEventDestroy(b->ev[i]);
    delete b;
}

wpbr_status wpbr_lbvh_set_order(wpbr_lbvh* b, int order) {
    if (!b || (order != WPBR_LBVH_ORDER_REFERENCE && order != WPBR_LBVH_ORDER_FULL)) return WPBR_ERR_INVALID_ARG;
    b->order = order;
    b->built = false;
    return WPBR_OK;
}

int wpbr_lbvh_get_order(const wpbr_lbvh* b) { return b ? b->order : -1; }

wpbr_status wpbr_set_default_lbvh_order(wpbr_ctx* ctx, int order) {
    if (!ctx || (order != WPBR_LBVH_ORDER_REFERENCE && order != WPBR_LBVH_ORDER_FULL)) return WPBR_ERR_INVALID_ARG;
    ctx->lbvh_order = order;
    return WPBR_OK;
}

wpbr_status wpbr_lbvh_build_stage(wpbr_ctx* ctx, const wpbr_geometry* geom, wpbr_lbvh* b, int stage) {
    if (!ctx || !geom || !b) return WPBR_ERR_INVALID_ARG;
    if (stage < 0 || stage >= LBVH_STAGES) return set_err(ctx, WPBR_ERR_INVALID_ARG, "stage (0 .. 5)");
    hipSetDevice(ctx->device);
    (void)hipGetLastError();  // launch checks below must not see stale errors
    if (stage == 0) {
        wpbr_status s = lbvh_prepare(ctx, geom, b);
        if (s) return s;
    } else {
        if (b->next_stage != stage)
            return set_err(ctx, WPBR_ERR_INVALID_ARG, "lbvh stages run in order 0 .. 5");
        GeomView g = to_view(geom);
        if (g.n_spheres != b->geom.n_spheres || g.n_triangles != b->geom.n_triangles ||
            g.sphere_xform != b->geom.sphere_xform || g.tri_verts != b->geom.tri_verts)
            return set_err(ctx, WPBR_ERR_INVALID_ARG, "geometry changed between lbvh stages");
    }
    wpbr_status s = lbvh_launch_stage(b->s, b->d_self, b->geom, b->n, ctx->stream, stage, b->ev, wpbr_lbvh::NEV,
                                      ctx->level_sync.host ? &ctx->level_sync : nullptr);
    if (s) {
        b->next_stage = 0;
        return set_err(ctx, s, "lbvh launch failed");
    }
    b->next_stage = stage + 1;
    if (stage < LBVH_STAGES - 1) return WPBR_OK;
    b->next_stage = 0;
    uint32_t info[5] = {0, 0, 0, 0, 0};
    uint32_t* src[5] = {b->s.n_nodes, b->s.n_treelets, b->s.err, b->s.node_nshapes, b->s.node_offset};
    for (int k = 0; k < 5 && !s; k++)
        s = hip_check(ctx, hipMemcpyAsync(&info[k], src[k], 4, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "lbvh build");
    if (s) return s;
    float ms = 0;
    hipEventElapsedTime(&ms, b->ev[0], b->ev[wpbr_lbvh::NEV - 1]);
    ctx->last_ms = ms;
    return lbvh_finish(ctx, b, info);
}

wpbr_status wpbr_lbvh_build_segmented(wpbr_ctx* ctx, uint32_t n_scenes, const wpbr_geometry* geoms,
                                      wpbr_lbvh* const* bvhs) {
    if (!ctx || !geoms || !bvhs || n_scenes == 0) return WPBR_ERR_INVALID_ARG;
    if (n_scenes > 65535) return set_err(ctx, WPBR_ERR_INVALID_ARG, "at most 65535 scenes per segmented build");
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    for (uint32_t i = 0; i < n_scenes; i++) {
        if (!bvhs[i]) return set_err(ctx, WPBR_ERR_INVALID_ARG, "null lbvh");
        for (uint32_t j = 0; j < i; j++)
            if (bvhs[j] == bvhs[i]) return set_err(ctx, WPBR_ERR_INVALID_ARG, "an lbvh appears twice");
        if (bvhs[i]->order != bvhs[0]->order)
            return set_err(ctx, WPBR_ERR_INVALID_ARG, "the scenes of a segmented build have one key order");
    }
    for (uint32_t i = 0; i < n_scenes; i++) {
        wpbr_status s = lbvh_prepare(ctx, &geoms[i], bvhs[i]);
        if (s) return set_err(ctx, s, "scene " + std::to_string(i) + ": " + ctx->last_error);
    }
    // host image of the batch descriptors: scratch structs, geometry, sizes,
    // then per launch kind the block -> scene map
    std::vector<LbvhScratch> scr(n_scenes);
    std::vector<GeomView> gv(n_scenes);
    std::vector<uint32_t> nv(n_scenes);
    std::vector<uint32_t> first[SEG_KINDS], count[SEG_KINDS];
    std::vector<uint16_t> scene_of[SEG_KINDS];
    uint32_t grid[SEG_KINDS] = {};
    for (uint32_t i = 0; i < n_scenes; i++) {
        scr[i] = bvhs[i]->s;
        gv[i] = bvhs[i]->geom;
        nv[i] = bvhs[i]->n;
        uint32_t c[SEG_KINDS];
        lbvh_seg_blocks(nv[i], scr[i].status_words, c);
        for (int k = 0; k < SEG_KINDS; k++) {
            first[k].push_back(grid[k]);
            count[k].push_back(c[k]);
            scene_of[k].insert(scene_of[k].end(), c[k], (uint16_t)i);
            grid[k] += c[k];
        }
    }
    size_t bytes = 0;
    auto reserve = [&](size_t b) { size_t o = (bytes + 255) & ~size_t(255); bytes = o + b; return o; };
    const size_t o_s = reserve(sizeof(LbvhScratch) * n_scenes), o_g = reserve(sizeof(GeomView) * n_scenes),
                 o_n = reserve(4 * (size_t)n_scenes), o_work = reserve(4 * ((size_t)n_scenes + 1)),
                 o_subwork = reserve(4 * ((size_t)n_scenes + 2)),  // + the range cursor
                 o_info = reserve(20 * (size_t)n_scenes + 4);  // + the level loop's range count
    size_t o_first[SEG_KINDS], o_count[SEG_KINDS], o_scene[SEG_KINDS];
    for (int k = 0; k < SEG_KINDS; k++) {
        o_first[k] = reserve(4 * (size_t)n_scenes);
        o_count[k] = reserve(4 * (size_t)n_scenes);
        o_scene[k] = reserve(2 * (size_t)grid[k]);
    }
    if (bytes > ctx->batch_cap) {
        if (ctx->batch_mem) { hipFree(ctx->batch_mem); ctx->batch_mem = nullptr; ctx->batch_cap = 0; }
        wpbr_status s = hip_check(ctx, hipMalloc(&ctx->batch_mem, bytes), "hipMalloc segmented build");
        if (s) return s;
        ctx->batch_cap = bytes;
    }
    std::vector<char> img(bytes, 0);
    auto put = [&](size_t o, const void* p, size_t len) { if (len) std::memcpy(img.data() + o, p, len); };
    put(o_s, scr.data(), sizeof(LbvhScratch) * n_scenes);
    put(o_g, gv.data(), sizeof(GeomView) * n_scenes);
    put(o_n, nv.data(), 4 * (size_t)n_scenes);
    char* d = (char*)ctx->batch_mem;
    LbvhBatch B;
    B.s = (const LbvhScratch*)(d + o_s);
    B.g = (const GeomView*)(d + o_g);
    B.n = (const uint32_t*)(d + o_n);
    B.n_scenes = n_scenes;
    B.work = (uint32_t*)(d + o_work);
    B.subwork = (uint32_t*)(d + o_subwork);
    B.info = (uint32_t*)(d + o_info);
    for (int k = 0; k < SEG_KINDS; k++) {
        put(o_first[k], first[k].data(), 4 * (size_t)n_scenes);
        put(o_count[k], count[k].data(), 4 * (size_t)n_scenes);
        put(o_scene[k], scene_of[k].data(), 2 * (size_t)grid[k]);
        B.map[k].first = (const uint32_t*)(d + o_first[k]);
        B.map[k].count = (const uint32_t*)(d + o_count[k]);
        B.map[k].scene = (const uint16_t*)(d + o_scene[k]);
    }
    wpbr_status s = hip_check(ctx, hipMemcpyAsync(ctx->batch_mem, img.data(), bytes, hipMemcpyHostToDevice,
                                                  ctx->stream), "segmented build upload");
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "segmented build upload");
    if (s) return s;
    hipEventRecord(ctx->ev0, ctx->stream);
    const uint32_t max_n = *std::max_element(nv.begin(), nv.end());
    s = lbvh_launch_segmented(B, grid, (uint32_t)bvhs[0]->order, max_n, ctx->stream,
                              ctx->level_sync.host ? &ctx->level_sync : nullptr);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("segmented lbvh launch failed: ") + hipGetErrorString(last_launch_error));
    std::vector<uint32_t> info(5 * (size_t)n_scenes);
    s = hip_check(ctx, hipMemcpyAsync(info.data(), B.info, 4 * info.size(), hipMemcpyDeviceToHost, ctx->stream), "d2h");
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "segmented lbvh build");
    if (s) return s;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) ms = 0;
    ctx->last_ms = ms;
    wpbr_status first_err = WPBR_OK;
    for (uint32_t i = 0; i < n_scenes; i++) {
        wpbr_status e = lbvh_finish(ctx, bvhs[i], &info[5 * (size_t)i]);
        if (e && !first_err) first_err = set_err(ctx, e, "scene " + std::to_string(i) + ": " + ctx->last_error);
    }
    return first_err;
}

wpbr_status wpbr_lbvh_build(wpbr_ctx* ctx, const wpbr_geometry* geom, wpbr_lbvh* b) {
    for (int stage = 0; stage < LBVH_STAGES; stage++) {
        wpbr_status s = wpbr_lbvh_build_stage(ctx, geom, b, stage);
        if (s) return s;
    }
    return WPBR_OK;
}

wpbr_status wpbr_lbvh_counts(wpbr_lbvh* b, uint32_t counts[4]) {
    if (!b || !counts) return WPBR_ERR_INVALID_ARG;
    counts[0] = b->n;
    counts[1] = b->n_nodes;
    counts[2] = b->n_treelets ? b->n_treelets - 1 : 0;
    counts[3] = b->n_treelets;
    return WPBR_OK;
}

int wpbr_lbvh_kernel_times(wpbr_lbvh* b, float* ms, int cap) {
    if (!b || !ms) return 0;
    int k = 0;
    for (int i = 1; i < wpbr_lbvh::NEV && k < cap; i++, k++) {
        float t = 0;
        hipEventElapsedTime(&t, b->ev[i - 1], b->ev[i]);
        ms[k] = t;
    }
    return k;
}

wpbr_status wpbr_lbvh_download(wpbr_ctx* ctx, wpbr_lbvh* b, int field, void* dst, size_t bytes) {
    if (!ctx || !b || !dst) return WPBR_ERR_INVALID_ARG;
    if (!b->built) return set_err(ctx, WPBR_ERR_NOT_BUILT, "lbvh not built");
    const void* src = nullptr;
    size_t need = 0;
    const size_t n = b->n, nn = b->n_nodes;
    switch (field) {
        case 0: src = b->s.morton; need = 4 * n; break;
        case 1: src = b->s.keys_a; need = 4 * n; break;
        case 2: src = b->s.vals_a; need = 4 * n; break;
        case 3: src = b->s.node_offset; need = 4 * nn; break;
        case 4: src = b->s.node_nshapes; need = 4 * nn; break;
        case 5: src = b->s.node_axis; need = nn; break;
        case 6: src = b->s.node_bounds; need = 24 * nn; break;
        case 7: src = b->s.cb_ord; need = 24; break;
        case 8: src = b->s.t_start; need = 4 * (size_t)b->n_treelets; break;
        default: return set_err(ctx, WPBR_ERR_INVALID_ARG, "unknown field");
    }
    if (bytes < need) return set_err(ctx, WPBR_ERR_INVALID_ARG, "destination too small");
    wpbr_status s = wpbr_download(ctx, dst, src, need);
    if (s || field != 7) return s;
    // centroid bounds are stored as order-preserving uints
    uint32_t* u = (uint32_t*)dst;
    for (int i = 0; i < 6; i++) {
        uint32_t v = u[i];
        v = (v & 0x80000000u) ? (v & 0x7fffffffu) : ~v;
        u[i] = v;
    }
    return WPBR_OK;
}

// ---- collision --------------------------------------------------------------
static wpbr_status collide_impl(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_lbvh* b,
                                uint32_t n, const float* d_rays, uint32_t* d_hit_shape,
                                float* d_hit_t, float* d_bary, bool wait) {
    if (!ctx || !geom || !b || (n && (!d_rays || !d_hit_shape || !d_hit_t)))
        return WPBR_ERR_INVALID_ARG;
    if (!b->built) return set_err(ctx, WPBR_ERR_NOT_BUILT, "lbvh not built");
    if ((uint64_t)geom->n_spheres + geom->n_triangles != b->n)
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "geometry does not match the lbvh");
    hipSetDevice(ctx->device);
    (void)hipGetLastError();  // launch checks below must not see stale errors
    wpbr_status s = ctx_ensure_trace(ctx, n);
    if (s) return s;
    TraceArgs a;
    a.g = to_view(geom);
    a.trav = (const float4*)b->s.trav;
    a.trav4 = (const float4*)b->s.trav4;
    a.prims = (const float4*)b->s.prims;
    a.root_box = b->s.node_bounds;
    a.root_word = b->root_word;
    a.n = n;
    a.n_dev = nullptr;
    a.stride = 0;
    a.rays = d_rays;
    a.hit_shape = d_hit_shape;
    a.hit_t = d_hit_t;
    a.bary = d_bary;
    a.target = nullptr;
    a.spill = ctx->spill;
    a.spill_depth = ctx->spill_depth;
    a.err = ctx->err_flag;
    const uint32_t dyn_blocks = ctx_trace_shape(ctx, a);
    hipEventRecord(ctx->ev0, ctx->stream);
    s = trace_launch(a, ctx->trace_block, ctx->trace_stack, ctx->stream, dyn_blocks);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, "trace launch failed");
    if (!wait) return WPBR_OK;
    s = ctx_finish_trace(ctx);
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

wpbr_status wpbr_collide(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_lbvh* b, uint32_t n,
                         const float* d_rays, uint32_t* d_hit_shape, float* d_hit_t, float* d_bary) {
    return collide_impl(ctx, geom, b, n, d_rays, d_hit_shape, d_hit_t, d_bary, true);
}

wpbr_status wpbr_collide_async(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_lbvh* b,
                               uint32_t n, const float* d_rays, uint32_t* d_hit_shape,
                               float* d_hit_t, float* d_bary) {
    return collide_impl(ctx, geom, b, n, d_rays, d_hit_shape, d_hit_t, d_bary, false);
}

wpbr_status wpbr_set_trace_config(wpbr_ctx* ctx, const wpbr_trace_config* cfg) {
    if (!ctx || !cfg) return WPBR_ERR_INVALID_ARG;
    const int sd = cfg->stack_depth;
    if ((cfg->block && cfg->block != 64 && cfg->block != 128 && cfg->block != 256 && cfg->block != 512) ||
        (sd && sd != 8 && sd != 16 && sd != 24 && sd != 32 && sd != 48 && sd != 64) ||
        (cfg->mapping && (cfg->mapping < 1 || cfg->mapping > 3)) ||
        (cfg->width && cfg->width != 2 && cfg->width != 4) ||
        (cfg->refill && (cfg->refill < 1 || cfg->refill > 64)) ||
        (cfg->chunk && (cfg->chunk < 64 || cfg->chunk > 65536)))
        return WPBR_ERR_INVALID_ARG;
    if (cfg->chunk) ctx->trace_chunk = cfg->chunk;
    if (cfg->block) ctx->trace_block = cfg->block;
    if (sd) ctx->trace_stack = sd;
    if (cfg->mapping) ctx->trace_mapping = cfg->mapping;
    if (cfg->width) ctx->trace_width = cfg->width;
    if (cfg->refill) ctx->trace_refill = cfg->refill;
    return WPBR_OK;
}

wpbr_status wpbr_get_trace_config(wpbr_ctx* ctx, wpbr_trace_config* cfg) {
    if (!ctx || !cfg) return WPBR_ERR_INVALID_ARG;
    cfg->block = ctx->trace_block;
    cfg->stack_depth = ctx->trace_stack;
    cfg->mapping = ctx->trace_mapping;
    cfg->width = ctx->trace_width;
    cfg->refill = ctx->trace_refill;
    cfg->chunk = ctx->trace_chunk;
    return WPBR_OK;
}

}  // extern "C"

namespace wpbr {

// Persistent grid: 32 waves per CU (8 per SIMD).
static uint32_t dyn_threads(const wpbr_ctx* ctx) { return (uint32_t)ctx->num_cus * 32u * 64u; }

uint32_t ctx_trace_shape(const wpbr_ctx* ctx, TraceArgs& a) {
    a.width = (uint32_t)ctx->trace_width;
    a.parts = ctx->trace_mapping == 2 ? (uint32_t)ctx->num_xcd : 1u;
    a.refill = (uint32_t)ctx->trace_refill;
    a.chunk = (uint32_t)ctx->trace_chunk;
    a.claim = ctx->err_flag + 2;
    return ctx->trace_mapping == 3 ? dyn_threads(ctx) / (uint32_t)ctx->trace_block : 0u;
}

wpbr_status ctx_ensure_trace(wpbr_ctx* ctx, uint32_t n) {
    // Spill capacity so that LDS + spill = 256 entries, the reference's stack:
    // one slot per ray (static kernel) or per persistent lane.
    uint32_t depth = 256 - (uint32_t)ctx->trace_stack;
    size_t need = (size_t)std::max(n, dyn_threads(ctx)) * depth;
    if (!ctx->err_flag) {
        wpbr_status s = hip_check(ctx, hipMalloc(&ctx->err_flag, 16), "hipMalloc");
        if (s) return s;
        s = hip_check(ctx, hipMemsetAsync(ctx->err_flag, 0, 16, ctx->stream), "memset");
        if (s) return s;
    }
    if (need > ctx->spill_words || depth != ctx->spill_depth) {
        if (ctx->spill) hipFree(ctx->spill);
        ctx->spill = nullptr;
        ctx->spill_words = 0;
        wpbr_status s = hip_check(ctx, hipMalloc(&ctx->spill, std::max<size_t>(need, 1) * 4), "hipMalloc spill");
        if (s) return s;
        ctx->spill_words = need;
        ctx->spill_depth = depth;
    }
    return WPBR_OK;
}

wpbr_status ctx_finish_trace(wpbr_ctx* ctx) {
    uint32_t err = 0;
    wpbr_status s = hip_check(ctx, hipMemcpyAsync(&err, ctx->err_flag, 4, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "collide");
    if (s) return s;
    if (err) {
        hipMemsetAsync(ctx->err_flag, 0, 16, ctx->stream);
        hipStreamSynchronize(ctx->stream);
        return set_err(ctx, WPBR_ERR_STACK_OVERFLOW, "traversal stack overflow (256 entries)");
    }
    return WPBR_OK;
}

}  // namespace wpbr

// ---- autotuner -------------------------------------------------------------
namespace wpbr {

std::vector<wpbr_trace_config> trace_candidates(bool small) {
    std::vector<wpbr_trace_config> c;
    const int blocks[2] = {64, 128};
    const int sds[3] = {8, 16, 24};
    for (int width : {4, 2})
        for (int b : blocks)
            for (int sd : sds) {
                if (small && sd != 16) continue;
                c.push_back({b, sd, 1, width, 16, 64});
                for (int rf : {8, 16}) {
                    if (small && rf != 16) continue;
                    c.push_back({b, sd, 3, width, rf, 64});
                }
            }
    return c;
}

wpbr_status autotune(wpbr_ctx* ctx, const std::vector<wpbr_trace_config>& cands,
                     const std::function<wpbr_status(float*)>& run, wpbr_trace_config* best) {
    wpbr_trace_config saved;
    wpbr_get_trace_config(ctx, &saved);
    float best_ms = 3.4e38f;
    wpbr_trace_config best_cfg = saved;
    for (const wpbr_trace_config& cfg : cands) {
        wpbr_status s = wpbr_set_trace_config(ctx, &cfg);
        float ms = 0, t = 0;
        if (!s) s = run(&t);  // warm-up (allocations, caches)
        for (int r = 0; r < 2 && !s; r++) {
            s = run(&t);
            ms += t;
        }
        if (s) {
            wpbr_set_trace_config(ctx, &saved);
            return s;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best_cfg = cfg;
        }
    }
    wpbr_set_trace_config(ctx, &best_cfg);
    if (best) *best = best_cfg;
    return WPBR_OK;
}

}  // namespace wpbr

extern "C" wpbr_status wpbr_autotune_collide(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_lbvh* b,
                                             uint32_t n, const float* d_rays, uint32_t* d_hit_shape,
                                             float* d_hit_t, wpbr_trace_config* best) {
    if (!ctx || !geom || !b) return WPBR_ERR_INVALID_ARG;
    return autotune(ctx, trace_candidates(false), [&](float* ms) {
        wpbr_status s = collide_impl(ctx, geom, b, n, d_rays, d_hit_shape, d_hit_t, nullptr, true);
        *ms = ctx->last_ms;
        return s;
    }, best);
}
