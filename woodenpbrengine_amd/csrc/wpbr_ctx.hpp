// Host-side context and object definitions behind the opaque C handles.
#pragma once
#include <hip/hip_runtime.h>
#include <functional>
#include <string>
#include <vector>
#include "wpbr_internal.hpp"

struct wpbr_render_state;

struct wpbr_ctx {
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    std::string last_error;
    // traversal
    int trace_block = 64;
    int trace_stack = 16;
    int num_xcd = 8;          // MI355X: 8 XCDs, workgroups dispatched round-robin over them
    int trace_mapping = 1;    // 1 linear, 2 one ray part per XCD, 3 persistent lanes + dynamic fetch
    int trace_width = 4;      // 2: binary records, 4: 4-wide records
    int trace_refill = 16;    // persistent kernel: idle lanes per wave that trigger a fetch
    int trace_chunk = 256;    // persistent kernel: rays claimed per atomic
    uint32_t* spill = nullptr;
    size_t spill_words = 0;
    uint32_t spill_depth = 0;
    uint32_t* err_flag = nullptr;
    wpbr_render_state* render = nullptr;
    int lbvh_order = WPBR_LBVH_ORDER_REFERENCE;  // key order of new LBVH objects
    int render_overlap = 1;   // wpbr_render: consecutive passes overlap on two streams
    float* loss_scratch = nullptr;  // this device's copy of wpbr_film_loss_noalloc's scratch
    uint32_t* bsdf_blk = nullptr;   // wpbr_bsdf_queues: 3 counts per 256 requests + 3 totals
    size_t bsdf_blk_words = 0;
    // segmented LBVH builds: per-scene descriptors and block maps
    void* batch_mem = nullptr;
    size_t batch_cap = 0;
    // the reference-order LBVH level loop's overlapped range-count read-back
    // (host null: the loop waits for each count)
    wpbr::LevelSync level_sync{nullptr, {nullptr, nullptr}};
};

struct wpbr_lbvh {
    static constexpr int NEV = 12;
    wpbr_ctx* ctx = nullptr;
    void* mem = nullptr;
    uint32_t cap = 0;
    uint32_t n = 0, n_nodes = 0, n_treelets = 0;
    bool root_is_leaf = false;
    uint32_t root_word = 0;
    bool built = false;
    int next_stage = 0;  // staged build: the stage wpbr_lbvh_build_stage accepts next
    int order = WPBR_LBVH_ORDER_REFERENCE;
    wpbr::LbvhScratch s{};
    wpbr::LbvhScratch* d_self = nullptr;  // device copy of s
    wpbr::GeomView geom{};
    hipEvent_t ev[NEV] = {};
};

namespace wpbr {
wpbr_status set_err(wpbr_ctx* ctx, wpbr_status s, const std::string& msg);
wpbr_status hip_check(wpbr_ctx* ctx, hipError_t e, const char* what);
GeomView to_view(const wpbr_geometry* g);
wpbr_status ctx_ensure_trace(wpbr_ctx* ctx, uint32_t n);
// Fills the launch-shape fields of a TraceArgs from the context's configuration
// and returns the persistent grid size (0: static kernel).
uint32_t ctx_trace_shape(const wpbr_ctx* ctx, TraceArgs& a);
// Autotuner: candidate configurations (small: a reduced set for expensive runs)
// and the timing loop (run() performs one measured call and returns its ms).
std::vector<wpbr_trace_config> trace_candidates(bool small);
wpbr_status autotune(wpbr_ctx* ctx, const std::vector<wpbr_trace_config>& cands,
                     const std::function<wpbr_status(float*)>& run, wpbr_trace_config* best);
wpbr_status ctx_finish_trace(wpbr_ctx* ctx);
void ctx_release_render(wpbr_ctx* ctx);
wpbr_status shape_bounds_launch(const GeomView& g, float* bounds, float* centroid, uint32_t* cb_ord,
                                hipStream_t st);
}  // namespace wpbr
