// Internal types shared by the HIP translation units of libwpbr_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <string>
#include <vector>
#include "../../include/wpbr.h"

namespace wpbr {

// Status of the launches since the last call: the HIP error (if any) is kept in
// last_launch_error so the C ABI can report it.
inline thread_local hipError_t last_launch_error = hipSuccess;
inline wpbr_status launch_status() {
    hipError_t e = hipGetLastError();
    last_launch_error = e;
    return e == hipSuccess ? WPBR_OK : WPBR_ERR_HIP;
}

struct GeomView {
    uint32_t n_spheres;
    const float* sphere_xform;
    const float* sphere_radius;
    uint32_t n_triangles;
    const float* tri_verts;
};

// Device buffers of one LBVH build (all pointers device memory).
// Radix sort tile: 256 threads x LBVH_SORT_ITEMS keys (one look-back status row
// of 256 digits per tile).
#ifndef LBVH_SORT_ITEMS
#define LBVH_SORT_ITEMS 20
#endif
constexpr uint32_t LBVH_SORT_TILE = 256 * LBVH_SORT_ITEMS;

struct LbvhScratch {
    // a1-a3
    float* slot_bounds;   // n x 6 (slot order: bounds of shape vals_a[p] at p)
    float* centroid;      // n x 3
    uint32_t* cb_ord;     // 6: centroid bounds as order-preserving uints
    uint32_t* morton;     // n (shape order)
    // a4 radix sort
    uint32_t* keys_a; uint32_t* vals_a;  // final sorted (code, shape index)
    uint32_t* keys_b; uint32_t* vals_b;
    uint32_t* hist;       // 4 x 256
    uint32_t* status0; uint32_t* status1;
    int status_words;
    uint32_t* tile_counter;  // 8
    uint32_t* err;
    // treelets (<= 4096)
    uint32_t* bin_start;  // 4096
    uint32_t* presence;   // 128 words
    uint32_t* t_start; uint32_t* t_count; uint32_t* t_slot; uint32_t* t_nodes;
    uint32_t* t_final; float* t_box;
    // emit records, indexed by the treelet's reserved slot range (2n)
    uint32_t* r_start; uint32_t* r_n; uint32_t* r_c0; uint32_t* r_meta;
    uint32_t* r_size; uint32_t* r_off; float* r_box;
    // upper SAH: tasks (ranges of final positions) and nodes
    uint32_t* task_i0; uint32_t* task_i1; uint32_t* task_dep; uint32_t* task_rt;
    uint32_t* u_i0; uint32_t* u_i1; uint32_t* u_mid; uint32_t* u_dep; uint32_t* u_rt;
    uint32_t* u_axis; float* u_box;
    uint32_t* upper_count;
    uint32_t* n_nodes; uint32_t* n_treelets;
    // SAH top in three kernels: final positions' treelets, leaf depth and right
    // turns, and the deferred ranges (i0, i1, depth, right turns)
    uint32_t* g_perm; uint32_t* g_la; uint32_t* g_lb;
    uint32_t* sub_i0; uint32_t* sub_i1; uint32_t* sub_dep; uint32_t* sub_rt; uint32_t* sub_n;
    uint32_t* sub_ub;     // first upper-node index of each deferred range
    // final CSceneTreeLBVH arrays (capacity 2n)
    uint32_t* node_offset; uint32_t* node_nshapes; uint8_t* node_axis; float* node_bounds;
    // traversal layout
    void* trav;           // 64 B per binary record
    void* trav4;          // 128 B per record: the record's grandchildren (4-wide)
    void* prims;          // 48 B per shape, at its slot
    uint32_t* big_list;   // reference leaves of more than 64 shapes
    uint32_t* big_count;
    // ---- key order and sizes
    uint32_t order;       // WPBR_LBVH_ORDER_REFERENCE or WPBR_LBVH_ORDER_FULL
    uint32_t n;           // shapes of this build
    uint32_t cap_t;       // capacity of the treelet, position, deferred-range and upper-node arrays
    uint32_t cap_r;       // level ranges per list
    uint32_t cap_c;       // level chunks per list
    // ---- reference order: treelets as a list (up to one per shape)
    uint32_t* head_blk;   // per 256 slots: treelet heads, then their exclusive prefix
    uint32_t* bigt;       // treelets of at least 30 shapes (one workgroup each)
    uint32_t* bigt_n;
    uint32_t* big_pre;    // bigt_n + 1: exclusive prefix of the big treelets' node counts (flatten)
    // level-synchronous upper SAH over the ranges of at least REF_CUT positions
    float* gp_box;        // cap_t x 6: root box of the treelet at each position
    uint8_t* gp_bucket;   // cap_t: SAH bucket of each position
    uint32_t* pl_a;       // cap_t: misplaced positions left of the split, by rank
    uint32_t* pl_b;       // cap_t: misplaced positions right of the split, by rank from the right
    uint32_t* lv_n;       // [k]: ranges of list k, [2 + k]: chunks of list k (k = level parity)
    uint32_t* lv_rec;     // 2 x cap_r x REF_RG_WORDS range records
    uint32_t* lv_stat;    // 2 x cap_r x REF_ST_WORDS bucket statistics
    uint32_t* ch_range;   // 2 x cap_c: range of each chunk
    uint32_t* ch_cnt;     // cap_c x REF_CH_WORDS: bucket counts of each chunk
    uint32_t* fin_p;      // cap_t + 1: prefix of treelet node counts over final positions
    uint32_t* fin_blk;    // per REF_FIN_BLOCK positions: node sums, then their exclusive prefix
};

// Reference key order: ranges of at least REF_CUT treelet roots are split by
// level-synchronous kernels over REF_CHUNK-position chunks; smaller ones by one
// workgroup each in LDS.
constexpr uint32_t REF_CUT = 256;
constexpr uint32_t REF_CHUNK = 2048;
constexpr uint32_t REF_FIN_BLOCK = 1024;
constexpr uint32_t REF_RG_WORDS = 20;
constexpr uint32_t REF_ST_WORDS = 164;
constexpr uint32_t REF_CH_WORDS = 12;  // per chunk: its positions' count in each SAH bucket
inline uint32_t ref_cap_r(uint32_t cap_t) { return cap_t / REF_CUT + 2; }
inline uint32_t ref_cap_c(uint32_t cap_t) { return cap_t / REF_CHUNK + ref_cap_r(cap_t) + 1; }

// Radix-sort passes of a build: 4 (8-bit digits over bits 0..31, the complete
// order) or 3 (bits 0..23: the order the reference's sort job leaves in
// shapeMortonCodes, SLBVH.cpp:245-291).
__host__ __device__ inline int sort_passes(const LbvhScratch& s) {
    return s.order == WPBR_LBVH_ORDER_REFERENCE ? 3 : 4;
}

// Buffers of radix-sort pass p: each pass reads what the previous one wrote,
// the first reads the Morton codes (no values: the value is the key's index)
// and the last writes keys_a / vals_a, so the outputs alternate ... B, A, B, A
// (4 passes: B A B A; 3 passes: A B A); look-back status alternates 0, 1, 0, ...
// and each pass clears the next pass's status.
struct SortPass {
    const uint32_t* kin; const uint32_t* vin; uint32_t* kout; uint32_t* vout;
    uint32_t* status; uint32_t* status_next;
};
__host__ __device__ inline SortPass sort_pass(const LbvhScratch& s, int p) {
    const int np = sort_passes(s);
    const bool in_a = p > 0 && ((np - p) & 1) == 0;   // pass p - 1 wrote A
    const bool out_a = ((np - 1 - p) & 1) == 0;
    SortPass r;
    r.kin = p == 0 ? s.morton : (in_a ? s.keys_a : s.keys_b);
    r.vin = p == 0 ? nullptr : (in_a ? s.vals_a : s.vals_b);
    r.kout = out_a ? s.keys_a : s.keys_b;
    r.vout = out_a ? s.vals_a : s.vals_b;
    r.status = (p & 1) ? s.status1 : s.status0;
    r.status_next = p == np - 1 ? nullptr : ((p & 1) ? s.status0 : s.status1);
    return r;
}

// Segmented build of several scenes (one launch per kernel for all of them).
// Block b of a launch of kind k works for scene map[k].scene[b].
enum SegKind { SEG_BOUNDS, SEG_MORTON, SEG_SORT, SEG_MARK, SEG_SUBTREE, SEG_SUBTREE_LEAF, SEG_SUBTREE_BIG, SEG_BVH4,
               SEG_KINDS };
struct SegMap {
    const uint16_t* scene;  // per block
    const uint32_t* first;  // per scene: its first block
    const uint32_t* count;  // per scene: its number of blocks
};
struct LbvhBatch {
    const LbvhScratch* s;   // per scene (device array)
    const GeomView* g;
    const uint32_t* n;
    uint32_t n_scenes;
    uint32_t* work;         // n_scenes + 1: treelets of the scenes before each scene
    uint32_t* subwork;      // n_scenes + 2: deferred SAH ranges of the scenes before each scene, then a cursor
    uint32_t* info;         // 5 per scene: n_nodes, n_treelets, err, nshapes[0], offset[0]
    SegMap map[SEG_KINDS];
};
void lbvh_seg_blocks(uint32_t n, int status_words, uint32_t* counts);
// The reference order's level loop reads back the next level's range count
// every few levels.  With a LevelSync (two words of pinned host memory and two
// events) the read-back of one block of levels overlaps the next block's
// kernels; without one the host waits for each count.
struct LevelSync {
    uint32_t* host;
    hipEvent_t ev[2];
};
// order: the key order of every scene of the batch; max_n: the largest scene.
wpbr_status lbvh_launch_segmented(const LbvhBatch& B, const uint32_t* grid, uint32_t order, uint32_t max_n,
                                  hipStream_t st, const LevelSync* ls = nullptr);

// wpbr_film_loss_noalloc: the file-scope scratch holds this many per-block loss
// partials (256 pixels each: films of up to 16 777 216 pixels), then their sum
// as a double; one copy per GPU, for up to this many devices.
constexpr uint32_t LOSS_NOALLOC_PARTIALS = 65536;
constexpr int LOSS_NOALLOC_DEVICES = 64;
float* noalloc_scratch(int device);

constexpr int LBVH_STAGES = 6;
// SAH top: task list of the one-workgroup top kernel; workgroups of the
// sub-range kernel
constexpr uint32_t SAH_TASKS = 8192;
constexpr uint32_t SAH_SUB_BLOCKS = 32;       // single scene
constexpr uint32_t SAH_SEG_SUB_BLOCKS = 256;  // segmented build, over all scenes' ranges
// d_self: a device copy of s (the reference order's kernels read the scratch
// through it).
wpbr_status lbvh_launch_stage(LbvhScratch& s, const LbvhScratch* d_self, const GeomView& g, uint32_t n,
                              hipStream_t st, int stage, hipEvent_t* ev, int nev, const LevelSync* ls = nullptr);

struct TraceArgs {
    GeomView g;
    const float4* trav;
    const float4* trav4;
    uint32_t width;       // 2: binary records, 4: 4-wide records
    const float4* prims;
    const float* root_box;
    uint32_t root_word;   // first traversal word: record 0 or a primitive range
    uint32_t n;           // ray count (upper bound when n_dev is set)
    const uint32_t* n_dev;  // optional device-side ray count (queue launches)
    uint32_t stride;      // SoA stride of rays (0: n)
    uint32_t parts;       // 1: linear; P: part x of the ray range to workgroups b with b % P == x
    uint32_t* claim;      // persistent kernel: [0] rays claimed, [1] waves finished (zero between launches)
    const uint32_t* target;  // visibility queries: per ray, the shape whose visibility is asked
                             // (0xFFFFFFFF: none); the result is the target, a miss or an occluder
    uint32_t refill;      // persistent kernel: idle lanes that trigger a fetch (1..64)
    uint32_t chunk;       // persistent kernel: rays claimed per atomic
    const float* rays;
    uint32_t* hit_shape;
    float* hit_t;
    float* bary;
    uint32_t* spill;      // per-thread overflow stack
    uint32_t spill_depth;
    uint32_t* err;        // bit 0: stack overflow
};

wpbr_status trace_launch(const TraceArgs& a, int block, int stack_depth, hipStream_t st, uint32_t dyn_blocks);
// Batched traversal: T holds one record per scene (device), map assigns blocks to scenes.
wpbr_status trace_launch_seg(const TraceArgs* T, const SegMap& map, uint32_t grid, int block, int stack_depth,
                             int width, hipStream_t st);

}  // namespace wpbr
