// LBVH build for gfx950 (MEngine::buildLBVH, MEngine.cpp:341-350).
//
// Kernel chain (one stream, no host round trip until the final node count):
//   k_init          zero histograms, tile counters, look-back status, treelet bits
//   k_bounds        a1: per-shape bounds + centroid (CSTriangle.cpp:190-216,
//                   CSSphere.h:136-152); a2: centroid bounds by wave reduction and
//                   order-preserving-uint atomics (SLBVH.cpp:164-181)
//   k_morton        a3: 30-bit Morton codes (SLBVH.cpp:186-227) + the 4x256 digit
//                   histograms of the radix sort
//   k_onesweep x4   a4: stable LSD radix sort, 8-bit digits, one pass each,
//                   decoupled look-back, wave-ballot ranking, LDS-staged scatter.
//                   A stable sort of (code, index) has a unique result, equal to
//                   the reference's 5 x 6-bit passes (SLBVH.cpp:245-283)
//   k_treelet_mark  a4: treelet = run of equal top-12 bits (SLBVH.cpp:299-319)
//   k_emit          a5: one workgroup per treelet builds the emitLBVH tree
//                   (SLBVH.cpp:353-406) breadth-first, then subtree sizes and
//                   bounds bottom-up and DFS-preorder offsets top-down
//   k_upper_sah     a6: one workgroup builds the 12-bucket SAH tree over the
//                   treelet roots (SLBVH.cpp:457-563), reproducing the
//                   libstdc++ std::partition order, and lays it out in preorder
//   k_scatter       a6: flatten (SLBVH.cpp:432-455): treelet nodes to their final
//                   preorder slots
//   k_trav_layout   derived traversal layout: per interior node both child boxes
//                   in one 64-byte record; leaf triangles gathered in leaf order
#include "wpbr_device.hpp"
#include "wpbr_internal.hpp"

namespace wpbr {

constexpr int SORT_BLOCK = 256;
constexpr int SORT_ITEMS = 15;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;
constexpr int SORT_WAVES = SORT_BLOCK / 64;
constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_INC = 2u << 30;
constexpr uint32_t FLAG_MASK = 3u << 30;
constexpr uint32_t SPIN_LIMIT = 1u << 20;
constexpr int MAX_PRIMS_IN_LEAF = 30;  // SLBVH.cpp:356

// ---------------------------------------------------------------------------
__global__ void k_init(LbvhScratch s) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t j = i; j < 4 * 256; j += stride) s.hist[j] = 0;
    for (uint32_t j = i; j < (uint32_t)s.status_words; j += stride) s.status0[j] = 0;
    for (uint32_t j = i; j < 128; j += stride) s.presence[j] = 0;
    if (i < 8) s.tile_counter[i] = 0;
    if (i < 3) { s.cb_ord[i] = f2ord(FLT_MAX_); s.cb_ord[3 + i] = f2ord(FLT_LOWEST_); }
    if (i == 0) { *s.err = 0; *s.upper_count = 0; }
}

// ---------------------------------------------------------------------------
WPBR_D void shape_box(const GeomView& g, uint32_t i, bbox& b) {
    if (i < g.n_spheres) {
        affine w = load_affine(g.sphere_xform + 24 * (size_t)i);
        float r = g.sphere_radius[i];
        b = box_empty();
#pragma unroll
        for (int c = 0; c < 8; c++) {
            f3 p = mk3((c & 1) ? r : -r, (c & 2) ? r : -r, (c & 4) ? r : -r);
            b = box_union_pt(b, xf_point(w, p));
        }
    } else {
        const float* t = g.tri_verts + 9 * (size_t)(i - g.n_spheres);
        b = box_empty();
        b = box_union_pt(b, mk3(t[0], t[1], t[2]));
        b = box_union_pt(b, mk3(t[3], t[4], t[5]));
        b = box_union_pt(b, mk3(t[6], t[7], t[8]));
    }
}

__global__ __launch_bounds__(256) void k_bounds(GeomView g, uint32_t n, float* bounds,
                                                float* centroid, uint32_t* cb_ord) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0, 0, 0};
    if (i < n) {
        bbox b;
        shape_box(g, i, b);
        store_box(bounds + 6 * (size_t)i, b);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float c = (float)((double)b.mx[k] * 0.5 + (double)b.mn[k] * 0.5);
            centroid[3 * (size_t)i + k] = c;
            mn[k] = f2ord(c);
            mx[k] = mn[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], off));
            mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], off));
        }
    }
    if (lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (mn[k] != 0xFFFFFFFFu) atomicMin(&cb_ord[k], mn[k]);
            if (mx[k] != 0) atomicMax(&cb_ord[3 + k], mx[k]);
        }
    }
}

// ---------------------------------------------------------------------------
WPBR_D uint32_t left_shift3(uint32_t x) {
    if (x == (1u << 10)) --x;
    x = (x | (x << 16)) & 0b00000011000000000000000011111111u;
    x = (x | (x << 8)) & 0b00000011000000001111000000001111u;
    x = (x | (x << 4)) & 0b00000011000011000011000011000011u;
    x = (x | (x << 2)) & 0b00001001001001001001001001001001u;
    return x;
}

constexpr int MORTON_BLOCK = 256;
constexpr int MORTON_ITEMS = 16;

__global__ __launch_bounds__(MORTON_BLOCK) void k_morton(const float* centroid, uint32_t n,
                                                         const uint32_t* cb_ord, uint32_t* codes,
                                                         uint32_t* hist) {
    __shared__ uint32_t h[4 * 256];
    for (int j = threadIdx.x; j < 4 * 256; j += MORTON_BLOCK) h[j] = 0;
    __syncthreads();
    float mn[3], mx[3];
#pragma unroll
    for (int k = 0; k < 3; k++) { mn[k] = ord2f(cb_ord[k]); mx[k] = ord2f(cb_ord[3 + k]); }
    uint32_t base = blockIdx.x * (MORTON_BLOCK * MORTON_ITEMS);
#pragma unroll 4
    for (int it = 0; it < MORTON_ITEMS; it++) {
        uint32_t i = base + it * MORTON_BLOCK + threadIdx.x;
        if (i >= n) break;
        const float* c = centroid + 3 * (size_t)i;
        float o[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            o[k] = c[k] - mn[k];
            if (mx[k] > mn[k]) o[k] = o[k] / (mx[k] - mn[k]);  // pbrt-v3 Bounds3::Offset
        }
        const float scale = 1024.0f;
        uint32_t code = (left_shift3((uint32_t)(o[2] * scale)) << 2) |
                        (left_shift3((uint32_t)(o[1] * scale)) << 1) |
                        left_shift3((uint32_t)(o[0] * scale));
        codes[i] = code;
        atomicAdd(&h[0 * 256 + (code & 255)], 1u);
        atomicAdd(&h[1 * 256 + ((code >> 8) & 255)], 1u);
        atomicAdd(&h[2 * 256 + ((code >> 16) & 255)], 1u);
        atomicAdd(&h[3 * 256 + ((code >> 24) & 255)], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < 4 * 256; j += MORTON_BLOCK)
        if (h[j]) atomicAdd(&hist[j], h[j]);
}

// ---------------------------------------------------------------------------
// One onesweep pass.  keys_in/vals_in -> keys_out/vals_out on digit `pass`.
// vals_in == nullptr means the value is the key's index (first pass).
__global__ __launch_bounds__(SORT_BLOCK) void k_onesweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, int pass,
    const uint32_t* __restrict__ hist, uint32_t* status, uint32_t* status_next,
    uint32_t* tile_counter, uint32_t* err, uint32_t num_tiles) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_whist[SORT_WAVES][256];
    __shared__ uint32_t s_goff[256];       // global exclusive digit offset
    __shared__ uint32_t s_texcl[256];      // exclusive prefix of previous tiles
    __shared__ uint32_t s_lstart[256];     // digit start inside this tile
    __shared__ uint32_t s_scan[256];
    __shared__ uint32_t s_keys[SORT_TILE];
    __shared__ uint32_t s_vals[SORT_TILE];

    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int shift = 8 * pass;
    if (tid == 0) s_tile = atomicAdd(&tile_counter[pass], 1u);
    for (int w = 0; w < SORT_WAVES; w++) s_whist[w][tid] = 0;
    // Global digit offsets: exclusive scan of this pass's histogram.
    s_scan[tid] = hist[pass * 256 + tid];
    __syncthreads();
    const uint32_t tile = s_tile;
    if (status_next) status_next[(size_t)tile * 256 + tid] = 0;
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = (tid >= off) ? s_scan[tid - off] : 0;
        __syncthreads();
        s_scan[tid] += v;
        __syncthreads();
    }
    s_goff[tid] = s_scan[tid] - hist[pass * 256 + tid];

    // Load and rank (wave-striped: item j of lane l is key wbase + j*64 + l).
    const uint32_t tbase = tile * SORT_TILE;
    const uint32_t wbase = tbase + wave * 64 * SORT_ITEMS;
    uint32_t key[SORT_ITEMS], val[SORT_ITEMS], rank[SORT_ITEMS];
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; j++) {
        uint32_t i = wbase + j * 64 + lane;
        bool valid = i < n;
        key[j] = valid ? keys_in[i] : 0xFFFFFFFFu;
        val[j] = valid ? (vals_in ? vals_in[i] : i) : 0;
    }
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; j++) {
        uint32_t i = wbase + j * 64 + lane;
        bool valid = i < n;
        uint32_t d = (key[j] >> shift) & 255u;
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bm = ballot64(bit);
            peers &= bit ? bm : ~bm;
        }
        uint32_t r = popc64(peers & lanemask_lt());
        uint32_t cnt = popc64(peers);
        uint32_t before = s_whist[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && r == 0) s_whist[wave][d] = before + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[j] = before + r;
    }
    __syncthreads();
    // Per-digit totals for the tile, per-wave exclusive offsets kept in s_whist.
    uint32_t total = 0;
    for (int w = 0; w < SORT_WAVES; w++) {
        uint32_t c = s_whist[w][tid];
        s_whist[w][tid] = total;
        total += c;
    }
    // Publish the aggregate, then look back over earlier tiles.
    uint32_t* my = status + (size_t)tile * 256 + tid;
    if (tile == 0) {
        __hip_atomic_store(my, FLAG_INC | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_texcl[tid] = 0;
    } else {
        __hip_atomic_store(my, FLAG_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t excl = 0;
        int t = (int)tile - 1;
        uint32_t spins = 0;
        while (t >= 0) {
            uint32_t v = __hip_atomic_load(status + (size_t)t * 256 + tid, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            uint32_t flag = v & FLAG_MASK;
            if (flag == 0) {
                if (++spins > SPIN_LIMIT) { atomicOr(err, 1u); break; }
                continue;
            }
            excl += v & ~FLAG_MASK;
            if (flag == FLAG_INC) break;
            t--;
        }
        __hip_atomic_store(my, FLAG_INC | (excl + total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        s_texcl[tid] = excl;
    }
    // Digit starts inside the tile.
    s_scan[tid] = total;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = (tid >= off) ? s_scan[tid - off] : 0;
        __syncthreads();
        s_scan[tid] += v;
        __syncthreads();
    }
    s_lstart[tid] = s_scan[tid] - total;
    __syncthreads();
    // Stage in digit order.
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; j++) {
        uint32_t i = wbase + j * 64 + lane;
        if (i < n) {
            uint32_t d = (key[j] >> shift) & 255u;
            uint32_t p = s_lstart[d] + s_whist[wave][d] + rank[j];
            s_keys[p] = key[j];
            s_vals[p] = val[j];
        }
    }
    __syncthreads();
    uint32_t cnt_tile = (tbase + SORT_TILE <= n) ? SORT_TILE : (n > tbase ? n - tbase : 0);
    for (uint32_t p = tid; p < cnt_tile; p += SORT_BLOCK) {
        uint32_t k = s_keys[p];
        uint32_t d = (k >> shift) & 255u;
        uint32_t dst = s_goff[d] + s_texcl[d] + (p - s_lstart[d]);
        keys_out[dst] = k;
        vals_out[dst] = s_vals[p];
    }
    (void)num_tiles;
}

// ---------------------------------------------------------------------------
__global__ void k_treelet_mark(const uint32_t* __restrict__ sorted, uint32_t n,
                               uint32_t* bin_start, uint32_t* presence) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t top = sorted[i] >> 18;
    if (i == 0 || top != (sorted[i - 1] >> 18)) {
        bin_start[top] = i;
        atomicOr(&presence[top >> 5], 1u << (top & 31));
    }
}

WPBR_D uint32_t presence_rank(const uint32_t* presence, uint32_t bin) {
    uint32_t r = 0;
    for (uint32_t w = 0; w < (bin >> 5); w++) r += __popc(presence[w]);
    uint32_t m = bin & 31;
    if (m) r += __popc(presence[bin >> 5] & ((1u << m) - 1));
    return r;
}

// ---------------------------------------------------------------------------
// Emit: one workgroup per occupied top-12-bit bin (= treelet).
constexpr int EMIT_BLOCK = 256;
constexpr int EMIT_MAX_LEVELS = 64;

__global__ __launch_bounds__(EMIT_BLOCK) void k_emit(LbvhScratch s, const uint32_t* __restrict__ codes,
                                                     const uint32_t* __restrict__ sorted_index,
                                                     const float* __restrict__ bounds, uint32_t n) {
    __shared__ uint32_t s_q;
    __shared__ uint32_t s_lvl[EMIT_MAX_LEVELS + 2];
    __shared__ int s_nlvl;
    __shared__ uint32_t s_tidx, s_start, s_cnt;
    const uint32_t bin = blockIdx.x;
    const int tid = threadIdx.x;
    if (!((s.presence[bin >> 5] >> (bin & 31)) & 1u)) return;
    if (tid == 0) {
        uint32_t t = presence_rank(s.presence, bin);
        uint32_t start = s.bin_start[bin];
        uint32_t end = n;
        for (uint32_t b = bin + 1; b < 4096; b++) {
            if ((s.presence[b >> 5] >> (b & 31)) & 1u) { end = s.bin_start[b]; break; }
        }
        s_tidx = t;
        s_start = start;
        s_cnt = end - start;
        s.t_start[t] = start;
        s.t_count[t] = end - start;
        uint32_t base = 2 * start - t;  // sum of (2 n_j - 1) over earlier treelets
        s.t_slot[t] = base;
        // root task
        s.r_start[base] = start;
        s.r_n[base] = end - start;
        s.r_meta[base] = 17u;  // first bit index = 30 - 12 - 1 (SLBVH.cpp:342)
        s_q = 1;
        s_lvl[0] = 0;
        s_lvl[1] = 1;
        s_nlvl = 1;
    }
    __syncthreads();
    const uint32_t base = s.t_slot[s_tidx];
    // Breadth-first construction.
    for (int L = 0; L < EMIT_MAX_LEVELS; L++) {
        uint32_t lb = s_lvl[L], le = s_lvl[L + 1];
        if (lb == le) break;
        for (uint32_t q = lb + tid; q < le; q += EMIT_BLOCK) {
            uint32_t rs = s.r_start[base + q], rn = s.r_n[base + q];
            int bit = (int)(s.r_meta[base + q] & 0xFF) - (((s.r_meta[base + q] & 0xFF) == 0xFF) ? 256 : 0);
            bool leaf = (bit < 0) || (rn < (uint32_t)MAX_PRIMS_IN_LEAF);
            uint32_t x = 0;
            if (!leaf) {
                x = codes[rs] ^ codes[rs + rn - 1];
                x &= (bit >= 31) ? 0xFFFFFFFFu : ((2u << bit) - 1u);
                if (x == 0) leaf = true;  // all remaining bits equal: recursion reaches bit -1
            }
            if (leaf) {
                s.r_meta[base + q] = 0x80000000u;
            } else {
                int b = 31 - __clz(x);
                uint32_t mask = 1u << b;
                uint32_t ss = rs, e = rs + rn - 1;
                while (ss + 1 != e) {  // SLBVH.cpp:383-391
                    uint32_t mid = (ss + e) / 2;
                    if ((codes[ss] & mask) == (codes[mid] & mask)) ss = mid;
                    else e = mid;
                }
                uint32_t split = e;
                uint32_t q0 = atomicAdd(&s_q, 2u);
                uint32_t cb = (uint32_t)(b - 1) & 0xFFu;  // b-1 in [-1, 16]
                s.r_start[base + q0] = rs;
                s.r_n[base + q0] = split - rs;
                s.r_meta[base + q0] = cb;
                s.r_start[base + q0 + 1] = split;
                s.r_n[base + q0 + 1] = rn - (split - rs);
                s.r_meta[base + q0 + 1] = cb;
                s.r_c0[base + q] = q0;
                s.r_meta[base + q] = (uint32_t)(b % 3) << 8;  // interior: axis = bit % 3
            }
        }
        __syncthreads();
        if (tid == 0) {
            s_lvl[L + 2] = s_q;
            s_nlvl = L + 1;
        }
        __syncthreads();
    }
    const int nlvl = s_nlvl;
    // Bottom-up: bounds and subtree sizes.
    for (int L = nlvl - 1; L >= 0; L--) {
        uint32_t lb = s_lvl[L], le = s_lvl[L + 1];
        for (uint32_t q = lb + tid; q < le; q += EMIT_BLOCK) {
            uint32_t meta = s.r_meta[base + q];
            bbox b;
            uint32_t size;
            if (meta & 0x80000000u) {
                uint32_t rs = s.r_start[base + q], rn = s.r_n[base + q];
                b = box_empty();
                for (uint32_t j = 0; j < rn; j++)
                    b = box_union(b, load_box(bounds + 6 * (size_t)sorted_index[rs + j]));
                size = 1;
            } else {
                uint32_t c0 = s.r_c0[base + q];
                b = box_union(load_box(s.r_box + 6 * (size_t)(base + c0)),
                              load_box(s.r_box + 6 * (size_t)(base + c0 + 1)));
                size = 1 + s.r_size[base + c0] + s.r_size[base + c0 + 1];
            }
            store_box(s.r_box + 6 * (size_t)(base + q), b);
            s.r_size[base + q] = size;
        }
        __syncthreads();
    }
    // Top-down: preorder offsets relative to the treelet root.
    if (tid == 0) s.r_off[base] = 0;
    __syncthreads();
    for (int L = 0; L < nlvl; L++) {
        uint32_t lb = s_lvl[L], le = s_lvl[L + 1];
        for (uint32_t q = lb + tid; q < le; q += EMIT_BLOCK) {
            uint32_t meta = s.r_meta[base + q];
            if (!(meta & 0x80000000u)) {
                uint32_t c0 = s.r_c0[base + q];
                uint32_t off = s.r_off[base + q];
                s.r_off[base + c0] = off + 1;
                s.r_off[base + c0 + 1] = off + 1 + s.r_size[base + c0];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        s.t_nodes[s_tidx] = s.r_size[base];
        bbox b = load_box(s.r_box + 6 * (size_t)base);
        store_box(s.t_box + 6 * (size_t)s_tidx, b);
    }
}

// ---------------------------------------------------------------------------
// Upper SAH over treelet roots: one workgroup, level-synchronous, one wave per
// task.  A task is a range [i0, i1] of the position array `perm` (treelet ids)
// together with the slot in its parent where the result is linked.
constexpr int SAH_BLOCK = 1024;
constexpr int SAH_WAVES = SAH_BLOCK / 64;
constexpr int NB = 12;

WPBR_D float wave_min_f(float v) {
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}
WPBR_D float wave_max_f(float v) {
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}
WPBR_D uint32_t wave_sum_u(uint32_t v) {
    for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
    return v;
}

__global__ __launch_bounds__(SAH_BLOCK) void k_upper_sah(LbvhScratch s, uint32_t n_total_shapes) {
    __shared__ uint32_t s_k;
    __shared__ uint32_t s_task_next;
    __shared__ uint32_t s_level_end;
    __shared__ uint32_t s_bcount[SAH_WAVES][NB];
    __shared__ uint32_t s_bmin[SAH_WAVES][NB][3];
    __shared__ uint32_t s_bmax[SAH_WAVES][NB][3];
    __shared__ uint32_t s_depth_max;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) {
        uint32_t k = 0;
        for (int w = 0; w < 128; w++) k += __popc(s.presence[w]);
        s_k = k;
        s_depth_max = 0;
    }
    __syncthreads();
    const uint32_t k = s_k;
    for (uint32_t i = tid; i < k; i += SAH_BLOCK) s.perm[i] = i;
    // Task arrays in global scratch: tasks of level L are [lvl_begin, lvl_end).
    if (tid == 0) {
        s.task_i0[0] = 0;
        s.task_i1[0] = k - 1;
        s.task_parent[0] = 0xFFFFFFFFu;  // root
        s.task_depth[0] = 0;
        *s.upper_count = 0;
        s_task_next = 1;
    }
    __syncthreads();
    uint32_t lb = 0, le = 1;
    uint32_t num_upper = 0;
    while (lb < le) {
        // Each wave takes tasks lb + wave, lb + wave + SAH_WAVES, ...
        for (uint32_t t = lb + wave; t < le; t += SAH_WAVES) {
            uint32_t i0 = s.task_i0[t], i1 = s.task_i1[t];
            uint32_t parent = s.task_parent[t];
            uint32_t depth = s.task_depth[t];
            uint32_t ref;
            if (i0 == i1) {
                ref = 0x80000000u | s.perm[i0];  // treelet root
            } else {
                // ---- bounds and centroid bounds over [i0, i1]
                float bmn[3] = {FLT_MAX_, FLT_MAX_, FLT_MAX_}, bmx[3] = {FLT_LOWEST_, FLT_LOWEST_, FLT_LOWEST_};
                float cmn[3] = {FLT_MAX_, FLT_MAX_, FLT_MAX_}, cmx[3] = {FLT_LOWEST_, FLT_LOWEST_, FLT_LOWEST_};
                for (uint32_t i = i0 + lane; i <= i1; i += 64) {
                    bbox rb = load_box(s.t_box + 6 * (size_t)s.perm[i]);
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        bmn[c] = smin(bmn[c], rb.mn[c]);
                        bmx[c] = smax(bmx[c], rb.mx[c]);
                        float cc = (float)((double)(rb.mn[c] + rb.mx[c]) * 0.5);
                        cmn[c] = smin(cmn[c], cc);
                        cmx[c] = smax(cmx[c], cc);
                    }
                }
                bbox bounds, cb;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    bounds.mn[c] = wave_min_f(bmn[c]);
                    bounds.mx[c] = wave_max_f(bmx[c]);
                    cb.mn[c] = wave_min_f(cmn[c]);
                    cb.mx[c] = wave_max_f(cmx[c]);
                }
                int dim = box_max_extent(cb);
                float dist = (dim == 0 ? cb.mx[0] - cb.mn[0] : (dim == 1 ? cb.mx[1] - cb.mn[1] : cb.mx[2] - cb.mn[2]));
                float cbmn = dim == 0 ? cb.mn[0] : (dim == 1 ? cb.mn[1] : cb.mn[2]);
                // ---- buckets
                if (lane < NB) {
                    s_bcount[wave][lane] = 0;
                    for (int c = 0; c < 3; c++) {
                        s_bmin[wave][lane][c] = f2ord(FLT_MAX_);
                        s_bmax[wave][lane][c] = f2ord(FLT_LOWEST_);
                    }
                }
                __builtin_amdgcn_wave_barrier();
                for (uint32_t i = i0 + lane; i <= i1; i += 64) {
                    bbox rb = load_box(s.t_box + 6 * (size_t)s.perm[i]);
                    float rmn = dim == 0 ? rb.mn[0] : (dim == 1 ? rb.mn[1] : rb.mn[2]);
                    float rmx = dim == 0 ? rb.mx[0] : (dim == 1 ? rb.mx[1] : rb.mx[2]);
                    float centroid = (float)((double)(rmn + rmx) * 0.5);
                    uint32_t ib = 0;
                    if (dist > 0.0f) {
                        float tt = (centroid - cbmn) / dist;
                        ib = (uint32_t)(uint8_t)(uint32_t)((float)NB * tt);
                        if (ib == NB) ib = NB - 1;
                    }
                    s.bucket[i] = (uint8_t)ib;
                    atomicAdd(&s_bcount[wave][ib], 1u);
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        atomicMin(&s_bmin[wave][ib][c], f2ord(rb.mn[c]));
                        atomicMax(&s_bmax[wave][ib][c], f2ord(rb.mx[c]));
                    }
                }
                __threadfence_block();
                __builtin_amdgcn_wave_barrier();
                // ---- costs (lane i computes cost i), first minimum
                float cost = INF;
                uint32_t cnt_le = 0;
                if (lane < NB - 1) {
                    bbox b0 = box_empty(), b1 = box_empty();
                    uint32_t c0 = 0, c1 = 0;
                    for (int j = 0; j < NB; j++) {
                        bbox bj;
                        for (int c = 0; c < 3; c++) {
                            bj.mn[c] = ord2f(s_bmin[wave][j][c]);
                            bj.mx[c] = ord2f(s_bmax[wave][j][c]);
                        }
                        if (j <= lane) { b0 = box_union(b0, bj); c0 += s_bcount[wave][j]; }
                        else { b1 = box_union(b1, bj); c1 += s_bcount[wave][j]; }
                    }
                    float l = (c0 == 0) ? 0.0f : (float)c0 * box_area(b0);
                    float r = (c1 == 0) ? 0.0f : (float)c1 * box_area(b1);
                    cost = (float)(0.125 + (double)((l + r) / box_area(bounds)));
                    cnt_le = c0;
                }
                int best = 0;
                float cmin = __shfl(cost, 0);
                for (int i = 1; i < NB - 1; i++) {
                    float ci = __shfl(cost, i);
                    if (ci < cmin) { best = i; cmin = ci; }
                }
                uint32_t T = __shfl((int)cnt_le, best);
                uint32_t size = i1 - i0 + 1;
                uint32_t mid;
                if (T == size) {
                    mid = size / 2 + i0;  // no element moves; midpoint fallback (SLBVH.cpp:550-553)
                } else {
                    // libstdc++ two-pointer partition == swap the r-th misplaced element
                    // of the left region [i0, i0+T) with the r-th (from the right)
                    // misplaced element of the right region.
                    uint32_t rl = 0;
                    for (uint32_t c = 0; c * 64 < T; c++) {
                        uint32_t i = i0 + c * 64 + lane;
                        bool bad = (c * 64 + lane < T) && (s.bucket[i] > (uint32_t)best);
                        uint64_t m = ballot64(bad);
                        if (bad) s.list_a[i0 + rl + popc64(m & lanemask_lt())] = i;
                        rl += popc64(m);
                    }
                    uint32_t rr = 0;
                    uint32_t rsz = size - T;
                    for (uint32_t c = 0; c * 64 < rsz; c++) {
                        uint32_t i = i1 - (c * 64 + lane);
                        bool bad = (c * 64 + lane < rsz) && (s.bucket[i] <= (uint32_t)best);
                        uint64_t m = ballot64(bad);
                        if (bad) s.list_b[i0 + rr + popc64(m & lanemask_lt())] = i;
                        rr += popc64(m);
                    }
                    __threadfence_block();
                    __builtin_amdgcn_wave_barrier();
                    for (uint32_t r = lane; r < rl; r += 64) {
                        uint32_t a = s.list_a[i0 + r], b = s.list_b[i0 + r];
                        uint32_t pa = s.perm[a], pb = s.perm[b];
                        s.perm[a] = pb;
                        s.perm[b] = pa;
                    }
                    __threadfence_block();
                    __builtin_amdgcn_wave_barrier();
                    mid = i0 + T;
                }
                // ---- create the node and two child tasks
                uint32_t node = 0;
                if (lane == 0) {
                    node = atomicAdd(s.upper_count, 1u);
                    uint32_t ct = atomicAdd(&s_task_next, 2u);
                    s.u_axis[node] = (uint32_t)dim;
                    s.u_depth[node] = depth;
                    s.task_i0[ct] = i0;
                    s.task_i1[ct] = mid - 1;
                    s.task_parent[ct] = node << 1;
                    s.task_depth[ct] = depth + 1;
                    s.task_i0[ct + 1] = mid;
                    s.task_i1[ct + 1] = i1;
                    s.task_parent[ct + 1] = (node << 1) | 1u;
                    s.task_depth[ct + 1] = depth + 1;
                    atomicMax(&s_depth_max, depth);
                }
                node = (uint32_t)__shfl((int)node, 0);
                ref = node;
            }
            if (lane == 0) {
                if (parent == 0xFFFFFFFFu) *s.upper_root = ref;
                else if (parent & 1u) s.u_c1[parent >> 1] = ref;
                else s.u_c0[parent >> 1] = ref;
            }
        }
        __threadfence_block();
        __syncthreads();
        lb = le;
        le = s_task_next;
        __syncthreads();
    }
    num_upper = *s.upper_count;
    __syncthreads();
    const uint32_t dmax = s_depth_max;
    // Bottom-up: sizes and bounds of upper nodes (children have larger depth).
    for (int d = (int)dmax; d >= 0; d--) {
        for (uint32_t u = tid; u < num_upper; u += SAH_BLOCK) {
            if (s.u_depth[u] != (uint32_t)d) continue;
            uint32_t sz = 1;
            bbox b[2];
            for (int c = 0; c < 2; c++) {
                uint32_t ref = c ? s.u_c1[u] : s.u_c0[u];
                if (ref & 0x80000000u) {
                    uint32_t t = ref & 0x7FFFFFFFu;
                    sz += s.t_nodes[t];
                    b[c] = load_box(s.t_box + 6 * (size_t)t);
                } else {
                    sz += s.u_size[ref];
                    b[c] = load_box(s.u_box + 6 * (size_t)ref);
                }
            }
            s.u_size[u] = sz;
            store_box(s.u_box + 6 * (size_t)u, box_union(b[0], b[1]));
        }
        __threadfence_block();
        __syncthreads();
    }
    // Top-down preorder offsets.
    uint32_t root = *s.upper_root;
    if (tid == 0) {
        if (root & 0x80000000u) s.t_final[root & 0x7FFFFFFFu] = 0;
        else s.u_off[root] = 0;
    }
    __threadfence_block();
    __syncthreads();
    for (uint32_t d = 0; d <= dmax && num_upper > 0; d++) {
        for (uint32_t u = tid; u < num_upper; u += SAH_BLOCK) {
            if (s.u_depth[u] != d) continue;
            uint32_t off = s.u_off[u];
            uint32_t r0 = s.u_c0[u], r1 = s.u_c1[u];
            uint32_t size0 = (r0 & 0x80000000u) ? s.t_nodes[r0 & 0x7FFFFFFFu] : s.u_size[r0];
            uint32_t o0 = off + 1, o1 = off + 1 + size0;
            if (r0 & 0x80000000u) s.t_final[r0 & 0x7FFFFFFFu] = o0; else s.u_off[r0] = o0;
            if (r1 & 0x80000000u) s.t_final[r1 & 0x7FFFFFFFu] = o1; else s.u_off[r1] = o1;
            // Final node record for the upper node.
            s.node_offset[off] = o1;
            s.node_nshapes[off] = 0;
            s.node_axis[off] = (uint8_t)s.u_axis[u];
            bbox b = load_box(s.u_box + 6 * (size_t)u);
            store_box(s.node_bounds + 6 * (size_t)off, b);
        }
        __threadfence_block();
        __syncthreads();
    }
    if (tid == 0) {
        uint32_t total = num_upper;
        for (uint32_t t = 0; t < k; t++) total += s.t_nodes[t];
        *s.n_nodes = total;
        *s.n_treelets = k;
    }
    (void)n_total_shapes;
}

// ---------------------------------------------------------------------------
// Flatten: treelet records to final preorder slots.  One workgroup per treelet.
__global__ __launch_bounds__(256) void k_scatter(LbvhScratch s) {
    const uint32_t t = blockIdx.x;
    if (t >= *s.n_treelets) return;
    const uint32_t base = s.t_slot[t];
    const uint32_t cnt = s.t_nodes[t];
    const uint32_t fbase = s.t_final[t];
    for (uint32_t q = threadIdx.x; q < cnt; q += blockDim.x) {
        uint32_t meta = s.r_meta[base + q];
        uint32_t fi = fbase + s.r_off[base + q];
        bbox b = load_box(s.r_box + 6 * (size_t)(base + q));
        store_box(s.node_bounds + 6 * (size_t)fi, b);
        if (meta & 0x80000000u) {
            s.node_offset[fi] = s.r_start[base + q];
            s.node_nshapes[fi] = s.r_n[base + q];
            s.node_axis[fi] = 0;
        } else {
            uint32_t c0 = s.r_c0[base + q];
            s.node_offset[fi] = fbase + s.r_off[base + c0 + 1];
            s.node_nshapes[fi] = 0;
            s.node_axis[fi] = (uint8_t)((meta >> 8) & 3u);
        }
    }
}

// ---------------------------------------------------------------------------
// Traversal layout.  Record i (64 bytes, one per reference node index):
//   interior: [0..5] child0 box, [6..11] child1 box, [12] child0 word,
//             [13] child1 word, [14] axis, [15] 0
//   leaf:     [12] first primitive slot, [13] primitive count, [15] = 1
// A child word is the child's node index, with bit 31 set when it is a leaf.
__global__ void k_trav_layout(LbvhScratch s) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nn = *s.n_nodes;
    if (i >= nn) return;
    float4* rec = reinterpret_cast<float4*>(s.trav) + 4 * (size_t)i;
    uint32_t ns = s.node_nshapes[i];
    if (ns > 0) {
        rec[3] = make_float4(__uint_as_float(s.node_offset[i]), __uint_as_float(ns), 0.0f,
                             __uint_as_float(1u));
        return;
    }
    uint32_t c0 = i + 1, c1 = s.node_offset[i];
    const float* b0 = s.node_bounds + 6 * (size_t)c0;
    const float* b1 = s.node_bounds + 6 * (size_t)c1;
    uint32_t w0 = c0 | (s.node_nshapes[c0] ? 0x80000000u : 0u);
    uint32_t w1 = c1 | (s.node_nshapes[c1] ? 0x80000000u : 0u);
    rec[0] = make_float4(b0[0], b0[1], b0[2], b0[3]);
    rec[1] = make_float4(b0[4], b0[5], b1[0], b1[1]);
    rec[2] = make_float4(b1[2], b1[3], b1[4], b1[5]);
    rec[3] = make_float4(__uint_as_float(w0), __uint_as_float(w1),
                         __uint_as_float((uint32_t)s.node_axis[i]), 0.0f);
}

// Leaf-ordered primitive records (48 bytes): v0.xyz + shape id, v1.xyz, v2.xyz.
__global__ void k_prim_gather(LbvhScratch s, GeomView g, const uint32_t* __restrict__ sorted_index,
                              uint32_t n) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t id = sorted_index[j];
    float4* rec = reinterpret_cast<float4*>(s.prims) + 3 * (size_t)j;
    if (id < g.n_spheres) {
        rec[0] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(id));
        return;
    }
    const float* t = g.tri_verts + 9 * (size_t)(id - g.n_spheres);
    rec[0] = make_float4(t[0], t[1], t[2], __uint_as_float(id));
    rec[1] = make_float4(t[3], t[4], t[5], 0.0f);
    rec[2] = make_float4(t[6], t[7], t[8], 0.0f);
}

// ---------------------------------------------------------------------------
__global__ void k_cb_init(uint32_t* cb_ord) {
    if (threadIdx.x < 3) {
        cb_ord[threadIdx.x] = f2ord(FLT_MAX_);
        cb_ord[3 + threadIdx.x] = f2ord(FLT_LOWEST_);
    }
}

wpbr_status shape_bounds_launch(const GeomView& g, float* bounds, float* centroid, uint32_t* cb_ord,
                                hipStream_t st) {
    uint32_t n = g.n_spheres + g.n_triangles;
    hipLaunchKernelGGL(k_cb_init, dim3(1), dim3(64), 0, st, cb_ord);
    if (n) hipLaunchKernelGGL(k_bounds, dim3((n + 255) / 256), dim3(256), 0, st, g, n, bounds, centroid, cb_ord);
    return hipGetLastError() == hipSuccess ? WPBR_OK : WPBR_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Host-side launch sequence.
wpbr_status lbvh_launch(LbvhScratch& s, const GeomView& g, uint32_t n, hipStream_t st,
                        hipEvent_t* ev, int nev) {
    auto rec = [&](int i) { if (ev && i < nev) hipEventRecord(ev[i], st); };
    const uint32_t num_tiles = (n + SORT_TILE - 1) / SORT_TILE;
    rec(0);
    hipLaunchKernelGGL(k_init, dim3(std::max<uint32_t>(1, (s.status_words + 255) / 256)), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_bounds, dim3((n + 255) / 256), dim3(256), 0, st, g, n, s.bounds,
                       s.centroid, s.cb_ord);
    rec(1);
    hipLaunchKernelGGL(k_morton, dim3((n + MORTON_BLOCK * MORTON_ITEMS - 1) / (MORTON_BLOCK * MORTON_ITEMS)),
                       dim3(MORTON_BLOCK), 0, st, s.centroid, n, s.cb_ord, s.morton, s.hist);
    rec(2);
    // ping-pong: morton -> B -> A -> B -> A
    const uint32_t* kin[4] = {s.morton, s.keys_b, s.keys_a, s.keys_b};
    const uint32_t* vin[4] = {nullptr, s.vals_b, s.vals_a, s.vals_b};
    uint32_t* kout[4] = {s.keys_b, s.keys_a, s.keys_b, s.keys_a};
    uint32_t* vout[4] = {s.vals_b, s.vals_a, s.vals_b, s.vals_a};
    for (int p = 0; p < 4; p++) {
        uint32_t* st_cur = (p & 1) ? s.status1 : s.status0;
        uint32_t* st_next = (p == 3) ? nullptr : ((p & 1) ? s.status0 : s.status1);
        hipLaunchKernelGGL(k_onesweep, dim3(num_tiles), dim3(SORT_BLOCK), 0, st, kin[p], vin[p],
                           kout[p], vout[p], n, p, s.hist, st_cur, st_next, s.tile_counter, s.err,
                           num_tiles);
        rec(3 + p);
    }
    hipLaunchKernelGGL(k_treelet_mark, dim3((n + 255) / 256), dim3(256), 0, st, s.keys_a, n,
                       s.bin_start, s.presence);
    rec(7);
    hipLaunchKernelGGL(k_emit, dim3(4096), dim3(EMIT_BLOCK), 0, st, s, s.keys_a, s.vals_a,
                       s.bounds, n);
    rec(8);
    hipLaunchKernelGGL(k_upper_sah, dim3(1), dim3(SAH_BLOCK), 0, st, s, n);
    rec(9);
    hipLaunchKernelGGL(k_scatter, dim3(4096), dim3(256), 0, st, s);
    rec(10);
    uint32_t max_nodes = 2 * n;
    hipLaunchKernelGGL(k_trav_layout, dim3((max_nodes + 255) / 256), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_prim_gather, dim3((n + 255) / 256), dim3(256), 0, st, s, g, s.vals_a, n);
    rec(11);
    return hipGetLastError() == hipSuccess ? WPBR_OK : WPBR_ERR_HIP;
}

}  // namespace wpbr
