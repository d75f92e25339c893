// LBVH build for gfx950 (MEngine::buildLBVH, MEngine.cpp:341-350).
//
// Kernel chain (one stream, no host round trip until the final node count):
//   k_init_bounds   zero histograms, tile counters, look-back status, treelet bits;
//                   a1: per-shape bounds + centroid (CSTriangle.cpp:116-142,
//                   CSSphere.h:136-152); a2: centroid bounds by wave reduction and
//                   order-preserving-uint atomics (SLBVH.cpp:164-181)
//   k_morton        a3: 30-bit Morton codes (SLBVH.cpp:186-227) + the 4x256 digit
//                   histograms of the radix sort
//   k_onesweep x4   a4: stable LSD radix sort, 8-bit digits, one pass each,
//                   decoupled look-back, wave-ballot ranking, LDS-staged scatter.
//                   This chain is the full key order (all 30 bits), a documented
//                   alternative.  The reference's 5 x 6-bit ping-pong
//                   (SLBVH.cpp:245-283) leaves pass 3's output in
//                   shapeMortonCodes, a stable sort on bits 0..23 only: the
//                   reference key order below (three passes) reproduces that
//   k_slot_prims    a4: treelet = run of equal top-12 bits (SLBVH.cpp:299-319);
//                   shape bounds gathered into slot order (SLBVH.cpp:231-324),
//                   with the traversal layout's primitive records
//   k_emit          a5: one workgroup per treelet builds the emitLBVH tree
//                   (SLBVH.cpp:353-406) breadth-first, then subtree sizes and
//                   bounds bottom-up and DFS-preorder offsets top-down
//   k_sah_top/sub/final  a6: the 12-bucket SAH tree over the treelet roots
//                   (SLBVH.cpp:457-563), reproducing the libstdc++
//                   std::partition order: one workgroup splits the large
//                   ranges, many workgroups finish the small ones, one lays the
//                   tree out in preorder
//   k_scatter       a6: flatten (SLBVH.cpp:432-455): treelet nodes to their final
//                   preorder slots
//   k_subtree(_big) derived traversal layout: per interior node both child boxes
//                   in one 64-byte record; leaves of more than 4 shapes get a
//                   sub-tree over their (Morton-ordered) slots
//   k_bvh4          4-wide records (a binary record's grandchildren)
#include "wpbr_device.hpp"
#include "wpbr_internal.hpp"

namespace wpbr {

constexpr int SORT_BLOCK = 256;
constexpr int SORT_ITEMS = LBVH_SORT_ITEMS;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;
static_assert(SORT_TILE == LBVH_SORT_TILE, "sort tile size");
constexpr int SORT_WAVES = SORT_BLOCK / 64;
constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_INC = 2u << 30;
constexpr uint32_t FLAG_MASK = 3u << 30;
constexpr uint32_t SPIN_LIMIT = 1u << 20;
constexpr int MAX_PRIMS_IN_LEAF = 30;  // SLBVH.cpp:356

// ---------------------------------------------------------------------------
WPBR_D void init_body(const LbvhScratch& s, uint32_t bid, uint32_t nblk) {
    uint32_t i = bid * blockDim.x + threadIdx.x;
    uint32_t stride = nblk * blockDim.x;
    for (uint32_t j = i; j < 4 * 256; j += stride) s.hist[j] = 0;
    for (uint32_t j = i; j < (uint32_t)s.status_words; j += stride) s.status0[j] = 0;
    for (uint32_t j = i; j < 128; j += stride) s.presence[j] = 0;
    if (i < 8) s.tile_counter[i] = 0;
    if (i == 0) {
        *s.err = 0; *s.upper_count = 0; *s.big_count = 0; *s.sub_n = 0;
    }
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

// Centroid bounds in two stages without atomics or device-wide fences: each
// k_bounds block (a grid-stride loop over a fixed number of blocks) writes its
// partial min/max (order-preserving uints) to cb_ord[8 + 6*b]; the next kernel
// reduces the partials (a kernel boundary orders the writes).  cb_ord[0..5]
// receive the final bounds, cb_ord[7] the number of partials.
constexpr int BOUNDS_BLOCKS = 512;

WPBR_D void bounds_body(const GeomView& g, uint32_t n, float* bounds, float* centroid, uint32_t* cb_ord,
                       uint32_t bid, uint32_t nblk) {
    __shared__ uint32_t s_red[4][6];
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, 0};  // min xyz, max xyz
    for (uint32_t i = bid * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
        bbox b;
        shape_box(g, i, b);
        if (bounds) store_box(bounds + 6 * (size_t)i, b);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float c = (float)((double)b.mx[k] * 0.5 + (double)b.mn[k] * 0.5);
            centroid[3 * (size_t)i + k] = c;
            uint32_t o = f2ord(c);
            v[k] = min(v[k], o);
            v[3 + k] = max(v[3 + k], o);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            v[k] = min(v[k], (uint32_t)__shfl_xor((int)v[k], off));
            v[3 + k] = max(v[3 + k], (uint32_t)__shfl_xor((int)v[3 + k], off));
        }
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
        for (int k = 0; k < 6; k++) s_red[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        uint32_t r = s_red[0][k];
        for (int ww = 1; ww < 4; ww++) r = (k < 3) ? min(r, s_red[ww][k]) : max(r, s_red[ww][k]);
        cb_ord[8 + 6 * (size_t)bid + k] = r;
        if (bid == 0 && k == 0) cb_ord[7] = nblk;
    }
}
__global__ __launch_bounds__(256) void k_bounds(GeomView g, uint32_t n, float* bounds, float* centroid,
                                                uint32_t* cb_ord) {
    bounds_body(g, n, bounds, centroid, cb_ord, blockIdx.x, gridDim.x);
}

// The build's first kernel: zero the build's counters, histograms, look-back
// status and treelet bits (grid-stride), then bounds and centroid partials.
__global__ __launch_bounds__(256) void k_init_bounds(LbvhScratch s, GeomView g, uint32_t n) {
    init_body(s, blockIdx.x, gridDim.x);
    bounds_body(g, n, nullptr, s.centroid, s.cb_ord, blockIdx.x, gridDim.x);
}

// Reduce the k_bounds partials (called by every thread of a consumer block;
// blockDim.x must be a multiple of 64 and at most 1024).
WPBR_D void reduce_cb(const uint32_t* cb_ord, uint32_t* s_cb /* LDS[6] */, uint32_t* out_or_null) {
    __shared__ uint32_t s_w[16][6];
    const uint32_t nb = cb_ord[7];
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
        const uint32_t* p = cb_ord + 8 + 6 * (size_t)b;
        for (int k = 0; k < 3; k++) { v[k] = min(v[k], p[k]); v[3 + k] = max(v[3 + k], p[3 + k]); }
    }
    for (int k = 0; k < 3; k++) {
        for (int off = 32; off >= 1; off >>= 1) {
            v[k] = min(v[k], (uint32_t)__shfl_xor((int)v[k], off));
            v[3 + k] = max(v[3 + k], (uint32_t)__shfl_xor((int)v[3 + k], off));
        }
    }
    if (lane_id() == 0)
        for (int k = 0; k < 6; k++) s_w[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        uint32_t r = s_w[0][k];
        for (uint32_t w = 1; w < blockDim.x / 64; w++) r = (k < 3) ? min(r, s_w[w][k]) : max(r, s_w[w][k]);
        s_cb[k] = r;
        if (out_or_null) out_or_null[k] = r;
    }
    __syncthreads();
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

WPBR_D void morton_body(const float* centroid, uint32_t n, const uint32_t* cb_ord, uint32_t* codes,
                       uint32_t* hist, uint32_t bid) {
    __shared__ uint32_t h[4 * 256];
    __shared__ uint32_t s_cb[6];
    for (int j = threadIdx.x; j < 4 * 256; j += MORTON_BLOCK) h[j] = 0;
    reduce_cb(cb_ord, s_cb, bid == 0 ? const_cast<uint32_t*>(cb_ord) : nullptr);
    float mn[3], mx[3];
#pragma unroll
    for (int k = 0; k < 3; k++) { mn[k] = ord2f(s_cb[k]); mx[k] = ord2f(s_cb[3 + k]); }
    uint32_t base = bid * (MORTON_BLOCK * MORTON_ITEMS);
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
__global__ __launch_bounds__(MORTON_BLOCK) void k_morton(const float* centroid, uint32_t n,
                                                         const uint32_t* cb_ord, uint32_t* codes,
                                                         uint32_t* hist) {
    morton_body(centroid, n, cb_ord, codes, hist, blockIdx.x);
}

// ---------------------------------------------------------------------------
// One onesweep pass.  keys_in/vals_in -> keys_out/vals_out on digit `pass`.
// vals_in == nullptr means the value is the key's index (first pass).
WPBR_D void onesweep_body(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, int pass,
    const uint32_t* __restrict__ hist, uint32_t* status, uint32_t* status_next,
    uint32_t* tile_counter, uint32_t* err) {
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
}
__global__ __launch_bounds__(SORT_BLOCK) void k_onesweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, int pass,
    const uint32_t* __restrict__ hist, uint32_t* status, uint32_t* status_next,
    uint32_t* tile_counter, uint32_t* err) {
    onesweep_body(keys_in, vals_in, keys_out, vals_out, n, pass, hist, status, status_next, tile_counter, err);
}

// ---------------------------------------------------------------------------
WPBR_D void treelet_mark_body(const uint32_t* __restrict__ sorted, uint32_t n, uint32_t* bin_start,
                             uint32_t* presence, uint32_t bid) {
    uint32_t i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t top = sorted[i] >> 18;
    if (i == 0 || top != (sorted[i - 1] >> 18)) {
        bin_start[top] = i;
        atomicOr(&presence[top >> 5], 1u << (top & 31));
    }
}

// Treelet starts (the first slot of each run of equal top-12 bits), and shape
// bounds in slot (sorted) order, as the reference's sort job gathers them with
// the entities (SLBVH.cpp:231-324), so that the treelet and leaf
// kernels read contiguous boxes; and the primitive records of the traversal
// layout at their slots: v0.xyz + shape id, v1.xyz + slot, v2.xyz + first slot
// of the reference leaf (written by the leaf kernels).  Spheres keep only the
// words.  One gather of the shape serves both: the box is recomputed from it
// (the same min/max sequence as k_bounds, so the same values).
WPBR_D void slot_prims_body(const LbvhScratch& s, const GeomView& g, uint32_t n, uint32_t bid) {
    const uint32_t p = bid * blockDim.x + threadIdx.x;
    if (p >= n) return;
    // occupied top-12-bit bins (the complete order's treelets); the reference
    // order finds its treelets by a scan of the run heads instead
    if (s.order != WPBR_LBVH_ORDER_REFERENCE) treelet_mark_body(s.keys_a, n, s.bin_start, s.presence, bid);
    const uint32_t id = s.vals_a[p];
    float v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    bbox b;
    if (id >= g.n_spheres) {
        const float* t = g.tri_verts + 9 * (size_t)(id - g.n_spheres);
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = t[k];
        b = box_empty();
        b = box_union_pt(b, mk3(v[0], v[1], v[2]));
        b = box_union_pt(b, mk3(v[3], v[4], v[5]));
        b = box_union_pt(b, mk3(v[6], v[7], v[8]));
    } else {
        shape_box(g, id, b);
    }
    float2* sb = reinterpret_cast<float2*>(s.slot_bounds + 6 * (size_t)p);
    sb[0] = make_float2(b.mn[0], b.mn[1]);
    sb[1] = make_float2(b.mn[2], b.mx[0]);
    sb[2] = make_float2(b.mx[1], b.mx[2]);
    float4* rec = reinterpret_cast<float4*>(s.prims) + 3 * (size_t)p;
    rec[0] = make_float4(v[0], v[1], v[2], __uint_as_float(id));
    rec[1] = make_float4(v[3], v[4], v[5], __uint_as_float(p));
    rec[2] = make_float4(v[6], v[7], v[8], 0.0f);
}
__global__ __launch_bounds__(256) void k_slot_prims(LbvhScratch s, GeomView g, uint32_t n) {
    slot_prims_body(s, g, n, blockIdx.x);
}

// Box of a slot (slot bounds are 8-byte aligned: 24 bytes per slot).
WPBR_D bbox slot_box(const float* sb, uint32_t q) {
    const float2* p = reinterpret_cast<const float2*>(sb + 6 * (size_t)q);
    const float2 x = p[0], y = p[1], z = p[2];
    bbox b;
    b.mn[0] = x.x; b.mn[1] = x.y; b.mn[2] = y.x;
    b.mx[0] = y.y; b.mx[1] = z.x; b.mx[2] = z.y;
    return b;
}

// Unions of the slot boxes of [a, m) and [m, e) (e > a), each in slot order;
// eight loads in flight at a time.
WPBR_D void slot_range_boxes(const float* sb, uint32_t a, uint32_t m, uint32_t e, bbox& b0, bbox& b1) {
    b0 = box_empty();
    b1 = box_empty();
    for (uint32_t q = a; q < e; q += 8) {
        bbox t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = slot_box(sb, min(q + u, e - 1));
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t j = q + u;
            if (j < m) b0 = box_union(b0, t[u]);
            else if (j < e) b1 = box_union(b1, t[u]);
        }
    }
}

// The leaf word of the primitive record at `slot`: its reference leaf's first slot.
WPBR_D void set_prim_leaf(const LbvhScratch& s, uint32_t slot, uint32_t leaf) {
    reinterpret_cast<uint32_t*>(s.prims)[12 * (size_t)slot + 11] = leaf;
}


// ---------------------------------------------------------------------------
// Emit: one workgroup per occupied top-12-bit bin (= treelet).
constexpr int EMIT_BLOCK = 64;  // one wave: a treelet level rarely has more nodes
constexpr int EMIT_MAX_LEVELS = 64;

constexpr uint32_t EMIT_LDS_CODES = 1024;  // a treelet's codes are staged in LDS up to this size

// The emitLBVH tree of treelet t (slots start .. start + cnt - 1) by one
// workgroup: breadth-first construction, then subtree sizes and bounds
// bottom-up and DFS-preorder offsets top-down.  Its records sit at the
// treelet's reserved slot range (2 cnt - 1 entries from 2 start - t: the sum of
// 2 n_j - 1 over the treelets before it, SLBVH.cpp:307-310).
WPBR_D void emit_tree(const LbvhScratch& s, const uint32_t* __restrict__ codes, uint32_t t, uint32_t start,
                      uint32_t cnt) {
    __shared__ uint32_t s_codes[EMIT_LDS_CODES];
    __shared__ uint32_t s_q;
    __shared__ uint32_t s_lvl[EMIT_MAX_LEVELS + 2];
    __shared__ int s_nlvl;
    const int tid = threadIdx.x;
    const uint32_t base = 2 * start - t;
    if (tid == 0) {
        s.t_start[t] = start;
        s.t_count[t] = cnt;
        s.t_slot[t] = base;
        // root task
        s.r_start[base] = start;
        s.r_n[base] = cnt;
        s.r_meta[base] = 17u;  // first bit index = 30 - 12 - 1 (SLBVH.cpp:342)
        s_q = 1;
        s_lvl[0] = 0;
        s_lvl[1] = 1;
        s_nlvl = 1;
    }
    __syncthreads();
    const uint32_t s_tidx = t;
    // the treelet's codes in LDS when they fit (the split searches read them)
    const uint32_t tstart = start, tcnt = cnt;
    const bool staged = tcnt <= EMIT_LDS_CODES;
    if (staged)
        for (uint32_t k = tid; k < tcnt; k += EMIT_BLOCK) s_codes[k] = codes[tstart + k];
    __syncthreads();
    auto code = [&](uint32_t i) -> uint32_t { return staged ? s_codes[i - tstart] : codes[i]; };
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
                x = code(rs) ^ code(rs + rn - 1);
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
                    if ((code(ss) & mask) == (code(mid) & mask)) ss = mid;
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
                bbox unused;
                slot_range_boxes(s.slot_bounds, rs, rs + rn, rs + rn, b, unused);
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
// Treelet of occupied top-12-bit bin `bin` (the complete sort: one treelet per
// occupied bin, its index the rank of the bin among the occupied ones).
WPBR_D void emit_body(const LbvhScratch& s, const uint32_t* __restrict__ codes, uint32_t n, uint32_t bin) {
    __shared__ uint32_t s_tidx, s_start, s_cnt;
    const int tid = threadIdx.x;
    if (!((s.presence[bin >> 5] >> (bin & 31)) & 1u)) return;
    // the treelet's rank (occupied bins before it) and the next occupied bin,
    // from the 128 presence words, two per lane (EMIT_BLOCK is one wave)
    const uint32_t bw = bin >> 5, below = (1u << (bin & 31)) - 1u, above = ~((2u << (bin & 31)) - 1u);
    const uint32_t q0 = 2 * (uint32_t)tid, q1 = q0 + 1;
    const uint32_t p0 = s.presence[q0], p1 = s.presence[q1];
    uint32_t cnt = (q0 < bw ? __popc(p0) : (q0 == bw ? __popc(p0 & below) : 0u)) +
                   (q1 < bw ? __popc(p1) : (q1 == bw ? __popc(p1 & below) : 0u));
    for (int d = 32; d >= 1; d >>= 1) cnt += (uint32_t)__shfl_xor((int)cnt, d);
    const uint32_t m0 = q0 > bw ? p0 : (q0 == bw ? p0 & above : 0u);
    const uint32_t m1 = q1 > bw ? p1 : (q1 == bw ? p1 & above : 0u);
    const uint64_t has = ballot64((m0 | m1) != 0u);
    const uint32_t nb_lane = m0 ? 32 * q0 + (uint32_t)(__ffs(m0) - 1) : 32 * q1 + (uint32_t)(__ffs(m1) - 1);
    const uint32_t next_bin = has ? (uint32_t)__shfl((int)nb_lane, (int)(__ffsll((unsigned long long)has) - 1)) : 0u;
    if (tid == 0) {
        uint32_t start = s.bin_start[bin];
        uint32_t end = has ? s.bin_start[next_bin] : n;
        s_tidx = cnt;
        s_start = start;
        s_cnt = end - start;
    }
    __syncthreads();
    emit_tree(s, codes, s_tidx, s_start, s_cnt);
}
__global__ __launch_bounds__(EMIT_BLOCK) void k_emit(LbvhScratch s, const uint32_t* __restrict__ codes, uint32_t n) {
    emit_body(s, codes, n, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Upper SAH over treelet roots (buildUpperSAH, SLBVH.cpp:457-563): one
// workgroup.  The treelet root boxes, the position array `perm` (treelet ids,
// permuted in place as std::partition does) and the per-position bucket live in
// LDS.  Ranges of at least BLOCK_TASK_MIN positions are split by the whole
// workgroup, ranges above LANE_TASK_MAX by one wave, smaller ones by one lane
// (depth first, the larger child waiting on the lane's LDS stack).
//
// No parent/child links are kept: every node covers a contiguous range of final
// positions [i0, i1], so with P = exclusive prefix sum of the treelet node
// counts over final positions, a node of depth d whose root path turns right r
// times sits at preorder offset d + (i0 - r) + P[i0] (the nodes before it are
// its d ancestors, and the r subtrees hanging left of its path, which hold the
// i0 treelets before it and i0 - r upper nodes).  The size of the subtree over
// [i0, m) is (m - 1 - i0) + P[m] - P[i0].  Node bounds are the union of the
// range's treelet boxes, computed when the range is split (the same value as
// the union of the two child boxes; min/max are exact).
constexpr int SAH_BLOCK = 512;   // 256 VGPRs per lane for the register-resident buckets
constexpr int SAH_WAVES = SAH_BLOCK / 64;
constexpr int NB = 12;
constexpr uint32_t LANE_TASK_MAX = 32;
constexpr int LANE_STACK = 6;               // smaller child first: depth <= log2(32) + 1
constexpr uint32_t BLOCK_TASK_MIN = 4096;   // ranges this large use the whole workgroup: smaller ones
                                            // run one per wave, in parallel (the LDS atomics of the
                                            // bucket statistics bound both ways alike)
constexpr uint32_t SAH_CUT = 256;           // sah_top defers ranges below this size
constexpr uint32_t LANE_TASK_MAX_SUB = 8;   // lane tasks of sah_sub
constexpr uint32_t SAH_SUB_TASKS = 2 * SAH_CUT;  // a deferred range makes fewer tasks than this
static_assert(LANE_STACK * SAH_BLOCK <= 4100, "lane stacks share the prefix-sum array");

// Bucket statistics rows: SC copies per wave, lane l of a wave adding to copy
// l % SC, so that fewer lanes of one LDS atomic meet on one address.
struct SahLds {
    static constexpr int SC = 4;
    alignas(16) float tbox[6 * 4096];  // treelet root boxes
    uint16_t perm[4096];
    uint8_t bucket[4096];
    uint16_t la[4096], lb[4096];            // partition lists; then leaf depth / right turns
    uint32_t aux[4100];         // lane stacks (LANE_STACK x SAH_BLOCK); then the prefix sums P[0..k]
    uint32_t bcount[SAH_WAVES * SC][NB];
    uint32_t bmin[SAH_WAVES * SC][NB][3];
    uint32_t bmax[SAH_WAVES * SC][NB][3];
    float red[SAH_WAVES][12];
    float fin[12];
    uint32_t wcnt[SAH_WAVES];
    uint32_t split[2];
    uint32_t k, task_next, nbig, nbig_next;
};

WPBR_D float wave_min_f(float v) {
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}
WPBR_D float wave_max_f(float v) {
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

// The split helpers below read an LDS layout LDS with members tbox (6 floats
// per treelet root), perm, bucket, la, lb (indexed by position) and bucket
// statistics rows bcount / bmin / bmax: the workgroup's SahLds or one wave's
// WaveRangeLds.
template <class LDS>
WPBR_D bbox tb_box(const LDS& L, uint32_t t) {
    const float* p = L.tbox + 6 * t;
    bbox b;
    b.mn[0] = p[0]; b.mn[1] = p[1]; b.mn[2] = p[2];
    b.mx[0] = p[3]; b.mx[1] = p[4]; b.mx[2] = p[5];
    return b;
}

// Treelet root centroid on one axis: (pMin + pMax) * 0.5 (SLBVH.cpp:472, 487).
template <class LDS>
WPBR_D float tb_cen(const LDS& L, uint32_t t, int dim) {
    const float* p = L.tbox + 6 * t;
    return (float)((double)(p[dim] + p[3 + dim]) * 0.5);
}

WPBR_D uint32_t sah_bucket(float c, float cmin, float dist) {
    uint32_t ib = 0;
    if (dist > 0.0f) {
        float tt = (c - cmin) / dist;
        ib = (uint32_t)(uint8_t)(uint32_t)((float)NB * tt);
        if (ib == NB) ib = NB - 1;
    }
    return ib;
}

// Records upper node [i0, i1] split at mid; returns its index.
// A task list (ranges i0..i1 of final positions with their depth and right
// turns) and the counter that numbers upper nodes: global memory for the top
// kernel, LDS for a workgroup of the sub-range kernel.
// Positions, depths and right turns of the tasks are relative to pos_base,
// dep_base and rt_base (a range run in LDS from local position 0).
struct SahTasks {
    uint32_t* i0; uint32_t* i1; uint32_t* dep; uint32_t* rt;
    uint32_t* uctr;
    uint32_t pos_base = 0, dep_base = 0, rt_base = 0;
};

WPBR_D uint32_t sah_node(LbvhScratch& s, const SahTasks& T, uint32_t i0, uint32_t i1, uint32_t mid, int dim,
                         const bbox& bounds, uint32_t dep, uint32_t rt) {
    uint32_t node = atomicAdd(T.uctr, 1u);
    s.u_i0[node] = T.pos_base + i0;
    s.u_i1[node] = T.pos_base + i1;
    s.u_mid[node] = T.pos_base + mid;
    s.u_dep[node] = T.dep_base + dep;
    s.u_rt[node] = T.rt_base + rt;
    s.u_axis[node] = (uint32_t)dim;
    store_box(s.u_box + 6 * (size_t)node, bounds);
    return node;
}

// Queues the two children of a split made by a wave or the workgroup.
WPBR_D void sah_children(const SahTasks& T, SahLds& L, uint32_t i0, uint32_t i1, uint32_t mid, uint32_t dep,
                         uint32_t rt) {
    uint32_t ct = atomicAdd(&L.task_next, 2u);
    T.i0[ct] = i0;
    T.i1[ct] = mid - 1;
    T.dep[ct] = dep + 1;
    T.rt[ct] = rt;
    T.i0[ct + 1] = mid;
    T.i1[ct + 1] = i1;
    T.dep[ct + 1] = dep + 1;
    T.rt[ct + 1] = rt + 1;
    uint32_t nb = (mid - i0 >= BLOCK_TASK_MIN) + (i1 + 1 - mid >= BLOCK_TASK_MIN);
    if (nb) atomicAdd(&L.nbig_next, nb);
}

// Treelet root centroid on one axis from its box: (pMin + pMax) * 0.5
// (SLBVH.cpp:472, 487), as tb_cen.
WPBR_D float box_cen(const bbox& b, int c) { return (float)((double)(b.mn[c] + b.mx[c]) * 0.5); }

// Inclusive scans over the 16 lanes of each row: UP from lane 0 (row_shr), or
// down from lane 15 (row_shl); lanes whose source lies outside the row take the
// identity.  All lanes must be active.
template <bool UP, bool MAX>
WPBR_D float row_scan(float v) {
    const int id = __float_as_int(MAX ? FLT_LOWEST_ : FLT_MAX_);
    auto op = [](float a, float b) { return MAX ? smax(a, b) : smin(a, b); };
#define WPBR_ROW_SCAN(d) v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), (UP ? 0x110 : 0x100) + (d), 0xF, 0xF, false)))
    WPBR_ROW_SCAN(1);
    WPBR_ROW_SCAN(2);
    WPBR_ROW_SCAN(4);
    WPBR_ROW_SCAN(8);
#undef WPBR_ROW_SCAN
    return v;
}
template <bool UP>
WPBR_D uint32_t row_scan_add(uint32_t v) {
#define WPBR_ROW_SCAN(d) v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, (UP ? 0x110 : 0x100) + (d), 0xF, 0xF, false)
    WPBR_ROW_SCAN(1);
    WPBR_ROW_SCAN(2);
    WPBR_ROW_SCAN(4);
    WPBR_ROW_SCAN(8);
#undef WPBR_ROW_SCAN
    return v;
}
// Lane value of lane + 1 in the row (identity for lane 15).
WPBR_D float row_next(float v, float id) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(id), __float_as_int(v), 0x101, 0xF, 0xF, false));
}
WPBR_D uint32_t row_next(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xF, 0xF, false); }

// Bucket j's statistics for lane j < NB of a wave, merged over ROWS
// consecutive statistics rows (count, and the box from its order-preserving
// min / max words); the other lanes get the identity (count 0, empty box).
template <int ROWS>
WPBR_D void bucket_lane_load(const uint32_t (*bc)[NB], const uint32_t (*bmn)[NB][3], const uint32_t (*bmx)[NB][3],
                             int lane, uint32_t& cnt, float (&mn)[3], float (&mx)[3]) {
    cnt = 0;
    uint32_t wn[3], wx[3];
#pragma unroll
    for (int c = 0; c < 3; c++) { wn[c] = f2ord(FLT_MAX_); wx[c] = f2ord(FLT_LOWEST_); }
    if (lane < NB)
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            cnt += bc[r][lane];
#pragma unroll
            for (int c = 0; c < 3; c++) { wn[c] = min(wn[c], bmn[r][lane][c]); wx[c] = max(wx[c], bmx[r][lane][c]); }
        }
#pragma unroll
    for (int c = 0; c < 3; c++) { mn[c] = ord2f(wn[c]); mx[c] = ord2f(wx[c]); }
}

// The reference's split cost (SLBVH.cpp:504-538) for split `lane` (< NB-1), lane
// j < NB holding bucket j (bucket_lane_load): the two sides of every split come
// from an inclusive prefix and an inclusive suffix over the row's lanes (DPP),
// lane s taking the prefix at s and the suffix at s + 1.  A side's box is the
// union of the same bucket boxes as the reference's loop (min/max are exact), so
// the cost is the reference's.  Lanes >= NB-1 return +inf.  All lanes must be
// active.
WPBR_D float sah_cost_lanes(uint32_t cnt, const float (&mn)[3], const float (&mx)[3], int lane, const bbox& bounds,
                            uint32_t* cnt_le) {
    bbox b0, b1;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        b0.mn[c] = row_scan<true, false>(mn[c]);
        b0.mx[c] = row_scan<true, true>(mx[c]);
        b1.mn[c] = row_next(row_scan<false, false>(mn[c]), FLT_MAX_);
        b1.mx[c] = row_next(row_scan<false, true>(mx[c]), FLT_LOWEST_);
    }
    const uint32_t c0 = row_scan_add<true>(cnt);
    const uint32_t c1 = row_next(row_scan_add<false>(cnt));
    float cost = INF;
    *cnt_le = 0;
    if (lane < NB - 1) {
        float l = (c0 == 0) ? 0.0f : (float)c0 * box_area(b0);
        float r = (c1 == 0) ? 0.0f : (float)c1 * box_area(b1);
        cost = (float)(0.125 + (double)((l + r) / box_area(bounds)));
        *cnt_le = c0;
    }
    return cost;
}

// First minimum over the 11 split costs held by lanes 0..10 of one wave.
WPBR_D int sah_best(float cost) {
    int best = 0;
    float cmin = __shfl(cost, 0);
    for (int i = 1; i < NB - 1; i++) {
        float ci = __shfl(cost, i);
        if (ci < cmin) { best = i; cmin = ci; }
    }
    return best;
}

// SAH split of perm[i0..i1] by one lane (the reference's loop, SLBVH.cpp:465-553).
// Bucket statistics stay in registers (statically indexed).  The split costs use
// one prefix and one suffix sweep of unions: min/max are exact, so these boxes
// equal the reference's per-split unions up to the sign of a zero coordinate,
// which no extent (and hence no area or cost) can see.
template <class LDS>
WPBR_D uint32_t lane_split(LDS& L, uint32_t i0, uint32_t i1, int* dimOut, bbox* boundsOut) {
    bbox bounds = box_empty(), cb = box_empty();
    for (uint32_t i = i0; i <= i1; i++) {
        uint32_t t = L.perm[i];
        bounds = box_union(bounds, tb_box(L, t));
        cb = box_union_pt(cb, mk3(tb_cen(L, t, 0), tb_cen(L, t, 1), tb_cen(L, t, 2)));
    }
    int dim = box_max_extent(cb);
    float cmin = dim == 0 ? cb.mn[0] : (dim == 1 ? cb.mn[1] : cb.mn[2]);
    float dist = (dim == 0 ? cb.mx[0] : (dim == 1 ? cb.mx[1] : cb.mx[2])) - cmin;
    bbox bb[NB];
    uint32_t bc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) { bb[b] = box_empty(); bc[b] = 0; }
    for (uint32_t i = i0; i <= i1; i++) {
        uint32_t t = L.perm[i];
        uint32_t ib = sah_bucket(tb_cen(L, t, dim), cmin, dist);
        L.bucket[i] = (uint8_t)ib;
        bbox tb = tb_box(L, t);
#pragma unroll
        for (int b = 0; b < NB; b++) {
            if (ib == (uint32_t)b) {
                bc[b]++;
                bb[b] = box_union(bb[b], tb);
            }
        }
    }
    float area = box_area(bounds);
    float la[NB - 1], ra[NB - 1];
    {
        bbox acc = box_empty();
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < NB - 1; i++) {
            acc = box_union(acc, bb[i]);
            c += bc[i];
            la[i] = (c == 0) ? 0.0f : (float)c * box_area(acc);
        }
        acc = box_empty();
        c = 0;
#pragma unroll
        for (int i = NB - 1; i >= 1; i--) {
            acc = box_union(acc, bb[i]);
            c += bc[i];
            ra[i - 1] = (c == 0) ? 0.0f : (float)c * box_area(acc);
        }
    }
    float cbest = 0.0f;
    int best = 0;
#pragma unroll
    for (int i = 0; i < NB - 1; i++) {
        float cost = (float)(0.125 + (double)((la[i] + ra[i]) / area));
        if (i == 0) cbest = cost;
        else if (cost < cbest) { best = i; cbest = cost; }
    }
    // libstdc++ two-pointer std::partition on perm[i0..i1]
    uint32_t first = i0, last = i1 + 1;
    while (true) {
        while (first != last && L.bucket[first] <= best) ++first;
        if (first == last) break;
        --last;
        while (first != last && !(L.bucket[last] <= best)) --last;
        if (first == last) break;
        uint16_t tp = L.perm[first]; L.perm[first] = L.perm[last]; L.perm[last] = tp;
        uint8_t tb = L.bucket[first]; L.bucket[first] = L.bucket[last]; L.bucket[last] = tb;
        ++first;
    }
    uint32_t mid = first;
    if (mid == i1 + 1) mid = (i1 - i0 + 1) / 2 + i0;
    *dimOut = dim;
    *boundsOut = bounds;
    return mid;
}

// One SAH split of perm[i0..i1] by the whole workgroup (every thread calls it
// with the same arguments): the same bucketing, costs and two-pointer
// partition, with block-wide reductions and ranks.
WPBR_D void sah_block_split(LbvhScratch& s, SahLds& L, const SahTasks& TL, uint32_t i0, uint32_t i1, uint32_t dep,
                            uint32_t rt) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    float v[12];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v[c] = FLT_MAX_; v[3 + c] = FLT_LOWEST_; v[6 + c] = FLT_MAX_; v[9 + c] = FLT_LOWEST_;
    }
    for (uint32_t i = i0 + tid; i <= i1; i += SAH_BLOCK) {
        uint32_t tr = L.perm[i];
        bbox rb = tb_box(L, tr);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v[c] = smin(v[c], rb.mn[c]);
            v[3 + c] = smax(v[3 + c], rb.mx[c]);
            float cc = tb_cen(L, tr, c);
            v[6 + c] = smin(v[6 + c], cc);
            v[9 + c] = smax(v[9 + c], cc);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v[c] = wave_min_f(v[c]); v[3 + c] = wave_max_f(v[3 + c]);
        v[6 + c] = wave_min_f(v[6 + c]); v[9 + c] = wave_max_f(v[9 + c]);
    }
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 12; q++) L.red[wave][q] = v[q];
    constexpr int SC = SahLds::SC;
    if (lane < NB * SC) {
        const int row = wave * SC + lane / NB, b = lane % NB;
        L.bcount[row][b] = 0;
        for (int c = 0; c < 3; c++) {
            L.bmin[row][b][c] = f2ord(FLT_MAX_);
            L.bmax[row][b][c] = f2ord(FLT_LOWEST_);
        }
    }
    __syncthreads();
    if (tid < 12) {
        bool mx = (tid % 6) >= 3;
        float r = L.red[0][tid];
        for (int w = 1; w < SAH_WAVES; w++) r = mx ? smax(r, L.red[w][tid]) : smin(r, L.red[w][tid]);
        L.fin[tid] = r;
    }
    __syncthreads();
    bbox bounds, cb;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        bounds.mn[c] = L.fin[c]; bounds.mx[c] = L.fin[3 + c];
        cb.mn[c] = L.fin[6 + c]; cb.mx[c] = L.fin[9 + c];
    }
    const int dim = box_max_extent(cb);
    const float cbmn = dim == 0 ? cb.mn[0] : (dim == 1 ? cb.mn[1] : cb.mn[2]);
    const float dist = (dim == 0 ? cb.mx[0] : (dim == 1 ? cb.mx[1] : cb.mx[2])) - cbmn;
    // each thread takes an odd-length run of positions: the lanes of one
    // instruction then hold treelets far apart, in different buckets, so fewer
    // of their LDS atomics meet on one address
    const uint32_t run = ((i1 - i0) / SAH_BLOCK + 1) | 1u;
    for (uint32_t j = 0; j < run; j++) {
        const uint32_t i = i0 + tid * run + j;
        if (i > i1) break;
        uint32_t tr = L.perm[i];
        bbox rb = tb_box(L, tr);
        uint32_t ib = sah_bucket(tb_cen(L, tr, dim), cbmn, dist);
        L.bucket[i] = (uint8_t)ib;
        const int row = wave * SC + (lane % SC);
        atomicAdd(&L.bcount[row][ib], 1u);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            atomicMin(&L.bmin[row][ib][c], f2ord(rb.mn[c]));
            atomicMax(&L.bmax[row][ib][c], f2ord(rb.mx[c]));
        }
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t cnt, cnt_le;
        float mn[3], mx[3];
        bucket_lane_load<SAH_WAVES * SC>(L.bcount, L.bmin, L.bmax, lane, cnt, mn, mx);
        float cost = sah_cost_lanes(cnt, mn, mx, lane, bounds, &cnt_le);
        int best = sah_best(cost);
        uint32_t T = (uint32_t)__shfl((int)cnt_le, best);
        if (lane == 0) { L.split[0] = (uint32_t)best; L.split[1] = T; }
    }
    __syncthreads();
    const uint32_t best = L.split[0], T = L.split[1], size = i1 - i0 + 1;
    uint32_t mid;
    if (T == size) {
        mid = size / 2 + i0;
    } else {
        // two-pointer partition: the r-th misplaced element of [i0, i0+T) swaps
        // with the r-th misplaced element of [i0+T, i1] counted from the right
        uint32_t run = 0;
        for (uint32_t c = 0; c < T; c += SAH_BLOCK) {
            uint32_t i = i0 + c + tid;
            bool bad = (c + tid < T) && (L.bucket[i] > best);
            uint64_t m = ballot64(bad);
            if (lane == 0) L.wcnt[wave] = popc64(m);
            __syncthreads();
            uint32_t before = 0, tot = 0;
            for (int w = 0; w < SAH_WAVES; w++) { uint32_t x = L.wcnt[w]; before += (w < wave) ? x : 0; tot += x; }
            if (bad) L.la[i0 + run + before + popc64(m & lanemask_lt())] = (uint16_t)i;
            run += tot;
            __syncthreads();
        }
        const uint32_t rl = run;
        run = 0;
        const uint32_t rsz = size - T;
        for (uint32_t c = 0; c < rsz; c += SAH_BLOCK) {
            uint32_t i = i1 - (c + tid);
            bool bad = (c + tid < rsz) && (L.bucket[i] <= best);
            uint64_t m = ballot64(bad);
            if (lane == 0) L.wcnt[wave] = popc64(m);
            __syncthreads();
            uint32_t before = 0, tot = 0;
            for (int w = 0; w < SAH_WAVES; w++) { uint32_t x = L.wcnt[w]; before += (w < wave) ? x : 0; tot += x; }
            if (bad) L.lb[i0 + run + before + popc64(m & lanemask_lt())] = (uint16_t)i;
            run += tot;
            __syncthreads();
        }
        for (uint32_t r = tid; r < rl; r += SAH_BLOCK) {
            uint32_t a = L.la[i0 + r], b = L.lb[i0 + r];
            uint16_t pa = L.perm[a], pb = L.perm[b];
            L.perm[a] = pb;
            L.perm[b] = pa;
        }
        mid = i0 + T;
    }
    if (tid == 0) {
        sah_node(s, TL, i0, i1, mid, dim, bounds, dep, rt);
        sah_children(TL, L, i0, i1, mid, dep, rt);
    }
    __syncthreads();
}

// Ranges up to this size get their split costs lane by lane (below); larger
// ones through bucket statistics.
constexpr uint32_t SPLIT_LANES_MAX = 64;

// One SAH split of perm[i0..i1] by one wave; returns the split position (every
// lane), its axis and the range's bounds.  Split costs (SLBVH.cpp:504-538):
//  - ranges of at most SPLIT_LANES_MAX treelets: lane 4s + q unions the boxes
//    of positions i0 + q, i0 + q + 4, ... on either side of split s, and the
//    four lanes of split s merge theirs; no bucket statistics, no atomics;
//  - larger ranges: bucket statistics in row `wave` (LDS atomics), then split
//    s's unions of buckets 0..s and s+1..NB-1 by lane s.
// Either way a side's box is the union of the same treelet boxes (min/max are
// exact), so the costs, and the first minimum, are those of the reference.
template <class LDS>
WPBR_D uint32_t wave_split_mid(LDS& L, int wave, uint32_t i0, uint32_t i1, int* dimOut, bbox* boundsOut) {
    const int lane = threadIdx.x & 63;
    float bmn[3] = {FLT_MAX_, FLT_MAX_, FLT_MAX_}, bmx[3] = {FLT_LOWEST_, FLT_LOWEST_, FLT_LOWEST_};
    float cmn[3] = {FLT_MAX_, FLT_MAX_, FLT_MAX_}, cmx[3] = {FLT_LOWEST_, FLT_LOWEST_, FLT_LOWEST_};
    auto acc = [&](const bbox& rb) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            bmn[c] = smin(bmn[c], rb.mn[c]);
            bmx[c] = smax(bmx[c], rb.mx[c]);
            const float cc = box_cen(rb, c);
            cmn[c] = smin(cmn[c], cc);
            cmx[c] = smax(cmx[c], cc);
        }
    };
    // range box and centroid bounds: four positions per lane per step, their
    // LDS loads issued before any is used (the walk is bound by LDS latency)
    uint32_t i = i0 + lane;
    for (; i + 192 <= i1; i += 256) {
        uint32_t tr[4];
        bbox rb[4];
#pragma unroll
        for (int q = 0; q < 4; q++) tr[q] = L.perm[i + 64 * q];
#pragma unroll
        for (int q = 0; q < 4; q++) rb[q] = tb_box(L, tr[q]);
#pragma unroll
        for (int q = 0; q < 4; q++) acc(rb[q]);
    }
    for (; i <= i1; i += 64) acc(tb_box(L, L.perm[i]));
    bbox bounds, cb;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        bounds.mn[c] = wave_min_dpp(bmn[c]);
        bounds.mx[c] = wave_max_dpp(bmx[c]);
        cb.mn[c] = wave_min_dpp(cmn[c]);
        cb.mx[c] = wave_max_dpp(cmx[c]);
    }
    int dim = box_max_extent(cb);
    float cbmn = dim == 0 ? cb.mn[0] : (dim == 1 ? cb.mn[1] : cb.mn[2]);
    float dist = (dim == 0 ? cb.mx[0] : (dim == 1 ? cb.mx[1] : cb.mx[2])) - cbmn;
    const uint32_t size = i1 - i0 + 1;
    const bool by_lane = size <= SPLIT_LANES_MAX;
    float cost = INF;
    uint32_t cnt_le = 0;
    if (by_lane) {
        for (uint32_t i = i0 + lane; i <= i1; i += 64)
            L.bucket[i] = (uint8_t)sah_bucket(tb_cen(L, L.perm[i], dim), cbmn, dist);
        __builtin_amdgcn_wave_barrier();
        const uint32_t sp = (uint32_t)lane >> 2, q = (uint32_t)lane & 3u;
        bbox b0 = box_empty(), b1 = box_empty();
        uint32_t c0 = 0, c1 = 0;
        if (sp < NB - 1)
            for (uint32_t i = i0 + q; i <= i1; i += 4) {
                const bbox tb = tb_box(L, L.perm[i]);
                if (L.bucket[i] <= sp) { b0 = box_union(b0, tb); c0++; }
                else { b1 = box_union(b1, tb); c1++; }
            }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            b0.mn[c] = smin(b0.mn[c], quad_xor1(b0.mn[c])); b0.mx[c] = smax(b0.mx[c], quad_xor1(b0.mx[c]));
            b1.mn[c] = smin(b1.mn[c], quad_xor1(b1.mn[c])); b1.mx[c] = smax(b1.mx[c], quad_xor1(b1.mx[c]));
        }
        c0 += quad_xor1(c0); c1 += quad_xor1(c1);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            b0.mn[c] = smin(b0.mn[c], quad_xor2(b0.mn[c])); b0.mx[c] = smax(b0.mx[c], quad_xor2(b0.mx[c]));
            b1.mn[c] = smin(b1.mn[c], quad_xor2(b1.mn[c])); b1.mx[c] = smax(b1.mx[c], quad_xor2(b1.mx[c]));
        }
        c0 += quad_xor2(c0); c1 += quad_xor2(c1);
        if (sp < NB - 1) {
            float l = (c0 == 0) ? 0.0f : (float)c0 * box_area(b0);
            float r = (c1 == 0) ? 0.0f : (float)c1 * box_area(b1);
            cost = (float)(0.125 + (double)((l + r) / box_area(bounds)));
            cnt_le = c0;
        }
    } else {
        // per-wave statistics in rows wave * SC .. wave * SC + SC - 1
        constexpr int SC = LDS::SC;
        const int row0 = wave * SC;
        if (lane < NB * SC) {
            const int row = row0 + lane / NB, b = lane % NB;
            L.bcount[row][b] = 0;
            for (int c = 0; c < 3; c++) {
                L.bmin[row][b][c] = f2ord(FLT_MAX_);
                L.bmax[row][b][c] = f2ord(FLT_LOWEST_);
            }
        }
        __builtin_amdgcn_wave_barrier();
        // odd-length runs per lane, as in the workgroup split (consecutive
        // positions of a lane often share a bucket: their atomics are kept
        // apart, one position per step)
        const uint32_t run = ((size - 1) / 64 + 1) | 1u;
        const int row = row0 + (lane % SC);
        for (uint32_t j = 0; j < run; j++) {
            const uint32_t i = i0 + (uint32_t)lane * run + j;
            if (i > i1) break;
            const bbox rb = tb_box(L, L.perm[i]);
            const uint32_t ib = sah_bucket(box_cen(rb, dim), cbmn, dist);
            L.bucket[i] = (uint8_t)ib;
            atomicAdd(&L.bcount[row][ib], 1u);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                atomicMin(&L.bmin[row][ib][c], f2ord(rb.mn[c]));
                atomicMax(&L.bmax[row][ib][c], f2ord(rb.mx[c]));
            }
        }
        __builtin_amdgcn_wave_barrier();
        uint32_t cnt;
        float mn[3], mx[3];
        bucket_lane_load<SC>(L.bcount + row0, L.bmin + row0, L.bmax + row0, lane, cnt, mn, mx);
        cost = sah_cost_lanes(cnt, mn, mx, lane, bounds, &cnt_le);
    }
    // first minimum over the NB - 1 costs (in lanes 4s, or lanes s)
    const int cs = by_lane ? 4 : 1;
    int best = 0;
    float cbest = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cost), 0));
    for (int sp = 1; sp < NB - 1; sp++) {
        const float c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cost), sp * cs));
        if (c < cbest) { best = sp; cbest = c; }
    }
    const uint32_t T = (uint32_t)__builtin_amdgcn_readlane((int)cnt_le, best * cs);
    uint32_t mid;
    if (T == size) {
        mid = size / 2 + i0;
    } else {
        // misplaced positions in rank order, four 64-position chunks per step
        const uint64_t lt = lanemask_lt();
        uint32_t rl = 0;
        for (uint32_t c = 0; c < T; c += 256) {
            uint32_t bk[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t k = c + 64 * q + lane;
                bk[q] = k < T ? L.bucket[i0 + k] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t k = c + 64 * q + lane;
                const bool bad = k < T && bk[q] > (uint32_t)best;
                const uint64_t m = ballot64(bad);
                if (bad) L.la[i0 + rl + popc64(m & lt)] = (uint16_t)(i0 + k);
                rl += popc64(m);
            }
        }
        uint32_t rr = 0;
        const uint32_t rsz = size - T;
        for (uint32_t c = 0; c < rsz; c += 256) {
            uint32_t bk[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t k = c + 64 * q + lane;
                bk[q] = k < rsz ? L.bucket[i1 - k] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t k = c + 64 * q + lane;
                const bool bad = k < rsz && bk[q] <= (uint32_t)best;
                const uint64_t m = ballot64(bad);
                if (bad) L.lb[i0 + rr + popc64(m & lt)] = (uint16_t)(i1 - k);
                rr += popc64(m);
            }
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t r = lane; r < rl; r += 64) {
            uint32_t a = L.la[i0 + r], b = L.lb[i0 + r];
            uint16_t pa = L.perm[a], pb = L.perm[b];
            L.perm[a] = pb;
            L.perm[b] = pa;
        }
        __builtin_amdgcn_wave_barrier();
        mid = i0 + T;
    }
    *dimOut = dim;
    *boundsOut = bounds;
    return mid;
}

// One SAH split of perm[i0..i1] by one wave, its node and its two child tasks.
WPBR_D void sah_wave_split(LbvhScratch& s, SahLds& L, const SahTasks& TL, uint32_t i0, uint32_t i1, uint32_t dep,
                           uint32_t rt) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int dim;
    bbox bounds;
    const uint32_t mid = wave_split_mid(L, wave, i0, i1, &dim, &bounds);
    if (lane == 0) {
        sah_node(s, TL, i0, i1, mid, dim, bounds, dep, rt);
        sah_children(TL, L, i0, i1, mid, dep, rt);
    }
}

// The SAH top runs as three kernels so that independent sub-ranges use many
// CUs: sah_top (one workgroup) splits the ranges of at least SAH_CUT treelets
// and defers the smaller ones to a list; sah_sub (many workgroups) finishes the
// deferred ranges, each in its own workgroup; sah_final lays the nodes out in
// preorder.  Every range is split by the same steps in any of them, and ranges
// are disjoint, so the result does not depend on the order.
// Generations of tasks lb .. le (in s.task_*): very large ranges by the whole
// workgroup, then ranges above lane_max by one wave each, then ranges of at
// most lane_max by one lane each, to completion; ranges below `cut` go to the
// deferred list instead.
WPBR_D void sah_run(LbvhScratch& s, SahLds& L, const SahTasks& T, uint32_t lb, uint32_t le, uint32_t lane_max,
                    uint32_t cut) {
    const int tid = threadIdx.x;
    while (lb < le) {
        // ---- very large ranges: the whole workgroup, one range at a time
        if (L.nbig) {
            for (uint32_t t = lb; t < le; t++) {
                const uint32_t i0 = T.i0[t], i1 = T.i1[t];
                if (i1 - i0 + 1 < BLOCK_TASK_MIN) continue;
                sah_block_split(s, L, T, i0, i1, T.dep[t], T.rt[t]);
            }
        }
        // ---- deferred ranges
        if (cut)
            for (uint32_t t = lb + tid; t < le; t += SAH_BLOCK) {
                const uint32_t i0 = T.i0[t], i1 = T.i1[t];
                if (i1 - i0 + 1 >= cut) continue;
                const uint32_t e = atomicAdd(s.sub_n, 1u);
                s.sub_i0[e] = i0;
                s.sub_i1[e] = i1;
                s.sub_dep[e] = T.dep[t];
                s.sub_rt[e] = T.rt[t];
                // a range of m treelets makes m - 1 upper nodes: number them now
                s.sub_ub[e] = atomicAdd(T.uctr, i1 - i0);
            }
        // ---- large ranges: one wave each
        for (uint32_t t = lb + (tid >> 6); t < le; t += SAH_WAVES) {
            const uint32_t i0 = T.i0[t], i1 = T.i1[t];
            const uint32_t m = i1 - i0 + 1;
            if (m <= lane_max || m >= BLOCK_TASK_MIN || m < cut) continue;
            sah_wave_split(s, L, T, i0, i1, T.dep[t], T.rt[t]);
        }
        // ---- small ranges: one lane each, to completion; the larger child of a
        // split waits on the lane's LDS stack (a, b, depth and right turns
        // relative to the task)
        for (uint32_t t = lb + tid; t < le; t += SAH_BLOCK) {
            const uint32_t t0 = T.i0[t], t1 = T.i1[t];
            if (t1 - t0 + 1 > lane_max || t1 - t0 + 1 < cut) continue;
            const uint32_t d0 = T.dep[t], r0 = T.rt[t];
            uint32_t a = t0, b = t1, dd = 0, dr = 0;
            int sp = 0;
            while (true) {
                if (a == b) {
                    L.la[a] = (uint16_t)(d0 + dd);
                    L.lb[a] = (uint16_t)(r0 + dr);
                    if (sp == 0) break;
                    uint32_t e = L.aux[(--sp) * SAH_BLOCK + tid];
                    a = t0 + (e & 31u);
                    b = t0 + ((e >> 5) & 31u);
                    dd = (e >> 10) & 31u;
                    dr = (e >> 15) & 31u;
                    continue;
                }
                int dim;
                bbox bounds;
                uint32_t mid = lane_split(L, a, b, &dim, &bounds);
                sah_node(s, T, a, b, mid, dim, bounds, d0 + dd, r0 + dr);
                uint32_t nl = mid - a, nr = b + 1 - mid;
                uint32_t pa, pb, pr;  // the child that waits (depth dd + 1)
                if (nl <= nr) {
                    pa = mid; pb = b; pr = dr + 1;
                    b = mid - 1;
                } else {
                    pa = a; pb = mid - 1; pr = dr;
                    a = mid; dr = dr + 1;
                }
                dd++;
                if (sp < LANE_STACK)
                    L.aux[(sp++) * SAH_BLOCK + tid] = (pa - t0) | ((pb - t0) << 5) | (dd << 10) | (pr << 15);
                else
                    atomicOr(s.err, 4u);
            }
        }
        __threadfence_block();
        __syncthreads();
        if (tid == 0) { L.nbig = L.nbig_next; L.nbig_next = 0; }
        lb = le;
        le = L.task_next;
        __syncthreads();
    }
}

// Number of treelets (occupied bins), into L.k.
WPBR_D uint32_t sah_treelets(const LbvhScratch& s, SahLds& L) {
    const int tid = threadIdx.x;
    if (tid == 0) L.k = 0;
    __syncthreads();
    if (tid < 128) atomicAdd(&L.k, (uint32_t)__popc(s.presence[tid]));
    __syncthreads();
    return L.k;
}

WPBR_D void sah_top_body(LbvhScratch& s) {
    __shared__ SahLds L;
    const int tid = threadIdx.x;
    const uint32_t k = sah_treelets(s, L);
    if (tid == 0) {
        L.task_next = 1;
        L.nbig_next = 0;
        s.task_i0[0] = 0;
        s.task_i1[0] = k - 1;
        s.task_dep[0] = 0;
        s.task_rt[0] = 0;
        L.nbig = k >= BLOCK_TASK_MIN;
    }
    for (uint32_t i = tid; i < k; i += SAH_BLOCK) L.perm[i] = (uint16_t)i;
    {
        // 6k floats as 3k float2 (t_box is 8-byte aligned), several loads in flight
        const float2* src = reinterpret_cast<const float2*>(s.t_box);
        float2* dst = reinterpret_cast<float2*>(L.tbox);
#pragma unroll 8
        for (uint32_t i = tid; i < 3 * k; i += SAH_BLOCK) dst[i] = src[i];
    }
    __threadfence_block();
    __syncthreads();
    const SahTasks T{s.task_i0, s.task_i1, s.task_dep, s.task_rt, s.upper_count};
    sah_run(s, L, T, 0, 1, LANE_TASK_MAX, SAH_CUT);
    // positions' treelets and root boxes, for the deferred ranges' waves
    for (uint32_t i = tid; i < k; i += SAH_BLOCK) {
        const uint32_t t = L.perm[i];
        s.g_perm[i] = t;
#pragma unroll
        for (int c = 0; c < 6; c++) s.gp_box[6 * (size_t)i + c] = L.tbox[6 * t + c];
    }
}

// Deferred range st of a scene, by the whole workgroup; its task list and node
// counter live in LDS.
struct SahSubLds {
    SahLds L;
    uint32_t r[4];
    uint32_t ti0[SAH_SUB_TASKS], ti1[SAH_SUB_TASKS], tdep[SAH_SUB_TASKS], trt[SAH_SUB_TASKS];
    uint32_t uctr;
};
WPBR_D void sah_sub_one(LbvhScratch& s, SahSubLds& S, uint32_t st) {
    SahLds& L = S.L;
    uint32_t* s_r = S.r;
    uint32_t* t_i0 = S.ti0;
    uint32_t* t_i1 = S.ti1;
    uint32_t* t_dep = S.tdep;
    uint32_t* t_rt = S.trt;
    uint32_t& s_uctr = S.uctr;
    const int tid = threadIdx.x;
    const SahTasks T{t_i0, t_i1, t_dep, t_rt, &s_uctr};
    const uint32_t base = 0;
    {
        if (tid == 0) {
            s_r[0] = s.sub_i0[st];
            s_r[1] = s.sub_i1[st];
            t_i0[base] = s_r[0];
            t_i1[base] = s_r[1];
            t_dep[base] = s.sub_dep[st];
            t_rt[base] = s.sub_rt[st];
            s_uctr = s.sub_ub[st];
            L.task_next = base + 1;
            L.nbig = 0;
            L.nbig_next = 0;
        }
        __syncthreads();
        const uint32_t i0 = s_r[0], i1 = s_r[1];
        for (uint32_t i = i0 + tid; i <= i1; i += SAH_BLOCK) {
            const uint32_t t = s.g_perm[i];
            L.perm[i] = (uint16_t)t;
            for (int c = 0; c < 6; c++) L.tbox[6 * t + c] = s.t_box[6 * (size_t)t + c];
        }
        __threadfence_block();
        __syncthreads();
        sah_run(s, L, T, base, base + 1, LANE_TASK_MAX_SUB, 0);
        for (uint32_t i = i0 + tid; i <= i1; i += SAH_BLOCK) {
            s.g_perm[i] = L.perm[i];
            s.g_la[i] = L.la[i];
            s.g_lb[i] = L.lb[i];
        }
        __syncthreads();
    }
}

// The deferred ranges of one scene, workgroup bid of nblk taking every nblk-th.
WPBR_D void sah_sub_body(LbvhScratch& s, uint32_t bid, uint32_t nblk) {
    __shared__ SahSubLds S;
    const uint32_t nsub = *s.sub_n;
    for (uint32_t st = bid; st < nsub; st += nblk) sah_sub_one(s, S, st);
}

// Preorder layout: P = exclusive prefix of treelet node counts over final
// positions (PER consecutive positions per thread), in aux[0..k]; then every
// treelet's first node and every upper node at their closed-form offsets.
// Every workgroup computes the prefix sums; workgroup bid of nblk writes every
// nblk-th slice of the treelets and upper nodes.
constexpr uint32_t SAH_FINAL_BLOCKS = 16;
WPBR_D void sah_final_body(LbvhScratch& s, uint32_t bid, uint32_t nblk) {
    __shared__ SahLds L;
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t k = sah_treelets(s, L);
#pragma unroll 8
    for (uint32_t i = tid; i < k; i += SAH_BLOCK) {
        L.perm[i] = (uint16_t)s.g_perm[i];
        L.la[i] = (uint16_t)s.g_la[i];
        L.lb[i] = (uint16_t)s.g_lb[i];
    }
    __threadfence_block();
    __syncthreads();
    constexpr int PER = 4096 / SAH_BLOCK;
    const uint32_t num_upper = *s.upper_count;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        uint32_t i = PER * tid + q;
        v[q] = (i < k) ? s.t_nodes[L.perm[i]] : 0u;
        sum += v[q];
    }
    uint32_t incl = sum;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t x = (uint32_t)__shfl_up((int)incl, d);
        if (lane >= d) incl += x;
    }
    if (lane == 63) L.wcnt[tid >> 6] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < (tid >> 6); w++) before += L.wcnt[w];
    uint32_t run = before + incl - sum;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        uint32_t i = PER * tid + q;
        if (i < k) L.aux[i] = run;
        run += v[q];
    }
    if (tid == 0) {
        uint32_t total = 0;
        for (int w = 0; w < SAH_WAVES; w++) total += L.wcnt[w];
        L.aux[k] = total;
    }
    __threadfence_block();
    __syncthreads();
    for (uint32_t i = bid * SAH_BLOCK + tid; i < k; i += nblk * SAH_BLOCK)
        s.t_final[L.perm[i]] = (uint32_t)L.la[i] + (i - (uint32_t)L.lb[i]) + L.aux[i];
#pragma unroll 2
    for (uint32_t u = bid * SAH_BLOCK + tid; u < num_upper; u += nblk * SAH_BLOCK) {
        const uint32_t i0 = s.u_i0[u], mid = s.u_mid[u];
        const uint32_t off = s.u_dep[u] + (i0 - s.u_rt[u]) + L.aux[i0];
        const uint32_t size0 = (mid - 1 - i0) + (L.aux[mid] - L.aux[i0]);
        s.node_offset[off] = off + 1 + size0;
        s.node_nshapes[off] = 0;
        s.node_axis[off] = (uint8_t)s.u_axis[u];
        store_box(s.node_bounds + 6 * (size_t)off, load_box(s.u_box + 6 * (size_t)u));
    }
    if (tid == 0 && bid == 0) {
        *s.n_nodes = num_upper + L.aux[k];
        *s.n_treelets = k;
    }
}

__global__ __launch_bounds__(SAH_BLOCK) void k_sah_top(LbvhScratch s) { sah_top_body(s); }
__global__ __launch_bounds__(SAH_BLOCK) void k_sah_sub(LbvhScratch s) { sah_sub_body(s, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(SAH_BLOCK) void k_sah_final(LbvhScratch s) { sah_final_body(s, blockIdx.x, gridDim.x); }

// ---------------------------------------------------------------------------
// Flatten: treelet records to final preorder slots.  One workgroup per treelet.
WPBR_D void scatter_body(const LbvhScratch& s, uint32_t t) {
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
__global__ __launch_bounds__(256) void k_scatter(LbvhScratch s) { scatter_body(s, blockIdx.x); }

// ---------------------------------------------------------------------------
// Traversal layout (derived from the reference node arrays; the reference arrays
// themselves are unchanged).  Record r (64 bytes):
//   [0..5] child0 box, [6..11] child1 box, [12] child0 word, [13] child1 word,
//   [14] split axis (child0 lies below child1 on it; the near child is child1
//   when dirIsNeg[axis])
// A child word is either a record index (bit 31 clear) or a primitive range:
// 0x80000000 | (count - 1) << 28 | first slot (count <= 4).
// Records 0 .. n_nodes-1 are the reference nodes: interior nodes hold their two
// children; a reference leaf with more than 4 shapes holds the root of a binary
// tree over its slots, whose other records are appended after n_nodes.
// The slots of a leaf are in Morton order, so halving the slot range is a
// spatial split: node (L, j) of a leaf (off, n) covers slots
// [off + j*n/2^L, off + (j+1)*n/2^L) and has children (L+1, 2j), (L+1, 2j+1).
// Primitive records sit at their slot (position in the reference's
// shapesEntities) and keep their leaf's first slot, so ties inside one
// reference leaf still go to the lowest slot, as the reference's candidate
// order does (SLBVH.cpp:110-126).
constexpr uint32_t RANGE_MAX = 4;
constexpr uint32_t BIG_LEAF = 64;  // larger leaves (all codes equal) get a workgroup

WPBR_D uint32_t range_word(uint32_t first, uint32_t count) {
    return 0x80000000u | ((count - 1) << 28) | first;
}

WPBR_D uint32_t sub_lo(uint32_t n, uint32_t L, uint32_t j) {
    return (uint32_t)(((uint64_t)j * n) >> L);
}

// Deepest level of a leaf's sub-tree that holds an interior node.
WPBR_D uint32_t sub_levels(uint32_t n) {
    uint32_t L = 0;
    while (((n + (2u << L) - 1) >> (L + 1)) > RANGE_MAX) L++;
    return L;
}

WPBR_D uint32_t sub_record(uint32_t root, uint32_t base, uint32_t L, uint32_t j) {
    return L == 0 ? root : base + (1u << L) - 2 + j;
}

WPBR_D bbox slots_box(const LbvhScratch& s, uint32_t a, uint32_t b) {
    bbox r, unused;
    slot_range_boxes(s.slot_bounds, a, b, b, r, unused);
    return r;
}

WPBR_D void write_rec(const LbvhScratch& s, uint32_t r, const bbox& b0, const bbox& b1, uint32_t w0,
                      uint32_t w1, uint32_t axis) {
    float4* rec = reinterpret_cast<float4*>(s.trav) + 4 * (size_t)r;
    rec[0] = make_float4(b0.mn[0], b0.mn[1], b0.mn[2], b0.mx[0]);
    rec[1] = make_float4(b0.mx[1], b0.mx[2], b1.mn[0], b1.mn[1]);
    rec[2] = make_float4(b1.mn[2], b1.mx[0], b1.mx[1], b1.mx[2]);
    rec[3] = make_float4(__uint_as_float(w0), __uint_as_float(w1), __uint_as_float(axis), 0.0f);
}

// Record r from two child boxes and words when the split has no axis of its
// own: children ordered along the axis on which their centres differ most.
WPBR_D void write_pair(const LbvhScratch& s, uint32_t r, bbox b0, bbox b1, uint32_t w0, uint32_t w1) {
    float d[3];
#pragma unroll
    for (int c = 0; c < 3; c++) d[c] = (b1.mn[c] + b1.mx[c]) - (b0.mn[c] + b0.mx[c]);
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    uint32_t axis = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    float da = axis == 0 ? d[0] : (axis == 1 ? d[1] : d[2]);
    if (da < 0.0f) write_rec(s, r, b1, b0, w1, w0, axis);
    else write_rec(s, r, b0, b1, w0, w1, axis);
}


// Record numbers: 0 .. nn - 1 are the reference nodes; the sub-tree records of
// a reference leaf at slots [off, off + n) are nn + off .. nn + off + n - 6
// (a sub-tree whose records each hold more than 4 slots has at most n - 4
// records, its root being the leaf's own record).  Leaves' slot ranges are
// disjoint, so no counter is needed; numbers of slots of small leaves stay
// unused.
//
// Leaves of at most 4 shapes are a primitive range of their parent's record;
// leaves of 5 .. BIG_LEAF shapes get one wave each (k_subtree_leaf); larger
// ones (only when all their codes are equal) a workgroup (k_subtree_big).
constexpr uint32_t SUBTREE_BLOCKS = 8192;

// One thread per reference node (grid-stride over the node count): interior
// nodes get their record (both child boxes and words; a child leaf of at most
// 4 shapes is a primitive range), small leaves write their first slot into
// their primitive records, the largest leaves are queued.
WPBR_D void subtree_body(const LbvhScratch& s, uint32_t bid, uint32_t nblk) {
    const uint32_t nn = *s.n_nodes;
    for (uint32_t i = bid * blockDim.x + threadIdx.x; i < nn; i += nblk * blockDim.x) {
        const uint32_t n = s.node_nshapes[i];
        if (n == 0) {
            uint32_t c[2] = {i + 1, s.node_offset[i]};
            uint32_t w[2];
            for (int k = 0; k < 2; k++) {
                uint32_t ns = s.node_nshapes[c[k]];
                w[k] = (ns > 0 && ns <= RANGE_MAX) ? range_word(s.node_offset[c[k]], ns) : c[k];
            }
            const float* b0 = s.node_bounds + 6 * (size_t)c[0];
            const float* b1 = s.node_bounds + 6 * (size_t)c[1];
            float4* rec = reinterpret_cast<float4*>(s.trav) + 4 * (size_t)i;
            rec[0] = make_float4(b0[0], b0[1], b0[2], b0[3]);
            rec[1] = make_float4(b0[4], b0[5], b1[0], b1[1]);
            rec[2] = make_float4(b1[2], b1[3], b1[4], b1[5]);
            rec[3] = make_float4(__uint_as_float(w[0]), __uint_as_float(w[1]),
                                 __uint_as_float((uint32_t)s.node_axis[i]), 0.0f);
        } else if (n <= RANGE_MAX) {
            const uint32_t off = s.node_offset[i];
            for (uint32_t k = 0; k < n; k++) set_prim_leaf(s, off + k, off);
        } else if (n > BIG_LEAF) {
            s.big_list[atomicAdd(s.big_count, 1u)] = i;
        }
    }
}
__global__ __launch_bounds__(256) void k_subtree(LbvhScratch s) { subtree_body(s, blockIdx.x, gridDim.x); }

// Component-wise choice of a box (no aggregate copies in divergent code).
WPBR_D bbox box_sel(bool c, const bbox& a, const bbox& b) {
    bbox r;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        r.mn[e] = c ? a.mn[e] : b.mn[e];
        r.mx[e] = c ? a.mx[e] : b.mx[e];
    }
    return r;
}

// Contents of a binary record: child boxes, child words, split axis.
struct BinRec {
    bbox b0, b1;
    uint32_t w0, w1, axis;
};

// Children of a split without an axis of its own, ordered along the axis on
// which their centres differ most (as write_pair); `swapped` when child 1 of
// the arguments comes first.
WPBR_D BinRec pair_rec(const bbox& b0, const bbox& b1, uint32_t w0, uint32_t w1, bool& swapped) {
    float d[3];
#pragma unroll
    for (int c = 0; c < 3; c++) d[c] = (b1.mn[c] + b1.mx[c]) - (b0.mn[c] + b0.mx[c]);
    const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    const uint32_t axis = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    const float da = axis == 0 ? d[0] : (axis == 1 ? d[1] : d[2]);
    swapped = da < 0.0f;
    BinRec q;
    q.b0 = box_sel(swapped, b1, b0);
    q.b1 = box_sel(swapped, b0, b1);
    q.w0 = swapped ? w1 : w0;
    q.w1 = swapped ? w0 : w1;
    q.axis = axis;
    return q;
}

WPBR_D BinRec load_rec(const LbvhScratch& s, uint32_t r) {
    const float4* g = reinterpret_cast<const float4*>(s.trav) + 4 * (size_t)r;
    const float4 g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
    BinRec q;
    q.b0.mn[0] = g0.x; q.b0.mn[1] = g0.y; q.b0.mn[2] = g0.z; q.b0.mx[0] = g0.w; q.b0.mx[1] = g1.x; q.b0.mx[2] = g1.y;
    q.b1.mn[0] = g1.z; q.b1.mn[1] = g1.w; q.b1.mn[2] = g2.x; q.b1.mx[0] = g2.y; q.b1.mx[1] = g2.z; q.b1.mx[2] = g2.w;
    q.w0 = __float_as_uint(g3.x);
    q.w1 = __float_as_uint(g3.y);
    q.axis = __float_as_uint(g3.z);
    return q;
}

// 4-wide record r: the grandchildren of binary record q (a child that is a
// primitive range stays one child; k0 / k1 are the binary records of q's
// children that are records).  Slots 0,1 come from child 0 and 2,3 from child
// 1; the last word packs the axes (bits 0-1 record r, 2-3 child 0, 4-5 child
// 1).  Visiting the pair of the near child first and, inside each pair, its
// near child first is the binary depth-first order.  Empty slots have the word
// 0xFFFFFFFF.
// One child's pair of slots: its two child boxes, words and axis when it is a
// record, else the child itself and an empty slot.
struct Rec4Pair {
    bbox x, y;
    uint32_t wx, wy, axis;
};
WPBR_D Rec4Pair rec4_pair(uint32_t cw, const bbox& cb, const BinRec& k) {
    const bool rec = !(cw & 0x80000000u);
    Rec4Pair p;
    bbox zero;
#pragma unroll
    for (int e = 0; e < 3; e++) { zero.mn[e] = 0.0f; zero.mx[e] = 0.0f; }
    p.x = box_sel(rec, k.b0, cb);
    p.y = box_sel(rec, k.b1, zero);
    p.wx = rec ? k.w0 : cw;
    p.wy = rec ? k.w1 : 0xFFFFFFFFu;
    p.axis = rec ? (k.axis & 3u) : 0u;
    return p;
}
WPBR_D void put_rec4(const LbvhScratch& s, uint32_t r, const BinRec& q, const BinRec& k0, const BinRec& k1) {
    const Rec4Pair p0 = rec4_pair(q.w0, q.b0, k0), p1 = rec4_pair(q.w1, q.b1, k1);
    const uint32_t axes = (q.axis & 3u) | (p0.axis << 2) | (p1.axis << 4);
    float4* o = reinterpret_cast<float4*>(s.trav4) + 8 * (size_t)r;
    o[0] = make_float4(p0.x.mn[0], p0.x.mn[1], p0.x.mn[2], p0.x.mx[0]);
    o[1] = make_float4(p0.x.mx[1], p0.x.mx[2], p0.y.mn[0], p0.y.mn[1]);
    o[2] = make_float4(p0.y.mn[2], p0.y.mx[0], p0.y.mx[1], p0.y.mx[2]);
    o[3] = make_float4(p1.x.mn[0], p1.x.mn[1], p1.x.mn[2], p1.x.mx[0]);
    o[4] = make_float4(p1.x.mx[1], p1.x.mx[2], p1.y.mn[0], p1.y.mn[1]);
    o[5] = make_float4(p1.y.mn[2], p1.y.mx[0], p1.y.mx[1], p1.y.mx[2]);
    o[6] = make_float4(__uint_as_float(p0.wx), __uint_as_float(p0.wy), __uint_as_float(p1.wx),
                       __uint_as_float(p1.wy));
    o[7] = make_float4(__uint_as_float(axes), 0.0f, 0.0f, 0.0f);
}

// Per wave of k_subtree_leaf: a sparse table of the slot boxes of the wave's
// leaf (or of its two leaves, one per 32-lane half, each half using its own
// lanes' entries): level k, entry j = union of slots j .. j + 2^k - 1 of the
// leaf, clipped to it; and per split the lanes of its children that are splits.
struct LeafLds {
    float2 st[6][64][3];
    uint8_t kid[64][2];
};

// Union of slots [x, y) (0 < y - x < 64) of the leaf whose table entries start
// at lane `base`: two overlapping power-of-two blocks (min/max are idempotent,
// so the overlap is exact).
WPBR_D bbox leaf_range_box(const LeafLds& W, uint32_t base, uint32_t x, uint32_t y) {
    const uint32_t k = 31 - __clz(y - x);
    const float2* p = W.st[k][base + x];
    const float2* q = W.st[k][base + y - (1u << k)];
    bbox a, b;
    a.mn[0] = p[0].x; a.mn[1] = p[0].y; a.mn[2] = p[1].x; a.mx[0] = p[1].y; a.mx[1] = p[2].x; a.mx[2] = p[2].y;
    b.mn[0] = q[0].x; b.mn[1] = q[0].y; b.mn[2] = q[1].x; b.mx[0] = q[1].y; b.mx[1] = q[2].x; b.mx[2] = q[2].y;
    return box_union(a, b);
}

// A range of equal codes [a0, a0 + len) (len > 4) is split in halves level by
// level: node (l, t) covers [a0 + lo(l, t), a0 + lo(l, t + 1)) with
// lo(l, t) = floor(t len / 2^l); the root is record `root`, node (l, t) of
// level l > 0 is record base + 2^l - 2 + t ((2 << levels) - 2 of them).
WPBR_D BinRec run_node(const LeafLds& W, uint32_t lb, uint32_t off, uint32_t a0, uint32_t len, uint32_t l,
                       uint32_t t, uint32_t root, uint32_t base, bool& swapped) {
    const uint32_t a = sub_lo(len, l, t), b = sub_lo(len, l, t + 1), m = sub_lo(len, l + 1, 2 * t + 1);
    const uint32_t w0 = (m - a <= RANGE_MAX) ? range_word(off + a0 + a, m - a) : sub_record(root, base, l + 1, 2 * t);
    const uint32_t w1 =
        (b - m <= RANGE_MAX) ? range_word(off + a0 + m, b - m) : sub_record(root, base, l + 1, 2 * t + 1);
    return pair_rec(leaf_range_box(W, lb, a0 + a, a0 + m), leaf_range_box(W, lb, a0 + m, a0 + b), w0, w1, swapped);
}

WPBR_D uint32_t leaf_run_count(uint32_t len) { return (2u << sub_levels(len)) - 2; }

// Binary and 4-wide records of a run's sub-tree (one lane).
WPBR_D void leaf_run_records(const LbvhScratch& s, const LeafLds& W, uint32_t lb, uint32_t off, uint32_t a0,
                             uint32_t len, uint32_t root, uint32_t base) {
    const uint32_t Lmax = sub_levels(len);
    for (uint32_t l = 0; l <= Lmax; l++) {
        for (uint32_t t = 0; t < (1u << l); t++) {
            const uint32_t a = sub_lo(len, l, t), b = sub_lo(len, l, t + 1);
            if (b - a <= RANGE_MAX) continue;
            const uint32_t m = sub_lo(len, l + 1, 2 * t + 1);
            bool sw, sw0, sw1;
            const BinRec q = run_node(W, lb, off, a0, len, l, t, root, base, sw);
            BinRec c0 = q, c1 = q;
            if (m - a > RANGE_MAX) c0 = run_node(W, lb, off, a0, len, l + 1, 2 * t, root, base, sw0);
            if (b - m > RANGE_MAX) c1 = run_node(W, lb, off, a0, len, l + 1, 2 * t + 1, root, base, sw1);
            const uint32_t r = sub_record(root, base, l, t);
            write_rec(s, r, q.b0, q.b1, q.w0, q.w1, q.axis);
            put_rec4(s, r, q, sw ? c1 : c0, sw ? c0 : c1);
        }
    }
}

// The sub-tree of a reference leaf of 5 .. WD shapes (slots off .. off + n - 1,
// Morton-ordered) by WD lanes, lane j holding slot j: binary and 4-wide records
// of all its nodes.  WD = 32 (16) runs two (four) leaves per wave, one per
// group of WD lanes (`i` is the group's leaf, or none).  It is the tree emitLBVH would continue below its leaf
// size (SLBVH.cpp:374-401): a range whose codes differ splits at its highest
// differing bit (child 0 below child 1 on axis bit % 3); a range of equal codes
// splits in halves; ranges of at most 4 slots are primitive ranges.  The splits
// at differing bits are found all at once: boundary j (between slots j - 1 and
// j) with d_j = clz(code_{j-1} ^ code_j) < 32 is the split of range [a, b), a
// and b the nearest boundaries on either side with a smaller d (or the leaf's
// ends), and its parent is the one of those two with the larger d.  The root is
// record i.  The sub-tree shape inside a reference leaf does not change any
// hit: candidates of one leaf are ranked by slot.
constexpr uint32_t NO_LEAF = 0xFFFFFFFFu;
template <int WD>
WPBR_D void subtree_leaf(const LbvhScratch& s, const uint32_t* __restrict__ codes, LeafLds& W, uint32_t i,
                         uint32_t nn) {
    constexpr uint64_t HMASK = WD == 64 ? ~0ull : ((1ull << WD) - 1ull);
    const uint32_t lane = lane_id(), j = lane % WD, lb = lane - j, sh = lb;  // lb: the half's first lane
    const bool has = i != NO_LEAF;
    const uint32_t n = has ? s.node_nshapes[i] : 0u, off = has ? s.node_offset[i] : 0u;
    const bool valid = j < n;
    const uint32_t c = valid ? codes[off + j] : 0u;
    bbox B = valid ? slot_box(s.slot_bounds, off + j) : box_empty();
    if (valid) set_prim_leaf(s, off + j, off);
#pragma unroll
    for (int k = 0; k < 6; k++) {
        if ((1 << k) >= 2 * WD) break;
        W.st[k][lane][0] = make_float2(B.mn[0], B.mn[1]);
        W.st[k][lane][1] = make_float2(B.mn[2], B.mx[0]);
        W.st[k][lane][2] = make_float2(B.mx[1], B.mx[2]);
        if ((2 << k) < WD) {
            const int d = 1 << k;
            bbox o;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                o.mn[q] = __shfl_down(B.mn[q], d, WD);
                o.mx[q] = __shfl_down(B.mx[q], d, WD);
            }
            if (j + d < (uint32_t)WD) B = box_union(B, o);
        }
    }
    const uint32_t cprev = (uint32_t)__shfl_up((int)c, 1, WD);
    const bool bnd = j >= 1 && j < n;
    const int delta = bnd ? ((c == cprev) ? 32 : __clz(c ^ cprev)) : 64;
    const bool kar = delta < 32;
    // nearest boundaries with a smaller delta: one ballot per delta value
    // present (values below the smallest one present have no lanes)
    uint32_t a = 0, b = n;
    int vmin = kar ? delta : 32;
    for (int d = 32; d >= 1; d >>= 1) vmin = min(vmin, __shfl_xor(vmin, d));
    for (int v = vmin; v < 32; v++) {
        const uint64_t same = ballot64(kar && delta == v);
        if (!same) continue;
        const uint64_t smaller = (ballot64(kar && delta < v) >> sh) & HMASK;
        if (kar && delta == v) {
            const uint64_t left = smaller & ((1ull << j) - 1ull);
            const uint64_t right = smaller & ~((2ull << j) - 1ull);
            a = left ? 63u - (uint32_t)__clzll((long long)left) : 0u;
            b = right ? (uint32_t)(__ffsll((unsigned long long)right) - 1) : n;
        }
    }
    const bool whole_run = has && n > RANGE_MAX && ((ballot64(kar) >> sh) & HMASK) == 0;
    const bool rec = kar && b - a > RANGE_MAX;
    const bool root = kar && a == 0 && b == n;
    // children that are runs of equal codes longer than a primitive range
    const uint32_t ca = (uint32_t)__shfl((int)c, (int)a, WD);
    const uint32_t cb1 = (uint32_t)__shfl((int)c, (int)((b - 1) & (WD - 1)), WD);
    const bool runL = rec && j - a > RANGE_MAX && ca == cprev;
    const bool runR = rec && b - j > RANGE_MAX && c == cb1;
    uint32_t cnt = (rec && !root) ? 1u : 0u;
    if (runL) cnt += 1 + leaf_run_count(j - a);
    if (runR) cnt += 1 + leaf_run_count(b - j);
    if (whole_run && j == 0) cnt = leaf_run_count(n);
    // record numbers: prefix of cnt over the leaf's lanes, from its block nn + off
    uint32_t excl;
    if (ballot64(cnt > 1u) == 0) {
        excl = (uint32_t)__popcll((long long)((ballot64(cnt != 0u) >> sh) & HMASK & ((1ull << j) - 1ull)));
    } else {
        uint32_t incl = cnt;
        for (int d = 1; d < WD; d <<= 1) {
            const uint32_t x = (uint32_t)__shfl_up((int)incl, d, WD);
            if (j >= (uint32_t)d) incl += x;
        }
        excl = incl - cnt;
    }
    uint32_t next = nn + off + excl;
    if (whole_run && j == 0) leaf_run_records(s, W, lb, off, 0, n, i, next);
    const uint32_t r = root ? i : (rec ? next++ : 0u);
    uint32_t rootL = 0, rootR = 0, baseL = 0, baseR = 0;
    if (runL) { rootL = next++; baseL = next; next += leaf_run_count(j - a); }
    if (runR) { rootR = next++; baseR = next; next += leaf_run_count(b - j); }
    // a split tells its parent split its lane
    const int sa = __shfl(delta, (int)(a & (WD - 1)), WD), sb = __shfl(delta, (int)(b & (WD - 1)), WD);
    W.kid[lane][0] = (uint8_t)j;
    W.kid[lane][1] = (uint8_t)j;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (rec && !root) {
        const int da = a == 0 ? -1 : sa;
        const int db = b == n ? -1 : sb;
        const bool parent_left = da > db;  // the parent splits at a: this is its child 1
        W.kid[lb + (parent_left ? a : b)][parent_left ? 1 : 0] = (uint8_t)j;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t k0 = W.kid[lane][0], k1 = W.kid[lane][1];
    // the child splits' record numbers, ranges and axes
    const uint32_t r0 = (uint32_t)__shfl((int)r, (int)k0, WD), r1 = (uint32_t)__shfl((int)r, (int)k1, WD);
    const uint32_t a0 = (uint32_t)__shfl((int)a, (int)k0, WD), b0 = (uint32_t)__shfl((int)b, (int)k0, WD);
    const uint32_t a1 = (uint32_t)__shfl((int)a, (int)k1, WD), b1 = (uint32_t)__shfl((int)b, (int)k1, WD);
    const int d0 = __shfl(delta, (int)k0, WD), d1 = __shfl(delta, (int)k1, WD);
    BinRec q;
    q.w0 = (j - a <= RANGE_MAX) ? range_word(off + a, j - a) : (runL ? rootL : r0);
    q.w1 = (b - j <= RANGE_MAX) ? range_word(off + j, b - j) : (runR ? rootR : r1);
    q.axis = (uint32_t)((31 - delta) % 3);
    // the child splits' words
    const uint32_t w00 = (uint32_t)__shfl((int)q.w0, (int)k0, WD), w01 = (uint32_t)__shfl((int)q.w1, (int)k0, WD);
    const uint32_t w10 = (uint32_t)__shfl((int)q.w0, (int)k1, WD), w11 = (uint32_t)__shfl((int)q.w1, (int)k1, WD);
    if (!rec) return;
    q.b0 = leaf_range_box(W, lb, a, j);
    q.b1 = leaf_range_box(W, lb, j, b);
    BinRec c0 = q, c1 = q;
    if (!(q.w0 & 0x80000000u)) {
        if (runL) {
            bool sw;
            c0 = run_node(W, lb, off, a, j - a, 0, 0, rootL, baseL, sw);
        } else {
            c0.b0 = leaf_range_box(W, lb, a0, k0);
            c0.b1 = leaf_range_box(W, lb, k0, b0);
            c0.w0 = w00;
            c0.w1 = w01;
            c0.axis = (uint32_t)((31 - d0) % 3);
        }
    }
    if (!(q.w1 & 0x80000000u)) {
        if (runR) {
            bool sw;
            c1 = run_node(W, lb, off, j, b - j, 0, 0, rootR, baseR, sw);
        } else {
            c1.b0 = leaf_range_box(W, lb, a1, k1);
            c1.b1 = leaf_range_box(W, lb, k1, b1);
            c1.w0 = w10;
            c1.w1 = w11;
            c1.axis = (uint32_t)((31 - d1) % 3);
        }
    }
    write_rec(s, r, q.b0, q.b1, q.w0, q.w1, q.axis);
    put_rec4(s, r, q, c0, c1);
    if (runL) leaf_run_records(s, W, lb, off, a, j - a, rootL, baseL);
    if (runR) leaf_run_records(s, W, lb, off, j, b - j, rootR, baseR);
}

// Reference leaves of 5 .. BIG_LEAF shapes: every wave scans LEAF_CHUNK
// consecutive nodes (grid-stride); leaves of at most 16 shapes go four per wave,
// of at most 32 two per wave, larger ones one per wave.
constexpr uint32_t LEAF_WAVES = 4;
constexpr uint32_t LEAF_BLOCKS = 2048;
constexpr uint32_t LEAF_CHUNK = 32;
WPBR_D void subtree_leaf_body(const LbvhScratch& s, const uint32_t* __restrict__ codes, uint32_t bid, uint32_t nblk) {
    __shared__ LeafLds W[LEAF_WAVES];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t nn = *s.n_nodes;
    for (uint32_t c0 = (bid * LEAF_WAVES + wave) * LEAF_CHUNK; c0 < nn; c0 += nblk * LEAF_WAVES * LEAF_CHUNK) {
        const uint32_t i = c0 + lane;
        const uint32_t ns = (lane < LEAF_CHUNK && i < nn) ? s.node_nshapes[i] : 0u;
        uint64_t m4 = ballot64(ns > RANGE_MAX && ns <= 16);
        uint64_t m2 = ballot64(ns > 16 && ns <= 32);
        uint64_t m1 = ballot64(ns > 32 && ns <= BIG_LEAF);
        while (m4) {
            uint32_t mine = NO_LEAF;
            for (uint32_t q = 0; q < 4 && m4; q++) {
                const uint32_t k = (uint32_t)(__ffsll((unsigned long long)m4) - 1);
                m4 &= m4 - 1;
                if ((lane >> 4) == q) mine = c0 + k;
            }
            subtree_leaf<16>(s, codes, W[wave], mine, nn);
            __builtin_amdgcn_wave_barrier();
        }
        while (m2) {
            const uint32_t ka = (uint32_t)(__ffsll((unsigned long long)m2) - 1);
            m2 &= m2 - 1;
            uint32_t kb = NO_LEAF;
            if (m2) {
                kb = (uint32_t)(__ffsll((unsigned long long)m2) - 1);
                m2 &= m2 - 1;
            }
            const uint32_t mine = lane < 32 ? c0 + ka : (kb == NO_LEAF ? NO_LEAF : c0 + kb);
            subtree_leaf<32>(s, codes, W[wave], mine, nn);
            __builtin_amdgcn_wave_barrier();
        }
        while (m1) {
            const uint32_t k = (uint32_t)(__ffsll((unsigned long long)m1) - 1);
            m1 &= m1 - 1;
            subtree_leaf<64>(s, codes, W[wave], c0 + k, nn);
            __builtin_amdgcn_wave_barrier();
        }
    }
}
__global__ __launch_bounds__(64 * LEAF_WAVES, 4) void k_subtree_leaf(LbvhScratch s, const uint32_t* __restrict__ codes) {
    subtree_leaf_body(s, codes, blockIdx.x, gridDim.x);
}

WPBR_D bbox record_union(const LbvhScratch& s, uint32_t r) {
    const float4* rec = reinterpret_cast<const float4*>(s.trav) + 4 * (size_t)r;
    float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
    bbox a, b;
    a.mn[0] = r0.x; a.mn[1] = r0.y; a.mn[2] = r0.z; a.mx[0] = r0.w; a.mx[1] = r1.x; a.mx[2] = r1.y;
    b.mn[0] = r1.z; b.mn[1] = r1.w; b.mn[2] = r2.x; b.mx[0] = r2.y; b.mx[1] = r2.z; b.mx[2] = r2.w;
    return box_union(a, b);
}

// Leaves of more than BIG_LEAF shapes (only when every Morton code in the leaf
// is equal): one workgroup each, levels bottom-up.
WPBR_D void subtree_big_body(const LbvhScratch& s, uint32_t bid,
                            uint32_t nblk) {
    const uint32_t count = *s.big_count;
    const uint32_t nn = *s.n_nodes;
    __shared__ uint32_t s_base;
    for (uint32_t e = bid; e < count; e += nblk) {
        const uint32_t i = s.big_list[e];
        const uint32_t n = s.node_nshapes[i], off = s.node_offset[i];
        const uint32_t Lmax = sub_levels(n);
        if (threadIdx.x == 0) s_base = nn + off;
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) set_prim_leaf(s, off + k, off);
        __syncthreads();
        const uint32_t base = s_base;
        for (int L = (int)Lmax; L >= 0; L--) {
            for (uint32_t j = threadIdx.x; j < (1u << L); j += blockDim.x) {
                uint32_t a = sub_lo(n, L, j), b = sub_lo(n, L, j + 1);
                if (b - a <= RANGE_MAX) continue;
                uint32_t m = sub_lo(n, L + 1, 2 * j + 1);
                uint32_t w0, w1;
                bbox b0, b1;
                if (m - a <= RANGE_MAX) {
                    w0 = range_word(off + a, m - a);
                    b0 = slots_box(s, off + a, off + m);
                } else {
                    w0 = sub_record(i, base, L + 1, 2 * j);
                    b0 = record_union(s, w0);
                }
                if (b - m <= RANGE_MAX) {
                    w1 = range_word(off + m, b - m);
                    b1 = slots_box(s, off + m, off + b);
                } else {
                    w1 = sub_record(i, base, L + 1, 2 * j + 1);
                    b1 = record_union(s, w1);
                }
                write_pair(s, sub_record(i, base, L, j), b0, b1, w0, w1);
            }
            __threadfence_block();
            __syncthreads();
        }
    }
}
__global__ __launch_bounds__(256) void k_subtree_big(LbvhScratch s) {
    subtree_big_body(s, blockIdx.x, gridDim.x);
}

// 4-wide records of the records not made by k_subtree_leaf: the interior
// reference nodes and the sub-trees of the largest leaves, from their binary
// records (the layout of put_rec4).
WPBR_D void bvh4_record(const LbvhScratch& s, uint32_t r) {
    const BinRec q = load_rec(s, r);
    BinRec k0 = q, k1 = q;
    if (!(q.w0 & 0x80000000u)) k0 = load_rec(s, q.w0);
    if (!(q.w1 & 0x80000000u)) k1 = load_rec(s, q.w1);
    put_rec4(s, r, q, k0, k1);
}
constexpr uint32_t BVH4_BLOCKS = 8192;
WPBR_D void bvh4_body(const LbvhScratch& s, uint32_t bid, uint32_t nblk) {
    const uint32_t nn = *s.n_nodes;
    for (uint32_t r = bid * blockDim.x + threadIdx.x; r < nn; r += nblk * blockDim.x) {
        const uint32_t ns = s.node_nshapes[r];
        if (ns == 0 || ns > BIG_LEAF) bvh4_record(s, r);
    }
    const uint32_t nbig = *s.big_count;
    for (uint32_t e = bid; e < nbig; e += nblk) {
        const uint32_t i = s.big_list[e];
        const uint32_t cnt = (2u << sub_levels(s.node_nshapes[i])) - 2, base = nn + s.node_offset[i];
        for (uint32_t q = threadIdx.x; q < cnt; q += blockDim.x) bvh4_record(s, base + q);
    }
}
__global__ __launch_bounds__(256) void k_bvh4(LbvhScratch s) { bvh4_body(s, blockIdx.x, gridDim.x); }

// ===========================================================================
// Reference key order (WPBR_LBVH_ORDER_REFERENCE).  The sort keeps bits 0..23
// (three 8-bit passes), so treelets are runs of equal top-12 bits in that order
// and number up to one per shape.  The kernels below take the build's scratch
// through a device array (one entry per scene; block b of a launch with G
// blocks per scene works for scene b / G), so the single-scene and the
// segmented build share them.
//   kr_heads / kr_head_scan / kr_treelets   treelet starts by a scan of the
//                   run heads (SLBVH.cpp:299-319), the upper SAH's first range
//   kr_emit_small / kr_emit_big   treelets below 30 shapes are one leaf
//                   (SLBVH.cpp:356-367), larger ones get emitLBVH by a workgroup
//   kr_lv_stats / kr_lv_ranges / kr_lv_lists / kr_lv_swap
//                   one level of the upper SAH (SLBVH.cpp:457-563) over every
//                   range of at least REF_CUT roots: bucket statistics (and each
//                   chunk's bucket counts), the split (cost, first minimum,
//                   midpoint fallback) one 16-lane row per range, and the
//                   libstdc++ two-pointer std::partition in rank form (the r-th
//                   misplaced position left of the split swaps with the r-th
//                   misplaced position right of it counted from the right)
//   kr_sub_order / kr_sub_wave   ranges below REF_CUT, largest first, one
//                   wave each in LDS
//   kr_fin_*        preorder layout with a device-wide prefix of the treelet
//                   node counts; kr_scatter_ref writes the treelet nodes
// Ranges and chunks of level k live in list k & 1; the level's split pushes
// the next level's ranges into the other list.

enum {
    RG_I0 = 0, RG_I1, RG_DEP, RG_RT,
    RG_CB,             // 6 order-preserving uints: centroid bounds min xyz, max xyz
    RG_CH = RG_CB + 6, // first chunk
    RG_NCH,            // chunks
    RG_BEST,           // split bucket
    RG_L,              // positions left of the split (= size: midpoint fallback)
    RG_ML,             // misplaced pairs
    RG_MID,
    RG_UB              // first upper node of the range's subtree (a range of m roots makes m - 1)
};
static_assert(RG_UB < (int)REF_RG_WORDS, "range record");
// Statistics: per bucket 13 words (count, box min xyz / max xyz, centroid min
// xyz / max xyz as order-preserving uints), then the range's box.
constexpr uint32_t ST_BOX = NB * 13;
static_assert(ST_BOX + 6 <= REF_ST_WORDS, "statistics record");
static_assert(REF_CH_WORDS == (uint32_t)NB, "a chunk's bucket counts");

#define REF_SCENE(S, G)                                   \
    const uint32_t sc = blockIdx.x / (G);                 \
    const uint32_t rj = blockIdx.x - sc * (G);            \
    const LbvhScratch& s = (S)[sc];

WPBR_D uint32_t* ref_rec(const LbvhScratch& s, uint32_t list, uint32_t r) {
    return s.lv_rec + ((size_t)list * s.cap_r + r) * REF_RG_WORDS;
}
WPBR_D uint32_t* ref_stat(const LbvhScratch& s, uint32_t list, uint32_t r) {
    return s.lv_stat + ((size_t)list * s.cap_r + r) * REF_ST_WORDS;
}
// Initial value of statistics word k: 0 for counts, the empty box otherwise.
WPBR_D uint32_t stat_init_word(uint32_t k) {
    if (k < ST_BOX) {
        const uint32_t j = k % 13;
        if (j == 0) return 0u;
        return (((j - 1) / 3) & 1u) ? f2ord(FLT_LOWEST_) : f2ord(FLT_MAX_);
    }
    if (k < ST_BOX + 3) return f2ord(FLT_MAX_);
    if (k < ST_BOX + 6) return f2ord(FLT_LOWEST_);
    return 0u;
}
WPBR_D bool stat_is_max(uint32_t k) {
    return k < ST_BOX ? ((((k % 13) - 1) / 3) & 1u) != 0 : k >= ST_BOX + 3;
}
WPBR_D bbox ord_box(const uint32_t* w) {
    bbox b;
#pragma unroll
    for (int c = 0; c < 3; c++) { b.mn[c] = ord2f(w[c]); b.mx[c] = ord2f(w[3 + c]); }
    return b;
}
// Root-box centroid of a treelet on every axis (pMin + pMax) * 0.5 (SLBVH.cpp:472).
WPBR_D f3 box_cen(const bbox& b) {
    return mk3((float)((double)(b.mn[0] + b.mx[0]) * 0.5), (float)((double)(b.mn[1] + b.mx[1]) * 0.5),
               (float)((double)(b.mn[2] + b.mx[2]) * 0.5));
}

// Block-wide exclusive scan of one value per thread (REF_BLOCK threads);
// *total receives the sum.  Every thread must call it.
constexpr int REF_BLOCK = 256;
WPBR_D uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t x = (uint32_t)__shfl_up((int)incl, d);
        if (lane >= d) incl += x;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
        const uint32_t x = s_w[k];
        before += (k < w) ? x : 0u;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return before + incl - v;
}

WPBR_D bool ref_head(const uint32_t* codes, uint32_t i, uint32_t n) {
    return i < n && (i == 0 || (codes[i] >> 18) != (codes[i - 1] >> 18));
}

// Treelet heads per 256 slots.
__global__ __launch_bounds__(REF_BLOCK) void kr_heads(const LbvhScratch* S, uint32_t G) {
    REF_SCENE(S, G);
    __shared__ uint32_t s_w[REF_BLOCK / 64];
    const uint32_t n = s.n, nblk = (n + REF_BLOCK - 1) / REF_BLOCK;
    for (uint32_t b = rj; b < nblk; b += G) {
        const bool h = ref_head(s.keys_a, b * REF_BLOCK + threadIdx.x, n);
        const uint64_t m = ballot64(h);
        if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = popc64(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t c = 0;
            for (int w = 0; w < REF_BLOCK / 64; w++) c += s_w[w];
            s.head_blk[b] = c;
        }
        __syncthreads();
    }
}

// One workgroup per scene: exclusive prefix of the head counts, the treelet
// count, the upper SAH's counters and its first range (the whole position
// range, as a level range, a deferred LDS range, or nothing for one treelet).
__global__ __launch_bounds__(1024) void kr_head_scan(const LbvhScratch* S) {
    const LbvhScratch& s = S[blockIdx.x];
    __shared__ uint32_t s_sum[1024];
    __shared__ uint32_t s_run;
    const uint32_t nblk = (s.n + REF_BLOCK - 1) / REF_BLOCK;
    const int tid = threadIdx.x;
    if (tid == 0) s_run = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nblk; base += 1024) {
        const uint32_t i = base + tid;
        const uint32_t v = i < nblk ? s.head_blk[i] : 0u;
        s_sum[tid] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t x = tid >= off ? s_sum[tid - off] : 0u;
            __syncthreads();
            s_sum[tid] += x;
            __syncthreads();
        }
        if (i < nblk) s.head_blk[i] = s_run + s_sum[tid] - v;
        __syncthreads();
        if (tid == 0) s_run += s_sum[1023];
        __syncthreads();
    }
    const uint32_t T = s_run;
    if (tid == 0) {
        *s.n_treelets = T;
        *s.bigt_n = 0;
        *s.sub_n = 0;
        // upper nodes are numbered by range: a range of m roots owns m - 1
        // consecutive numbers from its first (the root range's is 0)
        *s.upper_count = T > 1 ? T - 1 : 0;
        for (int k = 0; k < 4; k++) s.lv_n[k] = 0;
        if (T == 1) {
            s.g_la[0] = 0;
            s.g_lb[0] = 0;
        } else if (T < REF_CUT) {
            s.sub_i0[0] = 0;
            s.sub_i1[0] = T - 1;
            s.sub_dep[0] = 0;
            s.sub_rt[0] = 0;
            s.sub_ub[0] = 0;
            *s.sub_n = 1;
        } else {
            uint32_t* rec = ref_rec(s, 0, 0);
            rec[RG_I0] = 0;
            rec[RG_I1] = T - 1;
            rec[RG_DEP] = 0;
            rec[RG_RT] = 0;
            for (int c = 0; c < 3; c++) {
                rec[RG_CB + c] = f2ord(FLT_MAX_);
                rec[RG_CB + 3 + c] = f2ord(FLT_LOWEST_);
            }
            rec[RG_CH] = 0;
            rec[RG_NCH] = (T + REF_CHUNK - 1) / REF_CHUNK;
            rec[RG_UB] = 0;
            s.lv_n[0] = 1;
            s.lv_n[2] = rec[RG_NCH];
        }
    }
    if (T >= REF_CUT) {
        const uint32_t nch = (T + REF_CHUNK - 1) / REF_CHUNK;
        for (uint32_t c = tid; c < nch; c += 1024) s.ch_range[c] = 0;
        uint32_t* st = ref_stat(s, 0, 0);
        for (uint32_t k = tid; k < REF_ST_WORDS; k += 1024) st[k] = stat_init_word(k);
    }
}

// Treelet starts: t_start[rank of the head] = slot.
__global__ __launch_bounds__(REF_BLOCK) void kr_treelets(const LbvhScratch* S, uint32_t G) {
    REF_SCENE(S, G);
    __shared__ uint32_t s_w[REF_BLOCK / 64];
    const uint32_t n = s.n, nblk = (n + REF_BLOCK - 1) / REF_BLOCK;
    for (uint32_t b = rj; b < nblk; b += G) {
        const uint32_t i = b * REF_BLOCK + threadIdx.x;
        const bool h = ref_head(s.keys_a, i, n);
        uint32_t tot;
        const uint32_t r = block_excl_scan(h ? 1u : 0u, s_w, &tot);
        if (h) s.t_start[s.head_blk[b] + r] = i;
    }
}

// Centroid bounds of the treelet roots (the first level range's): every block
// of kr_emit_small and kr_emit_big writes its partial min/max (order-preserving
// uints) to cb_ord[8 + 6 k] (k < G for kr_emit_small's G blocks per scene,
// G + k for kr_emit_big's), kr_root_cb reduces the 2 G partials.  The k_bounds
// partials there are no longer needed by then.  G is at most REF_CB_PARTS.
constexpr uint32_t REF_CB_PARTS = 1024;
struct CbAcc {
    uint32_t v[6];
    WPBR_D CbAcc() { v[0] = v[1] = v[2] = 0xFFFFFFFFu; v[3] = v[4] = v[5] = 0u; }
    WPBR_D void add(const f3& c) {
        const uint32_t o[3] = {f2ord(c.x), f2ord(c.y), f2ord(c.z)};
#pragma unroll
        for (int k = 0; k < 3; k++) { v[k] = min(v[k], o[k]); v[3 + k] = max(v[3 + k], o[k]); }
    }
};
// Block reduction of the accumulators; thread 0 writes partial slot `part`.
WPBR_D void cb_block_write(const LbvhScratch& s, CbAcc a, uint32_t part, uint32_t* s_red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
#pragma unroll
    for (int k = 0; k < 3; k++)
        for (int off = 32; off >= 1; off >>= 1) {
            a.v[k] = min(a.v[k], (uint32_t)__shfl_xor((int)a.v[k], off));
            a.v[3 + k] = max(a.v[3 + k], (uint32_t)__shfl_xor((int)a.v[3 + k], off));
        }
    if (lane == 0)
        for (int k = 0; k < 6; k++) s_red[6 * w + k] = a.v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t r = s_red[k];
        for (int q = 1; q < nw; q++) r = k < 3 ? min(r, s_red[6 * q + k]) : max(r, s_red[6 * q + k]);
        s.cb_ord[8 + 6 * (size_t)part + k] = r;
    }
    __syncthreads();
}

// Treelets below 30 shapes: one leaf each (emitLBVH's first case,
// SLBVH.cpp:356-367); larger ones are listed for kr_emit_big.  Positions start
// as treelet order.
__global__ __launch_bounds__(REF_BLOCK, 7) void kr_emit_small(const LbvhScratch* S, uint32_t G) {
    REF_SCENE(S, G);
    __shared__ uint32_t s_red[6 * REF_BLOCK / 64];
    const uint32_t T = *s.n_treelets, n = s.n;
    CbAcc acc;
    for (uint32_t t = rj * REF_BLOCK + threadIdx.x; t < T; t += G * REF_BLOCK) {
        {
            const uint32_t start = s.t_start[t];
            const uint32_t end = t + 1 < T ? s.t_start[t + 1] : n;
            const uint32_t cnt = end - start;
            if (cnt < (uint32_t)MAX_PRIMS_IN_LEAF) {
                bbox b, unused;
                slot_range_boxes(s.slot_bounds, start, end, end, b, unused);
                const uint32_t base = 2 * start - t;
                s.r_start[base] = start;
                s.r_n[base] = cnt;
                s.r_meta[base] = 0x80000000u;
                s.r_size[base] = 1;
                s.r_off[base] = 0;
                store_box(s.r_box + 6 * (size_t)base, b);
                s.t_count[t] = cnt;
                s.t_slot[t] = base;
                s.t_nodes[t] = 1;
                store_box(s.t_box + 6 * (size_t)t, b);
                s.g_perm[t] = t;
                store_box(s.gp_box + 6 * (size_t)t, b);
                acc.add(box_cen(b));
            } else {
                s.bigt[atomicAdd(s.bigt_n, 1u)] = t;
            }
        }
    }
    cb_block_write(s, acc, rj, s_red);
}

// Treelets of at least 30 shapes: emitLBVH by one workgroup each.
// trees = 0: the trees were built by ks_emit_trees (segmented build); only the
// per-treelet epilogue and the centroid partials remain.
__global__ __launch_bounds__(EMIT_BLOCK) void kr_emit_big(const LbvhScratch* S, uint32_t G, uint32_t trees) {
    REF_SCENE(S, G);
    __shared__ uint32_t s_red[6];
    const uint32_t nbig = *s.bigt_n, T = *s.n_treelets, n = s.n;
    CbAcc acc;
    for (uint32_t q = rj; q < nbig; q += G) {
        const uint32_t t = s.bigt[q];
        const uint32_t start = s.t_start[t];
        const uint32_t end = t + 1 < T ? s.t_start[t + 1] : n;
        if (trees) emit_tree(s, s.keys_a, t, start, end - start);
        __syncthreads();
        if (threadIdx.x == 0) {
            const bbox b = load_box(s.t_box + 6 * (size_t)t);
            s.g_perm[t] = t;
            store_box(s.gp_box + 6 * (size_t)t, b);
            acc.add(box_cen(b));
        }
        __syncthreads();
    }
    cb_block_write(s, acc, G + rj, s_red);
}

// The first level range's centroid bounds from the 2 G partials.
__global__ __launch_bounds__(64) void kr_root_cb(const LbvhScratch* S, uint32_t G) {
    const LbvhScratch& s = S[blockIdx.x];
    if (*s.n_treelets < REF_CUT) return;
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (uint32_t p = threadIdx.x; p < 2 * G; p += 64)
        for (int k = 0; k < 6; k++) {
            const uint32_t x = s.cb_ord[8 + 6 * (size_t)p + k];
            v[k] = k < 3 ? min(v[k], x) : max(v[k], x);
        }
#pragma unroll
    for (int k = 0; k < 3; k++)
        for (int off = 32; off >= 1; off >>= 1) {
            v[k] = min(v[k], (uint32_t)__shfl_xor((int)v[k], off));
            v[3 + k] = max(v[3 + k], (uint32_t)__shfl_xor((int)v[3 + k], off));
        }
    if (threadIdx.x == 0) {
        uint32_t* rec = ref_rec(s, 0, 0);
        for (int k = 0; k < 6; k++) rec[RG_CB + k] = v[k];
    }
}

// ---- one level of the upper SAH over the large ranges ----------------------
struct RefRange {
    uint32_t r, i0, i1, ch0, nch, p0, p1;
    const uint32_t* rec;
};
WPBR_D RefRange ref_range(const LbvhScratch& s, uint32_t list, uint32_t c) {
    RefRange R;
    R.r = s.ch_range[(size_t)list * s.cap_c + c];
    R.rec = ref_rec(s, list, R.r);
    R.i0 = R.rec[RG_I0];
    R.i1 = R.rec[RG_I1];
    R.ch0 = R.rec[RG_CH];
    R.nch = R.rec[RG_NCH];
    R.p0 = R.i0 + (c - R.ch0) * REF_CHUNK;
    R.p1 = min(R.p0 + REF_CHUNK, R.i1 + 1);
    return R;
}

// Bucket of every position and the bucket statistics of every range
// (SLBVH.cpp:465-500): counts, bucket boxes, bucket centroid bounds (for the
// children's centroid bounds) and the range's box.
// The workgroup keeps ST_COPIES copies of the statistics, lane l adding to copy
// l % ST_COPIES: neighbouring positions mostly share a bucket, and the copies
// spread their LDS atomics over ST_COPIES addresses.
constexpr uint32_t ST_COPIES = 8;
__global__ __launch_bounds__(REF_BLOCK) void kr_lv_stats(const LbvhScratch* S, uint32_t G, uint32_t level) {
    REF_SCENE(S, G);
    __shared__ uint32_t st[ST_COPIES * REF_ST_WORDS];
    const uint32_t cur = level & 1u;
    const uint32_t nch = s.lv_n[2 + cur];
    if (rj == 0 && threadIdx.x == 0) { s.lv_n[cur ^ 1u] = 0; s.lv_n[2 + (cur ^ 1u)] = 0; }
    uint32_t* const mine = st + (lane_id() % ST_COPIES) * REF_ST_WORDS;
    for (uint32_t c = rj; c < nch; c += G) {
        const RefRange R = ref_range(s, cur, c);
        for (uint32_t k = threadIdx.x; k < ST_COPIES * REF_ST_WORDS; k += REF_BLOCK)
            st[k] = stat_init_word(k % REF_ST_WORDS);
        const bbox cb = ord_box(R.rec + RG_CB);
        const int dim = box_max_extent(cb);
        const float cmin = dim == 0 ? cb.mn[0] : (dim == 1 ? cb.mn[1] : cb.mn[2]);
        const float dist = (dim == 0 ? cb.mx[0] : (dim == 1 ? cb.mx[1] : cb.mx[2])) - cmin;
        __syncthreads();
        uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
        // the chunk's boxes first, all loads in flight at once
        constexpr uint32_t PER = REF_CHUNK / REF_BLOCK;
        bbox bx[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) {
            const uint32_t p = R.p0 + threadIdx.x + j * REF_BLOCK;
            if (p < R.p1) bx[j] = load_box(s.gp_box + 6 * (size_t)p);
        }
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) {
            const uint32_t p = R.p0 + threadIdx.x + j * REF_BLOCK;
            if (p >= R.p1) break;
            const bbox& b = bx[j];
            const f3 cc = box_cen(b);
            const uint32_t ib = sah_bucket(comp(cc, dim), cmin, dist);
            s.gp_bucket[p] = (uint8_t)ib;
            uint32_t* sb = mine + 13 * ib;
            atomicAdd(&sb[0], 1u);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const uint32_t lo = f2ord(b.mn[k]), hi = f2ord(b.mx[k]), ce = f2ord(comp(cc, k));
                atomicMin(&sb[1 + k], lo);
                atomicMax(&sb[4 + k], hi);
                atomicMin(&sb[7 + k], ce);
                atomicMax(&sb[10 + k], ce);
                v[k] = min(v[k], lo);
                v[3 + k] = max(v[3 + k], hi);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; k++)
            for (int off = 32; off >= 1; off >>= 1) {
                v[k] = min(v[k], (uint32_t)__shfl_xor((int)v[k], off));
                v[3 + k] = max(v[3 + k], (uint32_t)__shfl_xor((int)v[3 + k], off));
            }
        if (lane_id() == 0)
            for (int k = 0; k < 3; k++) {
                atomicMin(&st[ST_BOX + k], v[k]);
                atomicMax(&st[ST_BOX + 3 + k], v[3 + k]);
            }
        __syncthreads();
        uint32_t* gs = ref_stat(s, cur, R.r);
        for (uint32_t k = threadIdx.x; k < ST_BOX + 6; k += REF_BLOCK) {
            const bool cnt = k < ST_BOX && k % 13 == 0, mx = stat_is_max(k);
            uint32_t x = st[k];
            for (uint32_t q = 1; q < ST_COPIES; q++) {
                const uint32_t y = st[q * REF_ST_WORDS + k];
                x = cnt ? x + y : (mx ? max(x, y) : min(x, y));
            }
            // the chunk's own bucket counts, for the partition's ranks
            if (cnt) s.ch_cnt[(size_t)c * REF_CH_WORDS + k / 13] = x;
            if (x == stat_init_word(k)) continue;
            if (cnt) atomicAdd(&gs[k], x);
            else if (mx) atomicMax(&gs[k], x);
            else atomicMin(&gs[k], x);
        }
        __syncthreads();
    }
}

// Exclusive prefixes of three counts over the wave, with bases for the whole
// wave taken from three counters (lanes 0..2 issue the three device atomics at
// once).  All 64 lanes must call it.
WPBR_D void wave_alloc3(uint32_t* c0, uint32_t* c1, uint32_t* c2, const uint32_t (&want)[3], uint32_t (&got)[3]) {
    const int lane = (int)lane_id();
    uint32_t incl[3] = {want[0], want[1], want[2]};
    for (int d = 1; d < 64; d <<= 1)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint32_t x = (uint32_t)__shfl_up((int)incl[k], d);
            if (lane >= d) incl[k] += x;
        }
    uint32_t tot[3];
#pragma unroll
    for (int k = 0; k < 3; k++) tot[k] = (uint32_t)__builtin_amdgcn_readlane((int)incl[k], 63);
    uint32_t b = 0;
    if (lane < 3) {
        uint32_t* ctr = lane == 0 ? c0 : (lane == 1 ? c1 : c2);
        const uint32_t t = lane == 0 ? tot[0] : (lane == 1 ? tot[1] : tot[2]);
        if (t) b = atomicAdd(ctr, t);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) got[k] = (uint32_t)__builtin_amdgcn_readlane((int)b, k) + incl[k] - want[k];
}

// Order-preserving words' inclusive min (MAX false) or max scan over the 16
// lanes of each row, up from lane 0 or down from lane 15 (row_scan's integer
// form).  All lanes must be active.
template <bool UP, bool MAX>
WPBR_D uint32_t row_scan_u(uint32_t v) {
    const int id = MAX ? 0 : -1;
#define WPBR_ROW_SCAN(d)                                                                                          \
    {                                                                                                             \
        const uint32_t x = (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, (UP ? 0x110 : 0x100) + (d), 0xF, 0xF, \
                                                                 false);                                          \
        v = MAX ? max(v, x) : min(v, x);                                                                          \
    }
    WPBR_ROW_SCAN(1);
    WPBR_ROW_SCAN(2);
    WPBR_ROW_SCAN(4);
    WPBR_ROW_SCAN(8);
#undef WPBR_ROW_SCAN
    return v;
}

// The split of every range (SLBVH.cpp:504-553), one 16-lane row per range:
// lane j < NB of the row holds bucket j's statistics; the side boxes and counts
// of the eleven splits are inclusive prefix and suffix scans over the row (DPP;
// min/max are exact, so the sides' boxes, areas and costs are those of the
// reference's per-split unions), and every lane of the row takes the first
// minimum of the costs in order and the midpoint fallback.  The row's first
// lane then writes the upper node (numbered from the range's first number: the
// left subtree's follow it, then the right subtree's) and the children: single
// positions, deferred ranges below REF_CUT, or next-level ranges, whose
// centroid bounds are the same scans over their buckets' centroid bounds.  List
// slots and chunks are taken with one atomic per counter and wave.
constexpr uint32_t LV_RPB = 16;  // ranges per workgroup step: four per wave
__global__ __launch_bounds__(REF_BLOCK) void kr_lv_ranges(const LbvhScratch* S, uint32_t G, uint32_t level) {
    REF_SCENE(S, G);
    const uint32_t cur = level & 1u, nxt = cur ^ 1u;
    const uint32_t nr = s.lv_n[cur];
    const uint32_t lane = threadIdx.x & 63u, row = lane >> 4, j = lane & 15u, wave = threadIdx.x >> 6;
    const uint32_t rbase = lane & ~15u;
    for (uint32_t r0 = rj * LV_RPB; r0 < nr; r0 += G * LV_RPB) {
        const uint32_t r = r0 + 4 * wave + row;
        const bool valid = r < nr;
        const uint32_t* gs = ref_stat(s, cur, valid ? r : 0u);
        const uint32_t* rec = ref_rec(s, cur, valid ? r : 0u);
        // bucket j of the row's range (identities for the row's other lanes)
        const bool has = valid && j < (uint32_t)NB;
        const uint32_t* bj = gs + 13 * (has ? j : 0u);
        const uint32_t cnt = has ? bj[0] : 0u;
        float mn[3], mx[3];
        uint32_t cw[6];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            mn[c] = has ? ord2f(bj[1 + c]) : FLT_MAX_;
            mx[c] = has ? ord2f(bj[4 + c]) : FLT_LOWEST_;
        }
#pragma unroll
        for (int q = 0; q < 6; q++) cw[q] = (has && cnt != 0u) ? bj[7 + q] : (q < 3 ? 0xFFFFFFFFu : 0u);
        const bbox bounds = ord_box(gs + ST_BOX);
        bbox b0, b1;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            b0.mn[c] = row_scan<true, false>(mn[c]);
            b0.mx[c] = row_scan<true, true>(mx[c]);
            b1.mn[c] = row_next(row_scan<false, false>(mn[c]), FLT_MAX_);
            b1.mx[c] = row_next(row_scan<false, true>(mx[c]), FLT_LOWEST_);
        }
        const uint32_t c0 = row_scan_add<true>(cnt);
        const uint32_t c1 = row_next(row_scan_add<false>(cnt));
        const float la = (c0 == 0) ? 0.0f : (float)c0 * box_area(b0);
        const float ra = (c1 == 0) ? 0.0f : (float)c1 * box_area(b1);
        const float cost = (float)(0.125 + (double)((la + ra) / box_area(bounds)));
        // first minimum over splits 0 .. NB - 2, in order
        uint32_t best = 0;
        float cbest = __shfl(cost, (int)rbase);
        for (uint32_t i = 1; i < (uint32_t)NB - 1; i++) {
            const float ci = __shfl(cost, (int)(rbase + i));
            if (ci < cbest) { best = i; cbest = ci; }
        }
        const uint32_t L = (uint32_t)__shfl((int)c0, (int)(rbase + best));
        // the children's centroid bounds: buckets 0 .. best and best + 1 .. NB - 1
        uint32_t ccb[2][6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const uint32_t up = q < 3 ? row_scan_u<true, false>(cw[q]) : row_scan_u<true, true>(cw[q]);
            const uint32_t dn = q < 3 ? row_scan_u<false, false>(cw[q]) : row_scan_u<false, true>(cw[q]);
            ccb[0][q] = (uint32_t)__shfl((int)up, (int)(rbase + best));
            ccb[1][q] = (uint32_t)__shfl((int)dn, (int)(rbase + best + 1));
        }
        const bool lead = valid && j == 0;
        uint32_t i0 = 0, i1 = 0, mid = 0, ub = 0, dep = 0, rt = 0;
        bool fallback = false;
        if (lead) {
            i0 = rec[RG_I0];
            i1 = rec[RG_I1];
            dep = rec[RG_DEP];
            rt = rec[RG_RT];
            ub = rec[RG_UB];
            const uint32_t m = i1 - i0 + 1;
            fallback = L == m;
            mid = fallback ? m / 2 + i0 : i0 + L;
            s.u_i0[ub] = i0;
            s.u_i1[ub] = i1;
            s.u_mid[ub] = mid;
            s.u_dep[ub] = dep;
            s.u_rt[ub] = rt;
            s.u_axis[ub] = (uint32_t)box_max_extent(ord_box(rec + RG_CB));
            store_box(s.u_box + 6 * (size_t)ub, bounds);
            uint32_t* grec = ref_rec(s, cur, r);
            grec[RG_BEST] = best;
            grec[RG_L] = L;
            grec[RG_MID] = mid;
            grec[RG_ML] = 0;
        }
        // children: kind 0 none / single position, 1 deferred, 2 next level
        uint32_t kind[2] = {0, 0}, cn[2] = {0, 0};
        if (lead)
            for (int k = 0; k < 2; k++) {
                const uint32_t a = k == 0 ? i0 : mid, b = k == 0 ? mid - 1 : i1;
                const uint32_t size = b - a + 1;
                if (size == 1) {
                    s.g_la[a] = dep + 1;
                    s.g_lb[a] = rt + (uint32_t)k;
                } else if (size < REF_CUT) {
                    kind[k] = 1;
                } else {
                    kind[k] = 2;
                    cn[k] = (size + REF_CHUNK - 1) / REF_CHUNK;
                }
            }
        const uint32_t want[3] = {(uint32_t)(kind[0] == 1) + (uint32_t)(kind[1] == 1),
                                  (uint32_t)(kind[0] == 2) + (uint32_t)(kind[1] == 2), cn[0] + cn[1]};
        uint32_t got[3];
        wave_alloc3(s.sub_n, &s.lv_n[nxt], &s.lv_n[2 + nxt], want, got);
        if (lead) {
            uint32_t ne = got[0], nrr = got[1], nc = got[2];
            for (int k = 0; k < 2; k++) {
                if (!kind[k]) continue;
                const uint32_t a = k == 0 ? i0 : mid, b = k == 0 ? mid - 1 : i1;
                const uint32_t cub = k == 0 ? ub + 1 : ub + (mid - i0);
                const uint32_t cdep = dep + 1, crt = rt + (uint32_t)k;
                if (kind[k] == 1) {
                    s.sub_i0[ne] = a;
                    s.sub_i1[ne] = b;
                    s.sub_dep[ne] = cdep;
                    s.sub_rt[ne] = crt;
                    s.sub_ub[ne] = cub;
                    ne++;
                    continue;
                }
                uint32_t* cr = ref_rec(s, nxt, nrr);
                cr[RG_I0] = a;
                cr[RG_I1] = b;
                cr[RG_DEP] = cdep;
                cr[RG_RT] = crt;
                cr[RG_CH] = nc;
                cr[RG_NCH] = cn[k];
                cr[RG_UB] = cub;
                // the child's centroid bounds: the union of its buckets' (all
                // centroids coincide when the midpoint fallback splits)
#pragma unroll
                for (int q = 0; q < 6; q++) cr[RG_CB + q] = fallback ? rec[RG_CB + q] : ccb[k][q];
                nrr++;
                nc += cn[k];
            }
        }
    }
}

// The next level's chunk-to-range map and zeroed statistics (its ranges are
// written by kr_lv_ranges; its first kernel reads them).
WPBR_D void ref_next_level_init(const LbvhScratch& s, uint32_t nxt, uint32_t rj, uint32_t G) {
    const uint32_t nr2 = s.lv_n[nxt];
    for (uint32_t r = rj; r < nr2; r += G) {
        const uint32_t* cr = ref_rec(s, nxt, r);
        const uint32_t c0 = cr[RG_CH], cn = cr[RG_NCH];
        for (uint32_t q = threadIdx.x; q < cn; q += REF_BLOCK) s.ch_range[(size_t)nxt * s.cap_c + c0 + q] = r;
        uint32_t* st = ref_stat(s, nxt, r);
        for (uint32_t q = threadIdx.x; q < REF_ST_WORDS; q += REF_BLOCK) st[q] = stat_init_word(q);
    }
}


// Ranks of the misplaced positions: pl_a[i0 + r] = the r-th misplaced position
// left of the split (from the left), pl_b[i0 + r] = the r-th misplaced
// position right of it (from the right).
constexpr uint32_t REF_PER = REF_CHUNK / REF_BLOCK;
__global__ __launch_bounds__(REF_BLOCK) void kr_lv_lists(const LbvhScratch* S, uint32_t G, uint32_t level) {
    REF_SCENE(S, G);
    __shared__ uint32_t s_w[REF_BLOCK / 64];
    __shared__ uint32_t s_base[2];
    const uint32_t cur = level & 1u;
    const uint32_t nch = s.lv_n[2 + cur];
    for (uint32_t c = rj; c < nch; c += G) {
        const RefRange R = ref_range(s, cur, c);
        const uint32_t m = R.i1 - R.i0 + 1, L = R.rec[RG_L], best = R.rec[RG_BEST];
        if (L == m) continue;  // uniform over the workgroup
        // misplaced counts of the range's chunks before (left) and after (right)
        // this one: a chunk wholly left of the split counts its positions in
        // buckets above best, one wholly right of it those at or below best
        // (from its bucket counts); the chunk holding the split is counted
        // position by position
        const uint32_t split = R.i0 + L;
        uint32_t bl = 0, br = 0;
        for (uint32_t q = R.ch0 + threadIdx.x; q < R.ch0 + R.nch; q += REF_BLOCK) {
            if (q == c) continue;
            const uint32_t qp0 = R.i0 + (q - R.ch0) * REF_CHUNK, qp1 = min(qp0 + REF_CHUNK, R.i1 + 1);
            const uint32_t* h = s.ch_cnt + (size_t)q * REF_CH_WORDS;
            if (q < c && qp1 <= split)
                for (uint32_t b = best + 1; b < REF_CH_WORDS; b++) bl += h[b];
            if (q > c && qp0 >= split)
                for (uint32_t b = 0; b <= best; b++) br += h[b];
        }
        const uint32_t sc0 = R.ch0 + (split - R.i0) / REF_CHUNK;  // the chunk holding the split
        if (sc0 != c && sc0 < R.ch0 + R.nch) {
            const uint32_t sp0 = R.i0 + (sc0 - R.ch0) * REF_CHUNK, sp1 = min(sp0 + REF_CHUNK, R.i1 + 1);
            if (sp0 < split) {
                for (uint32_t p = sp0 + threadIdx.x; p < sp1; p += REF_BLOCK) {
                    const uint32_t ib = s.gp_bucket[p];
                    if (sc0 < c && p < split) bl += ib > best;
                    if (sc0 > c && p >= split) br += ib <= best;
                }
            }
        }
        uint32_t tbl, tbr;
        (void)block_excl_scan(bl, s_w, &tbl);
        (void)block_excl_scan(br, s_w, &tbr);
        const uint32_t q0 = R.p0 + threadIdx.x * REF_PER;
        uint32_t fl = 0, fr = 0;  // flags of this thread's positions
#pragma unroll
        for (uint32_t j = 0; j < REF_PER; j++) {
            const uint32_t p = q0 + j;
            if (p < R.p1) {
                const uint32_t ib = s.gp_bucket[p];
                if (p < split) fl |= (uint32_t)(ib > best) << j;
                else fr |= (uint32_t)(ib <= best) << j;
            }
        }
        uint32_t tl, tr;
        const uint32_t el = block_excl_scan(__popc(fl), s_w, &tl);
        const uint32_t er = block_excl_scan(__popc(fr), s_w, &tr);
        uint32_t kl = 0, kr = 0;
#pragma unroll
        for (uint32_t j = 0; j < REF_PER; j++) {
            const uint32_t p = q0 + j;
            if ((fl >> j) & 1u) s.pl_a[R.i0 + tbl + el + kl++] = p;
            if ((fr >> j) & 1u) {
                kr++;
                // right-misplaced positions after p: in later chunks, then in this one
                s.pl_b[R.i0 + tbr + (tr - (er + kr))] = p;
            }
        }
        if (c == R.ch0 + R.nch - 1 && threadIdx.x == 0) const_cast<uint32_t*>(R.rec)[RG_ML] = tbl + tl;
    }
}

// The swaps of the two-pointer partition: pair r exchanges pl_a[i0 + r] and
// pl_b[i0 + r] (treelet and root box).
__global__ __launch_bounds__(REF_BLOCK) void kr_lv_swap(const LbvhScratch* S, uint32_t G, uint32_t level) {
    REF_SCENE(S, G);
    const uint32_t cur = level & 1u;
    const uint32_t nch = s.lv_n[2 + cur];
    ref_next_level_init(s, cur ^ 1u, rj, G);
    for (uint32_t c = rj; c < nch; c += G) {
        const RefRange R = ref_range(s, cur, c);
        const uint32_t m = R.i1 - R.i0 + 1, L = R.rec[RG_L];
        if (L == m) continue;
        const uint32_t ml = R.rec[RG_ML];
        for (uint32_t q = R.p0 + threadIdx.x; q < R.p1; q += REF_BLOCK) {
            if (q - R.i0 >= ml) break;
            const uint32_t a = s.pl_a[q], b = s.pl_b[q];
            const uint32_t ta = s.g_perm[a], tb = s.g_perm[b];
            s.g_perm[a] = tb;
            s.g_perm[b] = ta;
            const bbox ba = load_box(s.gp_box + 6 * (size_t)a), bb = load_box(s.gp_box + 6 * (size_t)b);
            store_box(s.gp_box + 6 * (size_t)a, bb);
            store_box(s.gp_box + 6 * (size_t)b, ba);
        }
    }
}

// Sum of the ranges of the next level over all scenes (the host's loop ends
// when it is 0).
__global__ void kr_active(const LbvhScratch* S, uint32_t n_scenes, uint32_t list, uint32_t* out) {
    __shared__ uint32_t s_t;
    if (threadIdx.x == 0) s_t = 0;
    __syncthreads();
    uint32_t v = 0;
    for (uint32_t i = threadIdx.x; i < n_scenes; i += blockDim.x) v += S[i].lv_n[list];
    atomicAdd(&s_t, v);
    __syncthreads();
    if (threadIdx.x == 0) *out = s_t;
}


// ---- ranges below REF_CUT: one wave each ---------------------------------------
// Each wave of a workgroup takes its own deferred range (at most WR_CAP roots)
// in its own LDS: a stack of pending tasks, split by the whole wave while
// larger than WR_LANE_MAX roots; the small tasks then go to one lane each, to
// completion.  Tasks are packed as i0 | i1 << 8 | depth << 16 | right turns << 24,
// all relative to the range (a range of m roots has depth below m).
constexpr uint32_t WR_CAP = 256;
constexpr uint32_t WR_WAVES = 4;
constexpr uint32_t WR_LANE_MAX = 8;
struct WaveRangeLds {
    static constexpr int SC = 1;  // bucket statistics copies
    alignas(16) float tbox[6 * WR_CAP];
    uint16_t perm[WR_CAP];
    uint8_t bucket[WR_CAP];
    uint16_t la[WR_CAP], lb[WR_CAP];     // partition lists; then leaf depth / right turns
    uint32_t stack[WR_CAP];              // pending tasks
    uint32_t small[WR_CAP];              // tasks for the lanes
    uint32_t aux[LANE_STACK * 64];       // the lanes' DFS stacks; then the range's old g_perm
    uint32_t bcount[SC][NB];
    uint32_t bmin[SC][NB][3];
    uint32_t bmax[SC][NB][3];
    uint32_t mid[16];                    // tasks of ROW_MIN..ROW_MAX roots waiting for a row split
    float rcost[4][16];                  // row splits: split costs and left counts per row
    uint32_t rcnt[4][16];
    uint32_t rout[4][2];                 // row splits: each row's two children
    uint32_t uctr;
};
static_assert(LANE_STACK * 64 >= WR_CAP, "aux holds a range's old g_perm");
static_assert(REF_CUT <= WR_CAP, "a deferred range fits a wave's LDS");

WPBR_D uint32_t wr_pack(uint32_t a, uint32_t b, uint32_t d, uint32_t r) { return a | (b << 8) | (d << 16) | (r << 24); }

// One small task (at most WR_LANE_MAX roots) by one lane, to completion: the
// larger child of a split waits on the lane's stack.
WPBR_D void wr_lane_task(LbvhScratch& s, WaveRangeLds& L, const SahTasks& T, uint32_t task, int lane) {
    const uint32_t t0 = task & 255u, t1 = (task >> 8) & 255u, d0 = (task >> 16) & 255u, r0 = task >> 24;
    uint32_t a = t0, b = t1, dd = 0, dr = 0;
    int sp = 0;
    while (true) {
        if (a == b) {
            L.la[a] = (uint16_t)(d0 + dd);
            L.lb[a] = (uint16_t)(r0 + dr);
            if (sp == 0) break;
            const uint32_t e = L.aux[(--sp) * 64 + lane];
            a = t0 + (e & 31u);
            b = t0 + ((e >> 5) & 31u);
            dd = (e >> 10) & 31u;
            dr = (e >> 15) & 31u;
            continue;
        }
        int dim;
        bbox bounds;
        const uint32_t mid = lane_split(L, a, b, &dim, &bounds);
        sah_node(s, T, a, b, mid, dim, bounds, d0 + dd, r0 + dr);
        const uint32_t nl = mid - a, nr = b + 1 - mid;
        uint32_t pa, pb, pr;  // the child that waits (depth dd + 1)
        if (nl <= nr) {
            pa = mid; pb = b; pr = dr + 1;
            b = mid - 1;
        } else {
            pa = a; pb = mid - 1; pr = dr;
            a = mid; dr = dr + 1;
        }
        dd++;
        if (sp < LANE_STACK)
            L.aux[(sp++) * 64 + lane] = (pa - t0) | ((pb - t0) << 5) | (dd << 10) | (pr << 15);
        else
            atomicOr(s.err, 4u);
    }
}

// Tasks of ROW_MIN..ROW_MAX roots are split four at a time, one per 16-lane
// row of the wave (a wave split of one such task spends most of its time in
// work that does not depend on its size): lane j of a row holds root j of the
// row's task; the range box and centroid bounds are reduced over the row by
// DPP rotations; lane s < 11 of the row sums split s's two sides over the
// task's roots in position order; lane 0 of the row takes the first minimum of
// the eleven costs in order (SLBVH.cpp:540-548) and the partition is the
// libstdc++ two-pointer swap in rank form, over row-masked ballots.
constexpr uint32_t ROW_MIN = WR_LANE_MAX + 1, ROW_MAX = 16;
template <bool MAX>
WPBR_D float row_allreduce(float v) {
    auto op = [](float a, float b) { return MAX ? smax(a, b) : smin(a, b); };
#define WPBR_ROW_ROR(n) v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + (n), 0xF, 0xF, false)))
    WPBR_ROW_ROR(8);
    WPBR_ROW_ROR(4);
    WPBR_ROW_ROR(2);
    WPBR_ROW_ROR(1);
#undef WPBR_ROW_ROR
    return v;
}
// Splits L.mid[base .. base + nt) (nt <= 4); row k's children go to L.rout[k].
WPBR_D void wr_row_split(LbvhScratch& s, WaveRangeLds& L, const SahTasks& T, uint32_t base, uint32_t nt) {
    const uint32_t lane = threadIdx.x & 63, row = lane >> 4, j = lane & 15;
    const bool act = row < nt;
    const uint32_t task = act ? L.mid[base + row] : 0u;
    const uint32_t a = task & 255u, b = (task >> 8) & 255u, d = (task >> 16) & 255u, r = task >> 24;
    const uint32_t m = act ? b - a + 1 : 0u;
    const bool mine = j < m;
    float v[12];
#pragma unroll
    for (int c = 0; c < 3; c++) { v[c] = FLT_MAX_; v[3 + c] = FLT_LOWEST_; v[6 + c] = FLT_MAX_; v[9 + c] = FLT_LOWEST_; }
    if (mine) {
        const uint32_t t = L.perm[a + j];
        const bbox rb = tb_box(L, t);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v[c] = rb.mn[c];
            v[3 + c] = rb.mx[c];
            v[6 + c] = v[9 + c] = tb_cen(L, t, c);
        }
    }
    bbox bounds, cb;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        bounds.mn[c] = row_allreduce<false>(v[c]);
        bounds.mx[c] = row_allreduce<true>(v[3 + c]);
        cb.mn[c] = row_allreduce<false>(v[6 + c]);
        cb.mx[c] = row_allreduce<true>(v[9 + c]);
    }
    const int dim = box_max_extent(cb);
    const float cbmn = dim == 0 ? cb.mn[0] : (dim == 1 ? cb.mn[1] : cb.mn[2]);
    const float dist = (dim == 0 ? cb.mx[0] : (dim == 1 ? cb.mx[1] : cb.mx[2])) - cbmn;
    if (mine) L.bucket[a + j] = (uint8_t)sah_bucket(tb_cen(L, L.perm[a + j], dim), cbmn, dist);
    __builtin_amdgcn_wave_barrier();
    if (act && j < NB - 1) {
        bbox b0 = box_empty(), b1 = box_empty();
        uint32_t c0 = 0, c1 = 0;
        for (uint32_t k = 0; k < m; k++) {
            const bbox tb = tb_box(L, L.perm[a + k]);
            if (L.bucket[a + k] <= j) { b0 = box_union(b0, tb); c0++; }
            else { b1 = box_union(b1, tb); c1++; }
        }
        const float l = (c0 == 0) ? 0.0f : (float)c0 * box_area(b0);
        const float rr = (c1 == 0) ? 0.0f : (float)c1 * box_area(b1);
        L.rcost[row][j] = (float)(0.125 + (double)((l + rr) / box_area(bounds)));
        L.rcnt[row][j] = c0;
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t best = 0;
    if (act && j == 0) {
        float cbest = L.rcost[row][0];
        for (uint32_t i = 1; i < NB - 1; i++)
            if (L.rcost[row][i] < cbest) { best = i; cbest = L.rcost[row][i]; }
        L.rout[row][0] = best;
    }
    __builtin_amdgcn_wave_barrier();
    best = act ? L.rout[row][0] : 0u;
    const uint32_t Tn = act ? L.rcnt[row][best] : 0u;
    const uint64_t rmask = 0xFFFFull << (16 * row);
    // misplaced positions: left of the split with a bucket above best, right of
    // it (counted from the right) with a bucket at or below best
    const bool part = act && Tn != m;
    const bool badl = part && j < Tn && L.bucket[a + j] > best;
    const bool badr = part && j < m - Tn && L.bucket[b - j] <= best;
    const uint64_t ml = ballot64(badl) & rmask, mr = ballot64(badr) & rmask;
    const uint64_t below = lanemask_lt() & rmask;
    if (badl) L.la[a + popc64(ml & below)] = (uint16_t)(a + j);
    if (badr) L.lb[a + popc64(mr & below)] = (uint16_t)(b - j);
    __builtin_amdgcn_wave_barrier();
    if (part && j < popc64(ml)) {
        const uint32_t pa = L.la[a + j], pb = L.lb[a + j];
        const uint16_t ta = L.perm[pa], tb = L.perm[pb];
        L.perm[pa] = tb;
        L.perm[pb] = ta;
    }
    __builtin_amdgcn_wave_barrier();
    if (act && j == 0) {
        const uint32_t midp = (Tn == m) ? m / 2 + a : a + Tn;
        sah_node(s, T, a, b, midp, dim, bounds, d, r);
        L.rout[row][0] = wr_pack(a, midp - 1, d + 1, r);
        L.rout[row][1] = wr_pack(midp, b, d + 1, r + 1);
    }
    __builtin_amdgcn_wave_barrier();
}

WPBR_D void wr_range(LbvhScratch& s, WaveRangeLds& L, uint32_t st) {
    const int lane = threadIdx.x & 63;
    const uint32_t i0 = s.sub_i0[st], m = s.sub_i1[st] - i0 + 1;
    SahTasks T{nullptr, nullptr, nullptr, nullptr, &L.uctr};
    T.pos_base = i0;
    T.dep_base = s.sub_dep[st];
    T.rt_base = s.sub_rt[st];
    if (m > WR_CAP) {  // the level kernels defer only ranges below REF_CUT
        if (lane == 0) atomicOr(s.err, 4u);
        return;
    }
    for (uint32_t q = lane; q < m; q += 64) {
        L.perm[q] = (uint16_t)q;
        for (int c = 0; c < 6; c++) L.tbox[6 * q + c] = s.gp_box[6 * (size_t)(i0 + q) + c];
    }
    if (lane == 0) {
        L.uctr = s.sub_ub[st];
        L.stack[0] = wr_pack(0, m - 1, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t sp = 1, nsmall = 0, nmid = 0;  // uniform across the wave
    while (sp || nsmall || nmid) {
        if (nmid >= 4 || (nmid && !sp)) {
            // the last (up to) four tasks of ROW_MIN..ROW_MAX roots, one per
            // row; their children go to the lane list or back to this list (a
            // task of at most ROW_MAX roots has at most one child above
            // WR_LANE_MAX, so the list never grows past seven)
            const uint32_t nt = nmid < 4 ? nmid : 4u, base = nmid - nt;
            wr_row_split(s, L, T, base, nt);
            uint32_t ns = nsmall, nm = base;
            for (uint32_t k = 0; k < 2 * nt; k++) {
                const uint32_t c = L.rout[k >> 1][k & 1];
                const uint32_t sz = ((c >> 8) & 255u) - (c & 255u) + 1;
                if (sz <= WR_LANE_MAX) {
                    if (lane == 0) L.small[ns] = c;
                    ns++;
                } else {
                    if (lane == 0) L.mid[nm] = c;
                    nm++;
                }
            }
            nsmall = ns;
            nmid = nm;
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        if (sp && nsmall < 64) {
            const uint32_t task = L.stack[--sp];
            const uint32_t a = task & 255u, b = (task >> 8) & 255u, d = (task >> 16) & 255u, r = task >> 24;
            __builtin_amdgcn_wave_barrier();
            if (b - a + 1 <= WR_LANE_MAX) {
                if (lane == 0) L.small[nsmall] = task;
                nsmall++;
            } else if (b - a + 1 <= ROW_MAX) {
                if (lane == 0) L.mid[nmid] = task;
                nmid++;
            } else {
                int dim;
                bbox bounds;
                const uint32_t mid = wave_split_mid(L, 0, a, b, &dim, &bounds);
                if (lane == 0) {
                    sah_node(s, T, a, b, mid, dim, bounds, d, r);
                    L.stack[sp] = wr_pack(a, mid - 1, d + 1, r);
                    L.stack[sp + 1] = wr_pack(mid, b, d + 1, r + 1);
                }
                sp += 2;
            }
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        // a batch of up to 64 small tasks, one per lane
        const uint32_t nb = nsmall < 64 ? nsmall : 64;
        if ((uint32_t)lane < nb) wr_lane_task(s, L, T, L.small[nsmall - 1 - lane], lane);
        nsmall -= nb;
        __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t q = lane; q < m; q += 64) L.aux[q] = s.g_perm[i0 + q];
    __builtin_amdgcn_wave_barrier();
    for (uint32_t q = lane; q < m; q += 64) {
        s.g_perm[i0 + q] = L.aux[L.perm[q]];
        s.g_la[i0 + q] = T.dep_base + L.la[q];
        s.g_lb[i0 + q] = T.rt_base + L.lb[q];
    }
    __builtin_amdgcn_wave_barrier();
}

// The deferred ranges of a scene in order of decreasing size (a counting sort
// on the size, 2 .. REF_CUT - 1, into pl_a, free once the level loop is done),
// so that the range waves take the longest first and finish together.  Ranges
// are disjoint and numbered in advance, so the order changes no result.
__global__ __launch_bounds__(1024) void kr_sub_order(const LbvhScratch* S) {
    __shared__ uint32_t h[REF_CUT];
    const LbvhScratch& s = S[blockIdx.x];
    const uint32_t nsub = *s.sub_n;
    for (uint32_t b = threadIdx.x; b < REF_CUT; b += blockDim.x) h[b] = 0;
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < nsub; e += blockDim.x) {
        const uint32_t m = s.sub_i1[e] - s.sub_i0[e] + 1;
        atomicAdd(&h[REF_CUT - 1 - min(m, REF_CUT - 1)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // exclusive prefix over the REF_CUT bins, one wave (4 bins per lane)
        constexpr uint32_t PB = REF_CUT / 64;
        const uint32_t l = threadIdx.x;
        uint32_t v[PB], t = 0;
#pragma unroll
        for (uint32_t q = 0; q < PB; q++) { v[q] = h[PB * l + q]; t += v[q]; }
        uint32_t incl = t;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t x = (uint32_t)__shfl_up((int)incl, d);
            if ((int)l >= d) incl += x;
        }
        uint32_t run = incl - t;
#pragma unroll
        for (uint32_t q = 0; q < PB; q++) { h[PB * l + q] = run; run += v[q]; }
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < nsub; e += blockDim.x) {
        const uint32_t m = s.sub_i1[e] - s.sub_i0[e] + 1;
        s.pl_a[atomicAdd(&h[REF_CUT - 1 - min(m, REF_CUT - 1)], 1u)] = e;
    }
}
static_assert(REF_CUT % 64 == 0, "size bins per lane");

// G workgroups per scene, WR_WAVES ranges in flight per workgroup.
__global__ __launch_bounds__(64 * WR_WAVES) void kr_sub_wave(const LbvhScratch* S, uint32_t G) {
    __shared__ WaveRangeLds W[WR_WAVES];
    const uint32_t sc = blockIdx.x / G, rj = blockIdx.x - sc * G;
    LbvhScratch s = S[sc];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nsub = *s.sub_n;
    (void)rj;
    // each wave takes the next range in kr_sub_order's order when it has
    // finished one (the scene's cursor is tile_counter[7], zeroed with the
    // build's counters; the sort passes use words 0..3)
    for (;;) {
        uint32_t k = 0;
        if (lane_id() == 0) k = atomicAdd(s.tile_counter + 7, 1u);
        k = (uint32_t)__builtin_amdgcn_readfirstlane((int)k);
        if (k >= nsub) break;
        wr_range(s, W[wave], s.pl_a[k]);
    }
}
// ---- preorder layout ---------------------------------------------------------
// P = exclusive prefix of t_nodes over final positions; a treelet at final
// position i with leaf depth d and r right turns starts at d + (i - r) + P[i],
// an upper node [i0, i1] split at mid at dep + (i0 - rt) + P[i0] with its
// second child after (mid - 1 - i0) + P[mid] - P[i0] more nodes.
constexpr uint32_t REF_FIN_PER = REF_FIN_BLOCK / REF_BLOCK;
__global__ __launch_bounds__(REF_BLOCK) void kr_fin_sum(const LbvhScratch* S, uint32_t G) {
    REF_SCENE(S, G);
    __shared__ uint32_t s_w[REF_BLOCK / 64];
    const uint32_t T = *s.n_treelets, nb = (T + REF_FIN_BLOCK - 1) / REF_FIN_BLOCK;
    for (uint32_t b = rj; b < nb; b += G) {
        uint32_t v = 0;
        for (uint32_t j = 0; j < REF_FIN_PER; j++) {
            const uint32_t i = b * REF_FIN_BLOCK + threadIdx.x * REF_FIN_PER + j;
            if (i < T) v += s.t_nodes[s.g_perm[i]];
        }
        uint32_t tot;
        (void)block_excl_scan(v, s_w, &tot);
        if (threadIdx.x == 0) s.fin_blk[b] = tot;
    }
}
__global__ __launch_bounds__(1024) void kr_fin_scan(const LbvhScratch* S) {
    const LbvhScratch& s = S[blockIdx.x];
    __shared__ uint32_t s_sum[1024];
    __shared__ uint32_t s_run;
    const uint32_t T = *s.n_treelets, nb = (T + REF_FIN_BLOCK - 1) / REF_FIN_BLOCK;
    const int tid = threadIdx.x;
    if (tid == 0) s_run = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + tid;
        const uint32_t v = i < nb ? s.fin_blk[i] : 0u;
        s_sum[tid] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t x = tid >= off ? s_sum[tid - off] : 0u;
            __syncthreads();
            s_sum[tid] += x;
            __syncthreads();
        }
        if (i < nb) s.fin_blk[i] = s_run + s_sum[tid] - v;
        __syncthreads();
        if (tid == 0) s_run += s_sum[1023];
        __syncthreads();
    }
    if (tid == 0) {
        s.fin_p[T] = s_run;
        *s.n_nodes = *s.upper_count + s_run;
    }
}
__global__ __launch_bounds__(REF_BLOCK) void kr_fin_place(const LbvhScratch* S, uint32_t G) {
    REF_SCENE(S, G);
    __shared__ uint32_t s_w[REF_BLOCK / 64];
    const uint32_t T = *s.n_treelets, nb = (T + REF_FIN_BLOCK - 1) / REF_FIN_BLOCK;
    for (uint32_t b = rj; b < nb; b += G) {
        uint32_t v[REF_FIN_PER], t[REF_FIN_PER], sum = 0;
        const uint32_t i0 = b * REF_FIN_BLOCK + threadIdx.x * REF_FIN_PER;
        for (uint32_t j = 0; j < REF_FIN_PER; j++) {
            const uint32_t i = i0 + j;
            t[j] = i < T ? s.g_perm[i] : 0u;
            v[j] = i < T ? s.t_nodes[t[j]] : 0u;
            sum += v[j];
        }
        uint32_t tot;
        uint32_t run = s.fin_blk[b] + block_excl_scan(sum, s_w, &tot);
        for (uint32_t j = 0; j < REF_FIN_PER; j++) {
            const uint32_t i = i0 + j;
            if (i < T) {
                s.fin_p[i] = run;
                s.t_final[t[j]] = s.g_la[i] + (i - s.g_lb[i]) + run;
            }
            run += v[j];
        }
    }
}
__global__ __launch_bounds__(REF_BLOCK) void kr_fin_upper(const LbvhScratch* S, uint32_t G) {
    REF_SCENE(S, G);
    const uint32_t nu = *s.upper_count;
    for (uint32_t u = rj * REF_BLOCK + threadIdx.x; u < nu; u += G * REF_BLOCK) {
        const uint32_t i0 = s.u_i0[u], mid = s.u_mid[u];
        const uint32_t p0 = s.fin_p[i0];
        const uint32_t off = s.u_dep[u] + (i0 - s.u_rt[u]) + p0;
        const uint32_t size0 = (mid - 1 - i0) + (s.fin_p[mid] - p0);
        s.node_offset[off] = off + 1 + size0;
        s.node_nshapes[off] = 0;
        s.node_axis[off] = (uint8_t)s.u_axis[u];
        store_box(s.node_bounds + 6 * (size_t)off, load_box(s.u_box + 6 * (size_t)u));
    }
}
// Flatten (SLBVH.cpp:432-455) of the treelet nodes: record q of treelet t to
// its final position.
WPBR_D void scatter_record(const LbvhScratch& s, uint32_t t, uint32_t q) {
    const uint32_t base = s.t_slot[t], fbase = s.t_final[t];
    const uint32_t meta = s.r_meta[base + q];
    const uint32_t fi = fbase + s.r_off[base + q];
    store_box(s.node_bounds + 6 * (size_t)fi, load_box(s.r_box + 6 * (size_t)(base + q)));
    if (meta & 0x80000000u) {
        s.node_offset[fi] = s.r_start[base + q];
        s.node_nshapes[fi] = s.r_n[base + q];
        s.node_axis[fi] = 0;
    } else {
        const uint32_t c0 = s.r_c0[base + q];
        s.node_offset[fi] = fbase + s.r_off[base + c0 + 1];
        s.node_nshapes[fi] = 0;
        s.node_axis[fi] = (uint8_t)((meta >> 8) & 3u);
    }
}
// Treelets of one node, one thread each.
__global__ __launch_bounds__(REF_BLOCK) void kr_scatter_ref(const LbvhScratch* S, uint32_t G) {
    REF_SCENE(S, G);
    const uint32_t T = *s.n_treelets;
    for (uint32_t t = rj * REF_BLOCK + threadIdx.x; t < T; t += G * REF_BLOCK)
        if (s.t_nodes[t] == 1) scatter_record(s, t, 0);
}
// Exclusive prefix of the node counts of the big treelets with more than one
// node (one workgroup per scene; each thread sums a run of the list).
__global__ __launch_bounds__(1024) void kr_big_scan(const LbvhScratch* S) {
    const LbvhScratch& s = S[blockIdx.x];
    __shared__ uint32_t s_w[16];
    const uint32_t nb = *s.bigt_n, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t per = (nb + 1023) / 1024, b0 = min(nb, tid * per), b1 = min(nb, b0 + per);
    auto nodes = [&](uint32_t q) {
        const uint32_t c = s.t_nodes[s.bigt[q]];
        return c > 1 ? c : 0u;
    };
    uint32_t loc = 0;
    for (uint32_t q = b0; q < b1; q++) loc += nodes(q);
    uint32_t inc = loc;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t x = __shfl_up(inc, off, 64);
        if (lane >= (uint32_t)off) inc += x;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t k = 0; k < 16; k++) {
        before += k < w ? s_w[k] : 0u;
        total += s_w[k];
    }
    uint32_t run = before + inc - loc;
    for (uint32_t q = b0; q < b1; q++) {
        s.big_pre[q] = run;
        run += nodes(q);
    }
    if (tid == 0) s.big_pre[nb] = total;
}
// The big treelets' records, spread over every workgroup of the scene: record
// g of the flattened list belongs to the big treelet q with big_pre[q] <= g <
// big_pre[q + 1] (binary search).
__global__ __launch_bounds__(REF_BLOCK) void kr_scatter_big(const LbvhScratch* S, uint32_t G) {
    REF_SCENE(S, G);
    const uint32_t nb = *s.bigt_n, total = s.big_pre[nb];
    for (uint32_t g = rj * REF_BLOCK + threadIdx.x; g < total; g += G * REF_BLOCK) {
        uint32_t lo = 0, hi = nb;  // big_pre[lo] <= g < big_pre[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (s.big_pre[mid] <= g) lo = mid;
            else hi = mid;
        }
        scatter_record(s, s.bigt[lo], g - s.big_pre[lo]);
    }
}

// Host side of the reference order: blocks per scene of each kernel, the
// launch sequence of the sort-and-treelet, emit and upper-SAH stages.
struct RefGrid {
    uint32_t slots, treelets, emit_big, parts, level, ranges, sub, fin, upper;
};
static RefGrid ref_grid(uint32_t max_n, uint32_t n_scenes) {
    auto clampu = [](uint64_t v, uint32_t lo, uint32_t hi) { return (uint32_t)std::min<uint64_t>(hi, std::max<uint64_t>(lo, v)); };
    const uint32_t cap = n_scenes > 1 ? 256u : 2048u;
    RefGrid g;
    g.slots = clampu((max_n + REF_BLOCK - 1) / REF_BLOCK, 1, cap);
    g.treelets = g.slots;
    g.emit_big = clampu(max_n / 512 + 1, 1, n_scenes > 1 ? 64u : 1024u);
    // about one chunk per workgroup at every level (a level holds up to
    // n / REF_CUT ranges): the per-chunk steps are latency-bound
    g.level = clampu(max_n / REF_CUT + 1, 1, n_scenes > 1 ? 128u : 4096u);
    // one 16-lane row per range (a level holds up to n / REF_CUT ranges)
    g.ranges = clampu(max_n / (REF_CUT * LV_RPB) + 1, 1, 256u);
    g.sub = clampu(max_n / (64 * WR_WAVES) + 1, 1, n_scenes > 1 ? 64u : 1024u);
    g.fin = clampu(max_n / REF_FIN_BLOCK + 1, 1, n_scenes > 1 ? 64u : 1024u);
    // about one position or node per thread: these per-treelet and per-node
    // steps are chains of dependent loads
    g.upper = clampu(max_n / REF_BLOCK + 1, 1, n_scenes > 1 ? 64u : 4096u);
    g.parts = n_scenes > 1 ? 256u : REF_CB_PARTS;
    return g;
}

static void ref_launch_treelets(const LbvhScratch* S, uint32_t ns, const RefGrid& g, hipStream_t st) {
    hipLaunchKernelGGL(kr_heads, dim3(ns * g.slots), dim3(REF_BLOCK), 0, st, S, g.slots);
    hipLaunchKernelGGL(kr_head_scan, dim3(ns), dim3(1024), 0, st, S);
    hipLaunchKernelGGL(kr_treelets, dim3(ns * g.treelets), dim3(REF_BLOCK), 0, st, S, g.treelets);
}
static void ks_launch_emit_trees(const LbvhBatch& B, hipStream_t st);
static void ref_launch_emit(const LbvhScratch* S, uint32_t ns, const RefGrid& g, hipStream_t st,
                            const LbvhBatch* B = nullptr) {
    // exactly g.parts blocks per scene each: every partial slot is written
    hipLaunchKernelGGL(kr_emit_small, dim3(ns * g.parts), dim3(REF_BLOCK), 0, st, S, g.parts);
    // segmented: every scene's large treelets from one list over all waves
    if (B) ks_launch_emit_trees(*B, st);
    hipLaunchKernelGGL(kr_emit_big, dim3(ns * g.parts), dim3(EMIT_BLOCK), 0, st, S, g.parts, B ? 0u : 1u);
    hipLaunchKernelGGL(kr_root_cb, dim3(ns), dim3(64), 0, st, S, g.parts);
}
// The level loop: four levels between host checks of the next level's range
// count (kernels of an empty level return at once).
constexpr uint32_t SEG_REF_SUB_BLOCKS = 4096;  // segmented reference order: deferred-range workgroups
constexpr uint32_t REF_LEVELS_PER_CHECK = 2;  // measured against 4: 2.24 -> 2.22 ms at 1M shapes
constexpr uint32_t REF_MAX_LEVELS = 4096;
static void ks_launch_sub_wave(const LbvhBatch& B, uint32_t blocks, hipStream_t st, bool cursor);
static wpbr_status ref_launch_upper(const LbvhScratch* S, uint32_t ns, const RefGrid& g, uint32_t* d_active,
                                    hipStream_t st, const LevelSync* ls, const LbvhBatch* B = nullptr) {
    uint32_t level = 0;
    // REF_LEVELS_PER_CHECK levels, then the next level's range count into d_active
    auto block = [&]() {
        for (uint32_t k = 0; k < REF_LEVELS_PER_CHECK; k++, level++) {
            hipLaunchKernelGGL(kr_lv_stats, dim3(ns * g.level), dim3(REF_BLOCK), 0, st, S, g.level, level);
            hipLaunchKernelGGL(kr_lv_ranges, dim3(ns * g.ranges), dim3(REF_BLOCK), 0, st, S, g.ranges, level);
            hipLaunchKernelGGL(kr_lv_lists, dim3(ns * g.level), dim3(REF_BLOCK), 0, st, S, g.level, level);
            hipLaunchKernelGGL(kr_lv_swap, dim3(ns * g.level), dim3(REF_BLOCK), 0, st, S, g.level, level);
        }
        hipLaunchKernelGGL(kr_active, dim3(1), dim3(256), 0, st, S, ns, level & 1u, d_active);
    };
    // both children of a split are non-empty, so a level range shrinks every
    // level; deeper chains of ranges of at least REF_CUT roots are refused
    if (!ls) {
        for (;;) {
            block();
            uint32_t active = 0;
            if (hipMemcpyAsync(&active, d_active, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return WPBR_ERR_HIP;
            if (!active) break;
            if (level >= REF_MAX_LEVELS) return WPBR_ERR_HIP;
        }
    } else {
        // block k's count goes to host word k & 1 behind event k & 1; the host
        // reads it while block k + 1 runs (kernels of an empty level return at
        // once, so a block queued after the last active level changes nothing)
        block();
        if (hipMemcpyAsync(ls->host, d_active, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipEventRecord(ls->ev[0], st) != hipSuccess)
            return WPBR_ERR_HIP;
        for (uint32_t k = 0;; k++) {
            const uint32_t nx = (k + 1) & 1u;
            block();
            if (hipMemcpyAsync(ls->host + nx, d_active, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipEventRecord(ls->ev[nx], st) != hipSuccess || hipEventSynchronize(ls->ev[k & 1u]) != hipSuccess)
                return WPBR_ERR_HIP;
            if (!ls->host[k & 1u]) break;
            if (level >= REF_MAX_LEVELS) return WPBR_ERR_HIP;
        }
    }
    // the deferred ranges, one wave each: per scene from its cursor, largest
    // first, or for a segmented build from one list over all scenes (the
    // per-scene ordered form measured 2.75 -> 2.89 ms on the 64-scene batch)
    if (B) ks_launch_sub_wave(*B, SEG_REF_SUB_BLOCKS, st, false);
    else {
        hipLaunchKernelGGL(kr_sub_order, dim3(ns), dim3(1024), 0, st, S);
        hipLaunchKernelGGL(kr_sub_wave, dim3(ns * g.sub), dim3(64 * WR_WAVES), 0, st, S, g.sub);
    }
    hipLaunchKernelGGL(kr_fin_sum, dim3(ns * g.fin), dim3(REF_BLOCK), 0, st, S, g.fin);
    hipLaunchKernelGGL(kr_fin_scan, dim3(ns), dim3(1024), 0, st, S);
    hipLaunchKernelGGL(kr_fin_place, dim3(ns * g.fin), dim3(REF_BLOCK), 0, st, S, g.fin);
    hipLaunchKernelGGL(kr_fin_upper, dim3(ns * g.upper), dim3(REF_BLOCK), 0, st, S, g.upper);
    return launch_status();
}

// ---------------------------------------------------------------------------
__global__ void k_cb_final(uint32_t* cb_ord) {
    __shared__ uint32_t s_cb[6];
    reduce_cb(cb_ord, s_cb, cb_ord);
}

wpbr_status shape_bounds_launch(const GeomView& g, float* bounds, float* centroid, uint32_t* cb_ord,
                                hipStream_t st) {
    uint32_t n = g.n_spheres + g.n_triangles;
    if (!n) return WPBR_OK;
    uint32_t nb = std::min<uint32_t>(BOUNDS_BLOCKS, (n + 255) / 256);
    hipLaunchKernelGGL(k_bounds, dim3(nb), dim3(256), 0, st, g, n, bounds, centroid, cb_ord);
    hipLaunchKernelGGL(k_cb_final, dim3(1), dim3(64), 0, st, cb_ord);
    return launch_status();
}

// ---------------------------------------------------------------------------
// Segmented build: the LBVHs of many independent scenes, one launch per kernel
// for all of them.  Every scene keeps its own scratch (its own wpbr_lbvh), so
// each kernel body runs unchanged on its scene's data: block b of a launch
// works for scene map.scene[b] as block b - map.first[scene] of
// map.count[scene] (the grid the single-scene launch would use).  Per-treelet
// kernels run over a work list of (scene, treelet) pairs built on the device.
#define SEG_BLOCK(map)                                   \
    const uint32_t sc = (map).scene[blockIdx.x];         \
    const uint32_t bid = blockIdx.x - (map).first[sc];   \
    const uint32_t nblk = (map).count[sc];               \
    (void)nblk;


__global__ __launch_bounds__(256) void ks_bounds(LbvhBatch B) {
    SEG_BLOCK(B.map[SEG_BOUNDS]);
    const LbvhScratch s = B.s[sc];
    init_body(s, bid, nblk);
    const GeomView g = B.g[sc];
    bounds_body(g, B.n[sc], nullptr, s.centroid, s.cb_ord, bid, nblk);
}

__global__ void ks_cb_final(LbvhBatch B) {
    __shared__ uint32_t s_cb[6];
    uint32_t* cb = B.s[blockIdx.x].cb_ord;
    reduce_cb(cb, s_cb, cb);
}

__global__ __launch_bounds__(MORTON_BLOCK) void ks_morton(LbvhBatch B) {
    SEG_BLOCK(B.map[SEG_MORTON]);
    const LbvhScratch s = B.s[sc];
    morton_body(s.centroid, B.n[sc], s.cb_ord, s.morton, s.hist, bid);
}

// Pass p of the ping-pong: morton -> B -> A -> B -> A (as lbvh_launch_stage).
__global__ __launch_bounds__(SORT_BLOCK) void ks_onesweep(LbvhBatch B, int p) {
    SEG_BLOCK(B.map[SEG_SORT]);
    const LbvhScratch& s = B.s[sc];
    const SortPass q = sort_pass(s, p);
    onesweep_body(q.kin, q.vin, q.kout, q.vout, B.n[sc], p, s.hist, q.status, q.status_next, s.tile_counter, s.err);
}


__global__ __launch_bounds__(256) void ks_slot_prims(LbvhBatch B) {
    SEG_BLOCK(B.map[SEG_MARK]);
    const LbvhScratch s = B.s[sc];
    const GeomView g = B.g[sc];
    slot_prims_body(s, g, B.n[sc], bid);
}

// work[i] = number of treelets of scenes 0 .. i-1 (one workgroup).
__global__ void ks_treelet_count(LbvhBatch B) {
    __shared__ uint32_t s_sum[1024];
    uint32_t run = 0;
    for (uint32_t base = 0; base < B.n_scenes; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t c = 0;
        if (i < B.n_scenes)
            for (int w = 0; w < 128; w++) c += __popc(B.s[i].presence[w]);
        s_sum[threadIdx.x] = c;
        __syncthreads();
        for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
            uint32_t v = threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0;
            __syncthreads();
            s_sum[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < B.n_scenes) B.work[i + 1] = run + s_sum[threadIdx.x];
        if (base == 0 && threadIdx.x == 0) B.work[0] = 0;
        run += s_sum[blockDim.x - 1];
        __syncthreads();
    }
}

// Scene of work item g: the last i with work[i] <= g.
WPBR_D uint32_t work_scene(const uint32_t* work, uint32_t n_scenes, uint32_t g) {
    uint32_t lo = 0, hi = n_scenes;  // work[lo] <= g < work[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (work[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

constexpr uint32_t SEG_EMIT_BLOCKS = 8192;
__global__ __launch_bounds__(EMIT_BLOCK, 5) void ks_emit(LbvhBatch B) {
    const uint32_t total = B.work[B.n_scenes];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        // the (g - work[sc])-th occupied bin of scene sc, two presence words per lane
        const uint32_t sc = __builtin_amdgcn_readfirstlane(work_scene(B.work, B.n_scenes, g));
        const uint32_t t = g - B.work[sc];
        const uint32_t* pr = B.s[sc].presence;
        const uint32_t q0 = 2 * threadIdx.x;
        const uint32_t p0 = pr[q0], p1 = pr[q0 + 1];
        const uint32_t c = __popc(p0) + __popc(p1);
        uint32_t incl = c;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t x = (uint32_t)__shfl_up((int)incl, d);
            if (threadIdx.x >= (uint32_t)d) incl += x;
        }
        const bool mine = incl > t && incl - c <= t;
        uint32_t bin_lane = 0;
        if (mine) {
            uint32_t u = t - (incl - c);
            uint32_t w = p0, q = q0;
            if (u >= (uint32_t)__popc(p0)) { u -= __popc(p0); w = p1; q = q0 + 1; }
            for (uint32_t k = 0; k < u; k++) w &= w - 1;
            bin_lane = 32 * q + (uint32_t)(__ffs(w) - 1);
        }
        const uint64_t who = ballot64(mine);
        const uint32_t bin = __builtin_amdgcn_readfirstlane(
            (uint32_t)__shfl((int)bin_lane, (int)(__ffsll((unsigned long long)who) - 1)));
        const LbvhScratch s = B.s[sc];
        emit_body(s, s.keys_a, B.n[sc], bin);
        __syncthreads();
    }
}

__global__ __launch_bounds__(SAH_BLOCK) void ks_sah_top(LbvhBatch B) {
    LbvhScratch s = B.s[blockIdx.x];
    sah_top_body(s);
}
// subwork[i] = deferred SAH ranges of scenes 0 .. i-1 (one workgroup);
// subwork[n_scenes + 1] = 0, the range cursor of ks_sub_wave.
__global__ void ks_sah_sub_count(LbvhBatch B) {
    __shared__ uint32_t s_sum[1024];
    uint32_t run = 0;
    if (threadIdx.x == 0) B.subwork[B.n_scenes + 1] = 0;
    for (uint32_t base = 0; base < B.n_scenes; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        s_sum[threadIdx.x] = i < B.n_scenes ? *B.s[i].sub_n : 0u;
        __syncthreads();
        for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
            uint32_t v = threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0;
            __syncthreads();
            s_sum[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < B.n_scenes) B.subwork[i + 1] = run + s_sum[threadIdx.x];
        if (base == 0 && threadIdx.x == 0) B.subwork[0] = 0;
        run += s_sum[blockDim.x - 1];
        __syncthreads();
    }
}

// Every scene's deferred ranges from one work list over a fixed grid (few
// workgroups: each holds a CU's LDS, and launching them is not free).
// The deferred ranges of every scene, one wave each (the scenes' ranges in one
// list: range g belongs to the scene whose prefix of range counts holds it).
// One workgroup per range (as k_sah_sub) was measured slower here: the 64
// scenes' deep, unbalanced ranges keep every workgroup's barriers on their
// critical paths (segmented full-order build 1.76 -> 1.68 ms).
// CURSOR: each wave takes the next range of the list when it has finished one
// (the full order's few large ranges: segmented build 1.57 -> 1.56 ms);
// otherwise the waves stride over the list (the reference order's many small
// ranges: the shared cursor measured 2.75 -> 2.83 ms).
template <bool CURSOR>
__global__ __launch_bounds__(64 * WR_WAVES, 3) void ks_sub_wave(LbvhBatch B) {
    __shared__ WaveRangeLds W[WR_WAVES];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t total = B.subwork[B.n_scenes];
    for (uint32_t g = blockIdx.x * WR_WAVES + wave;; g += gridDim.x * WR_WAVES) {
        if (CURSOR) {
            uint32_t k = 0;
            if (lane_id() == 0) k = atomicAdd(B.subwork + B.n_scenes + 1, 1u);
            g = (uint32_t)__builtin_amdgcn_readfirstlane((int)k);
        }
        if (g >= total) break;
        // wave-uniform scene: its scratch pointers stay in scalar registers
        const uint32_t sc = __builtin_amdgcn_readfirstlane(work_scene(B.subwork, B.n_scenes, g));
        LbvhScratch s = B.s[sc];
        wr_range(s, W[wave], g - B.subwork[sc]);
    }
}
// Reference order, segmented: the scenes' treelets of at least 30 shapes in one
// list (prefix of the scenes' counts in subwork), one wave each.  A per-scene
// grid left the scenes with many large treelets to a few waves.
__global__ void ks_big_count(LbvhBatch B) {
    __shared__ uint32_t s_sum[1024];
    uint32_t run = 0;
    for (uint32_t base = 0; base < B.n_scenes; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        s_sum[threadIdx.x] = i < B.n_scenes ? *B.s[i].bigt_n : 0u;
        __syncthreads();
        for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
            uint32_t v = threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0;
            __syncthreads();
            s_sum[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < B.n_scenes) B.subwork[i + 1] = run + s_sum[threadIdx.x];
        if (base == 0 && threadIdx.x == 0) B.subwork[0] = 0;
        run += s_sum[blockDim.x - 1];
        __syncthreads();
    }
}
constexpr uint32_t SEG_EMIT_TREE_BLOCKS = 8192;
__global__ __launch_bounds__(EMIT_BLOCK, 7) void ks_emit_trees(LbvhBatch B) {
    const uint32_t total = B.subwork[B.n_scenes];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        const uint32_t sc = __builtin_amdgcn_readfirstlane(work_scene(B.subwork, B.n_scenes, g));
        const LbvhScratch& s = B.s[sc];
        const uint32_t t = s.bigt[g - B.subwork[sc]], T = *s.n_treelets;
        const uint32_t start = s.t_start[t];
        const uint32_t end = t + 1 < T ? s.t_start[t + 1] : s.n;
        emit_tree(s, s.keys_a, t, start, end - start);
        __syncthreads();
    }
}
static void ks_launch_sub_wave(const LbvhBatch& B, uint32_t blocks, hipStream_t st, bool cursor) {
    hipLaunchKernelGGL(ks_sah_sub_count, dim3(1), dim3(1024), 0, st, B);
    if (cursor) hipLaunchKernelGGL(ks_sub_wave<true>, dim3(blocks), dim3(64 * WR_WAVES), 0, st, B);
    else hipLaunchKernelGGL(ks_sub_wave<false>, dim3(blocks), dim3(64 * WR_WAVES), 0, st, B);
}
static void ks_launch_emit_trees(const LbvhBatch& B, hipStream_t st) {
    hipLaunchKernelGGL(ks_big_count, dim3(1), dim3(1024), 0, st, B);
    hipLaunchKernelGGL(ks_emit_trees, dim3(SEG_EMIT_TREE_BLOCKS), dim3(EMIT_BLOCK), 0, st, B);
}
constexpr uint32_t SEG_SAH_FINAL_BLOCKS = 4;  // per scene
__global__ __launch_bounds__(SAH_BLOCK) void ks_sah_final(LbvhBatch B) {
    LbvhScratch s = B.s[blockIdx.x / SEG_SAH_FINAL_BLOCKS];
    sah_final_body(s, blockIdx.x % SEG_SAH_FINAL_BLOCKS, SEG_SAH_FINAL_BLOCKS);
}

__global__ __launch_bounds__(256) void ks_scatter(LbvhBatch B) {
    const uint32_t total = B.work[B.n_scenes];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        const uint32_t sc = work_scene(B.work, B.n_scenes, g);
        const LbvhScratch s = B.s[sc];
        scatter_body(s, g - B.work[sc]);
    }
}

__global__ __launch_bounds__(256) void ks_subtree(LbvhBatch B) {
    SEG_BLOCK(B.map[SEG_SUBTREE]);
    const LbvhScratch s = B.s[sc];
    subtree_body(s, bid, nblk);
}

__global__ __launch_bounds__(64 * LEAF_WAVES, 4) void ks_subtree_leaf(LbvhBatch B) {
    SEG_BLOCK(B.map[SEG_SUBTREE_LEAF]);
    const LbvhScratch s = B.s[sc];
    subtree_leaf_body(s, s.keys_a, bid, nblk);
}

__global__ __launch_bounds__(256) void ks_subtree_big(LbvhBatch B) {
    SEG_BLOCK(B.map[SEG_SUBTREE_BIG]);
    const LbvhScratch s = B.s[sc];
    subtree_big_body(s, bid, nblk);
}

__global__ __launch_bounds__(256) void ks_bvh4(LbvhBatch B) {
    SEG_BLOCK(B.map[SEG_BVH4]);
    const LbvhScratch s = B.s[sc];
    bvh4_body(s, bid, nblk);
}

// Per scene: n_nodes, n_treelets, err, nshapes and offset of node 0.
__global__ void ks_info(LbvhBatch B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n_scenes) return;
    const LbvhScratch s = B.s[i];
    uint32_t* o = B.info + 5 * (size_t)i;
    o[0] = *s.n_nodes;
    o[1] = *s.n_treelets;
    o[2] = *s.err;
    o[3] = s.node_nshapes[0];
    o[4] = s.node_offset[0];
}

// Blocks per scene of each segmented launch: the single-scene grids.
void lbvh_seg_blocks(uint32_t n, int status_words, uint32_t* counts) {
    (void)status_words;
    counts[SEG_BOUNDS] = std::min<uint32_t>(BOUNDS_BLOCKS, (n + 255) / 256);
    counts[SEG_MORTON] = (n + MORTON_BLOCK * MORTON_ITEMS - 1) / (MORTON_BLOCK * MORTON_ITEMS);
    counts[SEG_SORT] = (n + SORT_TILE - 1) / SORT_TILE;
    counts[SEG_MARK] = (n + 255) / 256;
    // the node-level kernels: about one node per thread (nodes ~ n / 10)
    counts[SEG_SUBTREE] = std::min<uint32_t>(SUBTREE_BLOCKS, std::max<uint32_t>(1, (n + 2047) / 2048));
    counts[SEG_SUBTREE_LEAF] = std::min<uint32_t>(LEAF_BLOCKS, std::max<uint32_t>(1, (n + 511) / 512));
    counts[SEG_SUBTREE_BIG] = 4;
    counts[SEG_BVH4] = std::min<uint32_t>(BVH4_BLOCKS, std::max<uint32_t>(1, (n + 511) / 512));
}

wpbr_status lbvh_launch_segmented(const LbvhBatch& B, const uint32_t* grid, uint32_t order, uint32_t max_n,
                                  hipStream_t st, const LevelSync* ls) {
    const uint32_t S = B.n_scenes;
    const bool ref = order == WPBR_LBVH_ORDER_REFERENCE;
    hipLaunchKernelGGL(ks_bounds, dim3(grid[SEG_BOUNDS]), dim3(256), 0, st, B);
    hipLaunchKernelGGL(ks_cb_final, dim3(S), dim3(64), 0, st, B);
    hipLaunchKernelGGL(ks_morton, dim3(grid[SEG_MORTON]), dim3(MORTON_BLOCK), 0, st, B);
    for (int p = 0; p < (ref ? 3 : 4); p++)
        hipLaunchKernelGGL(ks_onesweep, dim3(grid[SEG_SORT]), dim3(SORT_BLOCK), 0, st, B, p);
    hipLaunchKernelGGL(ks_slot_prims, dim3(grid[SEG_MARK]), dim3(256), 0, st, B);
    if (ref) {
        const RefGrid g = ref_grid(max_n, S);
        ref_launch_treelets(B.s, S, g, st);
        ref_launch_emit(B.s, S, g, st, &B);
        wpbr_status e = ref_launch_upper(B.s, S, g, B.info + 5 * (size_t)S, st, ls, &B);
        if (e) return e;
        hipLaunchKernelGGL(kr_scatter_ref, dim3(S * g.upper), dim3(REF_BLOCK), 0, st, B.s, g.upper);
        hipLaunchKernelGGL(kr_big_scan, dim3(S), dim3(1024), 0, st, B.s);
        hipLaunchKernelGGL(kr_scatter_big, dim3(S * g.upper), dim3(REF_BLOCK), 0, st, B.s, g.upper);
    } else {
        hipLaunchKernelGGL(ks_treelet_count, dim3(1), dim3(1024), 0, st, B);
        hipLaunchKernelGGL(ks_emit, dim3(SEG_EMIT_BLOCKS), dim3(EMIT_BLOCK), 0, st, B);
        hipLaunchKernelGGL(ks_sah_top, dim3(S), dim3(SAH_BLOCK), 0, st, B);
        ks_launch_sub_wave(B, SAH_SEG_SUB_BLOCKS, st, true);
        hipLaunchKernelGGL(ks_sah_final, dim3(S * SEG_SAH_FINAL_BLOCKS), dim3(SAH_BLOCK), 0, st, B);
        hipLaunchKernelGGL(ks_scatter, dim3(4096), dim3(256), 0, st, B);
    }
    hipLaunchKernelGGL(ks_subtree, dim3(grid[SEG_SUBTREE]), dim3(256), 0, st, B);
    hipLaunchKernelGGL(ks_subtree_leaf, dim3(grid[SEG_SUBTREE_LEAF]), dim3(64 * LEAF_WAVES), 0, st, B);
    hipLaunchKernelGGL(ks_subtree_big, dim3(grid[SEG_SUBTREE_BIG]), dim3(256), 0, st, B);
    hipLaunchKernelGGL(ks_bvh4, dim3(grid[SEG_BVH4]), dim3(256), 0, st, B);
    hipLaunchKernelGGL(ks_info, dim3((S + 255) / 256), dim3(256), 0, st, B);
    return launch_status();
}

// ---------------------------------------------------------------------------
// Host-side launch sequence.
// One stage of the build (the reference's job boundaries, MEngine.cpp:341-350):
//   0 bounds + centroids      JobUpdateBoundsAndCentroidSphere / ...Triangle
//   1 centroid bounds         JobGenerateShapesCentroidBound
//   2 Morton codes            JobGenerateShapesMortonCode
//   3 sort + treelets         JobSortsShapesByMortonCode
//   4 treelet trees           JobEmitLBVH
//   5 upper SAH + flatten     JobBuildUpperSAH (+ the derived traversal layout)
// ev[0..11] bracket the kernels (the per-kernel times of wpbr_lbvh_kernel_times).
wpbr_status lbvh_launch_stage(LbvhScratch& s, const LbvhScratch* d_self, const GeomView& g, uint32_t n,
                              hipStream_t st, int stage, hipEvent_t* ev, int nev, const LevelSync* ls) {
    auto rec = [&](int i) { if (ev && i < nev) (void)hipEventRecord(ev[i], st); };
    const bool ref = s.order == WPBR_LBVH_ORDER_REFERENCE;
    const RefGrid rg = ref_grid(n, 1);
    switch (stage) {
        case 0:
            rec(0);
            hipLaunchKernelGGL(k_init_bounds, dim3(std::min<uint32_t>(BOUNDS_BLOCKS, (n + 255) / 256)), dim3(256), 0,
                               st, s, g, n);
            rec(1);
            break;
        case 1:
            hipLaunchKernelGGL(k_cb_final, dim3(1), dim3(64), 0, st, s.cb_ord);
            break;
        case 2:
            hipLaunchKernelGGL(k_morton, dim3((n + MORTON_BLOCK * MORTON_ITEMS - 1) / (MORTON_BLOCK * MORTON_ITEMS)),
                               dim3(MORTON_BLOCK), 0, st, s.centroid, n, s.cb_ord, s.morton, s.hist);
            rec(2);
            break;
        case 3: {
            const uint32_t num_tiles = (n + SORT_TILE - 1) / SORT_TILE;
            const int np = sort_passes(s);
            for (int p = 0; p < np; p++) {
                const SortPass q = sort_pass(s, p);
                hipLaunchKernelGGL(k_onesweep, dim3(num_tiles), dim3(SORT_BLOCK), 0, st, q.kin, q.vin, q.kout,
                                   q.vout, n, p, s.hist, q.status, q.status_next, s.tile_counter, s.err);
                rec(3 + p);
            }
            hipLaunchKernelGGL(k_slot_prims, dim3((n + 255) / 256), dim3(256), 0, st, s, g, n);
            if (ref) {
                ref_launch_treelets(d_self, 1, rg, st);
                rec(6);  // events 3..6: three sort passes, then the treelet scan
            }
            rec(7);
            break;
        }
        case 4:
            if (ref) ref_launch_emit(d_self, 1, rg, st);
            else hipLaunchKernelGGL(k_emit, dim3(4096), dim3(EMIT_BLOCK), 0, st, s, s.keys_a, n);
            rec(8);
            break;
        case 5: {
            if (ref) {
                wpbr_status e = ref_launch_upper(d_self, 1, rg, s.err + 3, st, ls);
                if (e) return e;
                rec(9);
                hipLaunchKernelGGL(kr_scatter_ref, dim3(rg.upper), dim3(REF_BLOCK), 0, st, d_self, rg.upper);
                hipLaunchKernelGGL(kr_big_scan, dim3(1), dim3(1024), 0, st, d_self);
                hipLaunchKernelGGL(kr_scatter_big, dim3(rg.upper), dim3(REF_BLOCK), 0, st, d_self, rg.upper);
            } else {
                hipLaunchKernelGGL(k_sah_top, dim3(1), dim3(SAH_BLOCK), 0, st, s);
                // the deferred ranges (16..32 of about 128..255 roots at 1M shapes):
                // one workgroup each (one wave each, as in the segmented build,
                // measured 191 -> 256 us for the upper SAH)
                hipLaunchKernelGGL(k_sah_sub, dim3(SAH_SUB_BLOCKS), dim3(SAH_BLOCK), 0, st, s);
                hipLaunchKernelGGL(k_sah_final, dim3(SAH_FINAL_BLOCKS), dim3(SAH_BLOCK), 0, st, s);
                rec(9);
                hipLaunchKernelGGL(k_scatter, dim3(4096), dim3(256), 0, st, s);
            }
            rec(10);
            const uint32_t max_nodes = 2 * n;
            hipLaunchKernelGGL(k_subtree, dim3(std::min<uint32_t>(SUBTREE_BLOCKS, (max_nodes + 255) / 256)), dim3(256),
                               0, st, s);
            hipLaunchKernelGGL(k_subtree_leaf, dim3(std::min<uint32_t>(LEAF_BLOCKS, std::max<uint32_t>(1, (n + 511) / 512))),
                               dim3(64 * LEAF_WAVES), 0, st, s, s.keys_a);
            hipLaunchKernelGGL(k_subtree_big, dim3(64), dim3(256), 0, st, s);
            const uint32_t max_records = 3 * std::max<uint32_t>(n, 1);
            hipLaunchKernelGGL(k_bvh4, dim3(std::min<uint32_t>(BVH4_BLOCKS, (max_records + 255) / 256)), dim3(256), 0,
                               st, s);
            rec(11);
            break;
        }
        default:
            return WPBR_ERR_INVALID_ARG;
    }
    return launch_status();
}

}  // namespace wpbr
