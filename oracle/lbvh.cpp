// Oracle: LBVH build (MEngine::buildLBVH, MEngine.cpp:341-350).
// TEST INFRASTRUCTURE ONLY - restates the reference's CPU algorithm, single thread.
//
//   shape bounds/centroids   CSTriangle.cpp:116-142, CSSphere.h:136-152, CCentroid.h:10-13
//   centroid bounds          SLBVH.cpp:164-181
//   Morton codes             SLBVH.cpp:186-227
//   6-bit LSD radix sort     SLBVH.cpp:231-295 (stable; 5 passes over 30 bits into
//                            two buffers that alternate; see lbvh_build for which
//                            buffer the rest of the job reads)
//   treelets (top 12 bits)   SLBVH.cpp:297-322
//   emitLBVH                 SLBVH.cpp:328-406
//   upper SAH + flatten      SLBVH.cpp:410-563
#include "oracle.h"
#include <vector>
#include <array>
#include <functional>

namespace orc {

void triangle_bounds(uint32_t n, const float* v, float* bounds, float* centroid) {
    for (uint32_t i = 0; i < n; i++) {
        const float* t = v + 9 * (size_t)i;
        // DBounds3f(p0, p1) then union with p2 (CSTriangle.cpp:137-139).
        Bounds b = bounds_empty();
        b = bounds_union_pt(b, v3(t[0], t[1], t[2]));
        b = bounds_union_pt(b, v3(t[3], t[4], t[5]));
        b = bounds_union_pt(b, v3(t[6], t[7], t[8]));
        std::memcpy(bounds + 6 * (size_t)i, &b, 24);
        // CCentroid: pMax*0.5 + pMin*0.5 with a double 0.5 (CCentroid.h:10-13).
        for (int k = 0; k < 3; k++)
            centroid[3 * (size_t)i + k] = (float)((double)b.mx[k] * 0.5 + (double)b.mn[k] * 0.5);
    }
}

void sphere_bounds(uint32_t n, const float* xform, const float* radius, float* bounds,
                   float* centroid) {
    for (uint32_t i = 0; i < n; i++) {
        Affine w;
        std::memcpy(&w, xform + 24 * (size_t)i, 48);
        float r = radius[i];
        Bounds loc;
        loc.mn[0] = loc.mn[1] = loc.mn[2] = -r;
        loc.mx[0] = loc.mx[1] = loc.mx[2] = r;
        Bounds b = xf_bounds(w, loc);  // SSphere::getBoundWorld, CSSphere.cpp:289-293
        std::memcpy(bounds + 6 * (size_t)i, &b, 24);
        for (int k = 0; k < 3; k++)
            centroid[3 * (size_t)i + k] = (float)((double)b.mx[k] * 0.5 + (double)b.mn[k] * 0.5);
    }
}

// leftShift3, SLBVH.cpp:219-227.
static uint32_t left_shift3(uint32_t x) {
    if (x == (1u << 10)) --x;
    x = (x | (x << 16)) & 0b00000011000000000000000011111111u;
    x = (x | (x << 8)) & 0b00000011000000001111000000001111u;
    x = (x | (x << 4)) & 0b00000011000011000011000011000011u;
    x = (x | (x << 2)) & 0b00001001001001001001001001001001u;
    return x;
}

Bounds centroid_bounds(uint32_t n, const float* centroid) {
    Bounds cb = bounds_empty();
    for (uint32_t i = 0; i < n; i++)
        cb = bounds_union_pt(cb, v3(centroid[3 * i], centroid[3 * i + 1], centroid[3 * i + 2]));
    return cb;
}

uint32_t morton_code(const Bounds& cb, const float* c) {
    V3 cn = bounds_offset(cb, v3(c[0], c[1], c[2]));
    const float scale = 1024;
    // float -> uint32_t argument conversion truncates (SLBVH.cpp:202-205).
    uint32_t mz = left_shift3((uint32_t)(cn.z * scale));
    uint32_t my = left_shift3((uint32_t)(cn.y * scale));
    uint32_t mx = left_shift3((uint32_t)(cn.x * scale));
    return (mz << 2) | (my << 1) | mx;
}

namespace {

struct BuildNode {
    uint32_t iStart = 0, n = 0;
    Bounds b;
    int child[2] = {-1, -1};  // indices into the owning node pool
    bool childIsSah[2] = {false, false};
    uint8_t axis = 0;
    bool leaf = false;
};

struct Treelet {
    uint32_t iStart, n;
    std::vector<BuildNode> nodes;  // DFS preorder as emitted
    uint8_t iBucket = 0;
};

struct Builder {
    const uint32_t* codes;        // sorted codes
    const float* sortedBounds;    // sorted bounds, 6 per shape
    std::vector<Treelet> treelets;
    std::vector<BuildNode> sah;   // upper nodes; children refer to sah[] or treelet roots
};

Bounds shape_bounds(const Builder& B, uint32_t i) {
    Bounds b;
    std::memcpy(&b, B.sortedBounds + 6 * (size_t)i, 24);
    return b;
}

// emitLBVH, SLBVH.cpp:353-406.  Returns the index of the emitted node in t.nodes.
int emit(Builder& B, Treelet& t, uint32_t iStart, uint32_t n, int bit) {
    const uint32_t maxPrims = 30;  // overridden at runtime, SLBVH.cpp:356
    if (bit == -1 || n < maxPrims) {
        int idx = (int)t.nodes.size();
        t.nodes.emplace_back();
        BuildNode& nd = t.nodes.back();
        Bounds b = bounds_empty();
        for (uint32_t i = 0; i < n; i++) b = bounds_union(b, shape_bounds(B, iStart + i));
        nd.b = b;
        nd.iStart = iStart;
        nd.n = n;
        nd.leaf = true;
        return idx;
    }
    uint32_t mask = 1u << bit;
    if ((B.codes[iStart] & mask) == (B.codes[iStart + n - 1] & mask))
        return emit(B, t, iStart, n, bit - 1);
    uint32_t s = iStart, e = iStart + n - 1;
    while (s + 1 != e) {
        uint32_t mid = (s + e) / 2;
        if ((B.codes[s] & mask) == (B.codes[mid] & mask)) s = mid;
        else e = mid;
    }
    uint32_t split = e;
    int idx = (int)t.nodes.size();
    t.nodes.emplace_back();
    int c0 = emit(B, t, iStart, split - iStart, bit - 1);
    int c1 = emit(B, t, split, n - (split - iStart), bit - 1);
    BuildNode& nd = t.nodes[idx];
    nd.child[0] = c0;
    nd.child[1] = c1;
    nd.axis = (uint8_t)(bit % 3);
    nd.leaf = false;
    nd.n = 0;
    nd.b = bounds_union(t.nodes[c0].b, t.nodes[c1].b);
    return idx;
}

// A reference to an upper-tree child: either an SAH node or a treelet root.
struct Ref { bool sah; int idx; };

// buildUpperSAH, SLBVH.cpp:457-563.  Operates on B.treelets in place.
Ref build_upper(Builder& B, uint32_t iStart, uint32_t iLast) {
    if (iStart == iLast) return Ref{false, (int)iStart};
    Bounds bounds = bounds_empty(), cbounds = bounds_empty();
    for (uint32_t i = iStart; i <= iLast; i++) {
        const Bounds& rb = B.treelets[i].nodes[0].b;
        bounds = bounds_union(bounds, rb);
        V3 c;
        for (int k = 0; k < 3; k++) c[k] = (float)((double)(rb.mn[k] + rb.mx[k]) * 0.5);
        cbounds = bounds_union_pt(cbounds, c);
    }
    int dim = bounds_max_extent(cbounds);
    const int NB = 12;
    std::array<Bounds, NB> bb;
    std::array<uint32_t, NB> bc;
    for (int i = 0; i < NB; i++) { bb[i] = bounds_empty(); bc[i] = 0; }
    for (uint32_t i = iStart; i <= iLast; i++) {
        float dist = cbounds.mx[dim] - cbounds.mn[dim];
        const Bounds& rb = B.treelets[i].nodes[0].b;
        float centroid = (float)((double)(rb.mn[dim] + rb.mx[dim]) * 0.5);
        uint8_t ib;
        if (dist > 0.0f) {
            float t = (centroid - cbounds.mn[dim]) / dist;
            ib = (uint8_t)(NB * t);
            if (ib == NB) ib = NB - 1;
        } else {
            ib = 0;  // all centroids coincide: (uint8_t)NaN is UB; defined as bucket 0
        }
        B.treelets[i].iBucket = ib;
        bc[ib]++;
        bb[ib] = bounds_union(bb[ib], rb);
    }
    float costs[NB - 1];
    float area = bounds_area(bounds);
    for (int i = 0; i < NB - 1; i++) {
        Bounds b0 = bounds_empty(), b1 = bounds_empty();
        uint32_t c0 = 0, c1 = 0;
        for (int j = 0; j <= i; j++) { b0 = bounds_union(b0, bb[j]); c0 += bc[j]; }
        for (int j = i + 1; j < NB; j++) { b1 = bounds_union(b1, bb[j]); c1 += bc[j]; }
        float l = (c0 == 0) ? 0 : (float)c0 * bounds_area(b0);
        float r = (c1 == 0) ? 0 : (float)c1 * bounds_area(b1);
        costs[i] = (float)(0.125 + (double)((l + r) / area));
    }
    float cmin = costs[0];
    int best = 0;
    for (int i = 1; i < NB - 1; i++)
        if (costs[i] < cmin) { best = i; cmin = costs[i]; }
    Treelet* first = B.treelets.data() + iStart;
    Treelet* last = B.treelets.data() + iLast + 1;
    Treelet* split = partition_bidir(first, last,
                                     [best](const Treelet& t) { return t.iBucket <= best; });
    uint32_t mid = (uint32_t)(split - B.treelets.data());
    if (mid == iLast + 1) mid = (iLast - iStart + 1) / 2 + iStart;
    int idx = (int)B.sah.size();
    B.sah.emplace_back();
    Ref a = build_upper(B, iStart, mid - 1);
    Ref b = build_upper(B, mid, iLast);
    BuildNode& nd = B.sah[idx];
    nd.child[0] = a.idx; nd.childIsSah[0] = a.sah;
    nd.child[1] = b.idx; nd.childIsSah[1] = b.sah;
    nd.axis = (uint8_t)dim;
    nd.leaf = false;
    const Bounds& ba = a.sah ? B.sah[a.idx].b : B.treelets[a.idx].nodes[0].b;
    const Bounds& bbnd = b.sah ? B.sah[b.idx].b : B.treelets[b.idx].nodes[0].b;
    nd.b = bounds_union(ba, bbnd);
    return Ref{true, idx};
}

}  // namespace

void lbvh_build(uint32_t n, const float* bounds, const float* centroid, const uint32_t* entities,
                LbvhResult& out, int order) {
    out = LbvhResult();
    if (n == 0) return;
    Bounds cb = centroid_bounds(n, centroid);
    out.centroid_bounds = cb;
    out.morton.resize(n);
    std::vector<uint32_t> code(n), index(n);
    for (uint32_t i = 0; i < n; i++) {
        out.morton[i] = morton_code(cb, centroid + 3 * (size_t)i);
        code[i] = out.morton[i];
        index[i] = i;
    }
    // 6-bit LSD radix sort, 5 passes (SLBVH.cpp:245-283).  Pass p reads `in`
    // and writes `out`, chosen afresh from the pass parity at the top of every
    // iteration (SLBVH.cpp:253-254): passes 0, 2, 4 go mortonCodes -> tmp and
    // passes 1, 3 go tmp -> mortonCodes.  The closing std::swap(in, out)
    // (SLBVH.cpp:283) exchanges the loop's locals only.  So the fully sorted
    // keys end in tmp, which the job never reads again, and mortonCodes -- the
    // array the gather (SLBVH.cpp:286-291), the treelet split (:297-319) and
    // emitLBVH (:374-387) read -- holds pass 3's output: a stable sort on bits
    // 0..23 only.  LBVH_ORDER_REFERENCE keeps that array; LBVH_ORDER_FULL takes
    // tmp (the fully sorted order the job's loop was written to produce).
    if (n > 1) {
        std::vector<uint32_t> tc(n), ti(n);
        for (int pass = 0; pass < 5; pass++) {
            uint32_t* in_c = (pass & 1) ? tc.data() : code.data();
            uint32_t* in_i = (pass & 1) ? ti.data() : index.data();
            uint32_t* out_c = (pass & 1) ? code.data() : tc.data();
            uint32_t* out_i = (pass & 1) ? index.data() : ti.data();
            int lo = pass * 6;
            uint32_t cnt[64] = {0};
            for (uint32_t i = 0; i < n; i++) cnt[(in_c[i] >> lo) & 63]++;
            uint32_t total = 0;
            for (int b = 0; b < 64; b++) { uint32_t t = cnt[b]; cnt[b] = total; total += t; }
            for (uint32_t i = 0; i < n; i++) {
                uint32_t d = (in_c[i] >> lo) & 63;
                uint32_t pos = cnt[d]++;
                out_c[pos] = in_c[i];
                out_i[pos] = in_i[i];
            }
        }
        if (order == LBVH_ORDER_FULL) {
            code.swap(tc);
            index.swap(ti);
        }
    }
    out.sorted_code = code;
    out.sorted_index = index;
    std::vector<float> sb(6 * (size_t)n);
    out.shape_entities.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        std::memcpy(&sb[6 * (size_t)i], bounds + 6 * (size_t)index[i], 24);
        out.shape_entities[i] = entities ? entities[index[i]] : index[i];
    }
    Builder B;
    B.codes = code.data();
    B.sortedBounds = sb.data();
    // Treelets: runs of equal top-12 bits (SLBVH.cpp:299-319).
    const uint32_t mask = 0b00111111111111000000000000000000u;
    for (uint32_t s = 0, e = 1; e <= n; e++) {
        if (e == n || (code[s] & mask) != (code[e] & mask)) {
            Treelet t;
            t.iStart = s;
            t.n = e - s;
            B.treelets.push_back(std::move(t));
            s = e;
        }
    }
    out.treelet_start.clear();
    for (auto& t : B.treelets) {
        out.treelet_start.push_back(t.iStart);
        emit(B, t, t.iStart, t.n, 30 - 12 - 1);
    }
    uint32_t k = (uint32_t)B.treelets.size();
    Ref root = build_upper(B, 0, k - 1);
    // flattenBVHTree, SLBVH.cpp:432-455 (DFS preorder; secondChildOffset = child 1).
    size_t total = B.sah.size();
    for (auto& t : B.treelets) total += t.nodes.size();
    out.node_offset.assign(total, 0);
    out.node_nshapes.assign(total, 0);
    out.node_axis.assign(total, 0);
    out.node_bounds.assign(6 * total, 0.0f);
    uint32_t offset = 0;
    std::function<uint32_t(const BuildNode&, Treelet*)> flatten =
        [&](const BuildNode& nd, Treelet* owner) -> uint32_t {
        uint32_t my = offset++;
        std::memcpy(&out.node_bounds[6 * (size_t)my], &nd.b, 24);
        if (nd.leaf) {
            out.node_offset[my] = nd.iStart;
            out.node_nshapes[my] = nd.n;
            out.node_axis[my] = 0;  // uninitialised in the reference; defined as 0
        } else {
            out.node_axis[my] = nd.axis;
            out.node_nshapes[my] = 0;
            uint32_t second = 0;
            for (int c = 0; c < 2; c++) {
                uint32_t o;
                if (owner) {
                    o = flatten(owner->nodes[nd.child[c]], owner);
                } else if (nd.childIsSah[c]) {
                    o = flatten(B.sah[nd.child[c]], nullptr);
                } else {
                    Treelet& t = B.treelets[nd.child[c]];
                    o = flatten(t.nodes[0], &t);
                }
                if (c == 1) second = o;
            }
            out.node_offset[my] = second;
        }
        return my;
    };
    if (root.sah) flatten(B.sah[root.idx], nullptr);
    else flatten(B.treelets[root.idx].nodes[0], &B.treelets[root.idx]);
    out.n_nodes = offset;
    out.n_sah_nodes = (uint32_t)B.sah.size();
    out.n_treelets = k;
}

}  // namespace orc
