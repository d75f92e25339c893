// Oracle: CPU restatement of WoodenPBREngine's ray-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg may load liboracle.so, and only as the checker / the timed CPU
// baseline.  The product library (woodenpbrengine_amd/) never includes, links or
// calls anything here.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   * Morton codes     - closed-form bit interleave, brute-force recomputation.
//   * radix sort       - std::stable_sort on (code, index) (a library routine).
//   * treelets / emit  - closed-form invariants (each prim in one leaf, leaf n < 30
//                        unless all codes equal, bounds = union, preorder layout,
//                        split = first index with the bit set).
//   * upper SAH        - restatement only; std::partition order is
//                        implementation-defined, fixed here to the libstdc++
//                        two-pointer form and checked against std::partition.
//   * traversal        - brute force over all primitives (same intersectors) gives
//                        the same minimum t; ids agree unless t ties.
//   * triangle test    - closed-form hits (axis-aligned / edge-on cases).
//   * sphere test      - analytic closed form.
//   * shading          - closed forms listed in shade.cpp.
#pragma once
#include "common.h"
#include <vector>

namespace orc {

struct Scene {
    uint32_t n_spheres = 0;
    const float* sphere_xform = nullptr;   // per sphere: world 3x4 then inverse 3x4 (24 floats)
    const float* sphere_radius = nullptr;
    uint32_t n_triangles = 0;
    const float* tri_verts = nullptr;      // per triangle: v0, v1, v2 (9 floats)
};

struct LbvhResult {
    Bounds centroid_bounds;
    std::vector<uint32_t> morton;          // per shape, input order
    std::vector<uint32_t> sorted_code, sorted_index;
    std::vector<uint32_t> shape_entities;  // CSceneTreeLBVH::shapesEntities
    std::vector<uint32_t> treelet_start;
    std::vector<uint32_t> node_offset, node_nshapes;
    std::vector<uint8_t> node_axis;
    std::vector<float> node_bounds;        // 6 per node: min xyz, max xyz
    uint32_t n_nodes = 0, n_sah_nodes = 0, n_treelets = 0;
};

// libstdc++ std::partition for bidirectional iterators (the two-pointer form).
template <class It, class P>
inline
It partition_bidir(It first, It last, P pred) {
    while (true) {
        while (true) {
            if (first == last) return first;
            if (pred(*first)) ++first;
            else break;
        }
        --last;
        while (true) {
            if (first == last) return first;
            if (!pred(*last)) --last;
            else break;
        }
        std::iter_swap(first, last);
        ++first;
    }
}

void triangle_bounds(uint32_t n, const float* v, float* bounds, float* centroid);
void sphere_bounds(uint32_t n, const float* xform, const float* radius, float* bounds,
                   float* centroid);
Bounds centroid_bounds(uint32_t n, const float* centroid);
uint32_t morton_code(const Bounds& cb, const float* c);
void lbvh_build(uint32_t n, const float* bounds, const float* centroid, const uint32_t* entities,
                LbvhResult& out);

bool box_hit(const float* bmn, const float* bmx, V3 o, V3 inv, const int* neg, float tMax,
             float* tNearOut);
float tri_intersect(const float* v, V3 o, V3 d, float* bary);
float sphere_intersect(const float* xform, float radius, V3 o, V3 d, V3* oL, V3* dL);
int64_t trace_ray(const Scene& sc, const LbvhResult& bvh, V3 o, V3 d, float* tOut,
                  uint64_t* nCandidates, int* stackOverflow);

}  // namespace orc
