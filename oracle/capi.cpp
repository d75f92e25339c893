// Oracle C entry points for the pytest checkers and bench.py's CPU baseline.
// TEST INFRASTRUCTURE ONLY.
#include "oracle.h"
#include <thread>
#include <atomic>

using namespace orc;

namespace {

template <class F>
void parallel_for(uint32_t n, int nthreads, F f) {
    if (nthreads <= 1 || n < 1024) {
        for (uint32_t i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<uint32_t> next(0);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        th.emplace_back([&]() {
            while (true) {
                uint32_t b = next.fetch_add(256);
                if (b >= n) break;
                uint32_t e = std::min(n, b + 256);
                for (uint32_t i = b; i < e; i++) f(i);
            }
        });
    }
    for (auto& x : th) x.join();
}

Scene make_scene(uint32_t ns, const float* sx, const float* sr, uint32_t nt, const float* tv) {
    Scene s;
    s.n_spheres = ns;
    s.sphere_xform = sx;
    s.sphere_radius = sr;
    s.n_triangles = nt;
    s.tri_verts = tv;
    return s;
}

}  // namespace

extern "C" {

// a1: bounds + centroids for spheres (first) then triangles.
void orc_shape_bounds(uint32_t ns, const float* sx, const float* sr, uint32_t nt, const float* tv,
                      float* bounds, float* centroid) {
    sphere_bounds(ns, sx, sr, bounds, centroid);
    triangle_bounds(nt, tv, bounds + 6 * (size_t)ns, centroid + 3 * (size_t)ns);
}

void orc_centroid_bounds(uint32_t n, const float* centroid, float* out6) {
    Bounds b = centroid_bounds(n, centroid);
    std::memcpy(out6, &b, 24);
}

void* orc_lbvh_build(uint32_t n, const float* bounds, const float* centroid,
                     const uint32_t* entities) {
    LbvhResult* r = new LbvhResult();
    lbvh_build(n, bounds, centroid, entities, *r);
    return r;
}

void orc_lbvh_free(void* h) { delete (LbvhResult*)h; }

// counts: n_nodes, n_sah_nodes, n_treelets
void orc_lbvh_info(void* h, uint32_t* counts) {
    LbvhResult* r = (LbvhResult*)h;
    counts[0] = r->n_nodes;
    counts[1] = r->n_sah_nodes;
    counts[2] = r->n_treelets;
}

// field: 0 morton, 1 sorted_code, 2 sorted_index, 3 shape_entities, 4 node_offset,
// 5 node_nshapes, 6 node_axis (u8), 7 node_bounds (f32 x6), 8 treelet_start,
// 9 centroid bounds (f32 x6)
int orc_lbvh_copy(void* h, int field, void* dst) {
    LbvhResult* r = (LbvhResult*)h;
    switch (field) {
        case 0: std::memcpy(dst, r->morton.data(), 4 * r->morton.size()); return 0;
        case 1: std::memcpy(dst, r->sorted_code.data(), 4 * r->sorted_code.size()); return 0;
        case 2: std::memcpy(dst, r->sorted_index.data(), 4 * r->sorted_index.size()); return 0;
        case 3: std::memcpy(dst, r->shape_entities.data(), 4 * r->shape_entities.size()); return 0;
        case 4: std::memcpy(dst, r->node_offset.data(), 4 * (size_t)r->n_nodes); return 0;
        case 5: std::memcpy(dst, r->node_nshapes.data(), 4 * (size_t)r->n_nodes); return 0;
        case 6: std::memcpy(dst, r->node_axis.data(), (size_t)r->n_nodes); return 0;
        case 7: std::memcpy(dst, r->node_bounds.data(), 24 * (size_t)r->n_nodes); return 0;
        case 8: std::memcpy(dst, r->treelet_start.data(), 4 * r->treelet_start.size()); return 0;
        case 9: std::memcpy(dst, &r->centroid_bounds, 24); return 0;
    }
    return -1;
}

// Closest hit for n rays (SoA: rays[0..6][n] = ox, oy, oz, dx, dy, dz).
// Writes the hit shape index (0xFFFFFFFF = miss) and t (+inf on miss).
// Returns the number of candidate primitive tests; *overflow set on stack overflow.
uint64_t orc_collide(uint32_t ns, const float* sx, const float* sr, uint32_t nt, const float* tv,
                     void* h, uint32_t n, const float* rays, uint32_t* hitShape, float* hitT,
                     int nthreads, int* overflow) {
    LbvhResult* r = (LbvhResult*)h;
    Scene sc = make_scene(ns, sx, sr, nt, tv);
    std::atomic<uint64_t> cand(0);
    std::atomic<int> ovf(0);
    parallel_for(n, nthreads, [&](uint32_t i) {
        V3 o = v3(rays[i], rays[(size_t)n + i], rays[2 * (size_t)n + i]);
        V3 d = v3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]);
        float t;
        uint64_t c = 0;
        int so = 0;
        int64_t slot = trace_ray(sc, *r, o, d, &t, &c, &so);
        hitShape[i] = slot < 0 ? 0xFFFFFFFFu : r->sorted_index[(size_t)slot];
        hitT[i] = t;
        cand += c;
        if (so) ovf = 1;
    });
    if (overflow) *overflow = ovf.load();
    return cand.load();
}

// Brute force: minimum t over every primitive in shape order (no BVH).
void orc_brute_force(uint32_t ns, const float* sx, const float* sr, uint32_t nt, const float* tv,
                     uint32_t n, const float* rays, uint32_t* hitShape, float* hitT, int nthreads) {
    Scene sc = make_scene(ns, sx, sr, nt, tv);
    parallel_for(n, nthreads, [&](uint32_t i) {
        V3 o = v3(rays[i], rays[(size_t)n + i], rays[2 * (size_t)n + i]);
        V3 d = v3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]);
        float best = std::numeric_limits<float>::infinity();
        uint32_t id = 0xFFFFFFFFu;
        for (uint32_t s = 0; s < sc.n_spheres + sc.n_triangles; s++) {
            float t = s < sc.n_spheres
                          ? sphere_intersect(sc.sphere_xform + 24 * (size_t)s, sc.sphere_radius[s],
                                             o, d, nullptr, nullptr)
                          : tri_intersect(sc.tri_verts + 9 * (size_t)(s - sc.n_spheres), o, d,
                                          nullptr);
            if (t == 0.0f) continue;
            if (best > t) { best = t; id = s; }
        }
        hitShape[i] = id;
        hitT[i] = best;
    });
}

float orc_tri_intersect(const float* v, const float* o, const float* d, float* bary) {
    return tri_intersect(v, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), bary);
}

float orc_sphere_intersect(const float* xform, float radius, const float* o, const float* d) {
    return sphere_intersect(xform, radius, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), nullptr,
                            nullptr);
}

int orc_box_hit(const float* b6, const float* o, const float* d, float tMax, float* tNear) {
    V3 oo = v3(o[0], o[1], o[2]);
    V3 inv = v3(1.0f / d[0], 1.0f / d[1], 1.0f / d[2]);
    int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
    return box_hit(b6, b6 + 3, oo, inv, neg, tMax, tNear) ? 1 : 0;
}

// libstdc++ std::partition reference used to pin the oracle's two-pointer form.
void orc_partition_check(uint32_t n, const uint8_t* buckets, int best, uint32_t* order_oracle,
                         uint32_t* order_std, uint32_t* mids) {
    std::vector<std::pair<uint8_t, uint32_t>> a(n), b(n);
    for (uint32_t i = 0; i < n; i++) a[i] = b[i] = {buckets[i], i};
    auto pred = [best](const std::pair<uint8_t, uint32_t>& x) { return x.first <= best; };
    auto ma = partition_bidir(a.data(), a.data() + n, pred);
    auto mb = std::partition(b.data(), b.data() + n, pred);
    for (uint32_t i = 0; i < n; i++) { order_oracle[i] = a[i].second; order_std[i] = b[i].second; }
    mids[0] = (uint32_t)(ma - a.data());
    mids[1] = (uint32_t)(mb - b.data());
}

}  // extern "C"
