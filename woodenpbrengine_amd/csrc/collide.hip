// Closest-hit traversal for gfx950 (runCollisionSystem, MEngine.cpp:352-361).
//
// The reference collects every primitive of every leaf whose box the ray hits
// (SLBVH.cpp:85-162), intersects them shape-major (CSTriangle.cpp:9-110,
// CSSphere.cpp:84-120) and keeps the first strict minimum of t in candidate order
// (SLBVH.cpp:27-37).  Here one lane owns one ray and walks the same depth-first
// order (near child first by dirIsNeg[axis], SLBVH.cpp:135-143), testing both
// child boxes at the parent from one 64-byte record, with closest-hit pruning:
// a box whose entry distance exceeds the best t found so far (plus 2^-12
// relative slack) holds no candidate that could win, so skipping it leaves the
// winner unchanged (later candidates only win with a strictly smaller t).
// The stack lives in LDS (lane-interleaved, conflict-free); entries beyond its
// depth spill to a per-thread global array.  The reference's stack holds 256
// entries and overflows silently; here an overflow is reported.
#include "wpbr_device.hpp"
#include "wpbr_internal.hpp"

namespace wpbr {

constexpr float PRUNE_SLACK = 1.0f + 1.0f / 4096.0f;

// pbrt-v3 Bounds3::IntersectP with a precomputed 1/dir and dirIsNeg.
WPBR_D bool slab_hit(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, f3 o,
                     f3 inv, bool nx, bool ny, bool nz, float tmax) {
    const float eps = 5.96046448e-08f;  // machine epsilon / 2
    const float g3 = (3 * eps) / (1 - 3 * eps);
    const float scale = 1 + 2 * g3;
    float tMin = ((nx ? mxx : mnx) - o.x) * inv.x;
    float tMx = ((nx ? mnx : mxx) - o.x) * inv.x;
    float tyMin = ((ny ? mxy : mny) - o.y) * inv.y;
    float tyMax = ((ny ? mny : mxy) - o.y) * inv.y;
    tMx *= scale;
    tyMax *= scale;
    if (tMin > tyMax || tyMin > tMx) return false;
    if (tyMin > tMin) tMin = tyMin;
    if (tyMax < tMx) tMx = tyMax;
    float tzMin = ((nz ? mxz : mnz) - o.z) * inv.z;
    float tzMax = ((nz ? mnz : mxz) - o.z) * inv.z;
    tzMax *= scale;
    if (tMin > tzMax || tzMin > tMx) return false;
    if (tzMin > tMin) tMin = tzMin;
    if (tzMax < tMx) tMx = tzMax;
    return (tMin < tmax) && (tMx > 0);
}

// CSTriangle.cpp:9-110 (watertight test without tMax and error bounds), with the
// pbrt-v3 cyclic axis permutation.  Returns 0 for a miss.
WPBR_D float tri_hit(float4 a, float4 b, float4 c, f3 o, f3 d, int kx, int ky, int kz, float sx,
                     float sy, float sz, float* bary) {
    f3 p0 = mk3(a.x - o.x, a.y - o.y, a.z - o.z);
    f3 p1 = mk3(b.x - o.x, b.y - o.y, b.z - o.z);
    f3 p2 = mk3(c.x - o.x, c.y - o.y, c.z - o.z);
    float q0x = comp(p0, kx), q0y = comp(p0, ky), q0z = comp(p0, kz);
    float q1x = comp(p1, kx), q1y = comp(p1, ky), q1z = comp(p1, kz);
    float q2x = comp(p2, kx), q2y = comp(p2, ky), q2z = comp(p2, kz);
    float x0 = q0x + sx * q0z, y0 = q0y + sy * q0z, z0 = q0z + 0.0f * q0z;
    float x1 = q1x + sx * q1z, y1 = q1y + sy * q1z, z1 = q1z + 0.0f * q1z;
    float x2 = q2x + sx * q2z, y2 = q2y + sy * q2z, z2 = q2z + 0.0f * q2z;
    float r0 = x0 * y1 - x1 * y0;
    float r1 = x1 * y2 - x2 * y1;
    float r2 = x2 * y0 - x0 * y2;
    if ((r0 < 0 || r1 < 0 || r2 < 0) && (r0 > 0 || r1 > 0 || r2 > 0)) return 0.0f;
    float det = (r0 + r1) + r2;
    if (det == 0) return 0.0f;
    float tScale = (r1 * (z0 * sz) + r2 * (z1 * sz)) + r0 * (z2 * sz);
    if (det < 0.0f && tScale >= 0) return 0.0f;
    if (det > 0.0f && tScale <= 0) return 0.0f;
    float invDet = 1.0f / det;
    bary[0] = r1 * invDet;
    bary[1] = r2 * invDet;
    bary[2] = r0 * invDet;
    return tScale * invDet;
}

// SSphere::intersect (CSSphere.cpp:84-120) with the pbrt-v3 FP64 quadratic.
WPBR_D float sphere_hit(const GeomView& g, uint32_t id, f3 o, f3 d) {
    affine inv = load_affine(g.sphere_xform + 24 * (size_t)id + 12);
    float radius = g.sphere_radius[id];
    f3 ro = xf_point(inv, o), rd = xf_vector(inv, d);
    float A = len2(rd);
    float B = 2 * dot3(rd, ro);
    float C = len2(ro) - radius * radius;
    double discrim = (double)B * (double)B - 4 * (double)A * (double)C;
    if (discrim < 0) return 0.0f;
    double rdsc = sqrt(discrim);
    double q = (B < 0) ? -0.5 * (B - rdsc) : -0.5 * (B + rdsc);
    float t0 = (float)(q / A), t1 = (float)(C / q);
    if (t0 > t1) { float tmp = t0; t0 = t1; t1 = tmp; }
    if (t1 <= 0) return 0.0f;
    return (t0 <= 0) ? t1 : t0;
}

// Per-lane traversal state of one ray.
struct RayState {
    f3 o, d, inv;
    bool nx, ny, nz;
    int kx, ky, kz;
    float sx, sy, sz;
    float tBest;
    uint32_t best, bestSlot, bestLeaf;
    float bb0, bb1, bb2;
    uint32_t node;  // next record or primitive range; NODE_DONE when finished
    uint32_t sp;
    // visibility mode (a.target set): is the first hit `target` (or nothing)?
    bool occ;
    uint32_t target;
    float tLim;     // t of the target along the ray (INF if it is missed)
};
constexpr uint32_t NODE_DONE = 0xFFFFFFFFu;

// Lane stack: SD entries in LDS (lane-interleaved), the rest in a global spill
// array (entry e of slot `base` at spill[e * stride + base]).
template <int BLOCK, int SD>
struct LaneStack {
    uint32_t* lds;
    uint32_t* spill;
    uint32_t depth;
    size_t stride;
    uint32_t base;
    uint32_t* err;
    WPBR_D void push(uint32_t& sp, uint32_t v) {
        if (sp < (uint32_t)SD) lds[sp * BLOCK + threadIdx.x] = v;
        else if (sp - SD < depth) spill[(size_t)(sp - SD) * stride + base] = v;
        else { atomicOr(err, 1u); return; }
        sp++;
    }
    WPBR_D uint32_t pop(uint32_t& sp) {
        if (sp == 0) return NODE_DONE;
        sp--;
        return sp < (uint32_t)SD ? lds[sp * BLOCK + threadIdx.x] : spill[(size_t)(sp - SD) * stride + base];
    }
};

WPBR_D void ray_begin(RayState& R, const TraceArgs& a, uint32_t i, size_t rs) {
    R.o = mk3(a.rays[i], a.rays[rs + i], a.rays[2 * rs + i]);
    R.d = mk3(a.rays[3 * rs + i], a.rays[4 * rs + i], a.rays[5 * rs + i]);
    R.inv = mk3(1.0f / R.d.x, 1.0f / R.d.y, 1.0f / R.d.z);
    R.nx = R.inv.x < 0; R.ny = R.inv.y < 0; R.nz = R.inv.z < 0;
    // Triangle setup shared by every candidate (pbrt-v3 MaxDimension, shear).
    float adx = fabsf(R.d.x), ady = fabsf(R.d.y), adz = fabsf(R.d.z);
    R.kz = (adx > ady) ? ((adx > adz) ? 0 : 2) : ((ady > adz) ? 1 : 2);
    R.kx = R.kz + 1; if (R.kx == 3) R.kx = 0;
    R.ky = R.kx + 1; if (R.ky == 3) R.ky = 0;
    float dkx = comp(R.d, R.kx), dky = comp(R.d, R.ky), dkz = comp(R.d, R.kz);
    R.sx = -dkx / dkz; R.sy = -dky / dkz; R.sz = 1.0f / dkz;
    R.tBest = INF;
    R.best = R.bestSlot = R.bestLeaf = 0xFFFFFFFFu;
    R.bb0 = R.bb1 = R.bb2 = 0.0f;
    R.occ = a.target != nullptr;
    R.target = 0xFFFFFFFFu;
    R.tLim = INF;
    if (R.occ) {
        R.target = a.target[i];
        if (R.target != 0xFFFFFFFFu) {
            float t;
            if (R.target < a.g.n_spheres) {
                t = sphere_hit(a.g, R.target, R.o, R.d);
            } else {
                const float* v = a.g.tri_verts + 9 * (size_t)(R.target - a.g.n_spheres);
                float b3[3];
                t = tri_hit(make_float4(v[0], v[1], v[2], 0.f), make_float4(v[3], v[4], v[5], 0.f),
                            make_float4(v[6], v[7], v[8], 0.f), R.o, R.d, R.kx, R.ky, R.kz, R.sx, R.sy,
                            R.sz, b3);
            }
            if (t != 0.0f) R.tLim = t;
        }
        // prune beyond the target: only closer candidates can occlude it
        R.tBest = R.tLim;
    }
    const float* rb = a.root_box;
    R.node = slab_hit(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], R.o, R.inv, R.nx, R.ny, R.nz, INF)
                 ? a.root_word : NODE_DONE;
    R.sp = 0;
}

WPBR_D void ray_end(const RayState& R, const TraceArgs& a, uint32_t i, size_t rs) {
    if (R.occ) {
        // occluded: the occluder; else the target when the ray reaches it, or a miss
        const bool occluded = R.best != 0xFFFFFFFFu;
        a.hit_shape[i] = occluded ? R.best : (R.tLim != INF ? R.target : 0xFFFFFFFFu);
        a.hit_t[i] = R.tBest;
        return;
    }
    a.hit_shape[i] = R.best;
    a.hit_t[i] = R.tBest;
    if (a.bary) {
        a.bary[i] = R.bb0;
        a.bary[rs + i] = R.bb1;
        a.bary[2 * rs + i] = R.bb2;
    }
}

// Speculative traversal (Aila and Laine 2009): a lane that reaches a primitive
// range while other lanes of its wave are still looking for one parks it and
// keeps walking records; the wave turns to the primitives once every lane holds
// a range (or has none left).  Ranges are still tested in depth-first order
// (the parked one first), and records walked while a range is parked are tested
// against a best t that is not yet lowered by it, so no record that could hold
// the winner is skipped: the closest hit is unchanged.
constexpr bool TRACE_SPECULATIVE = true;

// One step of the while-while loop: records until a primitive range (or the
// end), then primitive ranges until a record (or the end).
template <int BLOCK, int SD, int WIDTH>
WPBR_D void ray_step(RayState& R, const TraceArgs& a, LaneStack<BLOCK, SD>& S) {
    uint32_t node = R.node;
    const f3 o = R.o, inv = R.inv;
    const bool nx = R.nx, ny = R.ny, nz = R.nz;
    // the next node after record `nd`: its nearest hit child, the others pushed
    auto visit = [&](uint32_t nd) -> uint32_t {
        if (WIDTH == 2) {
            // Binary records: test both children, near first by dirIsNeg[axis].
            const float4* r = a.trav + 4 * (size_t)nd;
            float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
            float lim = R.tBest * PRUNE_SLACK;
            bool h0 = slab_hit(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, o, inv, nx, ny, nz, lim);
            bool h1 = slab_hit(r1.z, r1.w, r2.x, r2.y, r2.z, r2.w, o, inv, nx, ny, nz, lim);
            uint32_t w0 = __float_as_uint(r3.x), w1 = __float_as_uint(r3.y);
            uint32_t axis = __float_as_uint(r3.z);
            bool negAxis = (axis == 0) ? nx : ((axis == 1) ? ny : nz);
            uint32_t nearW = negAxis ? w1 : w0, farW = negAxis ? w0 : w1;
            bool hn = negAxis ? h1 : h0, hf = negAxis ? h0 : h1;
            if (hn && hf) { S.push(R.sp, farW); return nearW; }
            if (hn) return nearW;
            if (hf) return farW;
            return S.pop(R.sp);
        } else {
            // 4-wide records (the grandchildren of a binary record): the pair of the
            // near child first, the near grandchild first inside each pair; hits
            // after the first are pushed in reverse visiting order.
            const float4* r = a.trav4 + 8 * (size_t)nd;
            float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4], r5 = r[5], r6 = r[6], r7 = r[7];
            float lim = R.tBest * PRUNE_SLACK;
            uint32_t w[4] = {__float_as_uint(r6.x), __float_as_uint(r6.y), __float_as_uint(r6.z),
                             __float_as_uint(r6.w)};
            bool h[4];
            h[0] = slab_hit(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, o, inv, nx, ny, nz, lim);
            h[1] = w[1] != 0xFFFFFFFFu && slab_hit(r1.z, r1.w, r2.x, r2.y, r2.z, r2.w, o, inv, nx, ny, nz, lim);
            h[2] = slab_hit(r3.x, r3.y, r3.z, r3.w, r4.x, r4.y, o, inv, nx, ny, nz, lim);
            h[3] = w[3] != 0xFFFFFFFFu && slab_hit(r4.z, r4.w, r5.x, r5.y, r5.z, r5.w, o, inv, nx, ny, nz, lim);
            const uint32_t axes = __float_as_uint(r7.x);
            auto neg = [&](uint32_t ax) { return ax == 0 ? nx : (ax == 1 ? ny : nz); };
            const uint32_t pf = neg(axes & 3u) ? 1u : 0u;         // pair visited first
            const uint32_t i0 = neg((axes >> 2) & 3u) ? 1u : 0u;  // first inside pair 0
            const uint32_t i1 = neg((axes >> 4) & 3u) ? 1u : 0u;  // first inside pair 1
            const uint32_t f0 = pf ? 2 + i1 : i0;
            const uint32_t f1 = pf ? i0 : 2 + i1;
            const uint32_t seq[4] = {f0, f0 ^ 1u, f1, f1 ^ 1u};
            uint32_t next = NODE_DONE;
#pragma unroll
            for (int q = 3; q >= 0; q--) {
                const uint32_t sl = seq[q];
                bool hq = sl == 0 ? h[0] : (sl == 1 ? h[1] : (sl == 2 ? h[2] : h[3]));
                uint32_t wq = sl == 0 ? w[0] : (sl == 1 ? w[1] : (sl == 2 ? w[2] : w[3]));
                if (hq) {
                    if (next != NODE_DONE) S.push(R.sp, next);
                    next = wq;
                }
            }
            return (next != NODE_DONE) ? next : S.pop(R.sp);
        }
    };
    uint32_t parked = NODE_DONE;
    if (TRACE_SPECULATIVE) {
        while (true) {
            if (node != NODE_DONE && !(node & 0x80000000u)) node = visit(node);
            else if (node != NODE_DONE && parked == NODE_DONE) { parked = node; node = S.pop(R.sp); }
            if (__all(parked != NODE_DONE || node == NODE_DONE)) break;
        }
    } else {
        while (node != NODE_DONE && !(node & 0x80000000u)) node = visit(node);
    }
    // The primitives of range `leaf`; true when the ray is settled or restarts
    // (R.node set).
    auto leaf_prims = [&](uint32_t leaf_word) -> bool {
        uint32_t first = leaf_word & 0x0FFFFFFFu, cnt = ((leaf_word >> 28) & 7u) + 1u;
        for (uint32_t j = first; j < first + cnt; j++) {
            const float4* p = a.prims + 3 * (size_t)j;
            float4 v0 = p[0], v1 = p[1], v2 = p[2];
            uint32_t id = __float_as_uint(v0.w);
            uint32_t slot = __float_as_uint(v1.w), leaf = __float_as_uint(v2.w);
            if (R.occ && id == R.target) continue;  // the target never hides itself
            float t;
            float b3[3] = {0.0f, 0.0f, 0.0f};
            if (id < a.g.n_spheres) t = sphere_hit(a.g, id, R.o, R.d);
            else t = tri_hit(v0, v1, v2, R.o, R.d, R.kx, R.ky, R.kz, R.sx, R.sy, R.sz, b3);
            if (R.occ) {
                if (t != 0.0f && t <= R.tLim) {
                    if (t < R.tLim) {
                        // a candidate before the target: the first hit is not the target
                        R.best = id;
                        R.tBest = t;
                        R.node = NODE_DONE;
                        R.sp = 0;
                        return true;
                    }
                    // equal t: the candidate order decides; redo as a closest-hit query
                    R.occ = false;
                    R.tBest = INF;
                    R.best = R.bestSlot = R.bestLeaf = 0xFFFFFFFFu;
                    R.sp = 0;
                    R.node = a.root_word;
                    return true;
                }
                continue;
            }
            // first strict minimum in candidate order: leaves in depth-first
            // order, slots in order inside a leaf
            if (t != 0.0f && (t < R.tBest || (t == R.tBest && leaf == R.bestLeaf && slot < R.bestSlot))) {
                R.tBest = t;
                R.best = id;
                R.bestSlot = slot;
                R.bestLeaf = leaf;
                R.bb0 = b3[0]; R.bb1 = b3[1]; R.bb2 = b3[2];
            }
        }
        return false;
    };
    // Primitive ranges in depth-first order: the parked one, then the current run.
    if (parked != NODE_DONE && leaf_prims(parked)) return;
    while (node != NODE_DONE && (node & 0x80000000u)) {
        if (leaf_prims(node)) return;
        node = S.pop(R.sp);
    }
    R.node = node;
}

// One lane per ray, rays assigned to workgroups statically.
template <int BLOCK, int SD, int WIDTH>
__global__ __launch_bounds__(BLOCK) void k_trace(TraceArgs a) {
    __shared__ uint32_t stack[SD * BLOCK];
    const uint32_t n = a.n_dev ? *a.n_dev : a.n;
    uint32_t i;
    if (a.parts > 1) {
        // workgroup b runs on XCD b % parts: that XCD takes one contiguous part
        const uint32_t x = blockIdx.x % a.parts, j = blockIdx.x / a.parts;
        const uint32_t lo = (uint32_t)(((uint64_t)n * x) / a.parts);
        const uint32_t hi = (uint32_t)(((uint64_t)n * (x + 1)) / a.parts);
        i = lo + j * BLOCK + threadIdx.x;
        if (i >= hi) return;
    } else {
        i = blockIdx.x * BLOCK + threadIdx.x;
        if (i >= n) return;
    }
    const size_t rs = a.stride ? a.stride : n;
    LaneStack<BLOCK, SD> S{stack, a.spill, a.spill_depth, (size_t)a.n, i, a.err};
    RayState R;
    ray_begin(R, a, i, rs);
    while (R.node != NODE_DONE) ray_step<BLOCK, SD, WIDTH>(R, a, S);
    ray_end(R, a, i, rs);
}

// Persistent lanes with dynamic ray fetch: a lane whose ray is finished takes
// the next ray of its wave's pool once at least `refill` lanes of the wave are
// idle; the pool is refilled with `chunk` consecutive rays by one atomic on the
// global claim counter.  Every wave leaves when the rays are exhausted and its
// lanes are idle; the last wave out resets the claim counters for the next
// launch.
template <int BLOCK, int SD, int WIDTH>
__global__ __launch_bounds__(BLOCK) void k_trace_dyn(TraceArgs a) {
    __shared__ uint32_t stack[SD * BLOCK];
    const uint32_t n = a.n_dev ? *a.n_dev : a.n;
    const size_t rs = a.stride ? a.stride : n;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    LaneStack<BLOCK, SD> S{stack, a.spill, a.spill_depth, (size_t)gridDim.x * BLOCK, gtid, a.err};
    const uint32_t lane = __lane_id();
    RayState R;
    R.node = NODE_DONE;
    R.sp = 0;
    bool active = false;
    bool exhausted = false;                // wave-uniform: the claim counter has passed n
    uint32_t pool_next = 0, pool_end = 0;  // wave-uniform: unassigned rays of the wave
    uint32_t i = 0;
    while (true) {
        const uint64_t idle = ballot64(!active);
        const uint32_t nidle = popc64(idle);
        if ((!exhausted || pool_next < pool_end) && nidle >= a.refill) {
            // ranks below the pool's remainder take pool rays, the rest a new chunk
            const uint32_t rem = pool_end - pool_next;
            uint32_t nbase = 0, nend = 0;
            if (rem < nidle && !exhausted) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(a.claim, a.chunk);
                base = (uint32_t)__shfl((int)base, 0);
                if (base >= n) exhausted = true;
                else { nbase = base; nend = min(base + a.chunk, n); if (nend == n) exhausted = true; }
            }
            const uint32_t fresh = nend - nbase;
            const uint32_t take = min(nidle, rem + fresh);
            if (!active) {
                const uint32_t r = popc64(idle & lanemask_lt());
                if (r < take) {
                    i = (r < rem) ? pool_next + r : nbase + (r - rem);
                    ray_begin(R, a, i, rs);
                    active = true;
                }
            }
            if (take <= rem) {
                pool_next += take;
            } else {
                pool_next = nbase + (take - rem);
                pool_end = nend;
            }
        }
        if (ballot64(active) == 0) {
            if (exhausted && pool_next >= pool_end) break;
            continue;
        }
        if (active) {
            if (R.node != NODE_DONE) ray_step<BLOCK, SD, WIDTH>(R, a, S);
            if (R.node == NODE_DONE) {
                ray_end(R, a, i, rs);
                active = false;
            }
        }
    }
    if (lane == 0) {
        __threadfence();
        const uint32_t waves = gridDim.x * (BLOCK / 64);
        if (atomicAdd(a.claim + 1, 1u) == waves - 1) {
            a.claim[0] = 0;
            a.claim[1] = 0;
        }
    }
}

// Batched form of k_trace: block b traces for scene map.scene[b] (its launch
// record T[scene]: geometry, tree, rays, hits, spill slots) as its block
// b - map.first[scene].
template <int BLOCK, int SD, int WIDTH>
__global__ __launch_bounds__(BLOCK) void k_trace_seg(const TraceArgs* __restrict__ T, SegMap map) {
    __shared__ uint32_t stack[SD * BLOCK];
    const uint32_t sc = map.scene[blockIdx.x];
    const uint32_t bid = blockIdx.x - map.first[sc];
    const TraceArgs a = T[sc];
    const uint32_t n = a.n_dev ? *a.n_dev : a.n;
    const uint32_t i = bid * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t rs = a.stride ? a.stride : n;
    LaneStack<BLOCK, SD> S{stack, a.spill, a.spill_depth, (size_t)a.n, i, a.err};
    RayState R;
    ray_begin(R, a, i, rs);
    while (R.node != NODE_DONE) ray_step<BLOCK, SD, WIDTH>(R, a, S);
    ray_end(R, a, i, rs);
}

template <int BLOCK, int WIDTH>
static void launch_seg_sd(const TraceArgs* T, const SegMap& map, uint32_t grid, int sd, hipStream_t st) {
#define WPBR_TRACE_SEG(SDV) \
    case SDV: hipLaunchKernelGGL((k_trace_seg<BLOCK, SDV, WIDTH>), dim3(grid), dim3(BLOCK), 0, st, T, map); break;
    switch (sd) {
        WPBR_TRACE_SEG(8)
        WPBR_TRACE_SEG(16)
        WPBR_TRACE_SEG(24)
        WPBR_TRACE_SEG(32)
        WPBR_TRACE_SEG(48)
        default:
        WPBR_TRACE_SEG(64)
    }
#undef WPBR_TRACE_SEG
}

template <int BLOCK>
static void launch_seg_w(const TraceArgs* T, const SegMap& map, uint32_t grid, int sd, int width, hipStream_t st) {
    if (width == 4) launch_seg_sd<BLOCK, 4>(T, map, grid, sd, st);
    else launch_seg_sd<BLOCK, 2>(T, map, grid, sd, st);
}

// Traces the rays of several launch records in one launch (grid: the map's
// total blocks for this block size).
wpbr_status trace_launch_seg(const TraceArgs* T, const SegMap& map, uint32_t grid, int block, int sd, int width,
                             hipStream_t st) {
    if (grid == 0) return WPBR_OK;
    switch (block) {
        case 64: launch_seg_w<64>(T, map, grid, sd, width, st); break;
        case 128: launch_seg_w<128>(T, map, grid, sd, width, st); break;
        case 512: launch_seg_w<512>(T, map, grid, sd, width, st); break;
        default: launch_seg_w<256>(T, map, grid, sd, width, st); break;
    }
    return launch_status();
}

template <int BLOCK, int WIDTH, bool DYN>
static void launch_sd(const TraceArgs& a, int sd, hipStream_t st, uint32_t dyn_blocks) {
    dim3 grid;
    if (DYN) {
        grid = dim3(dyn_blocks);
    } else {
        const uint32_t P = a.parts > 1 ? a.parts : 1;
        const uint32_t per_part = (a.n + P - 1) / P;
        grid = dim3(P * ((per_part + BLOCK - 1) / BLOCK));
    }
#define WPBR_TRACE_CASE(SDV)                                                                         \
    case SDV:                                                                                        \
        if constexpr (DYN) hipLaunchKernelGGL((k_trace_dyn<BLOCK, SDV, WIDTH>), grid, dim3(BLOCK), 0, st, a); \
        else hipLaunchKernelGGL((k_trace<BLOCK, SDV, WIDTH>), grid, dim3(BLOCK), 0, st, a);             \
        break;
    switch (sd) {
        WPBR_TRACE_CASE(8)
        WPBR_TRACE_CASE(16)
        WPBR_TRACE_CASE(24)
        WPBR_TRACE_CASE(32)
        WPBR_TRACE_CASE(48)
        default:
        WPBR_TRACE_CASE(64)
    }
#undef WPBR_TRACE_CASE
}

template <int BLOCK>
static void launch_w(const TraceArgs& a, int sd, hipStream_t st, uint32_t dyn_blocks) {
    const bool dyn = dyn_blocks > 0;
    if (a.width == 4) {
        if (dyn) launch_sd<BLOCK, 4, true>(a, sd, st, dyn_blocks);
        else launch_sd<BLOCK, 4, false>(a, sd, st, 0);
    } else {
        if (dyn) launch_sd<BLOCK, 2, true>(a, sd, st, dyn_blocks);
        else launch_sd<BLOCK, 2, false>(a, sd, st, 0);
    }
}

// dyn_blocks > 0 selects the persistent dynamic-fetch kernel with that many
// workgroups (the spill array then has dyn_blocks * block slots).
wpbr_status trace_launch(const TraceArgs& a, int block, int sd, hipStream_t st, uint32_t dyn_blocks) {
    if (a.n == 0) return WPBR_OK;
    switch (block) {
        case 64: launch_w<64>(a, sd, st, dyn_blocks); break;
        case 128: launch_w<128>(a, sd, st, dyn_blocks); break;
        case 512: launch_w<512>(a, sd, st, dyn_blocks); break;
        default: launch_w<256>(a, sd, st, dyn_blocks); break;
    }
    return launch_status();
}

}  // namespace wpbr
