// Oracle: runCollisionSystem (MEngine.cpp:352-361).
// TEST INFRASTRUCTURE ONLY.
//
//   JobProcessRayCasts                 SLBVH.cpp:85-162  (unpruned DFS, stack of 256)
//   JobProcessTriangleInteractionReqs  CSTriangle.cpp:9-110 (watertight test)
//   SSphere::intersect                 CSSphere.cpp:84-120
//   JobProcessRayCastsResults          SLBVH.cpp:9-66    (argmin, strict '>', DFS order)
#include "oracle.h"

namespace orc {

// pbrt-v3 Bounds3::IntersectP(ray, invDir, dirIsNeg) - the textbook slab test the
// reference calls as DBounds3f::intersect (SLBVH.cpp:106).  tMax is the ray's
// extent (+inf: the reference sets none).
bool box_hit(const float* bmn, const float* bmx, V3 o, V3 inv, const int* neg, float tMax,
             float* tNearOut) {
    const float g3 = 3 * (std::numeric_limits<float>::epsilon() * 0.5f) /
                     (1 - 3 * (std::numeric_limits<float>::epsilon() * 0.5f));
    const float scale = 1 + 2 * g3;
    float tMin = ((neg[0] ? bmx[0] : bmn[0]) - o.x) * inv.x;
    float tMx = ((neg[0] ? bmn[0] : bmx[0]) - o.x) * inv.x;
    float tyMin = ((neg[1] ? bmx[1] : bmn[1]) - o.y) * inv.y;
    float tyMax = ((neg[1] ? bmn[1] : bmx[1]) - o.y) * inv.y;
    tMx *= scale;
    tyMax *= scale;
    if (tMin > tyMax || tyMin > tMx) return false;
    if (tyMin > tMin) tMin = tyMin;
    if (tyMax < tMx) tMx = tyMax;
    float tzMin = ((neg[2] ? bmx[2] : bmn[2]) - o.z) * inv.z;
    float tzMax = ((neg[2] ? bmn[2] : bmx[2]) - o.z) * inv.z;
    tzMax *= scale;
    if (tMin > tzMax || tzMin > tMx) return false;
    if (tzMin > tMin) tMin = tzMin;
    if (tzMax < tMx) tMx = tzMax;
    if (tNearOut) *tNearOut = tMin;
    return (tMin < tMax) && (tMx > 0);
}

// CSTriangle.cpp:9-110.  Returns tHitResult (0 = miss) and barycentrics.
// The axis permutation follows pbrt-v3 (kx = kz+1 mod 3, ky = kx+1 mod 3).  The
// reference writes ky = kz+2 (CSTriangle.cpp:36), which indexes component 3 of a
// 3-vector when kz == 1; it agrees with pbrt-v3 for kz = 0 and kz = 2.
float tri_intersect(const float* v, V3 o, V3 d, float* bary) {
    float ad[3] = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    // pbrt-v3 MaxDimension
    int kz = (ad[0] > ad[1]) ? ((ad[0] > ad[2]) ? 0 : 2) : ((ad[1] > ad[2]) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    V3 dp = v3(d[kx], d[ky], d[kz]);
    // shear = (-d.x, -d.y, 1) / d.z (component-wise IEEE division)
    float sx = -dp.x / dp.z, sy = -dp.y / dp.z, sz = 1.0f / dp.z;
    float px[3], py[3], pz[3];
    for (int i = 0; i < 3; i++) {
        V3 p = v3(v[3 * i] - o.x, v[3 * i + 1] - o.y, v[3 * i + 2] - o.z);
        V3 q = v3(p[kx], p[ky], p[kz]);
        // p += (sx, sy, 0) * p.z
        px[i] = q.x + sx * q.z;
        py[i] = q.y + sy * q.z;
        pz[i] = q.z + 0.0f * q.z;
    }
    // res = X * rot(Y) - rot(X) * Y with rot = (1,2,0)  (CSTriangle.cpp:60-63)
    float r0 = px[0] * py[1] - px[1] * py[0];
    float r1 = px[1] * py[2] - px[2] * py[1];
    float r2 = px[2] * py[0] - px[0] * py[2];
    if ((r0 < 0 || r1 < 0 || r2 < 0) && (r0 > 0 || r1 > 0 || r2 > 0)) return 0.0f;
    float det = (r0 + r1) + r2;
    if (det == 0) return 0.0f;
    // res = res.permute(1,2,0) -> (e0, e1, e2) = (r1, r2, r0)
    float e0 = r1, e1 = r2, e2 = r0;
    float tScale = (e0 * (pz[0] * sz) + e1 * (pz[1] * sz)) + e2 * (pz[2] * sz);
    if (det < 0.0f && tScale >= 0) return 0.0f;
    if (det > 0.0f && tScale <= 0) return 0.0f;
    float invDet = (float)(1.0 / (double)det);
    if (bary) { bary[0] = e0 * invDet; bary[1] = e1 * invDet; bary[2] = e2 * invDet; }
    return tScale * invDet;
}

// pbrt-v3 Quadratic (FP64), standing in for WoodenMath getRootsQuadraticEquation.
static bool quadratic(float a, float b, float c, float* t0, float* t1) {
    double discrim = (double)b * (double)b - 4 * (double)a * (double)c;
    if (discrim < 0) return false;
    double rd = std::sqrt(discrim);
    double q = (b < 0) ? -0.5 * (b - rd) : -0.5 * (b + rd);
    *t0 = (float)(q / a);
    *t1 = (float)(c / q);
    if (*t0 > *t1) std::swap(*t0, *t1);
    return true;
}

// SSphere::intersect, CSSphere.cpp:84-120.  Returns tHitResult (0 = miss).
float sphere_intersect(const float* xform, float radius, V3 o, V3 d, V3* oL, V3* dL) {
    Affine inv;
    std::memcpy(&inv, xform + 12, 48);
    V3 ro = xf_point(inv, o), rd = xf_vector(inv, d);
    float a = length2(rd);
    float b = 2 * dot3(rd, ro);
    float c = length2(ro) - radius * radius;
    float t0, t1;
    if (!quadratic(a, b, c, &t0, &t1)) return 0.0f;
    if (t1 <= 0) return 0.0f;
    float t = t0;
    if (t <= 0) t = t1;
    if (oL) *oL = ro;
    if (dL) *dL = rd;
    return t;
}

// One ray against the scene, following SLBVH.cpp:85-162 then :9-66.
// Returns the winning sorted-primitive slot (or -1) and t.
int64_t trace_ray(const Scene& sc, const LbvhResult& bvh, V3 o, V3 d, float* tOut,
                  uint64_t* nCandidates, int* stackOverflow) {
    V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
    int stack[256];
    int sp = 0;
    int cur = 0;
    float best = std::numeric_limits<float>::infinity();
    int64_t bestSlot = -1;
    uint64_t cand = 0;
    if (bvh.n_nodes == 0) { *tOut = best; return -1; }
    while (true) {
        const float* nb = &bvh.node_bounds[6 * (size_t)cur];
        if (box_hit(nb, nb + 3, o, inv, neg, std::numeric_limits<float>::infinity(), nullptr)) {
            uint32_t ns = bvh.node_nshapes[cur];
            if (ns > 0) {
                uint32_t off = bvh.node_offset[cur];
                for (uint32_t i = 0; i < ns; i++) {
                    uint32_t slot = off + i;
                    uint32_t shape = bvh.sorted_index[slot];
                    float t;
                    if (shape < sc.n_spheres) {
                        t = sphere_intersect(sc.sphere_xform + 24 * (size_t)shape,
                                             sc.sphere_radius[shape], o, d, nullptr, nullptr);
                    } else {
                        uint32_t tri = shape - sc.n_spheres;
                        t = tri_intersect(sc.tri_verts + 9 * (size_t)tri, o, d, nullptr);
                    }
                    cand++;
                    // JobProcessRayCastsResults: skip 0, strict '>' keeps the first.
                    if (t == 0.0f) continue;
                    if (best > t) { best = t; bestSlot = slot; }
                }
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                if (sp >= 256) { if (stackOverflow) *stackOverflow = 1; break; }
                if (neg[bvh.node_axis[cur]]) {
                    stack[sp++] = cur + 1;
                    cur = (int)bvh.node_offset[cur];
                } else {
                    stack[sp++] = (int)bvh.node_offset[cur];
                    cur = cur + 1;
                }
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    if (nCandidates) *nCandidates += cand;
    *tOut = best;
    return (best == std::numeric_limits<float>::infinity()) ? -1 : bestSlot;
}

}  // namespace orc
