// Oracle: the interactions runCollisionSystem leaves on each ray cast.
// TEST INFRASTRUCTURE ONLY - restates the reference's CPU jobs.
//
//   plain interaction of a ray cast        SLBVH.cpp:40-62 (JobProcessRayCastsResults):
//                                          miss -> hCollision INVALID; bSurfInteraction
//                                          false -> p = ray(t), wo = -dir, time = t
//   triangle surface interaction           CSTriangle.cpp:149-319
//                                          (JobProcessTriangleFullInteractionRequests),
//                                          the constructor and setShadingGeometry of
//                                          CSurfaceInteraction (CSufraceInteraction.h:54-196)
//   sphere surface interaction             SSphere::computeSurfInteraction, CSSphere.cpp:122-186,
//                                          then CTransform::operator()(SI), CTransform.h:41-66
//   differentials                          CSurfaceInteraction::computeDifferentials,
//                                          CSufraceInteraction.h:98-148, for the rays that
//                                          carry a CRayDifferential
//                                          (JobComputeDifferentialsForSurfInter, :200-225)
//
// Decisions where the reference reads something WoodenMath (absent) defines:
//   * a vector times a double scalar (`dpdu *= 2 * PI`) takes the scalar rounded
//     to float, as the render's restatement does;
//   * DNormal3f members (n, dndu, dndv, shading.n) transform by the inverse
//     transpose, vectors (dpdu, dpdv, wo, shading.dpdu/dpdv/dndu/dndv, which the
//     shading struct declares as DVector3f) by the matrix;
//   * default-constructed vectors, points and uv are zero: the sphere's world
//     interaction keeps uv = (0, 0) (CTransform.h:50 leaves the copy commented
//     out) and dpdx = dpdy = 0 (dpdy is set from dpdx, :52), and a triangle's
//     geometric dndu / dndv stay 0 (only the shading pair gets the normal
//     derivatives);
//   * solveLinearSystem2x2 is pbrt-v3's SolveLinearSystem2x2 (|det| < 1e-10 or a
//     NaN solution fails, and the reference then sets both unknowns to 0).
#include <thread>
#include <vector>

#include "oracle.h"

namespace orc {

namespace {

Affine load_aff(const float* p) { Affine a; std::memcpy(&a, p, 48); return a; }

struct FullSI {
    V3 p, wo;
    float time;
    V3 n;
    float uv[2];
    V3 dpdu, dpdv, dndu, dndv;
    V3 sn, sdpdu, sdpdv, sdndu, sdndv;
    V3 dpdx, dpdy;
    float dudx, dvdx, dudy, dvdy;
};

FullSI zero_si() {
    FullSI s;
    std::memset(&s, 0, sizeof(s));
    return s;
}

// CSurfaceInteraction's constructor (CSufraceInteraction.h:58-83).
void si_construct(FullSI& s, V3 p, const float* uv, V3 wo, V3 dpdu, V3 dpdv, V3 dndu, V3 dndv, float time) {
    s.p = p;
    s.uv[0] = uv[0];
    s.uv[1] = uv[1];
    s.wo = wo;
    s.time = time;
    s.dpdu = dpdu;
    s.dpdv = dpdv;
    s.dndu = dndu;
    s.dndv = dndv;
    s.n = normalize3(cross3(dpdu, dpdv));
    s.sn = s.n;
    s.sdpdu = dpdu;
    s.sdpdv = dpdv;
    s.sdndu = dndu;
    s.sdndv = dndv;
}

// CSTriangle.cpp:149-319.
FullSI si_triangle_full(const RenderScene& sc, uint32_t tri, V3 rd, const float* b, float t) {
    const float* v = sc.geom.tri_verts + 9 * (size_t)tri;
    const V3 P0 = v3(v[0], v[1], v[2]), P1 = v3(v[3], v[4], v[5]), P2 = v3(v[6], v[7], v[8]);
    const uint8_t flags = sc.tri_flags ? sc.tri_flags[tri] : 0;
    float uv[3][2] = {{0, 0}, {1, 0}, {1, 1}};
    if (flags & 1) {
        const float* u = sc.tri_uv + 6 * (size_t)tri;
        for (int i = 0; i < 3; i++) { uv[i][0] = u[2 * i]; uv[i][1] = u[2 * i + 1]; }
    }
    const V3 dp02 = P0 - P2, dp12 = P1 - P2;
    float duv02[2] = {uv[0][0] - uv[2][0], uv[0][1] - uv[2][1]};
    float duv12[2] = {uv[1][0] - uv[2][0], uv[1][1] - uv[2][1]};
    float det = duv02[0] * duv12[1] - duv02[1] * duv12[0];
    if (det == 0) {
        const float k = (float)(1.0 * 0.001);
        uv[0][0] = 0; uv[0][1] = 0; uv[1][0] = k; uv[1][1] = 0; uv[2][0] = k; uv[2][1] = k;
        duv02[0] = uv[0][0] - uv[2][0]; duv02[1] = uv[0][1] - uv[2][1];
        duv12[0] = uv[1][0] - uv[2][0]; duv12[1] = uv[1][1] - uv[2][1];
        det = duv02[0] * duv12[1] - duv02[1] * duv12[0];
    }
    const float detInv = 1 / det;
    const V3 dpdu = (dp02 * duv12[1] - dp12 * duv02[1]) * detInv;
    const V3 dpdv = ((-dp02) * duv12[0] + dp12 * duv02[0]) * detInv;
    const V3 pHit = (P0 * b[0] + P1 * b[1]) + P2 * b[2];
    const float uvHit[2] = {(uv[0][0] * b[0] + uv[1][0] * b[1]) + uv[2][0] * b[2],
                            (uv[0][1] * b[0] + uv[1][1] * b[1]) + uv[2][1] * b[2]};
    FullSI s = zero_si();
    si_construct(s, pHit, uvHit, -rd, dpdu, dpdv, v3(0, 0, 0), v3(0, 0, 0), t);
    s.n = s.sn = normalize3(cross3(dp02, dp12));
    if (flags & 2) {
        const float* nn = sc.tri_normal + 9 * (size_t)tri;
        const float* tt = sc.tri_tangent + 9 * (size_t)tri;
        const V3 N0 = v3(nn[0], nn[1], nn[2]), N1 = v3(nn[3], nn[4], nn[5]), N2 = v3(nn[6], nn[7], nn[8]);
        const V3 T0 = v3(tt[0], tt[1], tt[2]), T1 = v3(tt[3], tt[4], tt[5]), T2 = v3(tt[6], tt[7], tt[8]);
        V3 ns = (N0 * b[0] + N1 * b[1]) + N2 * b[2];
        ns = is_zero3(ns) ? s.n : normalize3(ns);
        V3 ss = (T0 * b[0] + T1 * b[1]) + T2 * b[2];
        ss = is_zero3(ss) ? normalize3(s.dpdu) : normalize3(ss);
        V3 ts = cross3(ss, ns);
        if (length2(ts) > 0) {
            ts = normalize3(ts);
            ss = cross3(ts, ns);
        }
        const V3 dn02 = N0 - N2, dn12 = N1 - N2;
        V3 dndu = v3(0, 0, 0), dndv = v3(0, 0, 0);
        const float det2 = duv02[0] * duv12[1] - duv02[1] * duv12[0];
        if (det2 != 0) {
            const float inv2 = 1 / det2;
            dndu = (dn02 * duv12[1] - dn12 * duv02[1]) * inv2;
            dndv = ((-dn02) * duv12[0] + dn12 * duv02[0]) * inv2;
        }
        // setShadingGeometry(ss, ts, dndu, dndv, true), CSufraceInteraction.h:171-196
        s.sn = normalize3(cross3(ss, ts));
        s.sdpdu = ss;
        s.sdpdv = ts;
        s.sdndu = dndu;
        s.sdndv = dndv;
        s.n = face_forward(s.n, s.sn);
    }
    return s;
}

// CSSphere.cpp:122-186 in the sphere's space, then CTransform.h:41-66.
FullSI si_sphere_full(const RenderScene& sc, uint32_t sid, V3 o, V3 d, float t) {
    const float* xf = sc.geom.sphere_xform + 24 * (size_t)sid;
    const Affine W = load_aff(xf), Inv = load_aff(xf + 12);
    const float radius = sc.geom.sphere_radius[sid];
    const V3 oL = xf_point(Inv, o), dL = xf_vector(Inv, d);
    const V3 pH = oL + dL * t;
    float phi = cr_atan2(pH.y, pH.x);
    if (phi < 0.0) phi = (float)((double)phi + 2 * PI_REF);
    const float u = (float)((double)phi / (2 * PI_REF));
    const float theta = cr_acos(pH.z / radius);
    const float v = (float)((double)theta / PI_REF);
    const float invZR = (float)(1.0 / (double)std::sqrt(pH.x * pH.x + pH.y * pH.y));
    const float cphi = pH.x * invZR, sphi = pH.y * invZR;
    const V3 dpdu = v3(-pH.y, pH.x, 0.0f) * (float)(2 * PI_REF);
    const V3 dpdv = v3(pH.z * cphi, pH.z * sphi, -radius * cr_sin(theta)) * (float)PI_REF;
    const V3 d2Pduu = v3(pH.x, pH.y, 0) * (float)(-2 * PI_REF * 2 * PI_REF);
    const V3 d2Pduv = (((v3(-sphi, cphi, 0.0f) * (float)PI_REF) * pH.z) * 2.0f) * (float)PI_REF;
    const V3 d2Pdvv = (v3(pH.x, pH.y, pH.z) * (float)(-PI_REF)) * (float)PI_REF;
    const float E = dot3(dpdu, dpdu), F = dot3(dpdu, dpdv), G = dot3(dpdv, dpdv);
    const V3 N = normalize3(cross3(dpdu, dpdv));
    const float e = dot3(N, d2Pduu), f = dot3(N, d2Pduv), g = dot3(N, d2Pdvv);
    const float invEGF2 = 1 / (E * G - F * F);
    const V3 dndu = dpdu * ((f * F - e * G) * invEGF2) + dpdv * ((e * F - f * E) * invEGF2);
    const V3 dndv = dpdu * ((g * F - f * G) * invEGF2) + dpdv * ((f * F - g * E) * invEGF2);
    FullSI l = zero_si();
    const float uv[2] = {u, v};
    si_construct(l, pH, uv, -dL, dpdu, dpdv, dndu, dndv, t);
    // CTransform::operator()(const CSurfaceInteraction&): a fresh interaction
    // (uv, dudx .. dvdy default) filled field by field
    FullSI s = zero_si();
    s.time = l.time;
    s.dndu = xf_normal(Inv, l.dndu);
    s.dndv = xf_normal(Inv, l.dndv);
    s.dpdu = xf_vector(W, l.dpdu);
    s.dpdv = xf_vector(W, l.dpdv);
    s.p = xf_point(W, l.p);
    s.dpdx = xf_vector(W, l.dpdx);
    s.dpdy = xf_vector(W, l.dpdx);
    s.n = normalize3(xf_normal(Inv, l.n));
    s.wo = normalize3(xf_vector(W, l.wo));
    s.sdndu = xf_vector(W, l.sdndu);
    s.sdndv = xf_vector(W, l.sdndv);
    s.sdpdu = xf_vector(W, l.sdpdu);
    s.sdpdv = xf_vector(W, l.sdpdv);
    s.sn = normalize3(xf_normal(Inv, l.sn));
    return s;
}

// pbrt-v3 SolveLinearSystem2x2.
bool solve2x2(const float A[2][2], const float B[2], float* x0, float* x1) {
    const float det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    if (std::fabs(det) < 1e-10f) return false;
    *x0 = (A[1][1] * B[0] - A[0][1] * B[1]) / det;
    *x1 = (A[0][0] * B[1] - A[1][0] * B[0]) / det;
    if (std::isnan(*x0) || std::isnan(*x1)) return false;
    return true;
}

// CSufraceInteraction.h:98-148.
void differentials(FullSI& s, V3 xo, V3 xd, V3 yo, V3 yd) {
    const float dd = dot3(s.n, s.p);
    const float tx = -(dot3(s.n, xo) - dd) / dot3(s.n, xd);
    const V3 px = xo + xd * tx;
    const float ty = -(dot3(s.n, yo) - dd) / dot3(s.n, yd);
    const V3 py = yo + yd * ty;
    s.dpdx = px - s.p;
    s.dpdy = py - s.p;
    int dim[2];
    if (std::fabs(s.n.x) > std::fabs(s.n.y) && std::fabs(s.n.x) > std::fabs(s.n.z)) { dim[0] = 1; dim[1] = 2; }
    else if (std::fabs(s.n.y) > std::fabs(s.n.z)) { dim[0] = 0; dim[1] = 2; }
    else { dim[0] = 0; dim[1] = 1; }
    const float A[2][2] = {{s.dpdu[dim[0]], s.dpdv[dim[0]]}, {s.dpdu[dim[1]], s.dpdv[dim[1]]}};
    const float Bx[2] = {px[dim[0]] - s.p[dim[0]], px[dim[1]] - s.p[dim[1]]};
    const float By[2] = {py[dim[0]] - s.p[dim[0]], py[dim[1]] - s.p[dim[1]]};
    if (!solve2x2(A, Bx, &s.dudx, &s.dvdx)) s.dudx = s.dvdx = 0;
    if (!solve2x2(A, By, &s.dudy, &s.dvdy)) s.dudy = s.dvdy = 0;
}

void put3(float* out, size_t n, size_t i, int f, V3 v) {
    out[(size_t)f * n + i] = v.x;
    out[(size_t)(f + 1) * n + i] = v.y;
    out[(size_t)(f + 2) * n + i] = v.z;
}

// The raster->camera transform of a point (the render's camera, render.cpp).
V3 xf_point4_cam(const float* m, V3 p) {
    float x = ((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3];
    float y = ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7];
    float z = ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11];
    float w = ((m[12] * p.x + m[13] * p.y) + m[14] * p.z) + m[15];
    if (w == 1.0f) return v3(x, y, z);
    return v3(x / w, y / w, z / w);
}

// pbrt-v3 SphericalTheta / SphericalPhi (phi wrapped by WoodenMath's own pi).
float sph_theta(V3 v) { return cr_acos(v.z < -1.0f ? -1.0f : (v.z > 1.0f ? 1.0f : v.z)); }
float sph_phi(V3 v) {
    const float p = cr_atan2(v.y, v.x);
    return (p < 0) ? (float)((double)p + 2.0 * PI_TRUE) : p;
}

// TextureMappingSphere::sphere (CTextureMapping.h:60-66).  SSphere::mapUV
// divides the inverse transform's scale by the radius first (CSSphere.cpp:
// 188-196); the point is normalised after it, so that scale drops out.
void sphere_st(const Affine& inv, V3 p, float* st) {
    const V3 v = normalize3(xf_point(inv, p));
    st[0] = (float)((double)sph_theta(v) * 1.0 / PI_REF);
    st[1] = (float)((double)sph_phi(v) / (2 * PI_REF));
}

}  // namespace

// The CTextureMappedPoint of a camera ray's hit (what JobMapUVRequestsGenerate
// and the shapes' ProcessMapUVRequests jobs give a textured material,
// MEngine.cpp:397-402): the full interaction, its differentials from the
// camera's ray differential (CSCameraPerspective.cpp:37-41, dxCamera /
// dyCamera = the raster->camera transform of the vectors (1, 0, 0) / (0, 1, 0)
// minus that of the zero vector, CSCameraPerspective.h:33-34: the matrix's first
// two columns), then UV mapping (triangles, CTextureMapping.h:39-55) or the
// sphere mapping with its finite differences (spheres, :68-90).
void camera_hit_texture_point(const RenderScene& sc, uint32_t shape, V3 o, V3 d, float t, const float* b,
                              const float* pFilm, float* q) {
    const bool sphere = shape < sc.geom.n_spheres;
    FullSI s = sphere ? si_sphere_full(sc, shape, o, d, t) : si_triangle_full(sc, shape - sc.geom.n_spheres, d, b, t);
    const float* rc = sc.raster_camera;
    const V3 pCam = xf_point4_cam(rc, v3(pFilm[0], pFilm[1], 0.0f));
    const V3 dxC = v3(rc[0], rc[4], rc[8]), dyC = v3(rc[1], rc[5], rc[9]);
    const Affine CW = load_aff(sc.camera_world);
    const V3 dx = normalize3(xf_vector(CW, normalize3(pCam + dxC)));
    const V3 dy = normalize3(xf_vector(CW, normalize3(pCam + dyC)));
    differentials(s, o, dx, o, dy);
    if (!sphere) {
        q[0] = s.uv[0]; q[1] = s.uv[1];
        q[2] = s.dudx; q[3] = s.dvdx;
        q[4] = s.dudy; q[5] = s.dvdy;
        return;
    }
    const Affine inv = load_aff(sc.geom.sphere_xform + 24 * (size_t)shape + 12);
    float st[2], sx[2], sy[2];
    sphere_st(inv, s.p, st);
    const float dt = 0.1f;
    sphere_st(inv, s.p + s.dpdx * dt, sx);
    sphere_st(inv, s.p + s.dpdy * dt, sy);
    float dstdx[2] = {(sx[0] - st[0]) / dt, (sx[1] - st[1]) / dt};
    float dstdy[2] = {(sy[0] - st[0]) / dt, (sy[1] - st[1]) / dt};
    if (dstdx[1] > .5) dstdx[1] = 1 - dstdx[1];
    else if (dstdx[1] < -.5f) dstdx[1] = -dstdx[1] + 1;
    if (dstdy[1] > .5) dstdy[1] = 1 - dstdy[1];
    else if (dstdy[1] < -.5f) dstdy[1] = -dstdy[1] + 1;
    q[0] = st[0]; q[1] = st[1];
    q[2] = dstdx[0]; q[3] = dstdx[1];
    q[4] = dstdy[0]; q[5] = dstdy[1];
}

void scene_collide(const RenderScene& sc, const LbvhResult& bvh, uint32_t n, const float* rays,
                   const uint8_t* surf, const float* diff, uint32_t* hit_shape, float* hit_t, float* si,
                   int nthreads) {
    auto one = [&](uint32_t i) {
        const V3 o = v3(rays[i], rays[(size_t)n + i], rays[2 * (size_t)n + i]);
        const V3 d = v3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]);
        float t;
        const int64_t slot = trace_ray(sc.geom, bvh, o, d, &t, nullptr, nullptr);
        const uint32_t shape = slot < 0 ? 0xFFFFFFFFu : bvh.sorted_index[(size_t)slot];
        hit_shape[i] = shape;
        hit_t[i] = t;
        if (!si) return;
        FullSI s = zero_si();
        if (shape != 0xFFFFFFFFu) {
            if (surf && !surf[i]) {
                // CInteraction of a plain ray cast (SLBVH.cpp:55-59)
                s.p = o + d * t;
                s.time = t;
                s.wo = -d;
            } else if (shape < sc.geom.n_spheres) {
                s = si_sphere_full(sc, shape, o, d, t);
            } else {
                const uint32_t tri = shape - sc.geom.n_spheres;
                float b[3];
                tri_intersect(sc.geom.tri_verts + 9 * (size_t)tri, o, d, b);
                s = si_triangle_full(sc, tri, d, b, t);
            }
            if (diff && (!surf || surf[i])) {
                auto g = [&](int k) { return diff[(size_t)k * n + i]; };
                differentials(s, v3(g(0), g(1), g(2)), v3(g(3), g(4), g(5)), v3(g(6), g(7), g(8)),
                              v3(g(9), g(10), g(11)));
            }
        }
        put3(si, n, i, SI_P, s.p);
        put3(si, n, i, SI_WO, s.wo);
        si[(size_t)SI_TIME * n + i] = s.time;
        put3(si, n, i, SI_N, s.n);
        si[(size_t)SI_UV * n + i] = s.uv[0];
        si[(size_t)(SI_UV + 1) * n + i] = s.uv[1];
        put3(si, n, i, SI_DPDU, s.dpdu);
        put3(si, n, i, SI_DPDV, s.dpdv);
        put3(si, n, i, SI_DNDU, s.dndu);
        put3(si, n, i, SI_DNDV, s.dndv);
        put3(si, n, i, SI_SHADING_N, s.sn);
        put3(si, n, i, SI_SHADING_DPDU, s.sdpdu);
        put3(si, n, i, SI_SHADING_DPDV, s.sdpdv);
        put3(si, n, i, SI_SHADING_DNDU, s.sdndu);
        put3(si, n, i, SI_SHADING_DNDV, s.sdndv);
        put3(si, n, i, SI_DPDX, s.dpdx);
        put3(si, n, i, SI_DPDY, s.dpdy);
        si[(size_t)SI_DUDX * n + i] = s.dudx;
        si[(size_t)SI_DVDX * n + i] = s.dvdx;
        si[(size_t)SI_DUDY * n + i] = s.dudy;
        si[(size_t)SI_DVDY * n + i] = s.dvdy;
    };
    const int nt = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int k = 0; k < nt; k++)
        th.emplace_back([&, k]() {
            for (uint32_t i = (uint32_t)k; i < n; i += (uint32_t)nt) one(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace orc
