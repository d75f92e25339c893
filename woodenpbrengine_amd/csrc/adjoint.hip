// Path-replay adjoint of the wavefront renderer (no counterpart in the
// reference, whose renderer is not differentiable).
//
// The backward pass replays every sample pass of the forward render from the
// same counter-based sample streams (camera, primary trace, shade1 with a
// record of the shading frame and the 2D samples, secondary trace), so the
// paths are identical and nothing per vertex is stored between the forward and
// the backward pass.  k_adjoint then differentiates each sample's radiance:
//
//   * the adjoint of the sample's XYZ is gathered from the loss gradient of
//     the film pixels its Mitchell splat covers (weight / pixel weight sum,
//     through the reference's xyzToRGB, CoefficientSpectrum.h:14-24);
//   * roughness alpha enters the GGX D, the BSDF pdfs and the MIS weights
//     (MicrofaceDistr.h:12-90, SScattering.cpp:163-167) and, for the BSDF
//     strategy, the sampled direction (forward-mode dual numbers along the
//     replayed sample: half vector, reflected direction, env lookup, sin theta);
//   * conductor eta/k enter the 64-wavelength Fresnel term (CBXDF.h:200-238)
//     through fromRGB(Reflectance) (CSpectrumSampled.h:110-158);
//   * env texels enter the env radiance of misses and of BSDF rays
//     (CLightInfiniteArea.cpp:72-92, 206-237) through the bilinear lookup and
//     fromRGB(Illuminant), and the env light pdf of the MIS weight through the
//     sampling distribution's cell value and its integral
//     (CLightInfiniteArea.cpp:44-70).
// Visibility, the light pick, rejected samples, env-map cells and fromRGB cases
// are held at their forward values (their changes are measure-zero events).
#include "wpbr_device.hpp"
#include "render_types.hpp"

namespace wpbr {

// ---- forward-mode dual numbers ----------------------------------------------
template <int N>
struct Dn {
    float v;
    float d[N];
};
template <int N> WPBR_D Dn<N> dc(float v) { Dn<N> r; r.v = v; for (int i = 0; i < N; i++) r.d[i] = 0.0f; return r; }
template <int N> WPBR_D Dn<N> dvar(float v, int k) { Dn<N> r = dc<N>(v); r.d[k] = 1.0f; return r; }
template <int N> WPBR_D Dn<N> operator+(Dn<N> a, Dn<N> b) { Dn<N> r; r.v = a.v + b.v; for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> WPBR_D Dn<N> operator-(Dn<N> a, Dn<N> b) { Dn<N> r; r.v = a.v - b.v; for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> WPBR_D Dn<N> operator-(Dn<N> a) { Dn<N> r; r.v = -a.v; for (int i = 0; i < N; i++) r.d[i] = -a.d[i]; return r; }
template <int N> WPBR_D Dn<N> operator*(Dn<N> a, Dn<N> b) { Dn<N> r; r.v = a.v * b.v; for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> WPBR_D Dn<N> operator/(Dn<N> a, Dn<N> b) {
    Dn<N> r;
    r.v = a.v / b.v;
    for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
template <int N> WPBR_D Dn<N> operator+(Dn<N> a, float b) { a.v += b; return a; }
template <int N> WPBR_D Dn<N> operator+(float b, Dn<N> a) { a.v += b; return a; }
template <int N> WPBR_D Dn<N> operator-(Dn<N> a, float b) { a.v -= b; return a; }
template <int N> WPBR_D Dn<N> operator-(float b, Dn<N> a) { return dc<N>(b) - a; }
template <int N> WPBR_D Dn<N> operator*(Dn<N> a, float b) { a.v *= b; for (int i = 0; i < N; i++) a.d[i] *= b; return a; }
template <int N> WPBR_D Dn<N> operator*(float b, Dn<N> a) { return a * b; }
template <int N> WPBR_D Dn<N> operator/(Dn<N> a, float b) { return a * (1.0f / b); }
template <int N> WPBR_D Dn<N> operator/(float a, Dn<N> b) { return dc<N>(a) / b; }
template <int N> WPBR_D Dn<N> dsqrt(Dn<N> a) {
    Dn<N> r;
    r.v = sqrtf(a.v);
    for (int i = 0; i < N; i++) r.d[i] = r.v > 0.0f ? a.d[i] / (2.0f * r.v) : 0.0f;
    return r;
}
template <int N> WPBR_D Dn<N> dabs(Dn<N> a) { return a.v < 0.0f ? -a : a; }
template <int N> WPBR_D Dn<N> dmax0(Dn<N> a) { return a.v > 0.0f ? a : dc<N>(0.0f); }
template <int N> WPBR_D Dn<N> dclamp1(Dn<N> a) {
    if (a.v < -1.0f) return dc<N>(-1.0f);
    if (a.v > 1.0f) return dc<N>(1.0f);
    return a;
}

using D1 = Dn<1>;
using D3 = Dn<3>;
struct V1 {
    D1 x, y, z;
};
WPBR_D V1 v1c(f3 a) { return {dc<1>(a.x), dc<1>(a.y), dc<1>(a.z)}; }
WPBR_D D1 vdot(V1 a, V1 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
WPBR_D D1 vdot(V1 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
WPBR_D V1 vscale(V1 a, D1 s) { return {a.x * s, a.y * s, a.z * s}; }

// ---- derivative helpers ------------------------------------------------------
// GGX D (isotropic alpha) with the half vector and alpha as duals.
WPBR_D D1 ggx_d1(V1 wh, D1 a) {
    D1 c2 = wh.z * wh.z;
    D1 s2 = dmax0(1.0f - c2);
    D1 t2 = s2 / c2;
    if (isinf(t2.v)) return dc<1>(0.0f);
    D1 cos4 = c2 * c2;
    D1 st = dsqrt(s2);
    D1 cp = st.v == 0.0f ? dc<1>(1.0f) : dclamp1(wh.x / st);
    D1 sp = st.v == 0.0f ? dc<1>(0.0f) : dclamp1(wh.y / st);
    D1 a2 = a * a;
    D1 e = (cp * cp / a2 + sp * sp / a2) * t2;
    D1 one_e = 1.0f + e;
    return 1.0f / (((a2 * cos4) * (one_e * one_e)) * (float)PI_REF);
}

template <int N>
WPBR_D Dn<N> fresnel_dielectric_n(Dn<N> cosI, float etaI, float etaT) {
    cosI = dclamp1(cosI);
    if (!(cosI.v > 0.0f)) { float t = etaI; etaI = etaT; etaT = t; cosI = dabs(cosI); }
    Dn<N> sinI = dsqrt(dmax0(1.0f - cosI * cosI));
    Dn<N> sinT = sinI * (etaI / etaT);
    if (sinT.v >= 1.0f) return dc<N>(1.0f);
    Dn<N> cosT = dsqrt(dmax0(1.0f - sinT * sinT));
    Dn<N> rpa = (cosI * etaT - cosT * etaI) / (cosI * etaT + cosT * etaI);
    Dn<N> rpe = (cosI * etaI - cosT * etaT) / (cosI * etaI + cosT * etaT);
    return (rpa * rpa + rpe * rpe) * 0.5f;
}

// FrConductor for one wavelength with partials by (cos, eta, k).
WPBR_D D3 fresnel_conductor3(float cosI, float eta, float k) {
    D3 c = dvar<3>(fminf(fmaxf(cosI, -1.0f), 1.0f), 0);
    if (cosI < -1.0f || cosI > 1.0f) c.d[0] = 0.0f;
    D3 e = dvar<3>(eta, 1), kk = dvar<3>(k, 2);
    D3 c2 = c * c, s2 = 1.0f - c2;
    D3 e2 = e * e, k2 = kk * kk;
    D3 t0 = (e2 - k2) - s2;
    D3 a2pb2 = dsqrt(t0 * t0 + (e2 * k2) * 4.0f);
    D3 t1 = a2pb2 + c2;
    D3 a = dsqrt((a2pb2 + t0) * 0.5f);
    D3 t2 = (a * c) * 2.0f;
    D3 Rs = (t1 - t2) / (t1 + t2);
    D3 t3 = a2pb2 * c2 + s2 * s2;
    D3 t4 = t2 * s2;
    D3 Rp = (Rs * (t3 - t4)) / (t3 + t4);
    return (Rp + Rs) * 0.5f;
}

// fromRGB ordering case (the branches of CSpectrumSampled.h:110-207): the
// channels holding the min / mid / max value and the bases of the three terms
// (basis 0 is white).  d spectrum / d rgb[ch_min] = B0 - B1, [ch_mid] = B1 - B2,
// [ch_max] = B2.
struct SmitsCase {
    int ch[3];  // min, mid, max channel
    int b1, b2;
};
WPBR_D SmitsCase smits_case(float r, float g, float b) {
    if (r <= g && r <= b) return (g <= b) ? SmitsCase{{0, 1, 2}, 1, 6} : SmitsCase{{0, 2, 1}, 1, 5};
    if (g <= r && g <= b) return (r <= b) ? SmitsCase{{1, 0, 2}, 2, 6} : SmitsCase{{1, 2, 0}, 2, 4};
    return (r <= g) ? SmitsCase{{2, 0, 1}, 3, 5} : SmitsCase{{2, 1, 0}, 3, 4};
}
WPBR_D SmitsCase smits_from_code(int code) {
    const SmitsCase t[6] = {{{0, 1, 2}, 1, 6}, {{0, 2, 1}, 1, 5}, {{1, 0, 2}, 2, 6},
                            {{1, 2, 0}, 2, 4}, {{2, 0, 1}, 3, 5}, {{2, 1, 0}, 3, 4}};
    return t[code];
}
// d spectrum_i / d rgb[c] for the three channels, in channel order
WPBR_D void smits_grad(const float* B, const SmitsCase& sc, int i, float* g) {
    const float B0 = B[i], B1 = B[sc.b1 * 64 + i], B2 = B[sc.b2 * 64 + i];
    g[sc.ch[0]] = B0 - B1;
    g[sc.ch[1]] = B1 - B2;
    g[sc.ch[2]] = B2;
}

// The env map's bilinear cell of (u, v) (as env_bilerp).
struct EnvCellD {
    int s0, t0, s0n, t0n;
    float ds, dt;
};
WPBR_D EnvCellD env_cell(const SceneView& sv, float u, float v) {
    int W = (int)sv.env_w, H = (int)sv.env_h;
    float s = u * (float)W - 0.5f, t = v * (float)H - 0.5f;
    int s0 = (int)floorf(s), t0 = (int)floorf(t);
    s0 = min(max(s0, 0), W - 1);
    t0 = min(max(t0, 0), H - 1);
    EnvCellD c;
    c.s0 = s0; c.t0 = t0;
    c.s0n = max(min(s0 + 1, W - 1), 0);
    c.t0n = max(min(t0 + 1, H - 1), 0);
    c.ds = s - (float)s0;
    c.dt = t - (float)t0;
    return c;
}
WPBR_D float texel(const SceneView& sv, int x, int y, int ch) {
    return sv.env_rgb[((size_t)y * sv.env_w + x) * 3 + ch];
}
WPBR_D void cell_rgb(const SceneView& sv, const EnvCellD& c, float* rgb, float* drds, float* drdt) {
    for (int ch = 0; ch < 3; ch++) {
        float a = texel(sv, c.s0, c.t0, ch), b = texel(sv, c.s0, c.t0n, ch);
        float cc = texel(sv, c.s0n, c.t0, ch), d = texel(sv, c.s0n, c.t0n, ch);
        rgb[ch] = ((a * (1 - c.ds)) * (1 - c.dt) + (b * (1 - c.ds)) * c.dt) + (cc * c.ds) * (1 - c.dt) +
                  (d * c.ds) * c.dt;
        drds[ch] = (cc - a) * (1 - c.dt) + (d - b) * c.dt;
        drdt[ch] = (b - a) * (1 - c.ds) + (d - cc) * c.ds;
    }
}
WPBR_D void scatter_cell(const SceneView& sv, float* grad_env, const EnvCellD& c, const float* g) {
    const int xs[4] = {c.s0, c.s0, c.s0n, c.s0n}, ys[4] = {c.t0, c.t0n, c.t0, c.t0n};
    const float w[4] = {(1 - c.ds) * (1 - c.dt), (1 - c.ds) * c.dt, c.ds * (1 - c.dt), c.ds * c.dt};
    for (int k = 0; k < 4; k++) {
        float* t = grad_env + ((size_t)ys[k] * sv.env_w + xs[k]) * 3;
        for (int ch = 0; ch < 3; ch++)
            if (g[ch] * w[k] != 0.0f) atomicAdd(t + ch, g[ch] * w[k]);
    }
}

// Env direction (world) -> (u, v, theta) as duals along alpha.
WPBR_D void env_uv1(const SceneView& sv, V1 dirW, D1* u, D1* v, D1* theta) {
    const float* m = sv.env_xform + 12;  // inverse 3x4
    V1 dl = {(dirW.x * m[0] + dirW.y * m[1]) + dirW.z * m[2], (dirW.x * m[4] + dirW.y * m[5]) + dirW.z * m[6],
             (dirW.x * m[8] + dirW.y * m[9]) + dirW.z * m[10]};
    D1 z = dclamp1(dl.z);
    D1 th;
    th.v = cr_acos(z.v);
    th.d[0] = (z.v > -1.0f && z.v < 1.0f) ? -z.d[0] / sqrtf(1.0f - z.v * z.v) : 0.0f;
    D1 ph;
    float p = cr_atan2(dl.y.v, dl.x.v);
    ph.v = (p < 0) ? (float)((double)p + 2.0 * 3.14159265358979323846) : p;
    float r2 = dl.x.v * dl.x.v + dl.y.v * dl.y.v;
    ph.d[0] = r2 > 0.0f ? (dl.x.v * dl.y.d[0] - dl.y.v * dl.x.d[0]) / r2 : 0.0f;
    u->v = (float)((double)ph.v / (2 * PI_REF));
    u->d[0] = ph.d[0] / (float)(2 * PI_REF);
    v->v = (float)((double)th.v / PI_REF);
    v->d[0] = th.d[0] / (float)PI_REF;
    *theta = th;
}

constexpr uint32_t AF_LIGHT = 1u, AF_BSDF = 2u, AF_ENV = 4u, AF_MISS = 16u;
constexpr int ADJ_BLOCK = 256;

WPBR_D float mitchell_w(float x) {
    const float B = (float)(1.0 / 3.0f), Cc = (float)(1.0 / 3.0f);
    x = fabsf(2 * x);
    if (x > 1)
        return ((-B - 6 * Cc) * x * x * x + (6 * B + 30 * Cc) * x * x + (-12 * B - 48 * Cc) * x +
                (8 * B + 24 * Cc)) * (1.f / 6.f);
    return ((12 - 9 * B - 6 * Cc) * x * x * x + (-18 + 12 * B + 6 * Cc) * x * x + (6 - 2 * B)) * (1.f / 6.f);
}

WPBR_D void adjoint_sample(const AdjArgs& a, uint32_t i, float* s_part) {
    const RenderLaunch& L = a.L;
    const SceneView& sv = L.sv;
    const WaveState& ws = L.ws;
    const uint32_t n = L.n, W = sv.width;
    const uint32_t nm = a.A.n_mat;
    const uint32_t px = i % W, py = i / W;
    // adjoint of the sample's XYZ from the pixels of its (tile-clipped) splat
    const float fx = ws.pfilm[i], fy = ws.pfilm[n + i];
    const int tx0 = (int)(px & ~31u), ty0 = (int)(py & ~31u);
    int x0 = (int)ceilf(fx - 0.5f - 2.0f), x1 = (int)floorf(fx - 0.5f + 2.0f);
    int y0 = (int)ceilf(fy - 0.5f - 2.0f), y1 = (int)floorf(fy - 0.5f + 2.0f);
    x0 = max(x0, tx0); y0 = max(y0, ty0);
    x1 = min(x1, min(tx0 + 31, (int)W - 1)); y1 = min(y1, min(ty0 + 31, (int)sv.height - 1));
    float gX = 0, gY = 0, gZ = 0;
    for (int y = y0; y <= y1; y++) {
        const float wy = mitchell_w(((float)y - fy) / 2.0f);
        for (int x = x0; x <= x1; x++) {
            const float w = mitchell_w(((float)x - fx) / 2.0f) * wy;
            const size_t p = (size_t)y * W + x;
            const float Wp = a.film[p].w;
            if (Wp == 0.0f || w == 0.0f) continue;
            const float s = w / Wp;
            const float d0 = a.dL[3 * p], d1 = a.dL[3 * p + 1], d2 = a.dL[3 * p + 2];
            // transpose of xyzToRGB (CoefficientSpectrum.h:14-24)
            gX += s * ((3.240479f * d0 + -0.969256f * d1) + 0.041556f * d2);
            gY += s * ((-1.537150f * d0 + 1.875991f * d1) + -0.204043f * d2);
            gZ += s * ((-0.498535f * d0 + 0.041556f * d1) + 1.057311f * d2);
        }
    }
    if (gX == 0.0f && gY == 0.0f && gZ == 0.0f) return;
    const float dl = 300.0f / 64.0f;
    auto gbar = [&](int k) { return ((gX * sv.xyz[k] + gY * sv.xyz[NSPEC + k]) + gZ * sv.xyz[2 * NSPEC + k]) * dl; };
    const uint32_t flags = ws.flags[i];
    if (flags & AF_MISS) {
        if (!sv.has_env) return;
        f3 d = mk3(ws.rays[3 * (size_t)n + i], ws.rays[4 * (size_t)n + i], ws.rays[5 * (size_t)n + i]);
        D1 u, v, th;
        env_uv1(sv, v1c(d), &u, &v, &th);
        EnvCellD c = env_cell(sv, u.v, v.v);
        float rgb[3], drds[3], drdt[3];
        cell_rgb(sv, c, rgb, drds, drdt);
        SmitsCase cs = smits_case(rgb[0], rgb[1], rgb[2]);
        float gr[3] = {0, 0, 0};
        for (int k = 0; k < NSPEC; k++) {
            float dg[3];
            smits_grad(sv.illum, cs, k, dg);
            const float t = gbar(k) * sv.env_emit[k];
            for (int ch = 0; ch < 3; ch++) gr[ch] += t * dg[ch];
        }
        scatter_cell(sv, a.grad_env, c, gr);
        return;
    }
    if (!(flags & (AF_LIGHT | AF_BSDF))) return;
    const uint32_t li = ws.light[i];
    const uint32_t mid = ws.mat[i];
    const DevMaterial m = sv.materials[mid];
    const bool envPick = flags & AF_ENV;
    const bool conductor = m.type == MAT_CONDUCTOR;
    const f3 woL = mk3(ws.r_wo[i], ws.r_wo[n + i], ws.r_wo[2 * (size_t)n + i]);
    const float* es = conductor ? sv.mat_spec + (size_t)m.spec * 2 * NSPEC : nullptr;
    SmitsCase ce{}, ck{};
    if (conductor) {
        ce = smits_from_code(a.A.mat_case[2 * mid]);
        ck = smits_from_code(a.A.mat_case[2 * mid + 1]);
    }
    float dAlpha = 0.0f;
    float dEta[3] = {0, 0, 0}, dK[3] = {0, 0, 0};
    auto add_fresnel_grads = [&](float dFe, float dFk, int k) {
        float ge[3], gk[3];
        smits_grad(a.A.refl, ce, k, ge);
        smits_grad(a.A.refl, ck, k, gk);
        for (int ch = 0; ch < 3; ch++) { dEta[ch] += dFe * ge[ch]; dK[ch] += dFk * gk[ch]; }
    };
    // ---- light strategy (sphere lights): alpha through D, eta/k through Fr
    if ((flags & AF_LIGHT) && m.type != MAT_SPECULAR) {
        const uint32_t h = ws.qhit[ws.l_ray[i]];
        const float* Le = sv.light_emit + (size_t)li * NSPEC;
        if (h == 0xFFFFFFFFu || h == sv.light_shape[li]) {
            const f3 wl = mk3(ws.r_wl[i], ws.r_wl[n + i], ws.r_wl[2 * (size_t)n + i]);
            const float pL = ws.r_pdfl[i];
            const V1 wh = v1c(normalize3(add3(woL, wl)));
            const D1 al = dvar<1>(m.alpha, 0);
            const D1 Dv = ggx_d1(wh, al);
            const float cO = fabsf(woL.z), cI = fabsf(wl.z);
            const float whdot = dot3(woL, mk3(wh.x.v, wh.y.v, wh.z.v));
            if (cI != 0.0f && cO != 0.0f && Dv.v != 0.0f) {
                const D1 pB = Dv * (fabsf(wh.z.v) / (4 * whdot));
                const float cf = (wl.x * wh.x.v + wl.y * wh.y.v) + wl.z * wh.z.v;
                D1 sc = Dv / (4 * cI * cO);
                if (!conductor) sc = sc * fresnel_dielectric_n<1>(dc<1>(cf), 1.0f, m.eta);
                const D1 wL = dc<1>(pL * pL) / (dc<1>(pL * pL) + pB * pB);
                const D1 S = sc * (cI / pL) * wL;
                float aS = 0.0f;
                for (int k = 0; k < NSPEC; k++) {
                    const float t = gbar(k) * Le[k];
                    if (conductor) {
                        D3 F = fresnel_conductor3(cf, es[k], es[NSPEC + k]);
                        aS += t * F.v;
                        add_fresnel_grads(t * S.v * F.d[1], t * S.v * F.d[2], k);
                    } else {
                        aS += t;
                    }
                }
                dAlpha += aS * S.d[0];
            }
        }
    }
    // ---- BSDF strategy: alpha along the sampled direction, eta/k, env
    if (flags & AF_BSDF) {
        const uint32_t h = ws.qhit[ws.b_ray[i]];
        if (h == 0xFFFFFFFFu || (!envPick && h == sv.light_shape[li])) {
            const f3 ss = mk3(ws.r_frame[i], ws.r_frame[n + i], ws.r_frame[2 * (size_t)n + i]);
            const f3 ts = mk3(ws.r_frame[3 * (size_t)n + i], ws.r_frame[4 * (size_t)n + i], ws.r_frame[5 * (size_t)n + i]);
            const f3 ns = mk3(ws.r_frame[6 * (size_t)n + i], ws.r_frame[7 * (size_t)n + i], ws.r_frame[8 * (size_t)n + i]);
            V1 wi, wh;
            D1 S, cf = dc<1>(0.0f);
            float sOverPB = 0.0f, pBv = 0.0f;  // for the pdf terms of the MIS weight
            D1 pB;
            bool ok = true;
            D1 scac;
            if (m.type == MAT_SPECULAR) {
                wi = v1c(mk3(-woL.x, -woL.y, woL.z));
                const float fr = fresnel_dielectric_n<1>(dc<1>(woL.z), 1.0f, m.eta).v;
                S = dc<1>(fr);
            } else {
                const D1 al = dvar<1>(m.alpha, 0);
                const float u0 = ws.r_ub[i], u1 = ws.r_ub[n + i];
                const float phi = (float)((2 * PI_REF) * (double)u1);
                const D1 tan2 = al * al * (u0 / (1.0f - u0));
                const D1 cosT = 1.0f / dsqrt(1.0f + tan2);
                const D1 sinT = dsqrt(dmax0(1.0f - cosT * cosT));
                wh = {sinT * cr_cos(phi), sinT * cr_sin(phi), cosT};
                if (!(woL.z * wh.z.v > 0)) wh = {-wh.x, -wh.y, -wh.z};
                const D1 Dv = ggx_d1(wh, al);
                const D1 wdh = vdot(wh, woL);
                wi = {wh.x * (2.0f * wdh) - woL.x, wh.y * (2.0f * wdh) - woL.y, wh.z * (2.0f * wdh) - woL.z};
                if (!(wi.z.v * woL.z > 0)) ok = false;
                const float cO = fabsf(woL.z);
                const D1 cI = dabs(wi.z);
                if (cI.v == 0.0f || cO == 0.0f || Dv.v == 0.0f) ok = false;
                if (ok) {
                    pB = Dv * dabs(wh.z) / (4.0f * wdh);
                    cf = vdot(wi, wh);
                    D1 sc = Dv / (cI * (4 * cO));
                    if (!conductor) sc = sc * fresnel_dielectric_n<1>(cf, 1.0f, m.eta);
                    scac = sc * cI;
                }
            }
            if (ok) {
                const V1 wiW = {(wi.x * ss.x + wi.y * ts.x) + wi.z * ns.x, (wi.x * ss.y + wi.y * ts.y) + wi.z * ns.y,
                                (wi.x * ss.z + wi.y * ts.z) + wi.z * ns.z};
                float rgb[3] = {0, 0, 0}, drds[3], drdt[3];
                D1 rgbd[3];
                EnvCellD c{};
                SmitsCase cs{};
                D1 pL2 = dc<1>(0.0f);
                float fcell = 0.0f, Ienv = 1.0f;
                int cellk = -1;
                bool black = false;
                if (envPick) {
                    D1 u, v, th;
                    env_uv1(sv, wiW, &u, &v, &th);
                    D1 sinT;
                    sinT.v = cr_sin(th.v);
                    sinT.d[0] = cr_cos(th.v) * th.d[0];
                    if (sinT.v == 0.0f) black = true;
                    c = env_cell(sv, u.v, v.v);
                    cell_rgb(sv, c, rgb, drds, drdt);
                    for (int ch = 0; ch < 3; ch++) {
                        rgbd[ch].v = rgb[ch];
                        rgbd[ch].d[0] = drds[ch] * (float)sv.env_w * u.d[0] + drdt[ch] * (float)sv.env_h * v.d[0];
                    }
                    cs = smits_case(rgb[0], rgb[1], rgb[2]);
                    int iu = (int)(u.v * (float)sv.env_w), iv = (int)(v.v * (float)sv.env_h);
                    iu = iu < 0 ? 0 : (iu > (int)sv.env_w - 1 ? (int)sv.env_w - 1 : iu);
                    iv = iv < 0 ? 0 : (iv > (int)sv.env_h - 1 ? (int)sv.env_h - 1 : iv);
                    cellk = iv * (int)sv.env_w + iu;
                    fcell = sv.env_cond_func[cellk];
                    Ienv = sv.env_marg_int[0];
                    pL2 = (fcell / Ienv) / (sinT * (float)(2 * PI_REF * PI_REF));
                } else {
                    // sphere light: constant pdf (CSSphere.cpp:221-246)
                    pL2 = dc<1>(0.0f);
                }
                if (!black) {
                    D1 wB;
                    float dwdpl2 = 0.0f;
                    if (m.type == MAT_SPECULAR) {
                        wB = dc<1>(1.0f);
                    } else {
                        float pl2v = pL2.v;
                        if (!envPick) {
                            // the light's pdf is a constant of the scene
                            const float* xf = sv.g.sphere_xform + 24 * (size_t)sv.light_shape[li];
                            f3 pc = mk3(xf[3], xf[7], xf[11]);
                            float r = sv.g.sphere_radius[sv.light_shape[li]];
                            float s2 = r * r / len2(pc);
                            pl2v = 1 / (2 * PI_TRUE_F * (1 - sqrtf(fmaxf(0.0f, 1 - s2))));
                            pL2 = dc<1>(pl2v);
                        }
                        wB = (pB * pB) / (pB * pB + pL2 * pL2);
                        const float den = pB.v * pB.v + pl2v * pl2v;
                        dwdpl2 = -2.0f * pB.v * pB.v * pl2v / (den * den);
                        S = scac * wB / pB;
                        sOverPB = scac.v / pB.v;
                        pBv = pB.v;
                    }
                    float aS = 0.0f, aCos = 0.0f, gr[3] = {0, 0, 0};
                    const float* Le = sv.light_emit + (size_t)li * NSPEC;
                    for (int k = 0; k < NSPEC; k++) {
                        float Li = 0.0f;
                        float dg[3] = {0, 0, 0};
                        if (envPick) {
                            smits_grad(sv.illum, cs, k, dg);
                            const float* B = sv.illum;
                            float w0 = rgb[cs.ch[0]], w1 = rgb[cs.ch[1]] - rgb[cs.ch[0]], w2 = rgb[cs.ch[2]] - rgb[cs.ch[1]];
                            Li = ((B[k] * w0 + B[cs.b1 * 64 + k] * w1) + B[cs.b2 * 64 + k] * w2) * sv.env_emit[k];
                        } else {
                            Li = Le[k];
                        }
                        const float t = gbar(k);
                        float F = 1.0f;
                        if (conductor) {
                            D3 F3 = fresnel_conductor3(cf.v, es[k], es[NSPEC + k]);
                            F = F3.v;
                            aCos += t * Li * F3.d[0];
                            add_fresnel_grads(t * Li * S.v * F3.d[1], t * Li * S.v * F3.d[2], k);
                        }
                        aS += t * Li * F;
                        if (envPick) {
                            const float q = t * F * S.v * sv.env_emit[k];
                            for (int ch = 0; ch < 3; ch++) gr[ch] += q * dg[ch];
                        }
                    }
                    dAlpha += aS * S.d[0] + aCos * S.v * cf.d[0];
                    if (envPick) {
                        for (int ch = 0; ch < 3; ch++) dAlpha += gr[ch] * rgbd[ch].d[0];
                        scatter_cell(sv, a.grad_env, c, gr);
                        if (m.type != MAT_SPECULAR && fcell != 0.0f) {
                            // MIS weight through the env pdf: cell value and integral
                            const float dLdp = aS * sOverPB * dwdpl2;
                            const float dLdf = dLdp * pL2.v / fcell;
                            const int* src = a.A.env_src + 4 * cellk;
                            const float* sw = a.A.env_srcw + 4 * cellk;
                            for (int q = 0; q < 4; q++)
                                if (src[q] >= 0 && sw[q] != 0.0f) atomicAdd(a.grad_env + 3 * (size_t)src[q] + 1, dLdf * sw[q]);
                            atomicAdd(&s_part[7 * nm], dLdp * (-pL2.v / Ienv));
                        }
                    }
                    (void)pBv;
                }
            }
        }
    }
    if (dAlpha != 0.0f) atomicAdd(&s_part[mid], dAlpha);
    if (conductor) {
        for (int ch = 0; ch < 3; ch++) {
            if (dEta[ch] != 0.0f) atomicAdd(&s_part[nm + 6 * mid + ch], dEta[ch]);
            if (dK[ch] != 0.0f) atomicAdd(&s_part[nm + 6 * mid + 3 + ch], dK[ch]);
        }
    }
}

__global__ __launch_bounds__(ADJ_BLOCK) void k_adjoint(AdjArgs a) {
    extern __shared__ float s_part[];
    const uint32_t ncol = 7 * a.A.n_mat + 1;
    for (uint32_t k = threadIdx.x; k < ncol; k += ADJ_BLOCK) s_part[k] = 0.0f;
    __syncthreads();
    const uint32_t i = blockIdx.x * ADJ_BLOCK + threadIdx.x;
    if (i < a.L.n) adjoint_sample(a, i, s_part);
    __syncthreads();
    float* row = a.part + (size_t)blockIdx.x * ncol;
    for (uint32_t k = threadIdx.x; k < ncol; k += ADJ_BLOCK) row[k] += s_part[k];
}

// Column sums of the per-block partials; column 7 n_mat (d loss / d integral)
// goes to *dI.
__global__ void k_adj_reduce(const float* part, uint32_t rows, uint32_t ncol, float* grad, float* dI) {
    const uint32_t c = blockIdx.x;
    __shared__ float s[256];
    float acc = 0.0f;
    for (uint32_t r = threadIdx.x; r < rows; r += blockDim.x) acc += part[(size_t)r * ncol + c];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (c + 1 == ncol) *dI = s[0];
        else grad[c] = s[0];
    }
}

// Dense part of the env gradient: every texel's green channel moves the
// distribution's integral.
__global__ void k_env_dense(float* grad_env, const float* env_dI, const float* dI, uint32_t ntex) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntex && env_dI[t] != 0.0f) grad_env[3 * (size_t)t + 1] += *dI * env_dI[t];
}

// Image (xyzToRGB of the film / weight), its L2 loss against a target and the
// loss gradient; per-block loss partials.
__global__ void k_film_loss(const float4* film, const float* target, uint32_t npix, float* rgb_out, float* dL,
                            float* partial) {
    __shared__ float s[256];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (p < npix) {
        const float4 f = film[p];
        const float rgb[3] = {((f.x * 3.240479f + f.y * -1.537150f) + f.z * -0.498535f) / f.w,
                              ((f.x * -0.969256f + f.y * 1.875991f) + f.z * 0.041556f) / f.w,
                              ((f.x * 0.041556f + f.y * -0.204043f) + f.z * 1.057311f) / f.w};
        // a pixel whose image or target is not finite (a NaN sample of the
        // estimator reached it) is left out of the loss: no loss, zero dL
        bool finite = true;
        for (int c = 0; c < 3; c++) {
            finite = finite && __builtin_isfinite(rgb[c]);
            if (target) finite = finite && __builtin_isfinite(target[3 * (size_t)p + c]);
        }
        for (int c = 0; c < 3; c++) {
            if (rgb_out) rgb_out[3 * (size_t)p + c] = rgb[c];
            if (target) {
                const float e = finite ? rgb[c] - target[3 * (size_t)p + c] : 0.0f;
                l += e * e;
                if (dL) dL[3 * (size_t)p + c] = 2.0f * e;
            }
        }
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && partial) partial[blockIdx.x] = s[0];
}

// The adjoint's conductor-microfacet derivatives for given local directions,
// one wavelength each: input SoA (n each) wo xyz, wi xyz, alpha, eta, k, u0, u1;
// output SoA D, dD/dalpha, pdf, dpdf/dalpha, f, df/dalpha (f = Fr D / (4 cos_i
// cos_o) |cos_i|), Fr, dFr/deta, dFr/dk at cos = wi.wh, then the half vector
// sampled from (u0, u1) and its reflected wi: wh_z, dwh_z/dalpha, wi_z,
// dwi_z/dalpha (CBSDFConductorMicroface.cpp:11-82, MicrofaceDistr.h:12-90,
// CBXDF.h:200-238).
constexpr int BSDF_GRAD_IN = 11, BSDF_GRAD_OUT = 13;
__global__ void k_bsdf_grad(const float* in, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[BSDF_GRAD_IN];
    for (int j = 0; j < BSDF_GRAD_IN; j++) x[j] = in[(size_t)j * n + i];
    const f3 wo = mk3(x[0], x[1], x[2]), wi = mk3(x[3], x[4], x[5]);
    const D1 al = dvar<1>(x[6], 0);
    const f3 h = normalize3(add3(wo, wi));
    const V1 wh = v1c(h);
    const D1 Dv = ggx_d1(wh, al);
    const D1 pdf = Dv * (fabsf(h.z) / (4 * dot3(wo, h)));
    const D3 F = fresnel_conductor3(dot3(wi, h), x[7], x[8]);
    const float cI = fabsf(wi.z), cO = fabsf(wo.z);
    const D1 f = (Dv / (4 * cI * cO)) * (F.v * cI);
    // sampled half vector and direction along alpha (as the adjoint's BSDF strategy)
    const float u0 = x[9], u1 = x[10];
    const float phi = (float)((2 * PI_REF) * (double)u1);
    const D1 tan2 = al * al * (u0 / (1.0f - u0));
    const D1 cosT = 1.0f / dsqrt(1.0f + tan2);
    const D1 sinT = dsqrt(dmax0(1.0f - cosT * cosT));
    V1 hs = {sinT * cr_cos(phi), sinT * cr_sin(phi), cosT};
    if (!(wo.z * hs.z.v > 0)) hs = {-hs.x, -hs.y, -hs.z};
    const D1 wdh = vdot(hs, wo);
    const D1 wiz = hs.z * (2.0f * wdh) - wo.z;
    const float o[BSDF_GRAD_OUT] = {Dv.v, Dv.d[0], pdf.v, pdf.d[0], f.v, f.d[0], F.v, F.d[1], F.d[2],
                                    hs.z.v, hs.z.d[0], wiz.v, wiz.d[0]};
    for (int j = 0; j < BSDF_GRAD_OUT; j++) out[(size_t)j * n + i] = o[j];
}

wpbr_status bsdf_grad_launch(const float* in, float* out, uint32_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_bsdf_grad, dim3((n + 255) / 256), dim3(256), 0, st, in, out, n);
    return launch_status();
}

wpbr_status adjoint_launch(const AdjArgs& a, hipStream_t st) {
    const uint32_t blocks = (a.L.n + ADJ_BLOCK - 1) / ADJ_BLOCK;
    const size_t lds = sizeof(float) * (7 * (size_t)a.A.n_mat + 1);
    hipLaunchKernelGGL(k_adjoint, dim3(blocks), dim3(ADJ_BLOCK), lds, st, a);
    return launch_status();
}

uint32_t adjoint_blocks(uint32_t n) { return (n + ADJ_BLOCK - 1) / ADJ_BLOCK; }

wpbr_status adjoint_finish(const float* part, uint32_t rows, uint32_t n_mat, float* grad, float* dI,
                           const float* env_dI, uint32_t ntex, hipStream_t st) {
    const uint32_t ncol = 7 * n_mat + 1;
    hipLaunchKernelGGL(k_adj_reduce, dim3(ncol), dim3(256), 0, st, part, rows, ncol, grad, dI);
    if (ntex) hipLaunchKernelGGL(k_env_dense, dim3((ntex + 255) / 256), dim3(256), 0, st, grad + 7 * n_mat, env_dI, dI, ntex);
    return launch_status();
}

wpbr_status film_loss_launch(const float4* film, const float* target, uint32_t npix, float* rgb, float* dL,
                             float* partial, hipStream_t st) {
    hipLaunchKernelGGL(k_film_loss, dim3((npix + 255) / 256), dim3(256), 0, st, film, target, npix, rgb, dL, partial);
    return launch_status();
}

// The loss from the per-block partials, summed as wpbr_film_loss sums them on
// the host: in block order, each partial widened to double (no contraction).
__global__ void k_loss_sum(const float* partial, uint32_t nb, double* out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double acc = 0.0;
    for (uint32_t b = 0; b < nb; b++) acc += (double)partial[b];
    *out = acc;
}

wpbr_status film_loss_sum_launch(const float* partial, uint32_t nb, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(64), 0, st, partial, nb, out);
    return launch_status();
}

}  // namespace wpbr
