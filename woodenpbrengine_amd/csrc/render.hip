// Wavefront renderer for gfx950: MEngine::render (MEngine.cpp:380-448), direct
// lighting with MIS, all film tiles in one launch per sample pass.
//
// One pass per sample index j (the per-pixel sample cursors of CSampler.h make
// pass j depend on pass j-1), each pass five kernels over all pixels:
//   k_camera  JobSamplerUpdateCameraSamples + JobCameraPerspGenerateRaysDifferential
//             (CSSampler.cpp:7-24, CSCameraPerspective.cpp:9-52)
//   trace     runCollisionSystem for primary rays (collide.hip)
//   k_shade1  emitted light of misses (SScattering.cpp:20-80), surface interaction
//             (CSTriangle.cpp:149-319 / CSSphere.cpp:122-186), BSDF request + frame
//             (CMaterial.h:30-38, CBSDF.cpp:12-24), light pick and sample
//             (SScattering.cpp:82-88, CSSphere.h:187-230), BSDF Compute toward the
//             light and BSDF Sample (CBSDF*.cpp); shadow and BSDF rays go to one
//             queue compacted with wave ballots (one atomic per wave)
//   trace     both secondary rays of every sample in one launch
//   k_radiance_splat  Li compute (CSSphere.h:232-267, CLightInfiniteArea.cpp:137-168),
//             both MIS integrations (SScattering.cpp:127-254) over the 64
//             wavelengths, XYZ, and the Mitchell splat into the sample's 32x32
//             tile (CSFilm.cpp:70-122), gathered per film pixel
// The reference's light sample for the env light stores a black Li
// (CLightInfiniteArea.cpp:106, 126-131): that strategy adds nothing, so its shadow
// ray is not cast (the 2D sample is still consumed).
#include "wpbr_device.hpp"
#include "render_types.hpp"

namespace wpbr {

// ---- sample streams (counter-based, lowbias32 chain) ------------------------
WPBR_D uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
WPBR_D float hash_unit(uint32_t seed, uint32_t pixel, uint32_t key) {
    uint32_t h = mix32(seed ^ 0x9E3779B9u);
    h = mix32(h ^ pixel);
    h = mix32(h ^ key);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// ---- stratified sample arrays ----------------------------------------------------
// JobSamplerStratifiedGenerateSampels1D/2D (CSSamplerStratified.cpp:8-90) give
// each pixel a 1D array (element k = min((k + xi_k) / n1, 1)) and a 2D array
// (element k = i n + j of an n x n grid = min((i + xi, j + xi') * inv2, 1), the
// n2 - n^2 tail left (0, 0)), then shuffle both.  Here no array is stored:
// position c of the shuffled array holds element permute_index(c), and an
// element's jitter is the hash stream's value at its index, so any position is
// computed on its own where it is consumed.

// A bijection of [0, n): a hash permutation of [0, 2^b), b = bits of n - 1
// (xor with a key, odd multipliers and right xor-shifts are each invertible
// modulo 2^b), restricted to [0, n) by cycle walking: the orbit of c under a
// permutation returns to c, so the walk ends; the construction of Kensler's
// permute() ("Correlated Multi-Jittered Sampling", Pixar 2013), with rounds of
// its own.
WPBR_D uint32_t permute_index(uint32_t c, uint32_t n, uint32_t key) {
    if (n <= 1u) return 0u;
    const uint32_t b = 32u - (uint32_t)__clz((int)(n - 1u));
    const uint32_t w = 0xFFFFFFFFu >> (32u - b);
    const uint32_t s = (b + 1u) >> 1;
    do {
        c = ((c ^ key) * 0x9E3779B1u) & w;
        c ^= c >> s;
        c = ((c + (key >> 11)) * 0x2C1B3C6Du) & w;
        c ^= c >> s;
        c = ((c ^ (key >> 21)) * 0x297A2D39u) & w;
        c ^= c >> s;
    } while (c >= n);
    return c;
}

WPBR_D uint32_t shuffle_key(uint32_t seed, uint32_t pixel, uint32_t dim) {
    return mix32(mix32(seed ^ (0x632BE5ABu + dim)) ^ pixel);
}

// Value at position c of a pixel's 1D array (positions past the end wrap).
WPBR_D float sample1_at(const SamplerDef& S, uint32_t pixel, uint32_t c) {
    if (S.mode == 0u) return hash_unit(S.seed, pixel, 3u * c);
    if (c >= S.n1) c %= S.n1;
    const uint32_t k = permute_index(c, S.n1, shuffle_key(S.seed, pixel, 1u));
    return fminf(((float)k + hash_unit(S.seed, pixel, 3u * k)) * S.inv1, 1.0f);
}

// Value at position c of a pixel's 2D array.
WPBR_D void sample2_at(const SamplerDef& S, uint32_t pixel, uint32_t c, float* u) {
    if (S.mode == 0u) {
        u[0] = hash_unit(S.seed, pixel, 3u * c + 1u);
        u[1] = hash_unit(S.seed, pixel, 3u * c + 2u);
        return;
    }
    if (c >= S.n2) c %= S.n2;
    const uint32_t k = permute_index(c, S.n2, shuffle_key(S.seed, pixel, 2u));
    if (k >= S.root * S.root) {
        u[0] = u[1] = 0.0f;
        return;
    }
    const uint32_t i = k / S.root, j = k - i * S.root;
    u[0] = fminf(((float)i + hash_unit(S.seed, pixel, 3u * k + 1u)) * S.inv2, 1.0f);
    u[1] = fminf(((float)j + hash_unit(S.seed, pixel, 3u * k + 2u)) * S.inv2, 1.0f);
}

WPBR_D float next1d(const SamplerDef& S, uint32_t pixel, uint32_t& c1) { return sample1_at(S, pixel, c1++); }
WPBR_D void next2d(const SamplerDef& S, uint32_t pixel, uint32_t& c2, float* u) { sample2_at(S, pixel, c2++, u); }

// The arrays themselves (wpbr_sample_arrays): block b writes elements
// [b per_block, (b + 1) per_block) of the npix x len array, each thread every
// 256th; a thread divides once for its first (pixel, position) and then steps
// by 256 = qs len + rs.
__global__ __launch_bounds__(256) void k_sample_arrays(SamplerDef S, uint32_t pixel0, uint64_t n_el,
                                                       uint32_t len, uint32_t dim, uint64_t per_block,
                                                       float* __restrict__ out) {
    const uint64_t e0 = (uint64_t)blockIdx.x * per_block + threadIdx.x;
    const uint64_t e1 = min(n_el, (uint64_t)(blockIdx.x + 1) * per_block);
    if (e0 >= e1) return;
    uint32_t p = (uint32_t)(e0 / len), c = (uint32_t)(e0 - (uint64_t)p * len);
    const uint32_t qs = 256u / len, rs = 256u % len;
    for (uint64_t e = e0; e < e1; e += 256) {
        if (dim == 1u) {
            out[e] = sample1_at(S, pixel0 + p, c);
        } else {
            float u[2];
            sample2_at(S, pixel0 + p, c, u);
            reinterpret_cast<float2*>(out)[e] = make_float2(u[0], u[1]);
        }
        c += rs;
        p += qs;
        if (c >= len) {
            c -= len;
            p++;
        }
    }
}

wpbr_status sample_arrays_launch(const SamplerDef& S, uint32_t pixel0, uint32_t npix, float* u1, float* u2,
                                 hipStream_t st) {
    for (uint32_t dim = 1; dim <= 2; dim++) {
        float* out = dim == 1 ? u1 : u2;
        const uint32_t len = dim == 1 ? S.n1 : S.n2;
        const uint64_t n_el = (uint64_t)npix * len;
        if (!out || n_el == 0) continue;
        const uint64_t blocks = std::min<uint64_t>((n_el + 255) / 256, 8192);
        const uint64_t per_block = (n_el + blocks * 256 - 1) / (blocks * 256) * 256;
        hipLaunchKernelGGL(k_sample_arrays, dim3((uint32_t)blocks), dim3(256), 0, st, S, pixel0, n_el, len, dim,
                           per_block, out);
    }
    return launch_status();
}

// ---- per-pixel state order ----------------------------------------------------------
// Per-pixel state (rays, hits, shading terms) is indexed like the film, row-major
// (py * W + px): a wavefront traces 64 consecutive pixels of a row.  (An 8x8
// block per wavefront measured slower for both the primary and the secondary
// rays of the 1024x1024 render.)
WPBR_D void state_pixel(uint32_t i, uint32_t W, uint32_t& px, uint32_t& py) {
    px = i % W;
    py = i / W;
}

// ---- math helpers ---------------------------------------------------------------
WPBR_D float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
WPBR_D f3 xf_point4(const float* m, f3 p) {
    float x = ((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3];
    float y = ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7];
    float z = ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11];
    float w = ((m[12] * p.x + m[13] * p.y) + m[14] * p.z) + m[15];
    if (w == 1.0f) return mk3(x, y, z);
    return mk3(x / w, y / w, z / w);
}
WPBR_D float spherical_theta(f3 v) { return cr_acos(clampf(v.z, -1.0f, 1.0f)); }
WPBR_D float spherical_phi(f3 v) {
    float p = cr_atan2(v.y, v.x);
    return (p < 0) ? (float)((double)p + 2.0 * 3.14159265358979323846) : p;
}
WPBR_D void coordinate_system(f3 v1, f3* v2, f3* v3o) {
    if (fabsf(v1.x) > fabsf(v1.y)) *v2 = div3(mk3(-v1.z, 0, v1.x), sqrtf(v1.x * v1.x + v1.z * v1.z));
    else *v2 = div3(mk3(0, v1.z, -v1.y), sqrtf(v1.y * v1.y + v1.z * v1.z));
    *v3o = cross3(v1, *v2);
}
WPBR_D float cone_pdf(float cosThetaMax) { return 1 / (2 * PI_TRUE_F * (1 - cosThetaMax)); }
WPBR_D bool is_zero(f3 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }

// ---- env light -------------------------------------------------------------------
WPBR_D void env_bilerp(const SceneView& sv, float s_, float t_, float* rgb) {
    int W = (int)sv.env_w, H = (int)sv.env_h;
    float s = s_ * (float)W - 0.5f;
    float t = t_ * (float)H - 0.5f;
    int s0 = (int)floorf(s), t0 = (int)floorf(t);
    s0 = max(s0, 0); t0 = max(t0, 0);
    s0 = min(s0, W - 1); t0 = min(t0, H - 1);
    float ds = s - (float)s0, dt = t - (float)t0;
    int s0n = max(min(s0 + 1, W - 1), 0);
    int t0n = max(min(t0 + 1, H - 1), 0);
    const float* a = sv.env_rgb + ((size_t)t0 * W + s0) * 3;
    const float* b = sv.env_rgb + ((size_t)t0n * W + s0) * 3;
    const float* c = sv.env_rgb + ((size_t)t0 * W + s0n) * 3;
    const float* d = sv.env_rgb + ((size_t)t0n * W + s0n) * 3;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float v = ((a[k] * (1 - ds)) * (1 - dt) + (b[k] * (1 - ds)) * dt) + (c[k] * ds) * (1 - dt);
        rgb[k] = v + (d[k] * ds) * dt;
    }
}
WPBR_D float env_pdf(const SceneView& sv, float u, float v) {
    int nu = (int)sv.env_w, nv = (int)sv.env_h;
    int iu = (int)(u * (float)nu), iv = (int)(v * (float)nv);
    iu = iu < 0 ? 0 : (iu > nu - 1 ? nu - 1 : iu);
    iv = iv < 0 ? 0 : (iv > nv - 1 ? nv - 1 : iv);
    return sv.env_cond_func[(size_t)iv * nu + iu] / sv.env_marg_int[0];
}
// Smits illuminant spectrum of rgb (CSpectrumSampled.h:159-207): three bases and weights.
WPBR_D void rgb_basis(const float* rgb, int* b, float* w) {
    float r = rgb[0], g = rgb[1], bl = rgb[2];
    if (r <= g && r <= bl) {
        b[0] = 0; w[0] = r;
        if (g <= bl) { b[1] = 1; w[1] = g - r; b[2] = 6; w[2] = bl - g; }
        else { b[1] = 1; w[1] = bl - r; b[2] = 5; w[2] = g - bl; }
    } else if (g <= r && g <= bl) {
        b[0] = 0; w[0] = g;
        if (r <= bl) { b[1] = 2; w[1] = r - g; b[2] = 6; w[2] = bl - r; }
        else { b[1] = 2; w[1] = bl - g; b[2] = 4; w[2] = r - bl; }
    } else {
        b[0] = 0; w[0] = bl;
        if (r <= g) { b[1] = 3; w[1] = r - bl; b[2] = 5; w[2] = g - r; }
        else { b[1] = 3; w[1] = g - bl; b[2] = 4; w[2] = r - g; }
    }
}
// fromRGB(rgb, Illuminant)[i]: ((0 + B0*w0) + B1*w1) + B2*w2
WPBR_D float illum_at(const SceneView& sv, const int* b, const float* w, int i) {
    float s = 0.0f;
    s += sv.illum[b[0] * NSPEC + i] * w[0];
    s += sv.illum[b[1] * NSPEC + i] * w[1];
    s += sv.illum[b[2] * NSPEC + i] * w[2];
    return s;
}
// Env direction -> (u, v, theta).
WPBR_D void env_uv(const SceneView& sv, f3 dirW, float* u, float* v, float* theta) {
    affine inv = load_affine(sv.env_xform + 12);
    f3 dL = xf_vector(inv, dirW);
    float th = spherical_theta(dL), ph = spherical_phi(dL);
    *u = (float)((double)ph / (2 * PI_REF));
    *v = (float)((double)th / PI_REF);
    *theta = th;
}

// ---- microfacet + Fresnel ---------------------------------------------------------
WPBR_D float cos2theta(f3 w) { return w.z * w.z; }
WPBR_D float sin2theta(f3 w) { return smax(0.0f, 1.0f - cos2theta(w)); }
WPBR_D float sintheta(f3 w) { return sqrtf(sin2theta(w)); }
WPBR_D float tan2theta(f3 w) { return sin2theta(w) / cos2theta(w); }
WPBR_D float cosphi(f3 w) { float s = sintheta(w); return (s == 0) ? 1 : clampf(w.x / s, -1, 1); }
WPBR_D float sinphi(f3 w) { float s = sintheta(w); return (s == 0) ? 0 : clampf(w.y / s, -1, 1); }
WPBR_D float abscos(f3 w) { return fabsf(w.z); }
WPBR_D bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0; }

WPBR_D float ggx_d(f3 wh, float ax, float ay) {
    float t2 = tan2theta(wh);
    if (isinf(t2)) return 0.0f;
    const float cos4 = cos2theta(wh) * cos2theta(wh);
    float c2p = cosphi(wh) * cosphi(wh), s2p = sinphi(wh) * sinphi(wh);
    float e = (c2p / (ax * ax) + s2p / (ay * ay)) * t2;
    float one_e = 1 + e;
    double den = PI_REF * (double)ax * (double)ay * (double)cos4 * (double)one_e * (double)one_e;
    return (float)(1 / den);
}

WPBR_D float fresnel_dielectric(float cosI, float etaI, float etaT) {
    cosI = clampf(cosI, -1.0f, 1.0f);
    bool entering = cosI > 0.0f;
    if (!entering) { float t = etaI; etaI = etaT; etaT = t; cosI = fabsf(cosI); }
    float sinI = sqrtf(smax(0.0f, (float)(1.0 - (double)(cosI * cosI))));
    float sinT = etaI / etaT * sinI;
    if (sinT >= 1.0f) return 1.0f;
    float cosT = sqrtf(smax(0.0f, (float)(1.0 - (double)(sinT * sinT))));
    float Rparl = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
    float Rperp = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
    return (Rparl * Rparl + Rperp * Rperp) / 2;
}

// FrConductor (CBXDF.h:200-238, pbrt-v3) for one wavelength.  This is the
// innermost spectral loop of the radiance kernel (64 wavelengths x 2 MIS
// strategies per sample); its two square roots and two quotients use the
// hardware approximations (v_sqrt_f32, v_rcp_f32: about 1 ulp), which keeps
// the per-sample radiance far inside the 1e-4 relative bound.
WPBR_D float fresnel_conductor1(float cosI, float etaT, float k) {
    cosI = clampf(cosI, -1.0f, 1.0f);
    float eta = etaT / 1.0f, etak = k / 1.0f;
    float c2 = cosI * cosI;
    float s2 = (float)(1. - (double)c2);
    float eta2 = eta * eta, etak2 = etak * etak;
    float t0 = eta2 - etak2 - s2;
    float a2pb22 = t0 * t0 + eta2 * etak2 * 4;
    float a2pb2 = __builtin_amdgcn_sqrtf(a2pb22);
    float t1 = a2pb2 + c2;
    float a = __builtin_amdgcn_sqrtf((a2pb2 + t0) * 0.5f);
    float t2 = a * cosI * 2.0f;
    float Rs = (t1 - t2) * __builtin_amdgcn_rcpf(t1 + t2);
    float t3 = a2pb2 * c2 + s2 * s2;
    float t4 = t2 * s2;
    float Rp = Rs * (t3 - t4) * __builtin_amdgcn_rcpf(t3 + t4);
    return (Rp + Rs) * 0.5f;
}

// Microfacet BSDF terms for (wo, wi) in the local frame: pdf and the scalar parts
// of f*|cos wi| = Fr(cos) * sc * ac (conductor) or (1 * sc) * ac (dielectric).
struct BsdfTerms { float cos_fr, sc, ac, pdf; uint32_t zero; };

WPBR_D BsdfTerms microfacet_terms(const DevMaterial& m, f3 wo, f3 wi, f3 wh, float whD) {
    BsdfTerms r;
    r.pdf = (whD * abscos(wh)) / (4 * dot3(wo, wh));
    float cO = abscos(wo), cI = abscos(wi);
    r.ac = abscos(wi);
    r.cos_fr = dot3(wi, wh);
    r.zero = (cI == 0 || cO == 0 || (wh.x == 0 && wh.y == 0 && wh.z == 0)) ? 1u : 0u;
    if (m.type == MAT_CONDUCTOR) r.sc = whD / (4 * cI * cO);
    else {
        float fr = fresnel_dielectric(r.cos_fr, 1.0f, m.eta);
        r.sc = fr * whD / (4 * cI * cO);
    }
    return r;
}

// fromRGB(Reflectance) of a textured conductor's metallic value, kept as its
// Smits case: bases b and weights w (CSpectrumSampled.h:110-158).
struct TexK {
    int b[3];
    float w[3];
};
WPBR_D float refl_at(const SceneView& sv, const TexK& t, int i) {
    float s = 0.0f;
    s += sv.refl[t.b[0] * NSPEC + i] * t.w[0];
    s += sv.refl[t.b[1] * NSPEC + i] * t.w[1];
    s += sv.refl[t.b[2] * NSPEC + i] * t.w[2];
    return s;
}

// val(lambda) of the BSDF*|cos| for one wavelength; tk: the k spectrum of a
// textured conductor hit (null: the material's).
WPBR_D float bsdf_val(const SceneView& sv, const DevMaterial& m, float cos_fr, float sc, float ac,
                      uint32_t zero, const TexK* tk, int i) {
    if (zero) return 0.0f * ac;
    if (m.type == MAT_CONDUCTOR) {
        const float* es = sv.mat_spec + (size_t)m.spec * 2 * NSPEC;
        const float k = tk ? refl_at(sv, *tk, i) : es[NSPEC + i];
        return ((fresnel_conductor1(cos_fr, es[i], k) * 1.0f) * sc) * ac;
    }
    return (1.0f * sc) * ac;
}

// ---- surface interactions -----------------------------------------------------------
struct SI { f3 p, n, ns, dpdu, wo; };

WPBR_D SI si_triangle(const SceneView& sv, uint32_t tri, f3 rd, const float* b) {
    const float* v = sv.g.tri_verts + 9 * (size_t)tri;
    f3 P0 = mk3(v[0], v[1], v[2]), P1 = mk3(v[3], v[4], v[5]), P2 = mk3(v[6], v[7], v[8]);
    uint8_t flags = sv.tri_flags[tri];
    float uv[3][2] = {{0, 0}, {1, 0}, {1, 1}};
    if (flags & 1) {
        const float* u = sv.tri_uv + 6 * (size_t)tri;
        for (int i = 0; i < 3; i++) { uv[i][0] = u[2 * i]; uv[i][1] = u[2 * i + 1]; }
    }
    f3 dp02 = sub3(P0, P2), dp12 = sub3(P1, P2);
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
    float detInv = 1 / det;
    f3 dpdu = mul3(sub3(mul3(dp02, duv12[1]), mul3(dp12, duv02[1])), detInv);
    SI s;
    s.p = add3(add3(mul3(P0, b[0]), mul3(P1, b[1])), mul3(P2, b[2]));
    s.wo = neg3(rd);
    s.dpdu = dpdu;
    f3 ng = normalize3(cross3(dp02, dp12));
    s.n = ng;
    s.ns = ng;
    if (flags & 2) {
        const float* nn = sv.tri_normal + 9 * (size_t)tri;
        const float* tt = sv.tri_tangent + 9 * (size_t)tri;
        f3 N0 = mk3(nn[0], nn[1], nn[2]), N1 = mk3(nn[3], nn[4], nn[5]), N2 = mk3(nn[6], nn[7], nn[8]);
        f3 T0 = mk3(tt[0], tt[1], tt[2]), T1 = mk3(tt[3], tt[4], tt[5]), T2 = mk3(tt[6], tt[7], tt[8]);
        f3 ns = add3(add3(mul3(N0, b[0]), mul3(N1, b[1])), mul3(N2, b[2]));
        ns = is_zero(ns) ? s.n : normalize3(ns);
        f3 ss = add3(add3(mul3(T0, b[0]), mul3(T1, b[1])), mul3(T2, b[2]));
        ss = is_zero(ss) ? normalize3(dpdu) : normalize3(ss);
        f3 ts = cross3(ss, ns);
        if (len2(ts) > 0) {
            ts = normalize3(ts);
            ss = cross3(ts, ns);
        }
        s.ns = normalize3(cross3(ss, ts));
        s.dpdu = ss;
        s.n = dot3(s.n, s.ns) < 0.0f ? neg3(s.n) : s.n;
    }
    return s;
}

WPBR_D SI si_sphere(const SceneView& sv, uint32_t sid, f3 o, f3 d, float t) {
    const float* xf = sv.g.sphere_xform + 24 * (size_t)sid;
    affine W = load_affine(xf), Inv = load_affine(xf + 12);
    float radius = sv.g.sphere_radius[sid];
    f3 oL = xf_point(Inv, o), dL = xf_vector(Inv, d);
    f3 pH = add3(oL, mul3(dL, t));
    float theta = cr_acos(pH.z / radius);
    float invZR = (float)(1.0 / (double)sqrtf(pH.x * pH.x + pH.y * pH.y));
    float cphi = pH.x * invZR, sphi = pH.y * invZR;
    f3 dpdu = mul3(mk3(-pH.y, pH.x, 0.0f), (float)(2 * PI_REF));
    f3 dpdv = mul3(mk3(pH.z * cphi, pH.z * sphi, -radius * cr_sin(theta)), (float)PI_REF);
    f3 n = normalize3(cross3(dpdu, dpdv));
    SI s;
    s.p = xf_point(W, pH);
    s.n = normalize3(xf_normal(Inv, n));
    s.ns = s.n;
    s.dpdu = xf_vector(W, dpdu);
    s.wo = normalize3(xf_vector(W, neg3(dL)));
    return s;
}

// ---- sphere lights ---------------------------------------------------------------------
WPBR_D f3 sphere_light_sample(const SceneView& sv, uint32_t sid, f3 p, const float* u, float* pdf) {
    const float* xf = sv.g.sphere_xform + 24 * (size_t)sid;
    affine W = load_affine(xf);
    float radius = sv.g.sphere_radius[sid];
    f3 pc = xf_point(W, mk3(0, 0, 0));
    f3 wc = normalize3(sub3(pc, p));
    f3 wcX, wcY;
    coordinate_system(wc, &wcX, &wcY);
    float sinThetaMax2 = radius * radius / len2(sub3(p, pc));
    float cosThetaMax = sqrtf(smax(0.0f, (float)(1.0 - (double)sinThetaMax2)));
    float cosTheta = (float)((double)u[0] * (1.0 - (double)cosThetaMax) + (double)cosThetaMax);
    float sinTheta = sqrtf(smax(0.0f, (float)(1.0 - (double)(cosTheta * cosTheta))));
    float phi = (float)((double)(u[1] * 2) * PI_REF);
    float dc = sqrtf(len2(sub3(p, pc)));
    float ds = dc * cosTheta - sqrtf(smax(0.0f, radius * radius - dc * dc * sinTheta * sinTheta));
    float cosAlpha = (dc * dc + radius * radius - ds * ds) / (2 * dc * radius);
    float sinAlpha = sqrtf(smax(0.0f, 1 - cosAlpha * cosAlpha));
    f3 nL = add3(add3(mul3(neg3(wcX), sinAlpha * cr_cos(phi)), mul3(neg3(wcY), sinAlpha * cr_sin(phi))),
                 mul3(neg3(wc), cosAlpha));
    f3 pL = mul3(nL, radius);
    *pdf = cone_pdf(cosThetaMax);
    return xf_point(W, pL);
}
WPBR_D float sphere_light_pdf(const SceneView& sv, uint32_t sid) {
    const float* xf = sv.g.sphere_xform + 24 * (size_t)sid;
    affine W = load_affine(xf);
    float radius = sv.g.sphere_radius[sid];
    f3 pc = xf_point(W, mk3(0, 0, 0));
    float sinThetaMax2 = radius * radius / len2(pc);
    float cosThetaMax = sqrtf(smax(0.0f, 1 - sinThetaMax2));
    return cone_pdf(cosThetaMax);
}

// Wave-aggregated append to the secondary ray queue.
// Queue slots for the wave's light rays and BSDF rays with one atomic: the
// wave's light rays first, then its BSDF rays, each in lane order.
WPBR_D void queue_append2(bool wantA, bool wantB, uint32_t* count, uint32_t& slotA, uint32_t& slotB) {
    const uint64_t ma = ballot64(wantA), mb = ballot64(wantB);
    const uint32_t na = popc64(ma), total = na + popc64(mb);
    uint32_t base = 0;
    if (total && lane_id() == 0) base = atomicAdd(count, total);
    base = (uint32_t)__shfl((int)base, 0);
    slotA = base + popc64(ma & lanemask_lt());
    slotB = base + na + popc64(mb & lanemask_lt());
}

WPBR_D uint32_t queue_append(bool want, uint32_t* count) {
    uint64_t m = ballot64(want);
    uint32_t total = popc64(m);
    uint32_t base = 0;
    uint32_t leader = m ? (uint32_t)__ffsll((long long)m) - 1 : 0;
    if (total && lane_id() == leader) base = atomicAdd(count, total);
    base = (uint32_t)__shfl((int)base, (int)leader);
    return base + popc64(m & lanemask_lt());
}

WPBR_D void put_ray(float* q, uint32_t cap, uint32_t slot, f3 o, f3 d) {
    q[slot] = o.x; q[cap + slot] = o.y; q[2 * (size_t)cap + slot] = o.z;
    q[3 * (size_t)cap + slot] = d.x; q[4 * (size_t)cap + slot] = d.y; q[5 * (size_t)cap + slot] = d.z;
}

// ---- full surface interactions (the collision boundary and textured hits) -----------
// The interaction runCollisionSystem attaches to each ray cast, in the SoA layout
// of WPBR_SI_* (wpbr.h): CSurfaceInteraction for bSurfInteraction rays
// (CSTriangle.cpp:149-319; CSSphere.cpp:122-186 then CTransform.h:41-66), the plain
// CInteraction fields for the others (SLBVH.cpp:55-59), zeros on a miss; the
// differentials of the rays that carry a CRayDifferential
// (CSufraceInteraction.h:98-148).  Default-constructed members are zero: the
// sphere's world interaction keeps uv = 0 and dpdx = dpdy = 0 (the transform
// leaves uv uncopied and sets dpdy from dpdx), and a triangle's geometric
// dndu / dndv stay 0.
struct FullSI {
    f3 p, wo, n, dpdu, dpdv, dndu, dndv, sn, sdpdu, sdpdv, sdndu, sdndv, dpdx, dpdy;
    float time, u, v, dudx, dvdx, dudy, dvdy;
};

WPBR_D FullSI fsi_zero() {
    FullSI s;
    const f3 z = mk3(0.0f, 0.0f, 0.0f);
    s.p = s.wo = s.n = s.dpdu = s.dpdv = s.dndu = s.dndv = z;
    s.sn = s.sdpdu = s.sdpdv = s.sdndu = s.sdndv = s.dpdx = s.dpdy = z;
    s.time = s.u = s.v = s.dudx = s.dvdx = s.dudy = s.dvdy = 0.0f;
    return s;
}

// CSurfaceInteraction's constructor (CSufraceInteraction.h:58-83).
WPBR_D void fsi_construct(FullSI& s, f3 p, float u, float v, f3 wo, f3 dpdu, f3 dpdv, f3 dndu, f3 dndv,
                          float time) {
    s.p = p; s.u = u; s.v = v; s.wo = wo; s.time = time;
    s.dpdu = dpdu; s.dpdv = dpdv; s.dndu = dndu; s.dndv = dndv;
    s.n = normalize3(cross3(dpdu, dpdv));
    s.sn = s.n; s.sdpdu = dpdu; s.sdpdv = dpdv; s.sdndu = dndu; s.sdndv = dndv;
}

WPBR_D FullSI si_full_triangle(const SceneView& sv, uint32_t tri, f3 rd, const float* b, float t) {
    const float* v = sv.g.tri_verts + 9 * (size_t)tri;
    const f3 P0 = mk3(v[0], v[1], v[2]), P1 = mk3(v[3], v[4], v[5]), P2 = mk3(v[6], v[7], v[8]);
    const uint8_t flags = sv.tri_flags ? sv.tri_flags[tri] : (uint8_t)0;
    float uv[3][2] = {{0, 0}, {1, 0}, {1, 1}};
    if (flags & 1) {
        const float* u = sv.tri_uv + 6 * (size_t)tri;
        for (int i = 0; i < 3; i++) { uv[i][0] = u[2 * i]; uv[i][1] = u[2 * i + 1]; }
    }
    const f3 dp02 = sub3(P0, P2), dp12 = sub3(P1, P2);
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
    const f3 dpdu = mul3(sub3(mul3(dp02, duv12[1]), mul3(dp12, duv02[1])), detInv);
    const f3 dpdv = mul3(add3(mul3(neg3(dp02), duv12[0]), mul3(dp12, duv02[0])), detInv);
    const f3 pHit = add3(add3(mul3(P0, b[0]), mul3(P1, b[1])), mul3(P2, b[2]));
    const float uh = (uv[0][0] * b[0] + uv[1][0] * b[1]) + uv[2][0] * b[2];
    const float vh = (uv[0][1] * b[0] + uv[1][1] * b[1]) + uv[2][1] * b[2];
    FullSI s = fsi_zero();
    const f3 z = mk3(0.0f, 0.0f, 0.0f);
    fsi_construct(s, pHit, uh, vh, neg3(rd), dpdu, dpdv, z, z, t);
    s.n = s.sn = normalize3(cross3(dp02, dp12));
    if (flags & 2) {
        const float* nn = sv.tri_normal + 9 * (size_t)tri;
        const float* tt = sv.tri_tangent + 9 * (size_t)tri;
        const f3 N0 = mk3(nn[0], nn[1], nn[2]), N1 = mk3(nn[3], nn[4], nn[5]), N2 = mk3(nn[6], nn[7], nn[8]);
        const f3 T0 = mk3(tt[0], tt[1], tt[2]), T1 = mk3(tt[3], tt[4], tt[5]), T2 = mk3(tt[6], tt[7], tt[8]);
        f3 ns = add3(add3(mul3(N0, b[0]), mul3(N1, b[1])), mul3(N2, b[2]));
        ns = is_zero(ns) ? s.n : normalize3(ns);
        f3 ss = add3(add3(mul3(T0, b[0]), mul3(T1, b[1])), mul3(T2, b[2]));
        ss = is_zero(ss) ? normalize3(s.dpdu) : normalize3(ss);
        f3 ts = cross3(ss, ns);
        if (len2(ts) > 0) {
            ts = normalize3(ts);
            ss = cross3(ts, ns);
        }
        const f3 dn02 = sub3(N0, N2), dn12 = sub3(N1, N2);
        f3 dndu = z, dndv = z;
        const float det2 = duv02[0] * duv12[1] - duv02[1] * duv12[0];
        if (det2 != 0) {
            const float inv2 = 1 / det2;
            dndu = mul3(sub3(mul3(dn02, duv12[1]), mul3(dn12, duv02[1])), inv2);
            dndv = mul3(add3(mul3(neg3(dn02), duv12[0]), mul3(dn12, duv02[0])), inv2);
        }
        s.sn = normalize3(cross3(ss, ts));
        s.sdpdu = ss;
        s.sdpdv = ts;
        s.sdndu = dndu;
        s.sdndv = dndv;
        s.n = dot3(s.n, s.sn) < 0.0f ? neg3(s.n) : s.n;
    }
    return s;
}

WPBR_D FullSI si_full_sphere(const SceneView& sv, uint32_t sid, f3 o, f3 d, float t) {
    const float* xf = sv.g.sphere_xform + 24 * (size_t)sid;
    const affine W = load_affine(xf), Inv = load_affine(xf + 12);
    const float radius = sv.g.sphere_radius[sid];
    const f3 oL = xf_point(Inv, o), dL = xf_vector(Inv, d);
    const f3 pH = add3(oL, mul3(dL, t));
    float phi = cr_atan2(pH.y, pH.x);
    if (phi < 0.0f) phi = (float)((double)phi + 2 * PI_REF);
    const float u = (float)((double)phi / (2 * PI_REF));
    const float theta = cr_acos(pH.z / radius);
    const float v = (float)((double)theta / PI_REF);
    const float invZR = (float)(1.0 / (double)sqrtf(pH.x * pH.x + pH.y * pH.y));
    const float cphi = pH.x * invZR, sphi = pH.y * invZR;
    const f3 dpdu = mul3(mk3(-pH.y, pH.x, 0.0f), (float)(2 * PI_REF));
    const f3 dpdv = mul3(mk3(pH.z * cphi, pH.z * sphi, -radius * cr_sin(theta)), (float)PI_REF);
    const f3 d2Pduu = mul3(mk3(pH.x, pH.y, 0.0f), (float)(-2 * PI_REF * 2 * PI_REF));
    const f3 d2Pduv = mul3(mul3(mul3(mul3(mk3(-sphi, cphi, 0.0f), (float)PI_REF), pH.z), 2.0f), (float)PI_REF);
    const f3 d2Pdvv = mul3(mul3(mk3(pH.x, pH.y, pH.z), (float)(-PI_REF)), (float)PI_REF);
    const float E = dot3(dpdu, dpdu), F = dot3(dpdu, dpdv), G = dot3(dpdv, dpdv);
    const f3 N = normalize3(cross3(dpdu, dpdv));
    const float e = dot3(N, d2Pduu), f = dot3(N, d2Pduv), g = dot3(N, d2Pdvv);
    const float invEGF2 = 1 / (E * G - F * F);
    const f3 dndu = add3(mul3(dpdu, (f * F - e * G) * invEGF2), mul3(dpdv, (e * F - f * E) * invEGF2));
    const f3 dndv = add3(mul3(dpdu, (g * F - f * G) * invEGF2), mul3(dpdv, (f * F - g * E) * invEGF2));
    FullSI l = fsi_zero();
    fsi_construct(l, pH, u, v, neg3(dL), dpdu, dpdv, dndu, dndv, t);
    // CTransform::operator()(const CSurfaceInteraction&) into a fresh interaction
    FullSI s = fsi_zero();
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

// pbrt-v3 SolveLinearSystem2x2 (WoodenMath's solveLinearSystem2x2).
WPBR_D bool solve2x2(float a00, float a01, float a10, float a11, float b0, float b1, float* x0, float* x1) {
    const float det = a00 * a11 - a01 * a10;
    if (fabsf(det) < 1e-10f) return false;
    *x0 = (a11 * b0 - a01 * b1) / det;
    *x1 = (a00 * b1 - a10 * b0) / det;
    return !(__builtin_isnan(*x0) || __builtin_isnan(*x1));
}

// CSurfaceInteraction::computeDifferentials (CSufraceInteraction.h:98-148).
WPBR_D void si_differentials(FullSI& s, f3 xo, f3 xd, f3 yo, f3 yd) {
    const float dd = dot3(s.n, s.p);
    const float tx = -(dot3(s.n, xo) - dd) / dot3(s.n, xd);
    const f3 px = add3(xo, mul3(xd, tx));
    const float ty = -(dot3(s.n, yo) - dd) / dot3(s.n, yd);
    const f3 py = add3(yo, mul3(yd, ty));
    s.dpdx = sub3(px, s.p);
    s.dpdy = sub3(py, s.p);
    int d0, d1;
    if (fabsf(s.n.x) > fabsf(s.n.y) && fabsf(s.n.x) > fabsf(s.n.z)) { d0 = 1; d1 = 2; }
    else if (fabsf(s.n.y) > fabsf(s.n.z)) { d0 = 0; d1 = 2; }
    else { d0 = 0; d1 = 1; }
    const float a00 = comp(s.dpdu, d0), a01 = comp(s.dpdv, d0), a10 = comp(s.dpdu, d1), a11 = comp(s.dpdv, d1);
    if (!solve2x2(a00, a01, a10, a11, comp(px, d0) - comp(s.p, d0), comp(px, d1) - comp(s.p, d1), &s.dudx, &s.dvdx))
        s.dudx = s.dvdx = 0;
    if (!solve2x2(a00, a01, a10, a11, comp(py, d0) - comp(s.p, d0), comp(py, d1) - comp(s.p, d1), &s.dudy, &s.dvdy))
        s.dudy = s.dvdy = 0;
}

// ---- kernels ----------------------------------------------------------------------------
WPBR_D void camera_body(const RenderLaunch& L, uint32_t i) {
    if (i >= L.n) return;
    const SceneView& sv = L.sv;
    uint32_t px, py;
    state_pixel(i, sv.width, px, py);
    uint32_t c2 = L.spp_index == 0 ? 0u : L.ws.c2[i];
    if (L.spp_index == 0) L.ws.c1[i] = 0;
    float u[2];
    next2d(L.smp, py * sv.width + px, c2, u);
    L.ws.c2[i] = c2;
    float fx = (float)px + u[0], fy = (float)py + u[1];
    L.ws.pfilm[i] = fx;
    L.ws.pfilm[L.n + i] = fy;
    f3 pCam = xf_point4(sv.raster_camera, mk3(fx, fy, 0.0f));
    f3 dC = normalize3(pCam);
    affine CW = load_affine(sv.camera_world);
    f3 o = xf_point(CW, mk3(0, 0, 0));
    f3 d = normalize3(xf_vector(CW, dC));
    const uint32_t n = L.n;
    L.ws.rays[i] = o.x; L.ws.rays[n + i] = o.y; L.ws.rays[2 * (size_t)n + i] = o.z;
    L.ws.rays[3 * (size_t)n + i] = d.x; L.ws.rays[4 * (size_t)n + i] = d.y; L.ws.rays[5 * (size_t)n + i] = d.z;
    if (i == 0) *L.ws.qcount = 0;
}
__global__ __launch_bounds__(256) void k_camera(RenderLaunch L) { camera_body(L, blockIdx.x * blockDim.x + threadIdx.x); }

constexpr uint32_t F_LIGHT = 1u, F_BSDF = 2u, F_ENV = 4u, F_DONE = 8u, F_MISS = 16u;
constexpr uint32_t F_TEXK = 32u;  // a textured conductor hit: its k RGB in WaveState::tk

// MicrofaceDistr.h:18-24.
WPBR_D float roughness_to_alpha(float roughness) {
    roughness = fmaxf(roughness, 1e-3f);
    const float x = cr_log(roughness);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

// TextureMappingSphere::sphere (CTextureMapping.h:60-66): the point in the
// sphere's object space, normalised, as (theta / PI, phi / (2 PI)).
WPBR_D void sphere_st(const affine& inv, f3 p, float* st) {
    const f3 v = normalize3(xf_point(inv, p));
    const float th = spherical_theta(v), ph = spherical_phi(v);
    st[0] = (float)((double)th * 1.0 / PI_REF);
    st[1] = (float)((double)ph / (2 * PI_REF));
}

// The CTextureMappedPoint (s, t, ds/dx, dt/dx, ds/dy, dt/dy) of a camera ray's
// hit: the hit's full interaction, its differentials from the camera's ray
// differential (CSCameraPerspective.cpp:37-41: the same origin, directions
// through pCamera + dxCamera / dyCamera, dxCamera the raster->camera transform
// of the vector (1, 0, 0), i.e. its first column), then the triangle's UV
// mapping (STriangle::mapUV -> TextureMappingUV, CTextureMapping.h:39-55) or
// the sphere's (SSphere::mapUV -> TextureMappingSphere, :68-90).
WPBR_D void camera_hit_texture_point(const RenderLaunch& L, uint32_t i, uint32_t shape, f3 o, f3 d, float t,
                                     const float* b, float* q) {
    const SceneView& sv = L.sv;
    const bool sphere = shape < sv.g.n_spheres;
    FullSI s = sphere ? si_full_sphere(sv, shape, o, d, t) : si_full_triangle(sv, shape - sv.g.n_spheres, d, b, t);
    const f3 pCam = xf_point4(sv.raster_camera, mk3(L.ws.pfilm[i], L.ws.pfilm[L.n + i], 0.0f));
    const float* rc = sv.raster_camera;
    const f3 dxC = mk3(rc[0], rc[4], rc[8]), dyC = mk3(rc[1], rc[5], rc[9]);
    const affine CW = load_affine(sv.camera_world);
    const f3 dx = normalize3(xf_vector(CW, normalize3(add3(pCam, dxC))));
    const f3 dy = normalize3(xf_vector(CW, normalize3(add3(pCam, dyC))));
    si_differentials(s, o, dx, o, dy);
    if (!sphere) {
        q[0] = s.u; q[1] = s.v;
        q[2] = s.dudx; q[3] = s.dvdx;
        q[4] = s.dudy; q[5] = s.dvdy;
        return;
    }
    const affine inv = load_affine(sv.g.sphere_xform + 24 * (size_t)shape + 12);
    float st[2], sx[2], sy[2];
    sphere_st(inv, s.p, st);
    const float dt = 0.1f;
    sphere_st(inv, add3(s.p, mul3(s.dpdx, dt)), sx);
    sphere_st(inv, add3(s.p, mul3(s.dpdy, dt)), sy);
    float dstdx[2] = {(sx[0] - st[0]) / dt, (sx[1] - st[1]) / dt};
    float dstdy[2] = {(sy[0] - st[0]) / dt, (sy[1] - st[1]) / dt};
    if (dstdx[1] > .5f) dstdx[1] = 1 - dstdx[1];
    else if (dstdx[1] < -.5f) dstdx[1] = -dstdx[1] + 1;
    if (dstdy[1] > .5f) dstdy[1] = 1 - dstdy[1];
    else if (dstdy[1] < -.5f) dstdy[1] = -dstdy[1] + 1;
    q[0] = st[0]; q[1] = st[1];
    q[2] = dstdx[0]; q[3] = dstdx[1];
    q[4] = dstdy[0]; q[5] = dstdy[1];
}

WPBR_D void shade1_body(const RenderLaunch& L, uint32_t i) {
    bool active = i < L.n;
    const SceneView& sv = L.sv;
    const WaveState& ws = L.ws;
    const uint32_t n = L.n, cap = 2 * L.n;
    uint32_t flags = F_DONE;
    bool wantL = false, wantB = false;
    uint32_t tgtL = 0xFFFFFFFFu, tgtB = 0xFFFFFFFFu;
    f3 lo = mk3(0, 0, 0), ld = lo, bo = lo, bd = lo;
    if (active) {
        f3 o = mk3(ws.rays[i], ws.rays[n + i], ws.rays[2 * (size_t)n + i]);
        f3 d = mk3(ws.rays[3 * (size_t)n + i], ws.rays[4 * (size_t)n + i], ws.rays[5 * (size_t)n + i]);
        uint32_t shape = ws.hit[i];
        uint32_t c1 = ws.c1[i], c2 = ws.c2[i];
        if (shape == 0xFFFFFFFFu) {
            flags = F_DONE | F_MISS;
            float X = 0, Y = 0, Z = 0;
            if (sv.has_env) {
                float u, v, th, rgb[3];
                env_uv(sv, d, &u, &v, &th);
                env_bilerp(sv, u, v, rgb);
                int b[3];
                float w[3];
                rgb_basis(rgb, b, w);
                for (int k = 0; k < NSPEC; k++) {
                    float Lk = illum_at(sv, b, w, k) * sv.env_emit[k];
                    X += Lk * sv.xyz[k];
                    Y += Lk * sv.xyz[NSPEC + k];
                    Z += Lk * sv.xyz[2 * NSPEC + k];
                }
            }
            const float dl = 300.0f / 64.0f;
            ws.le[i] = X * dl; ws.le[n + i] = Y * dl; ws.le[2 * (size_t)n + i] = Z * dl;
        } else {
            float t = ws.hit_t[i];
            SI si;
            if (shape < sv.g.n_spheres) {
                si = si_sphere(sv, shape, o, d, t);
            } else {
                float b[3] = {ws.bary[i], ws.bary[n + i], ws.bary[2 * (size_t)n + i]};
                si = si_triangle(sv, shape - sv.g.n_spheres, d, b);
            }
            const uint32_t nLights = sv.n_sphere_lights + (sv.has_env ? 1u : 0u);
            uint32_t px, py;
            state_pixel(i, sv.width, px, py);
            const uint32_t pix = py * sv.width + px;  // sample-stream id
            float u1 = next1d(L.smp, pix, c1);
            uint32_t li = (uint32_t)((u1 + 0.5f) * (float)(nLights - 1));
            if (li > nLights - 1) li = nLights - 1;
            bool envPick = sv.has_env && li == sv.n_sphere_lights;
            float ul[2];
            next2d(L.smp, pix, c2, ul);
            f3 wiL = mk3(0, 0, 0);
            float pdfL = 0;
            if (!envPick) {
                f3 sp = sphere_light_sample(sv, sv.light_shape[li], si.p, ul, &pdfL);
                wiL = normalize3(sub3(sp, si.p));
            }
            uint32_t mid = sv.shape_material[shape];
            flags = envPick ? F_ENV : 0u;
            ws.light[i] = li;
            if (!envPick) tgtL = tgtB = sv.light_shape[li];
            if (mid == 0xFFFFFFFFu) {
                flags |= F_DONE;
            } else {
                DevMaterial m = sv.materials[mid];
                ws.mat[i] = mid;
                if (m.tex_k >= 0 || m.tex_r >= 0) {
                    // a textured metal (JobGenerateBSDFMaterialMetal, CMaterialMetal.cpp:24-52)
                    float q[6];
                    const float bt3[3] = {ws.bary[i], ws.bary[n + i], ws.bary[2 * (size_t)n + i]};
                    camera_hit_texture_point(L, i, shape, o, d, t, bt3, q);
                    if (m.tex_r >= 0) {
                        const float3 r = tex_lookup(sv.textures[m.tex_r], q[0], q[1], q[2], q[3], q[4], q[5]);
                        m.alpha = roughness_to_alpha(clampf(r.x * 90, 0.0f, 90.0f));
                    }
                    if (m.tex_k >= 0) {
                        const float3 k = tex_lookup(sv.textures[m.tex_k], q[0], q[1], q[2], q[3], q[4], q[5]);
                        ws.tk[i] = k.x; ws.tk[n + i] = k.y; ws.tk[2 * (size_t)n + i] = k.z;
                        flags |= F_TEXK;
                    }
                }
                f3 ss = normalize3(si.dpdu);
                f3 ns = si.ns;
                f3 ts = cross3(ns, ss);
                f3 woL = mk3(dot3(ss, si.wo), dot3(ts, si.wo), dot3(ns, si.wo));
                if (L.record) {
                    // the adjoint replays this sample from its frame and samples
                    f3 wlr = mk3(dot3(ss, wiL), dot3(ts, wiL), dot3(ns, wiL));
                    const float rec3[5][3] = {{woL.x, woL.y, woL.z}, {ss.x, ss.y, ss.z}, {ts.x, ts.y, ts.z},
                                              {ns.x, ns.y, ns.z}, {wlr.x, wlr.y, wlr.z}};
                    for (int c = 0; c < 3; c++) {
                        ws.r_wo[(size_t)c * n + i] = rec3[0][c];
                        ws.r_frame[(size_t)c * n + i] = rec3[1][c];
                        ws.r_frame[(size_t)(3 + c) * n + i] = rec3[2][c];
                        ws.r_frame[(size_t)(6 + c) * n + i] = rec3[3][c];
                        ws.r_wl[(size_t)c * n + i] = rec3[4][c];
                    }
                    ws.r_pdfl[i] = pdfL;
                }
                // light strategy: BSDF Compute toward the light
                if (!envPick && m.type != MAT_SPECULAR) {
                    f3 wl = mk3(dot3(ss, wiL), dot3(ts, wiL), dot3(ns, wiL));
                    f3 wh = normalize3(add3(woL, wl));
                    float whD = ggx_d(wh, m.alpha, m.alpha);
                    BsdfTerms bt = microfacet_terms(m, woL, wl, wh, whD);
                    ws.l_cos[i] = bt.cos_fr; ws.l_sc[i] = bt.sc; ws.l_ac[i] = bt.ac;
                    ws.l_bpdf[i] = bt.pdf; ws.l_zero[i] = bt.zero; ws.l_pdf[i] = pdfL;
                    if (bt.pdf != 0.0f) {
                        wantL = true;
                        lo = add3(si.p, mul3(wiL, 0.1f));
                        ld = wiL;
                    }
                }
                // BSDF strategy: sample
                f3 wiBL;
                bool ok = true;
                if (m.type == MAT_SPECULAR) {
                    wiBL = mk3(-woL.x, -woL.y, woL.z);
                    float c = abscos(wiBL);
                    ws.b_pdf[i] = 1.0f;
                    ws.b_cos[i] = 0.0f;
                    ws.b_zero[i] = c == 0 ? 1u : 0u;
                    ws.b_sc[i] = (c == 0) ? 0.0f : (fresnel_dielectric(wiBL.z, 1.0f, m.eta) / c);
                    ws.b_ac[i] = c;
                } else {
                    float ub[2];
                    next2d(L.smp, pix, c2, ub);
                    if (L.record) { ws.r_ub[i] = ub[0]; ws.r_ub[n + i] = ub[1]; }
                    float phi = (float)((2 * PI_REF) * (double)ub[1]);
                    float tan2 = m.alpha * m.alpha * ub[0] / (1.0f - ub[0]);
                    float cosT = 1 / sqrtf(1 + tan2);
                    float sinT = sqrtf(smax(0.f, 1.f - cosT * cosT));
                    f3 wh = mk3(sinT * cr_cos(phi), sinT * cr_sin(phi), cosT);
                    if (!same_hemi(woL, wh)) wh = neg3(wh);
                    float whD = ggx_d(wh, m.alpha, m.alpha);
                    wiBL = add3(neg3(woL), mul3(wh, 2 * dot3(woL, wh)));
                    if (!same_hemi(wiBL, woL)) {
                        ok = false;
                    } else {
                        BsdfTerms bt = microfacet_terms(m, woL, wiBL, wh, whD);
                        ws.b_cos[i] = bt.cos_fr; ws.b_sc[i] = bt.sc; ws.b_ac[i] = bt.ac;
                        ws.b_pdf[i] = bt.pdf; ws.b_zero[i] = bt.zero;
                    }
                }
                if (ok) {
                    f3 wiB = mk3((ss.x * wiBL.x + ts.x * wiBL.y) + ns.x * wiBL.z,
                                 (ss.y * wiBL.x + ts.y * wiBL.y) + ns.y * wiBL.z,
                                 (ss.z * wiBL.x + ts.z * wiBL.y) + ns.z * wiBL.z);
                    ws.b_wi[i] = wiB.x; ws.b_wi[n + i] = wiB.y; ws.b_wi[2 * (size_t)n + i] = wiB.z;
                    wantB = true;
                    bo = add3(si.p, mul3(wiB, 0.01f));
                    bd = wiB;
                }
            }
            ws.c1[i] = c1;
            ws.c2[i] = c2;
        }
        if (wantL) flags |= F_LIGHT;
        if (wantB) flags |= F_BSDF;
        ws.flags[i] = flags;
    }
    uint32_t sl, sb;
    queue_append2(wantL, wantB, ws.qcount, sl, sb);
    // each queue ray asks whether its first hit is the sampled light (or nothing)
    if (wantL) { put_ray(ws.qrays, cap, sl, lo, ld); ws.l_ray[i] = sl; ws.qtarget[sl] = tgtL; }
    if (wantB) { put_ray(ws.qrays, cap, sb, bo, bd); ws.b_ray[i] = sb; ws.qtarget[sb] = tgtB; }
}
__global__ __launch_bounds__(256) void k_shade1(RenderLaunch L) { shade1_body(L, blockIdx.x * blockDim.x + threadIdx.x); }

WPBR_D float mitchell1(float x) {
    const float B = (float)(1.0 / 3.0f), Cc = (float)(1.0 / 3.0f);
    x = fabsf(2 * x);
    if (x > 1)
        return ((-B - 6 * Cc) * x * x * x + (6 * B + 30 * Cc) * x * x + (-12 * B - 48 * Cc) * x +
                (8 * B + 24 * Cc)) * (1.f / 6.f);
    return ((12 - 9 * B - 6 * Cc) * x * x * x + (-18 + 12 * B + 6 * Cc) * x * x + (6 - 2 * B)) *
           (1.f / 6.f);
}

// Radiance of pixel i's sample in this pass, as XYZ; black = every wavelength is 0.
WPBR_D void sample_radiance(const RenderLaunch& L, uint32_t i, float& X, float& Y, float& Z,
                            bool& black) {
    const SceneView& sv = L.sv;
    const WaveState& ws = L.ws;
    const uint32_t n = L.n;
    uint32_t flags = ws.flags[i];
    X = 0; Y = 0; Z = 0;
    black = true;
    if (flags & F_MISS) {
        X = ws.le[i]; Y = ws.le[n + i]; Z = ws.le[2 * (size_t)n + i];
        black = false;  // the splat tests L itself; an exactly black Le adds zeros
    } else if (flags & (F_LIGHT | F_BSDF)) {
        const uint32_t li = ws.light[i];
        const DevMaterial m = sv.materials[ws.mat[i]];
        bool envPick = flags & F_ENV;
        // light strategy
        bool useL = false;
        float sL = 0, lcos = 0, lsc = 0, lac = 0;
        uint32_t lzero = 0;
        const float* Le = nullptr;
        if (flags & F_LIGHT) {
            uint32_t h = ws.qhit[ws.l_ray[i]];
            if (h == 0xFFFFFFFFu || h == sv.light_shape[li]) {
                Le = sv.light_emit + (size_t)li * NSPEC;
                bool lb = true;
                for (int k = 0; k < NSPEC; k++) if (Le[k] != 0.0f) { lb = false; break; }
                if (!lb) {
                    float pdfL = ws.l_pdf[i], pdfB = ws.l_bpdf[i];
                    float w = (pdfL * pdfL) / (pdfL * pdfL + pdfB * pdfB);
                    sL = w / pdfL;
                    lcos = ws.l_cos[i]; lsc = ws.l_sc[i]; lac = ws.l_ac[i]; lzero = ws.l_zero[i];
                    useL = true;
                }
            }
        }
        // BSDF strategy
        bool useB = false;
        float sB = 0, bcos = 0, bsc = 0, bac = 0;
        uint32_t bzero = 0;
        int eb[3] = {0, 0, 0};
        float ew[3] = {0, 0, 0};
        const float* Lb = nullptr;
        if (flags & F_BSDF) {
            uint32_t h = ws.qhit[ws.b_ray[i]];
            bool reach = h == 0xFFFFFFFFu || (!envPick && h == sv.light_shape[li]);
            if (reach) {
                float pdfB = ws.b_pdf[i];
                float pdfL2 = 0;
                bool bb = true;
                if (envPick) {
                    f3 wi = mk3(ws.b_wi[i], ws.b_wi[n + i], ws.b_wi[2 * (size_t)n + i]);
                    float u, v, th;
                    env_uv(sv, wi, &u, &v, &th);
                    float sinT = cr_sin(th);
                    if (sinT != 0.0f) {
                        pdfL2 = (float)((double)env_pdf(sv, u, v) / (2 * PI_REF * PI_REF * (double)sinT));
                        float rgb[3];
                        env_bilerp(sv, u, v, rgb);
                        rgb_basis(rgb, eb, ew);
                        for (int k = 0; k < NSPEC; k++)
                            if (illum_at(sv, eb, ew, k) * sv.env_emit[k] != 0.0f) { bb = false; break; }
                    }
                } else {
                    pdfL2 = sphere_light_pdf(sv, sv.light_shape[li]);
                    Lb = sv.light_emit + (size_t)li * NSPEC;
                    for (int k = 0; k < NSPEC; k++) if (Lb[k] != 0.0f) { bb = false; break; }
                }
                if (!bb) {
                    float w = (pdfB == 1.0f) ? 1.0f : (pdfB * pdfB) / (pdfB * pdfB + pdfL2 * pdfL2);
                    sB = w / pdfB;
                    bcos = ws.b_cos[i]; bsc = ws.b_sc[i]; bac = ws.b_ac[i]; bzero = ws.b_zero[i];
                    useB = true;
                }
            }
        }
        if (useL || useB) {
            TexK tkb;
            const TexK* tkp = nullptr;
            if (flags & F_TEXK) {
                const float rgb[3] = {ws.tk[i], ws.tk[n + i], ws.tk[2 * (size_t)n + i]};
                rgb_basis(rgb, tkb.b, tkb.w);
                tkp = &tkb;
            }
            for (int k = 0; k < NSPEC; k++) {
                float Lk = 0.0f;
                if (useL) Lk += (Le[k] * bsdf_val(sv, m, lcos, lsc, lac, lzero, tkp, k)) * sL;
                if (useB) {
                    float li_k = envPick ? illum_at(sv, eb, ew, k) * sv.env_emit[k] : Lb[k];
                    float v = (m.type == MAT_SPECULAR) ? (bzero ? 0.0f * bac : (1.0f * bsc) * bac)
                                                       : bsdf_val(sv, m, bcos, bsc, bac, bzero, tkp, k);
                    Lk += (li_k * v) * sB;
                }
                if (Lk != 0.0f) black = false;
                X += Lk * sv.xyz[k];
                Y += Lk * sv.xyz[NSPEC + k];
                Z += Lk * sv.xyz[2 * NSPEC + k];
            }
            const float dl = 300.0f / 64.0f;
            X *= dl; Y *= dl; Z *= dl;
        }
    }
}

// One workgroup per 32x32 film tile, one lane per pixel: the radiance of the
// pixel's sample of this pass, then the Mitchell splat (CSFilm.cpp:70-122).
// The reference clips a sample's splat to its own tile (CSFilm.cpp:95-96), so
// every film pixel of the tile receives only the samples of the tile pixels
// within two pixels of it; each lane gathers those from LDS in a fixed order
// (deterministic sums, no atomics) and adds them to the film once.
constexpr int TILE_THREADS = 1024;
// Spectral tables the radiance of a sample reads 64 times (per wavelength):
// with TABLES, staged in LDS once per tile (the launch checks that they fit).
constexpr uint32_t SPLAT_MAX_LIGHTS = 16, SPLAT_MAX_SPEC = 8192;
template <bool TABLES>
WPBR_D void splat_body(const RenderLaunch& Lg, uint32_t tile) {
    __shared__ float s_fx[1024], s_fy[1024], s_x[1024], s_y[1024], s_z[1024];
    __shared__ uint32_t s_black[1024];
    __shared__ float t_illum[7 * NSPEC], t_xyz[3 * NSPEC], t_env[NSPEC], t_light[SPLAT_MAX_LIGHTS * NSPEC];
    __shared__ float t_spec[SPLAT_MAX_SPEC];
    RenderLaunch L = Lg;
    if (TABLES) {
        const SceneView& g = Lg.sv;
        for (uint32_t q = threadIdx.x; q < 7 * NSPEC; q += TILE_THREADS) t_illum[q] = g.illum[q];
        for (uint32_t q = threadIdx.x; q < 3 * NSPEC; q += TILE_THREADS) t_xyz[q] = g.xyz[q];
        if (g.has_env)
            for (uint32_t q = threadIdx.x; q < NSPEC; q += TILE_THREADS) t_env[q] = g.env_emit[q];
        for (uint32_t q = threadIdx.x; q < g.n_sphere_lights * NSPEC; q += TILE_THREADS) t_light[q] = g.light_emit[q];
        for (uint32_t q = threadIdx.x; q < g.n_spec_floats; q += TILE_THREADS) t_spec[q] = g.mat_spec[q];
        L.sv.illum = t_illum;
        L.sv.xyz = t_xyz;
        L.sv.env_emit = t_env;
        L.sv.light_emit = t_light;
        L.sv.mat_spec = t_spec;
        __syncthreads();
    }
    const SceneView& sv = L.sv;
    const uint32_t W = sv.width;
    const uint32_t ntx = W / 32;
    const uint32_t tx0 = (tile % ntx) * 32, ty0 = (tile / ntx) * 32;
    const uint32_t n = L.n;
    // secondary rays traced in this pass (statistics)
    if (tile == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(L.ws.qcount + 2), (unsigned long long)*L.ws.qcount);
    const uint32_t lp = threadIdx.x;  // local pixel, row-major in the tile
    const int lx = (int)(lp & 31), ly = (int)(lp >> 5);
    const uint32_t px = tx0 + lx, py = ty0 + ly;
    const uint32_t pix = py * W + px;
    const uint32_t i = pix;  // state index
    {
        float X, Y, Z;
        bool black;
        sample_radiance(L, i, X, Y, Z, black);
        if (L.sample_xyz) {
            size_t o = ((size_t)pix * L.spp + L.spp_index) * 3;
            L.sample_xyz[o] = X; L.sample_xyz[o + 1] = Y; L.sample_xyz[o + 2] = Z;
        }
        s_fx[lp] = L.ws.pfilm[i];
        s_fy[lp] = L.ws.pfilm[n + i];
        s_x[lp] = X; s_y[lp] = Y; s_z[lp] = Z;
        s_black[lp] = black ? 1u : 0u;
    }
    __syncthreads();
    float ax = 0.0f, ay = 0.0f, az = 0.0f, aw = 0.0f;
    const float fpx = (float)(int)px, fpy = (float)(int)py;
    for (int sy = max(ly - 2, 0); sy <= min(ly + 2, 31); sy++) {
        for (int sx = max(lx - 2, 0); sx <= min(lx + 2, 31); sx++) {
            const int q = sy * 32 + sx;
            const float fx = s_fx[q], fy = s_fy[q];
            // the sample's footprint (before the tile clip, which keeps this pixel)
            const float dx = fx - 0.5f, dy = fy - 0.5f;
            const int x0 = (int)ceilf(dx - 2.0f), x1 = (int)floorf(dx + 2.0f);
            const int y0 = (int)ceilf(dy - 2.0f), y1 = (int)floorf(dy + 2.0f);
            if ((int)px < x0 || (int)px > x1 || (int)py < y0 || (int)py > y1) continue;
            const float wy = mitchell1((fpy - fy) / 2.0f);
            const float w = mitchell1((fpx - fx) / 2.0f) * wy;
            if (!s_black[q]) {
                ax += s_x[q] * w;
                ay += s_y[q] * w;
                az += s_z[q] * w;
            }
            aw += w;
        }
    }
    float4* f = reinterpret_cast<float4*>(L.film) + pix;
    float4 v = *f;
    v.x += ax; v.y += ay; v.z += az; v.w += aw;
    *f = v;
}
template <bool TABLES>
__global__ __launch_bounds__(TILE_THREADS) void k_radiance_splat(RenderLaunch L) { splat_body<TABLES>(L, blockIdx.x); }

// The spectral tables of the scene fit the splat's LDS copy.
bool splat_tables_fit_h(const SceneView& sv) {
    return sv.n_sphere_lights <= SPLAT_MAX_LIGHTS && sv.n_spec_floats <= SPLAT_MAX_SPEC;
}

// ---- batched render: the stages of one sample pass for many scenes per launch.
// Block b works for scene map.scene[b] as its block b - map.first[scene]; the
// scene's launch record (scene view, wave state, film) is in a device table, the
// pass index a kernel argument.
#define SEG_BLOCK(map)                                   \
    const uint32_t sc = (map).scene[blockIdx.x];         \
    const uint32_t bid = blockIdx.x - (map).first[sc];

__global__ __launch_bounds__(256) void ks_camera(const RenderLaunch* __restrict__ T, SegMap map, uint32_t j) {
    SEG_BLOCK(map);
    RenderLaunch L = T[sc];
    L.spp_index = j;
    camera_body(L, bid * 256 + threadIdx.x);
}
__global__ __launch_bounds__(256) void ks_shade1(const RenderLaunch* __restrict__ T, SegMap map, uint32_t j) {
    SEG_BLOCK(map);
    RenderLaunch L = T[sc];
    L.spp_index = j;
    shade1_body(L, bid * 256 + threadIdx.x);
}
template <bool TABLES>
__global__ __launch_bounds__(TILE_THREADS) void ks_radiance_splat(const RenderLaunch* __restrict__ T, SegMap map,
                                                                  uint32_t j) {
    SEG_BLOCK(map);
    RenderLaunch L = T[sc];
    L.spp_index = j;
    splat_body<TABLES>(L, bid);
}
// Clears the films and the queue statistics of every scene (a kernel, so that
// the pass loop can be captured in a graph).
__global__ void ks_clear(const RenderLaunch* __restrict__ T, SegMap map) {
    SEG_BLOCK(map);
    const RenderLaunch& L = T[sc];
    const uint32_t i = bid * 256 + threadIdx.x;
    if (i < L.n) reinterpret_cast<float4*>(L.film)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < 16) L.ws.qcount[i] = 0;
}

// ---- film output -----------------------------------------------------------------------
// gammaCorrect (CSFilm.cpp:129-140), the reference's 12.98 slope included.
WPBR_D float gamma_correct(float v) {
    if (v <= 0.0031308f) return 12.98f * v;
    return 1.055f * cr_pow(v, (float)(1.f / 2.4f)) - 0.055f;
}

// JobOutputFilmTitles (CSFilm.cpp:141-170) on the film (sum XYZ w, sum w per
// pixel, row-major): black pixels 0; else rgb = xyzToRGB(sum XYZ w) /
// (sum w * 800000), clamped to [0, 1] (NaN to 0), gamma-corrected, * 255,
// truncated to 8 bits; three bytes per pixel.  HBM-bound: 16 B read and 3 B
// written per pixel.
__global__ __launch_bounds__(256) void k_film_rgb8(const float4* __restrict__ film, uint32_t npix,
                                                   uint8_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const float4 f = film[i];
    uint8_t* o = out + 3 * (size_t)i;
    if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) {
        o[0] = o[1] = o[2] = 0;
        return;
    }
    float rgb[3];
    rgb[0] = (f.x * 3.240479f + f.y * -1.537150f) + f.z * -0.498535f;
    rgb[1] = (f.x * -0.969256f + f.y * 1.875991f) + f.z * 0.041556f;
    rgb[2] = (f.x * 0.041556f + f.y * -0.204043f) + f.z * 1.057311f;
    const float d = f.w * 800000.f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float v = rgb[c] / d;
        v = fminf(fmaxf(v, 0.0f), 1.0f);
        v = gamma_correct(v) * 255.0f;
        o[c] = (uint8_t)v;
    }
}

// ---- interactions at the collision boundary (wpbr_scene_collide) ------------------
// (the interaction helpers are above the kernels; this is the boundary kernel)
WPBR_D void put3(float* si, size_t n, size_t i, int f, f3 v) {
    si[(size_t)f * n + i] = v.x;
    si[(size_t)(f + 1) * n + i] = v.y;
    si[(size_t)(f + 2) * n + i] = v.z;
}

// One thread per ray cast; the traversal left the barycentrics of its closest
// triangle hit in rows WPBR_SI_DUDX .. +2 of si (each thread reads its own
// ray's before writing them).
__global__ __launch_bounds__(256) void k_scene_si(SceneView sv, uint32_t n, const float* __restrict__ rays,
                                                  const uint8_t* __restrict__ surf, const float* __restrict__ diff,
                                                  const uint32_t* __restrict__ hit_shape,
                                                  const float* __restrict__ hit_t, float* __restrict__ si) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 o = mk3(rays[i], rays[(size_t)n + i], rays[2 * (size_t)n + i]);
    const f3 d = mk3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]);
    const uint32_t shape = hit_shape[i];
    const float t = hit_t[i];
    const bool want_si = !surf || surf[i];
    float b[3];
    for (int k = 0; k < 3; k++) b[k] = si[(size_t)(WPBR_SI_DUDX + k) * n + i];
    FullSI s = fsi_zero();
    if (shape != 0xFFFFFFFFu) {
        if (!want_si) {
            s.p = add3(o, mul3(d, t));
            s.time = t;
            s.wo = neg3(d);
        } else if (shape < sv.g.n_spheres) {
            s = si_full_sphere(sv, shape, o, d, t);
        } else {
            s = si_full_triangle(sv, shape - sv.g.n_spheres, d, b, t);
        }
        if (diff && want_si) {
            auto g = [&](int k) { return diff[(size_t)k * n + i]; };
            si_differentials(s, mk3(g(0), g(1), g(2)), mk3(g(3), g(4), g(5)), mk3(g(6), g(7), g(8)),
                             mk3(g(9), g(10), g(11)));
        }
    }
    put3(si, n, i, WPBR_SI_P, s.p);
    put3(si, n, i, WPBR_SI_WO, s.wo);
    si[(size_t)WPBR_SI_TIME * n + i] = s.time;
    put3(si, n, i, WPBR_SI_N, s.n);
    si[(size_t)WPBR_SI_UV * n + i] = s.u;
    si[(size_t)(WPBR_SI_UV + 1) * n + i] = s.v;
    put3(si, n, i, WPBR_SI_DPDU, s.dpdu);
    put3(si, n, i, WPBR_SI_DPDV, s.dpdv);
    put3(si, n, i, WPBR_SI_DNDU, s.dndu);
    put3(si, n, i, WPBR_SI_DNDV, s.dndv);
    put3(si, n, i, WPBR_SI_SHADING_N, s.sn);
    put3(si, n, i, WPBR_SI_SHADING_DPDU, s.sdpdu);
    put3(si, n, i, WPBR_SI_SHADING_DPDV, s.sdpdv);
    put3(si, n, i, WPBR_SI_SHADING_DNDU, s.sdndu);
    put3(si, n, i, WPBR_SI_SHADING_DNDV, s.sdndv);
    put3(si, n, i, WPBR_SI_DPDX, s.dpdx);
    put3(si, n, i, WPBR_SI_DPDY, s.dpdy);
    si[(size_t)WPBR_SI_DUDX * n + i] = s.dudx;
    si[(size_t)WPBR_SI_DVDX * n + i] = s.dvdx;
    si[(size_t)WPBR_SI_DUDY * n + i] = s.dudy;
    si[(size_t)WPBR_SI_DVDY * n + i] = s.dvdy;
}

wpbr_status scene_si_launch(const SceneView& sv, uint32_t n, const float* rays, const uint8_t* surf,
                            const float* diff, const uint32_t* hit_shape, const float* hit_t, float* si,
                            hipStream_t st) {
    if (!n) return WPBR_OK;
    hipLaunchKernelGGL(k_scene_si, dim3((n + 255) / 256), dim3(256), 0, st, sv, n, rays, surf, diff, hit_shape,
                       hit_t, si);
    return launch_status();
}

// ---- BSDF Compute / Sample jobs over request queues (wpbr_bsdf_*) ----------------
// One request: the shading frame (rows ss, ts, ns of CBSDFTransform, CBSDF.cpp:12-24),
// wo and wi (world), the 2D sample and the material.  The kind (conductor
// microfacet, dielectric microfacet, specular reflection) is the material's.
// value = f |cos wi| per wavelength, as the BSDF jobs store CSampledBSDFValue.
constexpr int BSDF_KIND_NONE = 3;  // no material (or an index beyond the table)

WPBR_D int bsdf_kind_of(const SceneView& sv, const BsdfArgs& a, uint32_t i) {
    const uint32_t mid = a.material[i];
    return mid >= a.n_materials ? BSDF_KIND_NONE : (int)sv.materials[mid].type;
}

// KIND: the request's kind when known at compile time (a queue launch), -1 to
// branch on it per lane (the fused launch).  BSDF_LANES lanes share a request
// (the per-request terms are recomputed by each): lane sub stores the float4s
// sub, sub + 4, sub + 8, sub + 12 of the request's 64-float spectrum (n x 64,
// request-major), so each store instruction covers 64 contiguous bytes of 16
// requests; lane 0 of the group writes pdf and wi.  Every output is stored once.
constexpr int BSDF_LANES = 4;
template <int KIND, bool SAMPLE>
WPBR_D void bsdf_request(const SceneView& sv, const BsdfArgs& a, uint32_t i, int sub) {
    const size_t n = a.n;
    float4* out = reinterpret_cast<float4*>(a.value + (size_t)i * NSPEC) + sub;
    // a queued request keeps its queue's kind; one without material gives zeros
    const int kind = KIND < 0 ? bsdf_kind_of(sv, a, i) : (a.material[i] < a.n_materials ? KIND : BSDF_KIND_NONE);
    float v = 0.0f;      // the value of every wavelength (all but a conductor)
    float pdf = 0.0f;
    f3 wiL = mk3(0.f, 0.f, 0.f), ss = wiL, ts = wiL, ns = wiL;
    bool conductor = false, has_wi = false;
    BsdfTerms bt{};
    DevMaterial m{};
    if (kind != BSDF_KIND_NONE && (SAMPLE || kind != MAT_SPECULAR)) {
        // (Compute of a specular reflection: f = 0, pdf = 0, CBSDFSpecularReflectance.cpp:42-49)
        m = sv.materials[a.material[i]];
        auto g = [&](const float* p, int k) { return p[(size_t)k * n + i]; };
        ss = mk3(g(a.frame, 0), g(a.frame, 1), g(a.frame, 2));
        ts = mk3(g(a.frame, 3), g(a.frame, 4), g(a.frame, 5));
        ns = mk3(g(a.frame, 6), g(a.frame, 7), g(a.frame, 8));
        const f3 wo = mk3(g(a.wo, 0), g(a.wo, 1), g(a.wo, 2));
        const f3 woL = mk3(dot3(ss, wo), dot3(ts, wo), dot3(ns, wo));
        if (kind == MAT_SPECULAR) {
            // Sample: mirror direction, pdf 1, R Fr(cos) / |cos| times |cos| (:26-61, 77-93)
            wiL = mk3(-woL.x, -woL.y, woL.z);
            const float c = abscos(wiL);
            v = (c == 0) ? 0.0f * c : (1.0f * (fresnel_dielectric(wiL.z, 1.0f, m.eta) / c)) * c;
            pdf = 1.0f;
            has_wi = true;
        } else {
            f3 wh;
            float whD;
            bool ok = true;
            if (!SAMPLE) {
                // f (CBSDFConductorMicroface.cpp:44-58): the half vector of wo and wi
                const f3 wi = mk3(g(a.wi, 0), g(a.wi, 1), g(a.wi, 2));
                wiL = mk3(dot3(ss, wi), dot3(ts, wi), dot3(ns, wi));
                wh = normalize3(add3(woL, wiL));
                whD = ggx_d(wh, m.alpha, m.alpha);
            } else {
                // sample_f (:60-82): the sampled half vector (MicrofaceDistr.h:49-85), its
                // reflection; a sample below the surface leaves everything 0
                const float u0 = g(a.u, 0), u1 = g(a.u, 1);
                const float phi = (float)((2 * PI_REF) * (double)u1);
                const float tan2 = m.alpha * m.alpha * u0 / (1.0f - u0);
                const float cosT = 1 / sqrtf(1 + tan2);
                const float sinT = sqrtf(smax(0.f, 1.f - cosT * cosT));
                wh = mk3(sinT * cr_cos(phi), sinT * cr_sin(phi), cosT);
                if (!same_hemi(woL, wh)) wh = neg3(wh);
                whD = ggx_d(wh, m.alpha, m.alpha);
                wiL = add3(neg3(woL), mul3(wh, 2 * dot3(woL, wh)));
                ok = has_wi = same_hemi(wiL, woL);
            }
            if (ok) {
                bt = microfacet_terms(m, woL, wiL, wh, whD);
                pdf = bt.pdf;
                conductor = (KIND == MAT_CONDUCTOR || (KIND < 0 && kind == MAT_CONDUCTOR));
                if (!conductor) v = bt.zero ? 0.0f * bt.ac : (1.0f * bt.sc) * bt.ac;
            }
        }
    }
    if (conductor) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 16 * j + 4 * sub;
            out[4 * j] = make_float4(bsdf_val(sv, m, bt.cos_fr, bt.sc, bt.ac, bt.zero, nullptr, k),
                                     bsdf_val(sv, m, bt.cos_fr, bt.sc, bt.ac, bt.zero, nullptr, k + 1),
                                     bsdf_val(sv, m, bt.cos_fr, bt.sc, bt.ac, bt.zero, nullptr, k + 2),
                                     bsdf_val(sv, m, bt.cos_fr, bt.sc, bt.ac, bt.zero, nullptr, k + 3));
        }
    } else {
        const float4 v4 = make_float4(v, v, v, v);
#pragma unroll
        for (int j = 0; j < 4; j++) out[4 * j] = v4;
    }
    if (sub) return;
    a.pdf[i] = pdf;
    if (SAMPLE) {
        a.wi_out[i] = has_wi ? (ss.x * wiL.x + ts.x * wiL.y) + ns.x * wiL.z : 0.0f;
        a.wi_out[n + i] = has_wi ? (ss.y * wiL.x + ts.y * wiL.y) + ns.y * wiL.z : 0.0f;
        a.wi_out[2 * n + i] = has_wi ? (ss.z * wiL.x + ts.z * wiL.y) + ns.z * wiL.z : 0.0f;
    }
}

// A queue launch (q requests listed in queue, all of kind KIND) or the fused
// launch (KIND -1, queue null: requests 0 .. q - 1).  A wave takes 16 requests:
// lane l serves request 16 w + (l & 15) with wavelength offset l >> 4, so each
// spectral store covers 16 consecutive requests in each of 4 rows.
constexpr uint32_t BSDF_PER_WAVE = 64 / BSDF_LANES;
template <int KIND, bool SAMPLE>
__global__ __launch_bounds__(256) void k_bsdf(SceneView sv, BsdfArgs a, const uint32_t* __restrict__ queue,
                                              uint32_t q) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t t = wave * BSDF_PER_WAVE + (lane % BSDF_PER_WAVE);
    if (t >= q) return;
    const uint32_t i = queue ? queue[t] : t;
    if (i < a.n) bsdf_request<KIND, SAMPLE>(sv, a, i, (int)(lane / BSDF_PER_WAVE));
}

// Per-kind queues in request order (JobGenerateBSDFRequests, CMaterial.h:30-38):
// per 256 requests the counts of each kind (wave ballots), a one-workgroup scan
// into block offsets (kinds one after another), then each request written at its
// kind's offset plus its rank in the block.
constexpr uint32_t BQ_BLOCK = 256;
WPBR_D void bq_ranks(const SceneView& sv, const BsdfArgs& a, uint32_t i, uint32_t* s_cnt /* [4 waves][3] */,
                     int* kind_out, uint32_t* rank_out) {
    const int kind = i < a.n ? bsdf_kind_of(sv, a, i) : BSDF_KIND_NONE;
    const int w = threadIdx.x >> 6;
    uint32_t rank = 0;
    for (int k = 0; k < 3; k++) {
        const uint64_t m = ballot64(kind == k);
        if (kind == k) rank = popc64(m & lanemask_lt());
        if (lane_id() == 0) s_cnt[3 * w + k] = popc64(m);
    }
    *kind_out = kind;
    *rank_out = rank;
}
__global__ __launch_bounds__(BQ_BLOCK) void k_bq_count(SceneView sv, BsdfArgs a, uint32_t* blk) {
    __shared__ uint32_t s_cnt[3 * (BQ_BLOCK / 64)];
    int kind;
    uint32_t rank;
    bq_ranks(sv, a, blockIdx.x * BQ_BLOCK + threadIdx.x, s_cnt, &kind, &rank);
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t c = 0;
        for (int w = 0; w < (int)(BQ_BLOCK / 64); w++) c += s_cnt[3 * w + threadIdx.x];
        blk[3 * (size_t)blockIdx.x + threadIdx.x] = c;
    }
}
// One workgroup: thread t takes a run of consecutive blocks, wave shuffles scan
// the runs' sums, a thread per kind scans the 16 wave totals; each block's entry
// becomes its offset in the kind-major queue and counts[] the kind totals.
__global__ __launch_bounds__(1024) void k_bq_scan(uint32_t* blk, uint32_t nb, uint32_t* counts) {
    __shared__ uint32_t s_w[16][3];
    __shared__ uint32_t s_tot[3];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t b0 = min(nb, tid * per), b1 = min(nb, b0 + per);
    uint32_t loc[3] = {0u, 0u, 0u}, inc[3];
    for (uint32_t b = b0; b < b1; b++)
        for (int k = 0; k < 3; k++) loc[k] += blk[3 * (size_t)b + k];
    for (int k = 0; k < 3; k++) {
        uint32_t v = loc[k];
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t x = __shfl_up(v, off, 64);
            if (lane >= (uint32_t)off) v += x;
        }
        inc[k] = v;
    }
    if (lane == 63)
        for (int k = 0; k < 3; k++) s_w[w][k] = inc[k];
    __syncthreads();
    if (tid < 3) {
        uint32_t run = 0;
        for (int ww = 0; ww < 16; ww++) {
            const uint32_t c = s_w[ww][tid];
            s_w[ww][tid] = run;
            run += c;
        }
        s_tot[tid] = run;
    }
    __syncthreads();
    const uint32_t kind_off[3] = {0u, s_tot[0], s_tot[0] + s_tot[1]};
    uint32_t base[3];
    for (int k = 0; k < 3; k++) base[k] = kind_off[k] + s_w[w][k] + inc[k] - loc[k];
    for (uint32_t b = b0; b < b1; b++)
        for (int k = 0; k < 3; k++) {
            const uint32_t c = blk[3 * (size_t)b + k];
            blk[3 * (size_t)b + k] = base[k];
            base[k] += c;
        }
    if (tid < 3) counts[tid] = s_tot[tid];
}
__global__ __launch_bounds__(BQ_BLOCK) void k_bq_scatter(SceneView sv, BsdfArgs a, const uint32_t* blk,
                                                         uint32_t* queue) {
    __shared__ uint32_t s_cnt[3 * (BQ_BLOCK / 64)];
    const uint32_t i = blockIdx.x * BQ_BLOCK + threadIdx.x;
    int kind;
    uint32_t rank;
    bq_ranks(sv, a, i, s_cnt, &kind, &rank);
    __syncthreads();
    if (kind < 3) {
        uint32_t before = 0;
        for (int w = 0; w < (int)(threadIdx.x >> 6); w++) before += s_cnt[3 * w + kind];
        queue[blk[3 * (size_t)blockIdx.x + kind] + before + rank] = i;
    }
}

wpbr_status bsdf_queues_launch(const SceneView& sv, const BsdfArgs& a, uint32_t* blk, uint32_t* counts,
                               uint32_t* queue, hipStream_t st) {
    const uint32_t nb = (a.n + BQ_BLOCK - 1) / BQ_BLOCK;
    if (!nb) return WPBR_OK;
    hipLaunchKernelGGL(k_bq_count, dim3(nb), dim3(BQ_BLOCK), 0, st, sv, a, blk);
    hipLaunchKernelGGL(k_bq_scan, dim3(1), dim3(1024), 0, st, blk, nb, counts);
    hipLaunchKernelGGL(k_bq_scatter, dim3(nb), dim3(BQ_BLOCK), 0, st, sv, a, blk, queue);
    return launch_status();
}

// kind: MAT_CONDUCTOR / MAT_DIELECTRIC / MAT_SPECULAR over q queued requests, or
// -1: the fused launch over all n.
wpbr_status bsdf_launch(const SceneView& sv, const BsdfArgs& a, int kind, bool sample, const uint32_t* queue,
                        uint32_t q, hipStream_t st) {
    if (!q) return WPBR_OK;
    const dim3 grid((uint32_t)(((uint64_t)q * BSDF_LANES + 255) / 256)), block(256);
#define WPBR_BSDF_LAUNCH(K)                                                                        \
    if (sample) hipLaunchKernelGGL((k_bsdf<K, true>), grid, block, 0, st, sv, a, queue, q);        \
    else hipLaunchKernelGGL((k_bsdf<K, false>), grid, block, 0, st, sv, a, queue, q);
    switch (kind) {
        case MAT_CONDUCTOR: WPBR_BSDF_LAUNCH(MAT_CONDUCTOR) break;
        case MAT_DIELECTRIC: WPBR_BSDF_LAUNCH(MAT_DIELECTRIC) break;
        case MAT_SPECULAR: WPBR_BSDF_LAUNCH(MAT_SPECULAR) break;
        default: WPBR_BSDF_LAUNCH(-1) break;
    }
#undef WPBR_BSDF_LAUNCH
    return launch_status();
}

wpbr_status film_rgb8_launch(const float* film, uint32_t npix, uint8_t* out, hipStream_t st) {
    if (npix) hipLaunchKernelGGL(k_film_rgb8, dim3((npix + 255) / 256), dim3(256), 0, st,
                                 reinterpret_cast<const float4*>(film), npix, out);
    return launch_status();
}

wpbr_status render_batch_launch(const RenderLaunch* T, const SegMap& map, uint32_t grid, int stage, uint32_t j,
                                hipStream_t st) {
    switch (stage) {
        case 0: hipLaunchKernelGGL(ks_camera, dim3(grid), dim3(256), 0, st, T, map, j); break;
        case 1: hipLaunchKernelGGL(ks_shade1, dim3(grid), dim3(256), 0, st, T, map, j); break;
        case 2: hipLaunchKernelGGL(ks_radiance_splat<false>, dim3(grid), dim3(TILE_THREADS), 0, st, T, map, j); break;
        case 4: hipLaunchKernelGGL(ks_radiance_splat<true>, dim3(grid), dim3(TILE_THREADS), 0, st, T, map, j); break;
        case 3: hipLaunchKernelGGL(ks_clear, dim3(grid), dim3(256), 0, st, T, map); break;
    }
    return launch_status();
}

wpbr_status render_pass_launch(RenderLaunch& L, int stage, hipStream_t st) {
    dim3 grid((L.n + 255) / 256), block(256);
    switch (stage) {
        case 0: hipLaunchKernelGGL(k_camera, grid, block, 0, st, L); break;
        case 1: hipLaunchKernelGGL(k_shade1, grid, block, 0, st, L); break;
        case 2:
            if (splat_tables_fit_h(L.sv))
                hipLaunchKernelGGL(k_radiance_splat<true>, dim3((L.sv.width / 32) * (L.sv.height / 32)),
                                   dim3(TILE_THREADS), 0, st, L);
            else
                hipLaunchKernelGGL(k_radiance_splat<false>, dim3((L.sv.width / 32) * (L.sv.height / 32)),
                                   dim3(TILE_THREADS), 0, st, L);
            break;
    }
    return launch_status();
}

}  // namespace wpbr
