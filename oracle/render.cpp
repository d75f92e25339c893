// Oracle: the render loop (MEngine::render, MEngine.cpp:380-448) for direct
// lighting with MIS.  TEST INFRASTRUCTURE ONLY.
//
// Per sample, in the reference's job order:
//   camera sample + ray          CSSampler.cpp:7-24, CSCameraPerspective.cpp:9-52
//   runCollisionSystem           collide.cpp (trace_ray)
//   emitted light for misses     SScattering.cpp:20-80, CLightInfiniteArea.cpp:72-92
//   surface interaction          CSTriangle.cpp:149-319, CSSphere.cpp:122-186, CTransform.h:41-66
//   BSDF requests / transform    CMaterial.h:30-38, CMaterial*.cpp, CBSDF.cpp:12-24
//   light pick                   SScattering.cpp:82-88
//   light sample                 CSSphere.h:187-230 + CSSphere.cpp:248-281,
//                                CLightInfiniteArea.cpp:94-135
//   shadow ray + BSDF compute    SScattering.cpp:99-125, CBSDF*.cpp Compute jobs
//   integrate (light strategy)   SScattering.cpp:127-167
//   BSDF sample + ray            SScattering.cpp:171-206, CBSDF*.cpp Sample jobs
//   light Li compute             CSSphere.h:232-267 + CSSphere.cpp:221-246,
//                                CLightInfiniteArea.cpp:137-168
//   integrate (BSDF strategy)    SScattering.cpp:209-254
//   film splat                   CSFilm.cpp:70-122 (Mitchell r=2, B=C=1/3), clipped to
//                                the sample's 32x32 tile
//
// Choices where the reference is undefined:
//   * light pick index clamped to n-1 (SScattering.cpp:85 overruns for n >= 3);
//   * a rejected microfacet sample gives f = 0, pdf = 0 and casts no BSDF ray
//     (CBSDFConductorMicroface.cpp:75, 99-115 leaves pdf and the ray unset);
//   * a hit on a shape with no material (a light sphere) gets no BSDF and no
//     emission (CMaterial.h:34 is undefined there; emission on hit is disabled);
//   * the env-light sampling strategy stores a black Li (CLightInfiniteArea.cpp:106, 126-131),
//     so its shadow ray cannot contribute and is skipped;
//   * bilinear texel indices are clamped to the texture (CSTextureProcess.h:64-74
//     clamps only below).
#include "oracle.h"
#include <thread>
#include <atomic>
#include <mutex>

namespace orc {

// ---- sample streams -------------------------------------------------------
// Counter-based: value = lowbias32 hash chain of (seed, pixel, slot) -> 24 bits.
static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
static inline float hash_unit(uint32_t seed, uint32_t pixel, uint32_t key) {
    uint32_t h = mix32(seed ^ 0x9E3779B9u);
    h = mix32(h ^ pixel);
    h = mix32(h ^ key);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

Sampler make_sampler(uint32_t seed, uint32_t mode, uint32_t n1, uint32_t n2) {
    Sampler S;
    S.seed = seed;
    S.mode = mode;
    S.n1 = n1;
    S.n2 = n2;
    // CSSamplerStratified.cpp:63-80: `int n = std::sqrt((float)size)`;
    // `float invNSamples = 1.0 / size` (1D, and the inner 2D one that is used),
    // the outer `1.0 / n` for mode 2
    S.root = (uint32_t)(int)std::sqrt((float)n2);
    S.inv1 = (float)(1.0 / n1);
    S.inv2 = mode == 2 ? (float)(1.0 / (int)S.root) : (float)(1.0 / n2);
    return S;
}

// Shuffle: position c of a shuffled array holds element permute_index(c) of the
// array in generation order.  permute_index is a bijection of [0, n) (this
// build's definition; the reference's std::shuffle draws from a random_device):
// a hash permutation of [0, 2^b), b = bits of n - 1 (key xor, odd multipliers and
// right xor-shifts are invertible modulo 2^b), restricted to [0, n) by cycle
// walking, after Kensler's permute() (Correlated Multi-Jittered Sampling, 2013).
uint32_t permute_index(uint32_t c, uint32_t n, uint32_t key) {
    if (n <= 1u) return 0u;
    const uint32_t b = 32u - (uint32_t)__builtin_clz(n - 1u);
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
static uint32_t shuffle_key(uint32_t seed, uint32_t pixel, uint32_t dim) {
    return mix32(mix32(seed ^ (0x632BE5ABu + dim)) ^ pixel);
}

// JobSamplerStratifiedGenerateSampels1D / 2D (CSSamplerStratified.cpp:8-46,
// 49-90) for one pixel, the jitter of element k being the hash stream's value
// at index k (the reference draws uniform(gen) in the same order).
void stratified_arrays(const Sampler& S, uint32_t pixel, float* u1, float* u2) {
    std::vector<float> d1(S.n1);
    for (uint32_t k = 0; k < S.n1; k++)  // (iData + u) * invNSamples, minv(.., 1)
        d1[k] = std::min(((float)k + hash_unit(S.seed, pixel, 3u * k)) * S.inv1, 1.0f);
    std::vector<float> d2(2 * (size_t)S.n2, 0.0f);  // resize(): value-initialised tail
    for (uint32_t i = 0, k = 0; i < S.root; i++) {
        for (uint32_t j = 0; j < S.root; j++, k++) {
            const float dtX = hash_unit(S.seed, pixel, 3u * k + 1u);
            const float dtY = hash_unit(S.seed, pixel, 3u * k + 2u);
            d2[2 * k] = std::min(((float)i + dtX) * S.inv2, 1.0f);
            d2[2 * k + 1] = std::min(((float)j + dtY) * S.inv2, 1.0f);
        }
    }
    const uint32_t key1 = shuffle_key(S.seed, pixel, 1u), key2 = shuffle_key(S.seed, pixel, 2u);
    for (uint32_t c = 0; c < S.n1; c++) u1[c] = d1[permute_index(c, S.n1, key1)];
    for (uint32_t c = 0; c < S.n2; c++) {
        const uint32_t k = permute_index(c, S.n2, key2);
        u2[2 * c] = d2[2 * k];
        u2[2 * c + 1] = d2[2 * k + 1];
    }
}

// Value at position `index` of a pixel's 1D array, one element at a time
// (positions past the end of a stratified array wrap).
float sample_1d(const Sampler& S, uint32_t pixel, uint32_t index) {
    if (S.mode == 0) return hash_unit(S.seed, pixel, 3u * index);
    const uint32_t k = permute_index(index % S.n1, S.n1, shuffle_key(S.seed, pixel, 1u));
    return std::min(((float)k + hash_unit(S.seed, pixel, 3u * k)) * S.inv1, 1.0f);
}
void sample_2d(const Sampler& S, uint32_t pixel, uint32_t index, float* u) {
    if (S.mode == 0) {
        u[0] = hash_unit(S.seed, pixel, 3u * index + 1u);
        u[1] = hash_unit(S.seed, pixel, 3u * index + 2u);
        return;
    }
    const uint32_t k = permute_index(index % S.n2, S.n2, shuffle_key(S.seed, pixel, 2u));
    if (k >= S.root * S.root) {
        u[0] = u[1] = 0.0f;
        return;
    }
    const uint32_t i = k / S.root, j = k % S.root;
    u[0] = std::min(((float)i + hash_unit(S.seed, pixel, 3u * k + 1u)) * S.inv2, 1.0f);
    u[1] = std::min(((float)j + hash_unit(S.seed, pixel, 3u * k + 2u)) * S.inv2, 1.0f);
}

// ---- helpers -------------------------------------------------------------
static Affine load_affine(const float* p) { Affine a; std::memcpy(&a, p, 48); return a; }

static V3 xf_point4(const float* m, V3 p) {
    float x = ((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3];
    float y = ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7];
    float z = ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11];
    float w = ((m[12] * p.x + m[13] * p.y) + m[14] * p.z) + m[15];
    if (w == 1.0f) return v3(x, y, z);
    return v3(x / w, y / w, z / w);
}

// pbrt-v3 SphericalTheta / SphericalPhi (WoodenMath's own pi).
static float spherical_theta(V3 v) {
    float z = v.z < -1.0f ? -1.0f : (v.z > 1.0f ? 1.0f : v.z);
    return cr_acos(z);
}
static float spherical_phi(V3 v) {
    float p = cr_atan2(v.y, v.x);
    return (p < 0) ? (float)((double)p + 2.0 * PI_TRUE) : p;
}
// pbrt-v3 CoordinateSystem ("makeBasisByVector").
static void coordinate_system(V3 v1, V3* v2, V3* v3o) {
    if (std::fabs(v1.x) > std::fabs(v1.y))
        *v2 = v3(-v1.z, 0, v1.x) / std::sqrt(v1.x * v1.x + v1.z * v1.z);
    else
        *v2 = v3(0, v1.z, -v1.y) / std::sqrt(v1.y * v1.y + v1.z * v1.z);
    *v3o = cross3(v1, *v2);
}
// pbrt-v3 UniformConePdf.
static float cone_pdf(float cosThetaMax) { return 1 / (2 * PI_TRUE_F * (1 - cosThetaMax)); }

static inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// ---- env texture, bilinear at mip 0 (CSTextureProcess.h:58-81) ------------
static void env_bilerp(const RenderScene& sc, float s_, float t_, float* rgb) {
    int W = (int)sc.env_w, H = (int)sc.env_h;
    float s = s_ * (float)W - 0.5f;
    float t = t_ * (float)H - 0.5f;
    int s0 = (int)std::floor(s), t0 = (int)std::floor(t);
    s0 = std::max(s0, 0); t0 = std::max(t0, 0);
    s0 = std::min(s0, W - 1); t0 = std::min(t0, H - 1);
    float ds = s - (float)s0, dt = t - (float)t0;
    int s0n = std::max(std::min(s0 + 1, W - 1), 0);
    int t0n = std::max(std::min(t0 + 1, H - 1), 0);
    auto tex = [&](int x, int y, int c) { return sc.env_rgb[((size_t)y * W + x) * 3 + c]; };
    for (int c = 0; c < 3; c++) {
        rgb[c] = ((tex(s0, t0, c) * (1 - ds)) * (1 - dt) + (tex(s0, t0n, c) * (1 - ds)) * dt) +
                 (tex(s0n, t0, c) * ds) * (1 - dt);
        rgb[c] = rgb[c] + (tex(s0n, t0n, c) * ds) * dt;
    }
}

// ---- pbrt-v3 Distribution1D / Distribution2D --------------------------------
struct Dist1D {
    std::vector<float> func, cdf;
    float funcInt = 0;
    void build(const float* f, int n) {
        func.assign(f, f + n);
        cdf.assign(n + 1, 0.0f);
        cdf[0] = 0;
        for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / n;
        funcInt = cdf[n];
        if (funcInt == 0) {
            for (int i = 1; i < n + 1; ++i) cdf[i] = float(i) / float(n);
        } else {
            for (int i = 1; i < n + 1; ++i) cdf[i] /= funcInt;
        }
    }
    int count() const { return (int)func.size(); }
    int find_interval(float u) const {
        int size = (int)cdf.size();
        int first = 0, len = size;
        while (len > 0) {
            int half = len >> 1, middle = first + half;
            if (cdf[middle] <= u) { first = middle + 1; len -= half + 1; }
            else len = half;
        }
        int r = first - 1;
        return r < 0 ? 0 : (r > size - 2 ? size - 2 : r);
    }
    float sample_continuous(float u, float* pdf, int* off) const {
        int offset = find_interval(u);
        if (off) *off = offset;
        float du = u - cdf[offset];
        if ((cdf[offset + 1] - cdf[offset]) > 0) du /= (cdf[offset + 1] - cdf[offset]);
        if (pdf) *pdf = (funcInt > 0) ? func[offset] / funcInt : 0;
        return ((float)offset + du) / (float)count();
    }
};

struct EnvDist {
    std::vector<Dist1D> cond;
    Dist1D marg;
    void sample(const float* u, float* uv, float* pdf) const {
        float pdfs[2];
        int v;
        float d1 = marg.sample_continuous(u[1], &pdfs[1], &v);
        float d0 = cond[v].sample_continuous(u[0], &pdfs[0], nullptr);
        *pdf = pdfs[0] * pdfs[1];
        uv[0] = d0;
        uv[1] = d1;
    }
    float pdf(float u, float v) const {
        int nu = cond[0].count(), nv = marg.count();
        int iu = (int)(u * (float)nu), iv = (int)(v * (float)nv);
        iu = iu < 0 ? 0 : (iu > nu - 1 ? nu - 1 : iu);
        iv = iv < 0 ? 0 : (iv > nv - 1 ? nv - 1 : iv);
        return cond[iv].func[iu] / marg.funcInt;
    }
};

// JobLightInfiniteAreaPreprocessSampling, CLightInfiniteArea.cpp:44-70: the
// green channel of a bilinear lookup at (x/W, y/H) times sin(theta), written to
// data[x + y * H] (the reference's index) and distributed over W x H.
static void build_env_dist(const RenderScene& sc, EnvDist& d) {
    uint32_t W = sc.env_w, H = sc.env_h;
    std::vector<float> data((size_t)W * H, 0.0f);
    for (uint32_t y = 0; y < H; y++) {
        float yN = (float)y / (float)H;
        float sinTheta = (float)std::sin(PI_REF * (double)float(y + 0.5f) / (double)float(H));
        for (uint32_t x = 0; x < W; x++) {
            float xN = (float)x / (float)W;
            float rgb[3];
            env_bilerp(sc, xN, yN, rgb);
            data[x + (size_t)y * H] = rgb[1];
            data[x + (size_t)y * H] *= sinTheta;
        }
    }
    d.cond.resize(H);
    std::vector<float> marg(H);
    for (uint32_t v = 0; v < H; v++) {
        d.cond[v].build(&data[(size_t)v * W], (int)W);
        marg[v] = d.cond[v].funcInt;
    }
    d.marg.build(marg.data(), (int)H);
}

// ---- microfacet + Fresnel (MicrofaceDistr.h:12-90, CBXDF.h:23-62, 200-238) ----
static float cos2theta(V3 w) { return w.z * w.z; }
static float sin2theta(V3 w) { return std::max(0.0f, 1.0f - cos2theta(w)); }
static float sintheta(V3 w) { return std::sqrt(sin2theta(w)); }
static float tan2theta(V3 w) { return sin2theta(w) / cos2theta(w); }
static float cosphi(V3 w) { float s = sintheta(w); return (s == 0) ? 1 : clampf(w.x / s, -1, 1); }
static float sinphi(V3 w) { float s = sintheta(w); return (s == 0) ? 0 : clampf(w.y / s, -1, 1); }
static float abscos(V3 w) { return std::fabs(w.z); }
static bool same_hemi(V3 a, V3 b) { return a.z * b.z > 0; }

static float ggx_d(V3 wh, float ax, float ay) {
    float t2 = tan2theta(wh);
    if (std::isinf(t2)) return 0.;
    const float cos4 = cos2theta(wh) * cos2theta(wh);
    float c2p = cosphi(wh) * cosphi(wh), s2p = sinphi(wh) * sinphi(wh);
    float e = (c2p / (ax * ax) + s2p / (ay * ay)) * t2;
    float one_e = 1 + e;
    double den = PI_REF * (double)ax * (double)ay * (double)cos4 * (double)one_e * (double)one_e;
    return (float)(1 / den);
}

static V3 ggx_sample(V3 wo, const float* u, float ax, float ay, float* whD) {
    float cosT, phi = (float)((2 * PI_REF) * (double)u[1]);
    // alphax == alphay for every material here (MicrofaceDistr.h:58-62)
    float tan2 = ax * ax * u[0] / (1.0f - u[0]);
    cosT = 1 / std::sqrt(1 + tan2);
    float sinT = std::sqrt(std::max(0.f, 1.f - cosT * cosT));
    V3 wh = v3(sinT * cr_cos(phi), sinT * cr_sin(phi), cosT);
    if (!same_hemi(wo, wh)) wh = -wh;
    *whD = ggx_d(wh, ax, ay);
    return wh;
}

static float fresnel_dielectric(float cosI, float etaI, float etaT) {
    cosI = clampf(cosI, -1.0f, 1.0f);
    bool entering = cosI > 0.0f;
    if (!entering) { std::swap(etaI, etaT); cosI = std::fabs(cosI); }
    float sinI = std::sqrt(std::max(0.0f, (float)(1.0 - (double)(cosI * cosI))));
    float sinT = etaI / etaT * sinI;
    if (sinT >= 1.0) return 1.0;
    float cosT = std::sqrt(std::max(0.0f, (float)(1.0 - (double)(sinT * sinT))));
    float Rparl = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
    float Rperp = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
    return (Rparl * Rparl + Rperp * Rperp) / 2;
}

// One wavelength of FresnelConductor::f with etaI = 1 (CBXDF.h:209-234).
static float fresnel_conductor1(float cosI, float etaT, float k) {
    cosI = clampf(cosI, -1.0f, 1.0f);
    float eta = etaT / 1.0f, etak = k / 1.0f;
    float c2 = cosI * cosI;
    float s2 = (float)(1. - (double)c2);
    float eta2 = eta * eta, etak2 = etak * etak;
    float t0 = eta2 - etak2 - s2;
    float a2pb22 = t0 * t0 + eta2 * etak2 * 4;
    float a2pb2 = std::sqrt(a2pb22);
    float t1 = a2pb2 + c2;
    float a = std::sqrt((a2pb2 + t0) * 0.5f);
    float t2 = a * cosI * 2.0f;
    float Rs = (t1 - t2) / (t1 + t2);
    float t3 = a2pb2 * c2 + s2 * s2;
    float t4 = t2 * s2;
    float Rp = Rs * (t3 - t4) / (t3 + t4);
    return (Rp + Rs) * 0.5f;
}

struct MatSpec {
    Spec eta, k;  // conductor
};

// A conductor's spectra: fromRGB(Reflectance) of its eta and k RGB; a textured
// one (CMaterialMetal.cpp:24-52) takes the constant spectrum of its eta (the
// reference's Spectrum(1.5f)).
static MatSpec material_spectra(const RenderScene& sc, uint32_t m) {
    MatSpec s;
    const Material& mt = sc.materials[m];
    if (sc.tex_index(m, 0) >= 0 || sc.tex_index(m, 1) >= 0) s.eta.fill(mt.eta);
    else s.eta = from_rgb(mt.eta_rgb, false);
    s.k = from_rgb(mt.k_rgb, false);
    return s;
}

// MicrofaceDistr.h:18-24.
static float roughness_to_alpha(float roughness) {
    roughness = std::max(roughness, 1e-3f);
    const float x = cr_log(roughness);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

// BSDF value (already multiplied by |cos theta_i|) and pdf for a direction pair
// in the local frame: the Compute jobs (CBSDFConductorMicroface.cpp:84-97,
// CBSDFDielectricMicroface.cpp:84-97).  Returns false when f is exactly 0.
static void bsdf_microfacet_f(const Material& m, const MatSpec& ms, V3 wo, V3 wi, Spec& val,
                              float* pdf) {
    V3 wh = normalize3(wo + wi);
    float whD = ggx_d(wh, m.alpha, m.alpha);
    *pdf = (whD * abscos(wh)) / (4 * dot3(wo, wh));
    float cO = abscos(wo), cI = abscos(wi);
    float ac = abscos(wi);
    if (cI == 0 || cO == 0 || (wh.x == 0 && wh.y == 0 && wh.z == 0)) {
        val.fill(0.0f * ac);
        return;
    }
    if (m.type == MAT_CONDUCTOR) {
        float sc = whD / (4 * cI * cO);
        float c = dot3(wi, wh);
        for (int i = 0; i < NSPEC; i++) val[i] = ((fresnel_conductor1(c, ms.eta[i], ms.k[i]) * 1.0f) * sc) * ac;
    } else {
        float fr = fresnel_dielectric(dot3(wi, wh), 1.0f, m.eta);
        float sc = fr * whD / (4 * cI * cO);
        for (int i = 0; i < NSPEC; i++) val[i] = (1.0f * sc) * ac;
    }
}

// ---- surface interactions ---------------------------------------------------
struct SI {
    V3 p, n, ns, dpdu, wo;
};

// JobProcessTriangleFullInteractionRequests, CSTriangle.cpp:149-319.
static SI si_triangle(const RenderScene& sc, uint32_t tri, V3 rd, const float* b) {
    const float* v = sc.geom.tri_verts + 9 * (size_t)tri;
    V3 P0 = v3(v[0], v[1], v[2]), P1 = v3(v[3], v[4], v[5]), P2 = v3(v[6], v[7], v[8]);
    uint8_t flags = sc.tri_flags ? sc.tri_flags[tri] : 0;
    float uv[3][2] = {{0, 0}, {1, 0}, {1, 1}};
    if (flags & 1) {
        const float* u = sc.tri_uv + 6 * (size_t)tri;
        for (int i = 0; i < 3; i++) { uv[i][0] = u[2 * i]; uv[i][1] = u[2 * i + 1]; }
    }
    V3 dp02 = P0 - P2, dp12 = P1 - P2;
    float duv02[2] = {uv[0][0] - uv[2][0], uv[0][1] - uv[2][1]};
    float duv12[2] = {uv[1][0] - uv[2][0], uv[1][1] - uv[2][1]};
    float det = duv02[0] * duv12[1] - duv02[1] * duv12[0];
    if (det == 0) {
        float d[3][2] = {{0, 0}, {(float)(1.0 * 0.001), 0}, {(float)(1.0 * 0.001), (float)(1.0 * 0.001)}};
        for (int i = 0; i < 3; i++) { uv[i][0] = d[i][0]; uv[i][1] = d[i][1]; }
        duv02[0] = uv[0][0] - uv[2][0]; duv02[1] = uv[0][1] - uv[2][1];
        duv12[0] = uv[1][0] - uv[2][0]; duv12[1] = uv[1][1] - uv[2][1];
        det = duv02[0] * duv12[1] - duv02[1] * duv12[0];
    }
    float detInv = 1 / det;
    V3 dpdu = (dp02 * duv12[1] - dp12 * duv02[1]) * detInv;
    SI s;
    s.p = (P0 * b[0] + P1 * b[1]) + P2 * b[2];
    s.wo = -rd;
    s.dpdu = dpdu;  // shading.dpdu
    V3 ng = normalize3(cross3(dp02, dp12));
    s.n = ng;
    s.ns = ng;
    if (flags & 2) {
        const float* nn = sc.tri_normal + 9 * (size_t)tri;
        const float* tt = sc.tri_tangent + 9 * (size_t)tri;
        V3 N0 = v3(nn[0], nn[1], nn[2]), N1 = v3(nn[3], nn[4], nn[5]), N2 = v3(nn[6], nn[7], nn[8]);
        V3 T0 = v3(tt[0], tt[1], tt[2]), T1 = v3(tt[3], tt[4], tt[5]), T2 = v3(tt[6], tt[7], tt[8]);
        V3 ns = (N0 * b[0] + N1 * b[1]) + N2 * b[2];
        ns = is_zero3(ns) ? s.n : normalize3(ns);
        V3 ss = (T0 * b[0] + T1 * b[1]) + T2 * b[2];
        ss = is_zero3(ss) ? normalize3(dpdu) : normalize3(ss);
        V3 ts = cross3(ss, ns);
        if (length2(ts) > 0) {
            ts = normalize3(ts);
            ss = cross3(ts, ns);
        }
        // setShadingGeometry(ss, ts, dndu, dndv, true), CSufraceInteraction.h:171-196
        s.ns = normalize3(cross3(ss, ts));
        s.dpdu = ss;
        s.n = face_forward(s.n, s.ns);
    }
    return s;
}

// SSphere::computeSurfInteraction + CTransform::operator()(SI), CSSphere.cpp:122-186.
// Vector-by-double-scalar products take the scalar rounded to float.
static SI si_sphere(const RenderScene& sc, uint32_t sid, V3 o, V3 d, float t) {
    const float* xf = sc.geom.sphere_xform + 24 * (size_t)sid;
    Affine W = load_affine(xf), Inv = load_affine(xf + 12);
    float radius = sc.geom.sphere_radius[sid];
    V3 oL = xf_point(Inv, o), dL = xf_vector(Inv, d);
    V3 pH = oL + dL * t;
    float phi = cr_atan2(pH.y, pH.x);
    (void)phi;
    float theta = cr_acos(pH.z / radius);
    float invZR = (float)(1.0 / (double)std::sqrt(pH.x * pH.x + pH.y * pH.y));
    float cphi = pH.x * invZR, sphi = pH.y * invZR;
    V3 dpdu = v3(-pH.y, pH.x, 0.0f) * (float)(2 * PI_REF);
    V3 dpdv = v3(pH.z * cphi, pH.z * sphi, -radius * cr_sin(theta)) * (float)PI_REF;
    V3 n = normalize3(cross3(dpdu, dpdv));
    SI s;
    s.p = xf_point(W, pH);
    s.n = normalize3(xf_normal(Inv, n));
    s.ns = normalize3(xf_normal(Inv, n));
    s.dpdu = xf_vector(W, dpdu);
    s.wo = normalize3(xf_vector(W, -dL));
    return s;
}

// ---- lights -------------------------------------------------------------------
struct LightSample {
    V3 wi;
    float pdf;
    bool black;  // Li is black: no light-strategy contribution
};

// SSphere::sample, CSSphere.cpp:248-281.
static V3 sphere_light_sample(const RenderScene& sc, uint32_t sid, V3 p, const float* u, float* pdf) {
    const float* xf = sc.geom.sphere_xform + 24 * (size_t)sid;
    Affine W = load_affine(xf);
    float radius = sc.geom.sphere_radius[sid];
    V3 pc = xf_point(W, v3(0, 0, 0));
    V3 wc = normalize3(pc - p);
    V3 wcX, wcY;
    coordinate_system(wc, &wcX, &wcY);
    float sinThetaMax2 = radius * radius / length2(p - pc);
    float cosThetaMax = std::sqrt(std::max(0.0f, (float)(1.0 - (double)sinThetaMax2)));
    float cosTheta = (float)((double)u[0] * (1.0 - (double)cosThetaMax) + (double)cosThetaMax);
    float sinTheta = std::sqrt(std::max(0.0f, (float)(1.0 - (double)(cosTheta * cosTheta))));
    float phi = (float)((double)(u[1] * 2) * PI_REF);
    float dc = std::sqrt(length2(p - pc));
    float ds = dc * cosTheta - std::sqrt(std::max(0.0f, radius * radius - dc * dc * sinTheta * sinTheta));
    float cosAlpha = (dc * dc + radius * radius - ds * ds) / (2 * dc * radius);
    float sinAlpha = std::sqrt(std::max(0.0f, 1 - cosAlpha * cosAlpha));
    // pbrt-v3 SphericalDirection(sinA, cosA, phi, -wcX, -wcY, -wc)
    V3 nL = ((-wcX) * (sinAlpha * cr_cos(phi)) + (-wcY) * (sinAlpha * cr_sin(phi))) + (-wc) * cosAlpha;
    V3 pL = nL * radius;
    *pdf = cone_pdf(cosThetaMax);
    return xf_point(W, pL);
}

// SSphere::pdf with the comma operator (CSSphere.cpp:221-246): distance from the origin.
static float sphere_light_pdf(const RenderScene& sc, uint32_t sid) {
    const float* xf = sc.geom.sphere_xform + 24 * (size_t)sid;
    Affine W = load_affine(xf);
    float radius = sc.geom.sphere_radius[sid];
    V3 pc = xf_point(W, v3(0, 0, 0));
    float sinThetaMax2 = radius * radius / length2(pc);
    float cosThetaMax = std::sqrt(std::max(0.0f, 1 - sinThetaMax2));
    return cone_pdf(cosThetaMax);
}

// ---- one sample -----------------------------------------------------------------
struct Ctx {
    const RenderScene& sc;
    const LbvhResult& bvh;
    EnvDist env;
    std::vector<MatSpec> mats;
    Spec env_emit;
    std::vector<Spec> light_emit;
    std::vector<float> tex_filter;
};

static bool trace(const Ctx& C, V3 o, V3 d, uint32_t* shape, float* t) {
    int64_t slot = trace_ray(C.sc.geom, C.bvh, o, d, t, nullptr, nullptr);
    if (slot < 0) { *shape = 0xFFFFFFFFu; return false; }
    *shape = C.bvh.sorted_index[(size_t)slot];
    return true;
}

static Spec env_le(const Ctx& C, V3 dirW, bool* black, float* pdfOut, bool wantPdf,
                   EnvCell* cell = nullptr) {
    const RenderScene& sc = C.sc;
    Affine Inv = load_affine(sc.env_xform + 12);
    V3 dL = xf_vector(Inv, dirW);
    float theta = spherical_theta(dL), phi = spherical_phi(dL);
    float u = (float)((double)phi / (2 * PI_REF)), v = (float)((double)theta / PI_REF);
    if (cell) {
        const int W = (int)sc.env_w, H = (int)sc.env_h;
        int s0 = (int)std::floor(u * (float)W - 0.5f), t0 = (int)std::floor(v * (float)H - 0.5f);
        s0 = std::min(std::max(s0, 0), W - 1);
        t0 = std::min(std::max(t0, 0), H - 1);
        cell->s0 = s0; cell->t0 = t0;
        cell->s0n = std::max(std::min(s0 + 1, W - 1), 0);
        cell->t0n = std::max(std::min(t0 + 1, H - 1), 0);
        int iu = (int)(u * (float)W), iv = (int)(v * (float)H);
        cell->iu = iu < 0 ? 0 : (iu > W - 1 ? W - 1 : iu);
        cell->iv = iv < 0 ? 0 : (iv > H - 1 ? H - 1 : iv);
        float rgbc[3];
        env_bilerp(sc, u, v, rgbc);
        cell->rgb_case = rgb_case(rgbc);
    }
    Spec L;
    L.fill(0.0f);
    *black = true;
    if (wantPdf) {
        float sinT = cr_sin(theta);
        if (sinT == 0.0f) return L;
        *pdfOut = (float)((double)C.env.pdf(u, v) / (2 * PI_REF * PI_REF * (double)sinT));
    }
    float rgb[3];
    env_bilerp(sc, u, v, rgb);
    Spec s = from_rgb(rgb, true);
    for (int i = 0; i < NSPEC; i++) L[i] = s[i] * C.env_emit[i];
    *black = true;
    for (int i = 0; i < NSPEC; i++) if (L[i] != 0.0f) { *black = false; break; }
    return L;
}

static float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }

static Spec run_sample(const Ctx& C, uint32_t px, uint32_t py, uint32_t& c1, uint32_t& c2,
                       float* pFilm, uint64_t* counters, SampleRec* rec) {
    const RenderScene& sc = C.sc;
    const uint32_t pixel = py * sc.width + px;
    Spec L;
    L.fill(0.0f);
    float u2[2];
    sample_2d(sc.smp, pixel, c2++, u2);
    pFilm[0] = (float)px + u2[0];
    pFilm[1] = (float)py + u2[1];
    if (rec) { rec->pfilm[0] = pFilm[0]; rec->pfilm[1] = pFilm[1]; }
    V3 pCam = xf_point4(sc.raster_camera, v3(pFilm[0], pFilm[1], 0.0f));
    V3 dC = normalize3(pCam);
    Affine CW = load_affine(sc.camera_world);
    V3 o = xf_point(CW, v3(0, 0, 0));
    V3 d = normalize3(xf_vector(CW, dC));
    uint32_t shape;
    float t;
    if (!trace(C, o, d, &shape, &t)) {
        counters[1]++;
        if (rec) { rec->kind = 0; rec->dir[0] = d.x; rec->dir[1] = d.y; rec->dir[2] = d.z; }
        if (sc.has_env) {
            bool black;
            L = env_le(C, d, &black, nullptr, false, rec ? &rec->env : nullptr);
        }
        return L;
    }
    if (rec) rec->kind = 1;
    counters[0]++;
    SI si;
    float b[3] = {0.0f, 0.0f, 0.0f};
    if (shape < sc.geom.n_spheres) {
        si = si_sphere(sc, shape, o, d, t);
    } else {
        uint32_t tri = shape - sc.geom.n_spheres;
        tri_intersect(sc.geom.tri_verts + 9 * (size_t)tri, o, d, b);
        si = si_triangle(sc, tri, d, b);
    }
    // light pick (1D), clamped
    const uint32_t nLights = sc.n_sphere_lights + (sc.has_env ? 1u : 0u);
    float u1 = sample_1d(sc.smp, pixel, c1++);
    uint32_t li = (uint32_t)((u1 + 0.5f) * (float)(nLights - 1));
    if (li > nLights - 1) { li = nLights - 1; counters[5]++; }
    bool envPick = sc.has_env && li == sc.n_sphere_lights;
    if (envPick) counters[6]++;
    // light sample (2D)
    float ul[2];
    sample_2d(sc.smp, pixel, c2++, ul);
    V3 wiL = v3(0, 0, 0);
    float pdfL = 0;
    if (!envPick) {
        V3 sp = sphere_light_sample(sc, sc.light_shape[li], si.p, ul, &pdfL);
        wiL = normalize3(sp - si.p);
    } else {
        float uv[2], mapPdf;
        C.env.sample(ul, uv, &mapPdf);  // wi/pdf unused: Li is black for this strategy
    }
    uint32_t mid = sc.shape_material ? sc.shape_material[shape] : 0xFFFFFFFFu;
    if (mid == 0xFFFFFFFFu) return L;
    if (rec) { rec->kind = 2; rec->mid = mid; rec->li = li; rec->env_pick = envPick; rec->pdfL = pdfL; }
    Material m = sc.materials[mid];
    MatSpec ms = C.mats[mid];
    if (m.type == MAT_SPECULAR) counters[4]++;
    const int32_t texK = sc.tex_index(mid, 0), texR = sc.tex_index(mid, 1);
    if (texK >= 0 || texR >= 0) {
        // a textured metal (JobGenerateBSDFMaterialMetal, CMaterialMetal.cpp:24-52): both
        // textures at the hit's mapped point, EWA with maxAnisotropy 32
        float q[6];
        camera_hit_texture_point(sc, shape, o, d, t, b, pFilm, q);
        if (texR >= 0) {
            const TexRGB r = texture_eval(q, sc.textures[texR], C.tex_filter);
            m.alpha = roughness_to_alpha(clampf(r.v[0] * 90, 0.0f, 90.0f));
        }
        if (texK >= 0) {
            const TexRGB k = texture_eval(q, sc.textures[texK], C.tex_filter);
            ms.k = from_rgb(k.v, false);
        }
    }
    // BSDF frame (CBSDF.cpp:12-24)
    V3 ss = normalize3(si.dpdu);
    V3 ns = si.ns;
    V3 ts = cross3(ns, ss);
    auto toLocal = [&](V3 w) { return v3(dot3(ss, w), dot3(ts, w), dot3(ns, w)); };
    auto toWorld = [&](V3 w) {
        return v3((ss.x * w.x + ts.x * w.y) + ns.x * w.z, (ss.y * w.x + ts.y * w.y) + ns.y * w.z,
                  (ss.z * w.x + ts.z * w.y) + ns.z * w.z);
    };
    V3 woL = toLocal(si.wo);
    if (rec) {
        V3 wl = toLocal(wiL);
        for (int c = 0; c < 3; c++) {
            rec->wo[c] = (&woL.x)[c]; rec->ss[c] = (&ss.x)[c]; rec->ts[c] = (&ts.x)[c];
            rec->ns[c] = (&ns.x)[c]; rec->wl[c] = (&wl.x)[c];
        }
    }
    // ---- light strategy (sphere lights only: env Li is black)
    if (!envPick) {
        counters[2]++;
        V3 so = si.p + wiL * 0.1f;
        uint32_t hs;
        float ht;
        bool hit = trace(C, so, wiL, &hs, &ht);
        if (!hit || hs == sc.light_shape[li]) {
            Spec val;
            float pdfB;
            if (m.type == MAT_SPECULAR) {
                val.fill(0.0f);
                pdfB = 0.0f;
            } else {
                bsdf_microfacet_f(m, ms, woL, toLocal(wiL), val, &pdfB);
            }
            const float* Le = &C.light_emit[li][0];
            bool black = true;
            for (int i = 0; i < NSPEC; i++) if (Le[i] != 0.0f) { black = false; break; }
            if (!black && pdfB != 0.0f) {
                if (rec) rec->light_on = 1;
                float w = power_heuristic(pdfL, pdfB);
                float s = w / pdfL;
                for (int i = 0; i < NSPEC; i++) L[i] += (Le[i] * val[i]) * s;
            }
        }
    }
    // ---- BSDF strategy
    Spec valB;
    float pdfB = 0;
    V3 wiBL;
    if (m.type == MAT_SPECULAR) {
        wiBL = v3(-woL.x, -woL.y, woL.z);
        pdfB = 1.0f;
        float c = abscos(wiBL);
        if (c == 0) valB.fill(0.0f * c);
        else {
            float fr = fresnel_dielectric(wiBL.z, 1.0f, m.eta);
            for (int i = 0; i < NSPEC; i++) valB[i] = (1.0f * (fr / c)) * c;
        }
    } else {
        float ub[2];
        sample_2d(sc.smp, pixel, c2++, ub);
        if (rec) { rec->ub[0] = ub[0]; rec->ub[1] = ub[1]; }
        float whD;
        V3 wh = ggx_sample(woL, ub, m.alpha, m.alpha, &whD);
        wiBL = (-woL) + wh * (2 * dot3(woL, wh));
        if (!same_hemi(wiBL, woL)) return L;  // rejected sample
        pdfB = (whD * abscos(wh)) / (4 * dot3(woL, wh));
        float cO = abscos(woL), cI = abscos(wiBL);
        float ac = abscos(wiBL);
        if (cI == 0 || cO == 0 || (wh.x == 0 && wh.y == 0 && wh.z == 0)) {
            valB.fill(0.0f * ac);
        } else if (m.type == MAT_CONDUCTOR) {
            float sc_ = whD / (4 * cI * cO);
            float c = dot3(wiBL, wh);
            for (int i = 0; i < NSPEC; i++) valB[i] = ((fresnel_conductor1(c, ms.eta[i], ms.k[i]) * 1.0f) * sc_) * ac;
        } else {
            float fr = fresnel_dielectric(dot3(wiBL, wh), 1.0f, m.eta);
            float sc_ = fr * whD / (4 * cI * cO);
            for (int i = 0; i < NSPEC; i++) valB[i] = (1.0f * sc_) * ac;
        }
    }
    V3 wiB = toWorld(wiBL);
    counters[3]++;
    uint32_t hs;
    float ht;
    V3 bo = si.p + wiB * 0.01f;
    bool hit = trace(C, bo, wiB, &hs, &ht);
    bool reach = !hit || (!envPick && hs == sc.light_shape[li]);
    if (!reach) return L;
    Spec Li;
    float pdfL2 = 0;
    bool black = true;
    if (envPick) {
        Li = env_le(C, wiB, &black, &pdfL2, true, rec ? &rec->env : nullptr);
    } else {
        pdfL2 = sphere_light_pdf(sc, sc.light_shape[li]);
        for (int i = 0; i < NSPEC; i++) Li[i] = C.light_emit[li][i];
        for (int i = 0; i < NSPEC; i++) if (Li[i] != 0.0f) { black = false; break; }
    }
    if (black) return L;
    if (rec) { rec->bsdf_on = 1; rec->pdfL2s = envPick ? 0.0f : pdfL2; }
    float w = (pdfB == 1.0f) ? 1.0f : power_heuristic(pdfB, pdfL2);
    float s = w / pdfB;
    for (int i = 0; i < NSPEC; i++) L[i] += (Li[i] * valB[i]) * s;
    return L;
}

// Mitchell filter, CSFilm.h:20-44 (radius 2, B = C = 1/3 passed as 1.0/3.0f).
static float mitchell1(float x) {
    const float B = (float)(1.0 / 3.0f), Cc = (float)(1.0 / 3.0f);
    x = std::fabs(2 * x);
    if (x > 1)
        return ((-B - 6 * Cc) * x * x * x + (6 * B + 30 * Cc) * x * x + (-12 * B - 48 * Cc) * x +
                (8 * B + 24 * Cc)) * (1.f / 6.f);
    return ((12 - 9 * B - 6 * Cc) * x * x * x + (-18 + 12 * B + 6 * Cc) * x * x + (6 - 2 * B)) *
           (1.f / 6.f);
}

float mitchell_weight(float x) { return mitchell1(x); }

// The BSDF Compute / Sample jobs for n requests (SoA, world space): the
// microfacet conductor and dielectric families (CBSDFConductorMicroface.cpp:11-115,
// CBSDFDielectricMicroface.cpp:11-115) and the specular reflection
// (CBSDFSpecularReflectance.cpp:26-93), chosen by each request's material, in the
// shading frame of the request (CBSDF.cpp:12-24: rows ss, ts, ns).  value is
// f * |cos wi| per wavelength (n x 64: a request's spectrum contiguous, as the
// dense CSampledBSDFValue components of CBSDF.h:24-33), pdf per request, wi
// (Sample: 3 x n, world).
// A rejected microfacet sample (wi below the surface; the reference leaves pdf and
// wi unset) gives value 0, pdf 0, wi 0; a request without material gives zeros.
void bsdf_eval(const RenderScene& sc, int sample, uint32_t n, const float* frame, const float* wo,
               const float* wi, const float* u, const uint32_t* material, float* value, float* pdf,
               float* wi_out) {
    std::vector<MatSpec> mats(sc.n_materials);
    for (uint32_t m = 0; m < sc.n_materials; m++) mats[m] = material_spectra(sc, m);
    for (uint32_t i = 0; i < n; i++) {
        auto g = [&](const float* a, int k) { return a[(size_t)k * n + i]; };
        for (int k = 0; k < NSPEC; k++) value[(size_t)i * NSPEC + k] = 0.0f;
        pdf[i] = 0.0f;
        if (sample)
            for (int k = 0; k < 3; k++) wi_out[(size_t)k * n + i] = 0.0f;
        const uint32_t mid = material[i];
        if (mid >= sc.n_materials) continue;
        const Material& m = sc.materials[mid];
        const V3 ss = v3(g(frame, 0), g(frame, 1), g(frame, 2));
        const V3 ts = v3(g(frame, 3), g(frame, 4), g(frame, 5));
        const V3 ns = v3(g(frame, 6), g(frame, 7), g(frame, 8));
        auto toLocal = [&](V3 w) { return v3(dot3(ss, w), dot3(ts, w), dot3(ns, w)); };
        const V3 woL = toLocal(v3(g(wo, 0), g(wo, 1), g(wo, 2)));
        Spec val;
        float p = 0.0f;
        V3 wiL = v3(0, 0, 0);
        if (!sample) {
            if (m.type == MAT_SPECULAR) continue;  // Compute: f = 0, pdf = 0 (:42-49)
            wiL = toLocal(v3(g(wi, 0), g(wi, 1), g(wi, 2)));
            bsdf_microfacet_f(m, mats[mid], woL, wiL, val, &p);
        } else if (m.type == MAT_SPECULAR) {
            wiL = v3(-woL.x, -woL.y, woL.z);
            p = 1.0f;
            const float c = abscos(wiL);
            if (c == 0) val.fill(0.0f * c);
            else {
                const float fr = fresnel_dielectric(wiL.z, 1.0f, m.eta);
                for (int k = 0; k < NSPEC; k++) val[k] = (1.0f * (fr / c)) * c;
            }
        } else {
            const float uu[2] = {g(u, 0), g(u, 1)};
            float whD;
            const V3 wh = ggx_sample(woL, uu, m.alpha, m.alpha, &whD);
            wiL = (-woL) + wh * (2 * dot3(woL, wh));
            if (!same_hemi(wiL, woL)) continue;
            // sample_f evaluates f with the sampled half vector and its D
            // (CBSDFConductorMicroface.cpp:60-82, f_impl :20-41)
            p = (whD * abscos(wh)) / (4 * dot3(woL, wh));
            const float cO = abscos(woL), cI = abscos(wiL), ac = abscos(wiL);
            if (cI == 0 || cO == 0 || (wh.x == 0 && wh.y == 0 && wh.z == 0)) {
                val.fill(0.0f * ac);
            } else if (m.type == MAT_CONDUCTOR) {
                const float sc_ = whD / (4 * cI * cO);
                const float c = dot3(wiL, wh);
                for (int k = 0; k < NSPEC; k++)
                    val[k] = ((fresnel_conductor1(c, mats[mid].eta[k], mats[mid].k[k]) * 1.0f) * sc_) * ac;
            } else {
                const float fr = fresnel_dielectric(dot3(wiL, wh), 1.0f, m.eta);
                const float sc_ = fr * whD / (4 * cI * cO);
                for (int k = 0; k < NSPEC; k++) val[k] = (1.0f * sc_) * ac;
            }
        }
        for (int k = 0; k < NSPEC; k++) value[(size_t)i * NSPEC + k] = val[k];
        pdf[i] = p;
        if (sample) {
            const V3 w = v3((ss.x * wiL.x + ts.x * wiL.y) + ns.x * wiL.z, (ss.y * wiL.x + ts.y * wiL.y) + ns.y * wiL.z,
                            (ss.z * wiL.x + ts.z * wiL.y) + ns.z * wiL.z);
            wi_out[i] = w.x;
            wi_out[n + i] = w.y;
            wi_out[2 * (size_t)n + i] = w.z;
        }
    }
}

void render(const RenderScene& sc, const LbvhResult& bvh, RenderOut& out, int nthreads,
            const uint32_t* tile_subset, uint32_t n_tiles_subset, SampleRec* recs) {
    Ctx C{sc, bvh, {}, {}, {}, {}, {}};
    if (sc.has_env) {
        build_env_dist(sc, C.env);
        for (int i = 0; i < NSPEC; i++) C.env_emit[i] = sc.env_emit[i];
    }
    C.mats.resize(sc.n_materials);
    for (uint32_t m = 0; m < sc.n_materials; m++) C.mats[m] = material_spectra(sc, m);
    C.tex_filter = texture_filter_table();
    C.light_emit.resize(sc.n_sphere_lights);
    for (uint32_t l = 0; l < sc.n_sphere_lights; l++)
        for (int i = 0; i < NSPEC; i++) C.light_emit[l][i] = sc.light_emit[(size_t)l * NSPEC + i];

    const uint32_t W = sc.width, H = sc.height, spp = sc.spp;
    const uint32_t TS = 32;
    const uint32_t ntx = (W + TS - 1) / TS, nty = (H + TS - 1) / TS;
    std::vector<uint32_t> tiles;
    if (tile_subset) tiles.assign(tile_subset, tile_subset + n_tiles_subset);
    else for (uint32_t t = 0; t < ntx * nty; t++) tiles.push_back(t);
    out.sample_xyz.assign((size_t)W * H * spp * 3, 0.0f);
    out.sample_pfilm.assign((size_t)W * H * spp * 2, 0.0f);
    out.film_rgb.assign((size_t)W * H * 3, 0.0f);
    out.film_xyz.assign((size_t)W * H * 3, 0.0f);
    out.film_w.assign((size_t)W * H, 0.0f);
    std::atomic<uint32_t> next(0);
    std::mutex cm;
    auto worker = [&]() {
        uint64_t cnt[8] = {0};
        std::vector<Spec> pixL;
        std::vector<float> pixW;
        while (true) {
            uint32_t ti = next.fetch_add(1);
            if (ti >= tiles.size()) break;
            uint32_t tile = tiles[ti];
            uint32_t tx0 = (tile % ntx) * TS, ty0 = (tile / ntx) * TS;
            uint32_t tx1 = std::min(tx0 + TS, W) - 1, ty1 = std::min(ty0 + TS, H) - 1;
            uint32_t tw = tx1 - tx0 + 1, th = ty1 - ty0 + 1;
            pixL.assign((size_t)tw * th, Spec());
            for (auto& s : pixL) s.fill(0.0f);
            pixW.assign((size_t)tw * th, 0.0f);
            std::vector<uint32_t> c1((size_t)tw * th, 0), c2((size_t)tw * th, 0);
            for (uint32_t j = 0; j < spp; j++) {
                for (uint32_t py = ty0; py <= ty1; py++) {
                    for (uint32_t px = tx0; px <= tx1; px++) {
                        size_t lp = (size_t)(py - ty0) * tw + (px - tx0);
                        float pf[2];
                        size_t si_ = ((size_t)(py * W + px) * spp + j);
                        Spec L = run_sample(C, px, py, c1[lp], c2[lp], pf, cnt, recs ? recs + si_ : nullptr);
                        to_xyz(L, &out.sample_xyz[3 * si_]);
                        out.sample_pfilm[2 * si_] = pf[0];
                        out.sample_pfilm[2 * si_ + 1] = pf[1];
                        // splat (CSFilm.cpp:92-117)
                        float dx = pf[0] - 0.5f, dy = pf[1] - 0.5f;
                        int x0 = (int)std::ceil(dx - 2.0f), x1 = (int)std::floor(dx + 2.0f);
                        int y0 = (int)std::ceil(dy - 2.0f), y1 = (int)std::floor(dy + 2.0f);
                        x0 = std::max(x0, (int)tx0); y0 = std::max(y0, (int)ty0);
                        x1 = std::min(x1, (int)tx1); y1 = std::min(y1, (int)ty1);
                        bool black = true;
                        for (int i = 0; i < NSPEC; i++) if (L[i] != 0.0f) { black = false; break; }
                        for (int y = y0; y <= y1; y++) {
                            for (int x = x0; x <= x1; x++) {
                                float w = mitchell1(((float)x - pf[0]) / 2.0f) *
                                          mitchell1(((float)y - pf[1]) / 2.0f);
                                size_t q = (size_t)(y - ty0) * tw + (x - tx0);
                                if (!black)
                                    for (int i = 0; i < NSPEC; i++) pixL[q][i] += L[i] * w;
                                pixW[q] += w;
                            }
                        }
                    }
                }
            }
            for (uint32_t py = ty0; py <= ty1; py++) {
                for (uint32_t px = tx0; px <= tx1; px++) {
                    size_t q = (size_t)(py - ty0) * tw + (px - tx0);
                    size_t o = (size_t)py * W + px;
                    float xyz[3], rgb[3];
                    to_xyz(pixL[q], xyz);
                    xyz_to_rgb(xyz, rgb);
                    for (int c = 0; c < 3; c++) {
                        out.film_xyz[3 * o + c] = xyz[c];
                        out.film_rgb[3 * o + c] = rgb[c] / pixW[q];
                    }
                    out.film_w[o] = pixW[q];
                }
            }
        }
        std::lock_guard<std::mutex> g(cm);
        for (int i = 0; i < 8; i++) out.counters[i] += cnt[i];
    };
    int nt = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int i = 0; i < nt; i++) th.emplace_back(worker);
    for (auto& x : th) x.join();
}

}  // namespace orc

namespace orc {

// One shading function at a time, for the pinning tests (tests/test_oracle_shade.py):
//   0 fresnel_dielectric(cosI, etaI, etaT)       1 fresnel_conductor1(cosI, eta, k)
//   2 ggx_d(wh.xyz, ax, ay)                      3 ggx_sample(wo.xyz, u0, u1, ax, ay) -> wh.xyz, D
//   4 mitchell1(x)                               5 cone_pdf(cosThetaMax)
//   6 coordinate_system(v.xyz) -> v2.xyz, v3.xyz 7 spherical theta, phi of v.xyz
//   8 Dist1D over in[2..2+n) (n = in[0]): sample_continuous(in[1]) -> x, pdf, offset, funcInt
void shade_probe(int fn, const float* in, float* out) {
    switch (fn) {
        case 0: out[0] = fresnel_dielectric(in[0], in[1], in[2]); break;
        case 1: out[0] = fresnel_conductor1(in[0], in[1], in[2]); break;
        case 2: out[0] = ggx_d(v3(in[0], in[1], in[2]), in[3], in[4]); break;
        case 3: {
            float whD;
            V3 wh = ggx_sample(v3(in[0], in[1], in[2]), in + 3, in[5], in[6], &whD);
            out[0] = wh.x; out[1] = wh.y; out[2] = wh.z; out[3] = whD;
            break;
        }
        case 4: out[0] = mitchell1(in[0]); break;
        case 5: out[0] = cone_pdf(in[0]); break;
        case 6: {
            V3 a, b;
            coordinate_system(v3(in[0], in[1], in[2]), &a, &b);
            out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = b.x; out[4] = b.y; out[5] = b.z;
            break;
        }
        case 7: {
            V3 v = v3(in[0], in[1], in[2]);
            out[0] = spherical_theta(v);
            out[1] = spherical_phi(v);
            break;
        }
        case 8: {
            Dist1D d;
            d.build(in + 2, (int)in[0]);
            float pdf;
            int off;
            out[0] = d.sample_continuous(in[1], &pdf, &off);
            out[1] = pdf;
            out[2] = (float)off;
            out[3] = d.funcInt;
            break;
        }
        case 9: out[0] = roughness_to_alpha(in[0]); break;
        default: out[0] = 0.0f;
    }
}

// The texture-mapped points of camera rays through film positions pfilm (2 x n):
// the hit shape (0xFFFFFFFF: miss) and q (6 x n, zeros on a miss), as the
// render's textured materials see them.
void texture_points(const RenderScene& sc, const LbvhResult& bvh, uint32_t n, const float* pfilm, uint32_t* shape,
                    float* q) {
    Ctx C{sc, bvh, {}, {}, {}, {}, {}};
    for (uint32_t i = 0; i < n; i++) {
        const float pf[2] = {pfilm[i], pfilm[n + i]};
        V3 pCam = xf_point4(sc.raster_camera, v3(pf[0], pf[1], 0.0f));
        V3 dC = normalize3(pCam);
        Affine CW = load_affine(sc.camera_world);
        V3 o = xf_point(CW, v3(0, 0, 0));
        V3 d = normalize3(xf_vector(CW, dC));
        float t, qi[6] = {0, 0, 0, 0, 0, 0};
        uint32_t sh;
        if (trace(C, o, d, &sh, &t)) {
            float b[3] = {0.0f, 0.0f, 0.0f};
            if (sh >= sc.geom.n_spheres)
                tri_intersect(sc.geom.tri_verts + 9 * (size_t)(sh - sc.geom.n_spheres), o, d, b);
            camera_hit_texture_point(sc, sh, o, d, t, b, pf, qi);
        }
        shape[i] = sh;
        for (int k = 0; k < 6; k++) q[(size_t)k * n + i] = qi[k];
    }
}

}  // namespace orc
