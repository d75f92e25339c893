// Oracle: finite-difference reference for the render's parameter gradients.
// TEST INFRASTRUCTURE ONLY.
//
// The float oracle renders once at the base parameters (render.cpp, the
// reference's path) and records every discrete decision of every sample: hit
// shapes, visibility of the shadow and BSDF rays, the light pick, rejected
// samples, the env-map cells and the fromRGB case of each lookup
// (SampleRec).  diff_loss() re-evaluates the same estimator in double
// precision at any parameter values with those decisions frozen, then splats
// (Mitchell, clipped to the tile, CSFilm.cpp:70-122), converts to RGB
// (CoefficientSpectrum.h:14-24) and returns the L2 loss against a target image.
// Central differences of diff_loss are the derivative that a path-replay
// adjoint of this estimator computes (visibility and cell changes excluded).
//
// Parameters: per material alpha, conductor eta/k as RGB (fromRGB(Reflectance),
// CMaterialMetal.cpp:46) and the env map texels (RGB), which enter the env
// light's radiance (CLightInfiniteArea.cpp:72-92, 206-237) and its sampling
// distribution (CLightInfiniteArea.cpp:44-70, pdf in the BSDF strategy's MIS
// weight).
#include <atomic>
#include <cmath>
#include <thread>
#include "oracle.h"

namespace orc {

namespace {

struct D3 {
    double x, y, z;
};
D3 d3(double x, double y, double z) { return {x, y, z}; }
D3 d3f(const float* p) { return {p[0], p[1], p[2]}; }
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
D3 operator-(D3 a) { return {-a.x, -a.y, -a.z}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 normalize(D3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

using SpecD = std::array<double, NSPEC>;

// fromRGB with the ordering case frozen (CSpectrumSampled.h:106-209).
void from_rgb_d(const double* rgb, int cs, const float (*B)[NSPEC], SpecD& out) {
    double r = rgb[0], g = rgb[1], b = rgb[2];
    int b0 = 0, b1, b2;
    double w0, w1, w2;
    switch (cs) {
        case 0: w0 = r; b1 = 1; w1 = g - r; b2 = 6; w2 = b - g; break;
        case 1: w0 = r; b1 = 1; w1 = b - r; b2 = 5; w2 = g - b; break;
        case 2: w0 = g; b1 = 2; w1 = r - g; b2 = 6; w2 = b - r; break;
        case 3: w0 = g; b1 = 2; w1 = b - g; b2 = 4; w2 = r - b; break;
        case 4: w0 = b; b1 = 3; w1 = r - b; b2 = 5; w2 = g - r; break;
        default: w0 = b; b1 = 3; w1 = g - b; b2 = 4; w2 = r - g; break;
    }
    for (int i = 0; i < NSPEC; i++) out[i] = B[b0][i] * w0 + B[b1][i] * w1 + B[b2][i] * w2;
}

double ggx_d_d(D3 wh, double a) {
    double c2 = wh.z * wh.z, s2 = std::max(0.0, 1.0 - c2);
    double t2 = s2 / c2;
    if (std::isinf(t2)) return 0.0;
    double st = std::sqrt(s2);
    double cp = st == 0 ? 1.0 : std::min(1.0, std::max(-1.0, wh.x / st));
    double sp = st == 0 ? 0.0 : std::min(1.0, std::max(-1.0, wh.y / st));
    double e = (cp * cp / (a * a) + sp * sp / (a * a)) * t2;
    return 1.0 / (PI_REF * a * a * c2 * c2 * (1 + e) * (1 + e));
}

double fresnel_dielectric_d(double cosI, double etaI, double etaT) {
    cosI = std::min(1.0, std::max(-1.0, cosI));
    if (cosI <= 0) { std::swap(etaI, etaT); cosI = std::fabs(cosI); }
    double sinI = std::sqrt(std::max(0.0, 1 - cosI * cosI));
    double sinT = etaI / etaT * sinI;
    if (sinT >= 1) return 1.0;
    double cosT = std::sqrt(std::max(0.0, 1 - sinT * sinT));
    double rpa = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
    double rpe = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
    return (rpa * rpa + rpe * rpe) / 2;
}

double fresnel_conductor_d(double cosI, double eta, double k) {
    cosI = std::min(1.0, std::max(-1.0, cosI));
    double c2 = cosI * cosI, s2 = 1 - c2;
    double eta2 = eta * eta, k2 = k * k;
    double t0 = eta2 - k2 - s2;
    double a2pb2 = std::sqrt(t0 * t0 + 4 * eta2 * k2);
    double t1 = a2pb2 + c2;
    double a = std::sqrt(0.5 * (a2pb2 + t0));
    double t2 = 2 * cosI * a;
    double Rs = (t1 - t2) / (t1 + t2);
    double t3 = c2 * a2pb2 + s2 * s2;
    double t4 = t2 * s2;
    double Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5 * (Rp + Rs);
}

}  // namespace

struct DiffState {
    uint32_t W = 0, H = 0, spp = 0, env_w = 0, env_h = 0, n_mat = 0, n_lights = 0;
    bool has_env = false;
    std::vector<uint32_t> mat_type;
    std::vector<float> mat_eta;
    std::vector<int> eta_case, k_case;
    std::vector<float> light_emit, env_emit, env_xform;
    std::vector<SampleRec> recs;
};

DiffState* diff_record(const RenderScene& sc, const LbvhResult& bvh, int nthreads) {
    DiffState* D = new DiffState();
    D->W = sc.width; D->H = sc.height; D->spp = sc.spp;
    D->env_w = sc.env_w; D->env_h = sc.env_h; D->has_env = sc.has_env != 0;
    D->n_mat = sc.n_materials;
    D->n_lights = sc.n_sphere_lights;
    for (uint32_t m = 0; m < sc.n_materials; m++) {
        D->mat_type.push_back(sc.materials[m].type);
        D->mat_eta.push_back(sc.materials[m].eta);
        D->eta_case.push_back(rgb_case(sc.materials[m].eta_rgb));
        D->k_case.push_back(rgb_case(sc.materials[m].k_rgb));
    }
    D->light_emit.assign(sc.light_emit, sc.light_emit + (size_t)sc.n_sphere_lights * NSPEC);
    if (sc.has_env) {
        D->env_emit.assign(sc.env_emit, sc.env_emit + NSPEC);
        D->env_xform.assign(sc.env_xform, sc.env_xform + 24);
    }
    D->recs.resize((size_t)sc.width * sc.height * sc.spp);
    RenderOut out;
    render(sc, bvh, out, nthreads, nullptr, 0, D->recs.data());
    return D;
}

void diff_free(DiffState* D) { delete D; }

double diff_loss(const DiffState& D, const double* alpha, const double* etak, const double* env,
                 const float* target, float* rgb_out, int nthreads) {
    const Tables& T = tables();
    const uint32_t EW = D.env_w, EH = D.env_h;
    // env distribution (CLightInfiniteArea.cpp:44-70) in double: the green
    // channel of a bilinear lookup at (x/W, y/H) times sin(theta), stored at the
    // reference's index x + y * H
    std::vector<double> func((size_t)EW * EH, 0.0);
    double I = 0;
    auto tex = [&](int x, int y, int c) { return env[((size_t)y * EW + x) * 3 + c]; };
    auto bilerp = [&](int s0, int t0, int s0n, int t0n, double ds, double dt, int c) {
        return ((tex(s0, t0, c) * (1 - ds)) * (1 - dt) + (tex(s0, t0n, c) * (1 - ds)) * dt) +
               (tex(s0n, t0, c) * ds) * (1 - dt) + (tex(s0n, t0n, c) * ds) * dt;
    };
    if (D.has_env) {
        for (uint32_t y = 0; y < EH; y++) {
            float sinTheta = (float)std::sin(PI_REF * (double)float(y + 0.5f) / (double)float(EH));
            float t = ((float)y / (float)EH) * (float)EH - 0.5f;
            int t0 = std::min(std::max((int)std::floor(t), 0), (int)EH - 1);
            int t0n = std::max(std::min(t0 + 1, (int)EH - 1), 0);
            double dt = (double)t - t0;
            for (uint32_t x = 0; x < EW; x++) {
                float s = ((float)x / (float)EW) * (float)EW - 0.5f;
                int s0 = std::min(std::max((int)std::floor(s), 0), (int)EW - 1);
                int s0n = std::max(std::min(s0 + 1, (int)EW - 1), 0);
                double ds = (double)s - s0;
                func[x + (size_t)y * EH] = bilerp(s0, t0, s0n, t0n, ds, dt, 1) * (double)sinTheta;
            }
        }
        for (double f : func) I += f;
        I /= (double)EW * EH;
    }
    // env radiance of a direction with the lookup cell frozen; returns theta
    auto env_spec = [&](D3 dirW, const EnvCell& cell, SpecD& L) {
        const float* inv = D.env_xform.data() + 12;
        D3 dl = d3((inv[0] * dirW.x + inv[1] * dirW.y) + inv[2] * dirW.z,
                   (inv[4] * dirW.x + inv[5] * dirW.y) + inv[6] * dirW.z,
                   (inv[8] * dirW.x + inv[9] * dirW.y) + inv[10] * dirW.z);
        double th = std::acos(std::min(1.0, std::max(-1.0, dl.z)));
        double ph = std::atan2(dl.y, dl.x);
        if (ph < 0) ph += 2 * 3.14159265358979323846;
        double u = ph / (2 * PI_REF), v = th / PI_REF;
        double ds = u * EW - 0.5 - cell.s0, dt = v * EH - 0.5 - cell.t0;
        double rgb[3];
        for (int c = 0; c < 3; c++) rgb[c] = bilerp(cell.s0, cell.t0, cell.s0n, cell.t0n, ds, dt, c);
        from_rgb_d(rgb, cell.rgb_case, T.illum, L);
        for (int i = 0; i < NSPEC; i++) L[i] *= D.env_emit[i];
        return th;
    };
    std::vector<SpecD> eta(D.n_mat), kk(D.n_mat);
    for (uint32_t m = 0; m < D.n_mat; m++) {
        from_rgb_d(etak + 6 * (size_t)m, D.eta_case[m], T.refl, eta[m]);
        from_rgb_d(etak + 6 * (size_t)m + 3, D.k_case[m], T.refl, kk[m]);
    }
    const double dl = (700.0 - 400.0) / NSPEC;
    auto eval = [&](const SampleRec& r, double* xyz) {
        SpecD L;
        L.fill(0.0);
        if (r.kind == 0) {
            if (D.has_env) env_spec(d3f(r.dir), r.env, L);
        } else if (r.kind == 2) {
            const uint32_t m = r.mid;
            const uint32_t type = D.mat_type[m];
            const double a = alpha[m];
            const D3 wo = d3f(r.wo);
            auto val = [&](D3 wi, D3 wh, double Dv, SpecD& v) {
                double cO = std::fabs(wo.z), cI = std::fabs(wi.z), cf = dot(wi, wh);
                if (type == MAT_CONDUCTOR) {
                    double s = Dv / (4 * cI * cO);
                    for (int i = 0; i < NSPEC; i++) v[i] = fresnel_conductor_d(cf, eta[m][i], kk[m][i]) * s * cI;
                } else {
                    double s = fresnel_dielectric_d(cf, 1.0, D.mat_eta[m]) * Dv / (4 * cI * cO);
                    for (int i = 0; i < NSPEC; i++) v[i] = s * cI;
                }
            };
            if (r.light_on) {
                D3 wl = d3f(r.wl);
                D3 wh = normalize(wo + wl);
                double Dv = ggx_d_d(wh, a);
                double pdfB = Dv * std::fabs(wh.z) / (4 * dot(wo, wh));
                SpecD v;
                val(wl, wh, Dv, v);
                double pl = r.pdfL, w = pl * pl / (pl * pl + pdfB * pdfB);
                for (int i = 0; i < NSPEC; i++) L[i] += D.light_emit[(size_t)r.li * NSPEC + i] * v[i] * (w / pl);
            }
            if (r.bsdf_on) {
                SpecD v;
                double pdfB;
                D3 wi;
                if (type == MAT_SPECULAR) {
                    wi = d3(-wo.x, -wo.y, wo.z);
                    pdfB = 1.0;
                    double fr = fresnel_dielectric_d(wi.z, 1.0, D.mat_eta[m]);
                    v.fill(fr);
                } else {
                    double phi = 2 * PI_REF * r.ub[1];
                    double tan2 = a * a * r.ub[0] / (1.0 - r.ub[0]);
                    double cT = 1 / std::sqrt(1 + tan2), sT = std::sqrt(std::max(0.0, 1 - cT * cT));
                    D3 wh = d3(sT * std::cos(phi), sT * std::sin(phi), cT);
                    if (!(wo.z * wh.z > 0)) wh = -wh;
                    double Dv = ggx_d_d(wh, a);
                    wi = -wo + wh * (2 * dot(wo, wh));
                    pdfB = Dv * std::fabs(wh.z) / (4 * dot(wo, wh));
                    val(wi, wh, Dv, v);
                }
                D3 ss = d3f(r.ss), ts = d3f(r.ts), ns = d3f(r.ns);
                D3 wiW = d3((ss.x * wi.x + ts.x * wi.y) + ns.x * wi.z, (ss.y * wi.x + ts.y * wi.y) + ns.y * wi.z,
                            (ss.z * wi.x + ts.z * wi.y) + ns.z * wi.z);
                SpecD Li;
                double pdfL2;
                if (r.env_pick) {
                    double th = env_spec(wiW, r.env, Li);
                    pdfL2 = func[(size_t)r.env.iv * EW + r.env.iu] / I / (2 * PI_REF * PI_REF * std::sin(th));
                } else {
                    for (int i = 0; i < NSPEC; i++) Li[i] = D.light_emit[(size_t)r.li * NSPEC + i];
                    pdfL2 = r.pdfL2s;
                }
                double w = (type == MAT_SPECULAR) ? 1.0 : pdfB * pdfB / (pdfB * pdfB + pdfL2 * pdfL2);
                for (int i = 0; i < NSPEC; i++) L[i] += Li[i] * v[i] * (w / pdfB);
            }
        }
        double x = 0, y = 0, z = 0;
        for (int i = 0; i < NSPEC; i++) { x += L[i] * T.X[i]; y += L[i] * T.Y[i]; z += L[i] * T.Z[i]; }
        xyz[0] = x * dl; xyz[1] = y * dl; xyz[2] = z * dl;
    };
    const uint32_t W = D.W, H = D.H, spp = D.spp, TS = 32;
    const uint32_t ntx = (W + TS - 1) / TS, nty = (H + TS - 1) / TS;
    std::vector<double> loss_t((size_t)ntx * nty, 0.0);
    std::atomic<uint32_t> next(0);
    auto worker = [&]() {
        std::vector<double> acc(TS * TS * 4);
        while (true) {
            uint32_t tile = next.fetch_add(1);
            if (tile >= ntx * nty) break;
            const uint32_t tx0 = (tile % ntx) * TS, ty0 = (tile / ntx) * TS;
            const uint32_t tx1 = std::min(tx0 + TS, W), ty1 = std::min(ty0 + TS, H);  // clipped tile
            std::fill(acc.begin(), acc.end(), 0.0);
            for (uint32_t py = ty0; py < ty1; py++)
                for (uint32_t px = tx0; px < tx1; px++)
                    for (uint32_t j = 0; j < spp; j++) {
                        const SampleRec& r = D.recs[((size_t)py * W + px) * spp + j];
                        double xyz[3];
                        eval(r, xyz);
                        const float fx = r.pfilm[0], fy = r.pfilm[1];
                        int x0 = (int)std::ceil(fx - 0.5f - 2.0f), x1 = (int)std::floor(fx - 0.5f + 2.0f);
                        int y0 = (int)std::ceil(fy - 0.5f - 2.0f), y1 = (int)std::floor(fy - 0.5f + 2.0f);
                        x0 = std::max(x0, (int)tx0); y0 = std::max(y0, (int)ty0);
                        x1 = std::min(x1, (int)tx1 - 1); y1 = std::min(y1, (int)ty1 - 1);
                        for (int y = y0; y <= y1; y++)
                            for (int x = x0; x <= x1; x++) {
                                double w = (double)mitchell_weight(((float)x - fx) / 2.0f) *
                                           (double)mitchell_weight(((float)y - fy) / 2.0f);
                                double* a = &acc[4 * ((y - ty0) * TS + (x - tx0))];
                                a[0] += xyz[0] * w; a[1] += xyz[1] * w; a[2] += xyz[2] * w; a[3] += w;
                            }
                    }
            double lsum = 0;
            for (uint32_t q = 0; q < TS * TS; q++) {
                const uint32_t px = tx0 + (q % TS), py = ty0 + (q / TS);
                if (px >= tx1 || py >= ty1) continue;
                const double* a = &acc[4 * q];
                // xyzToRGB (CoefficientSpectrum.h:14-24), then / sum of weights
                double rgb[3] = {((a[0] * 3.240479 + a[1] * -1.537150) + a[2] * -0.498535) / a[3],
                                 ((a[0] * -0.969256 + a[1] * 1.875991) + a[2] * 0.041556) / a[3],
                                 ((a[0] * 0.041556 + a[1] * -0.204043) + a[2] * 1.057311) / a[3]};
                // pixels with a non-finite image or target value are left out of the loss
                bool finite = true;
                for (int c = 0; c < 3; c++) {
                    const size_t o = ((size_t)py * W + px) * 3 + c;
                    finite = finite && std::isfinite(rgb[c]) && (!target || std::isfinite(target[o]));
                }
                for (int c = 0; c < 3; c++) {
                    const size_t o = ((size_t)py * W + px) * 3 + c;
                    if (rgb_out) rgb_out[o] = (float)rgb[c];
                    if (target && finite) {
                        double e = rgb[c] - (double)target[o];
                        lsum += e * e;
                    }
                }
            }
            loss_t[tile] = lsum;
        }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < std::max(1, nthreads); i++) th.emplace_back(worker);
    for (auto& t : th) t.join();
    double loss = 0;
    for (double l : loss_t) loss += l;
    return loss;
}

}  // namespace orc
