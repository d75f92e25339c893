// Oracle: 64-bin sampled spectrum (CSpectrumSampled.h, CSpectrumSampled.cpp,
// CoefficientSpectrum.h).  TEST INFRASTRUCTURE ONLY.
//
//   averageSpectrumSamples  CSpectrumSampled.h:46-89
//   SampledSpectrum::init   CSpectrumSampled.h:236-281 (64 bins over 400-700 nm)
//   fromRGB                 CSpectrumSampled.h:106-209 (Smits, Reflectance/Illuminant)
//   toXYZ / toRGB           CSpectrumSampled.cpp:27-45, xyzToRGB CoefficientSpectrum.h:14-24
//
// Pinned by: the CIE/Smits tables the reference holds (SpectrumData.h), the
// closed form Y(fromRGB(c, c, c)) ~ c * integral(Y) and the round trip
// toRGB(fromRGB(rgb)) ~ rgb (tests/test_oracle_shade.py).
#include "oracle.h"
#include "../woodenpbrengine_amd/csrc/spectrum_tables.inc"

namespace orc {

static float lerpf(float a, float b, float t) { return (1 - t) * a + t * b; }

float average_spectrum_samples(const float* lambda, const float* vals, int n, float l0, float l1) {
    if (l1 <= lambda[0]) return vals[0];
    if (l0 >= lambda[n - 1]) return vals[n - 1];
    if (n == 1) return vals[0];
    float sum = 0;
    if (l0 < lambda[0]) sum += vals[0] * (lambda[0] - l0);
    if (l1 > lambda[n - 1]) sum += vals[n - 1] * (l1 - lambda[n - 1]);
    int j = 0;
    while (l0 > lambda[j + 1]) j++;
    auto interp = [&](float w, int i) {
        return lerpf(vals[i], vals[i + 1], (w - lambda[i]) / (lambda[i + 1] - lambda[i]));
    };
    for (; j + 1 < n && l1 >= lambda[j]; j++) {
        float s0 = std::max(l0, lambda[j]);
        float s1 = std::min(l1, lambda[j + 1]);
        // (a + b) * (s1 - s0) in float, then "* 0.5" and "+=" in double (CSpectrumSampled.h:85-86)
        float prod = (interp(s0, j) + interp(s1, j)) * (s1 - s0);
        sum = (float)((double)sum + (double)prod * 0.5);
    }
    return sum / (l1 - l0);
}

Tables make_tables() {
    Tables t;
    const float* refl[7] = {WPBR_RGBRefl2SpectWhite, WPBR_RGBRefl2SpectCyan, WPBR_RGBRefl2SpectMagenta,
                            WPBR_RGBRefl2SpectYellow, WPBR_RGBRefl2SpectRed, WPBR_RGBRefl2SpectGreen,
                            WPBR_RGBRefl2SpectBlue};
    const float* illum[7] = {WPBR_RGBIllum2SpectWhite, WPBR_RGBIllum2SpectCyan,
                             WPBR_RGBIllum2SpectMagenta, WPBR_RGBIllum2SpectYellow,
                             WPBR_RGBIllum2SpectRed, WPBR_RGBIllum2SpectGreen,
                             WPBR_RGBIllum2SpectBlue};
    for (int i = 0; i < NSPEC; i++) {
        float wl0 = lerpf(400.0f, 700.0f, float(i) / float(NSPEC));
        float wl1 = lerpf(400.0f, 700.0f, float(i + 1) / float(NSPEC));
        t.X[i] = average_spectrum_samples(WPBR_CIE_lambda, WPBR_CIE_X, WPBR_N_CIE, wl0, wl1);
        t.Y[i] = average_spectrum_samples(WPBR_CIE_lambda, WPBR_CIE_Y, WPBR_N_CIE, wl0, wl1);
        t.Z[i] = average_spectrum_samples(WPBR_CIE_lambda, WPBR_CIE_Z, WPBR_N_CIE, wl0, wl1);
        for (int b = 0; b < 7; b++) {
            t.refl[b][i] = average_spectrum_samples(WPBR_RGB2SpectLambda, refl[b], WPBR_N_RGB2SPECT, wl0, wl1);
            t.illum[b][i] = average_spectrum_samples(WPBR_RGB2SpectLambda, illum[b], WPBR_N_RGB2SPECT, wl0, wl1);
        }
    }
    return t;
}

const Tables& tables() {
    static Tables t = make_tables();
    return t;
}

int rgb_case(const float* rgb) {
    float r = rgb[0], g = rgb[1], b = rgb[2];
    if (r <= g && r <= b) return g <= b ? 0 : 1;
    if (g <= r && g <= b) return r <= b ? 2 : 3;
    return r <= g ? 4 : 5;
}

// Basis order: 0 white, 1 cyan, 2 magenta, 3 yellow, 4 red, 5 green, 6 blue.
Spec from_rgb(const float* rgb, bool illuminant) {
    const Tables& T = tables();
    const float(*B)[NSPEC] = illuminant ? T.illum : T.refl;
    Spec s;
    s.fill(0.0f);
    auto add = [&](int b, float k) { for (int i = 0; i < NSPEC; i++) s[i] += B[b][i] * k; };
    float r = rgb[0], g = rgb[1], b = rgb[2];
    if (r <= g && r <= b) {
        add(0, r);
        if (g <= b) { add(1, g - r); add(6, b - g); }
        else { add(1, b - r); add(5, g - b); }
    } else if (g <= r && g <= b) {
        add(0, g);
        if (r <= b) { add(2, r - g); add(6, b - r); }
        else { add(2, b - g); add(4, r - b); }
    } else {
        add(0, b);
        if (r <= g) { add(3, r - b); add(5, g - r); }
        else { add(3, g - b); add(4, r - g); }
    }
    return s;
}

void to_xyz(const Spec& s, float* xyz) {
    const Tables& T = tables();
    float x = 0, y = 0, z = 0;
    for (int i = 0; i < NSPEC; i++) { x += s[i] * T.X[i]; y += s[i] * T.Y[i]; z += s[i] * T.Z[i]; }
    const float dl = (float)(700 - 400) / float(NSPEC);
    xyz[0] = x * dl;
    xyz[1] = y * dl;
    xyz[2] = z * dl;
}

void xyz_to_rgb(const float* xyz, float* rgb) {
    // xyz (row vector) * M with the matrix of CoefficientSpectrum.h:16-21
    rgb[0] = (xyz[0] * 3.240479f + xyz[1] * -1.537150f) + xyz[2] * -0.498535f;
    rgb[1] = (xyz[0] * -0.969256f + xyz[1] * 1.875991f) + xyz[2] * 0.041556f;
    rgb[2] = (xyz[0] * 0.041556f + xyz[1] * -0.204043f) + xyz[2] * 1.057311f;
}

// gammaCorrect, CSFilm.cpp:129-140 (slope 12.98 as written there, not sRGB's 12.92).
static float gamma_correct(float v) {
    if (v <= 0.0031308f) return 12.98f * v;
    return 1.055f * cr_pow(v, (float)(1.f / 2.4f)) - 0.055f;
}

// JobOutputFilmTitles::update, CSFilm.cpp:141-170.  The pixel is black when its
// accumulated XYZ is zero (the reference tests the accumulated spectrum; the XYZ
// of a non-zero spectrum is non-zero because the Y table is positive on every
// bin).  rgb / (w * 800000) is a divide per component; the clamp maps NaN to 0.
void film_rgb8(uint32_t npix, const float* xyz, const float* w, uint8_t* out) {
    for (uint32_t i = 0; i < npix; i++) {
        const float* p = xyz + 3 * (size_t)i;
        uint8_t* o = out + 3 * (size_t)i;
        if (p[0] == 0.0f && p[1] == 0.0f && p[2] == 0.0f) {
            o[0] = o[1] = o[2] = 0;
            continue;
        }
        float rgb[3];
        xyz_to_rgb(p, rgb);
        const float d = w[i] * 800000.f;
        for (int c = 0; c < 3; c++) {
            float v = rgb[c] / d;
            v = std::fmin(std::fmax(v, 0.0f), 1.0f);
            v = gamma_correct(v) * 255.0f;
            o[c] = (uint8_t)v;
        }
    }
}

}  // namespace orc
