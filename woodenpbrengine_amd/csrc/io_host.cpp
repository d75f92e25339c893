// The data formats either side of the render: the per-pixel sample arrays
// (JobSamplerStratifiedGenerateSampels1D/2D, CSSamplerStratified.cpp:8-90) and
// the film output, the tonemap of JobOutputFilmTitles on the GPU and the PNG
// file of JobOutputFilm (CSFilm.cpp:129-226).  The reference encodes with lodepng; this writer stores
// the same 8-bit RGB image in a PNG with uncompressed deflate blocks (any PNG
// reader decodes both to the same pixels).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "render_host.hpp"

using namespace wpbr;

namespace {

uint32_t crc32_of(const uint8_t* p, size_t n, uint32_t crc) {
    // built once (a function-local static's initialisation is thread-safe)
    struct Table { uint32_t t[256]; };
    static const Table table = [] {
        Table tb;
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            tb.t[i] = c;
        }
        return tb;
    }();
    crc = ~crc;
    for (size_t i = 0; i < n; i++) crc = table.t[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put32(out, crc32_of(out.data() + start, out.size() - start, 0));
}

}  // namespace

namespace wpbr {

SamplerDef make_sampler(uint32_t seed, uint32_t mode, uint32_t n1, uint32_t n2) {
    SamplerDef S;
    S.seed = seed;
    S.mode = mode;
    S.n1 = n1;
    S.n2 = n2;
    // CSSamplerStratified.cpp:63-80: int n = sqrt((float)size); the 1D scale and the
    // reference's 2D scale are 1.0 / size rounded to float, the intended 2D one 1.0 / n
    S.root = (uint32_t)(int)std::sqrt((float)n2);
    S.inv1 = (float)(1.0 / n1);
    S.inv2 = mode == WPBR_SAMPLER_STRATIFIED_N ? (float)(1.0 / (int)S.root) : (float)(1.0 / n2);
    return S;
}

}  // namespace wpbr

extern "C" {

wpbr_status wpbr_sample_arrays(wpbr_ctx* ctx, uint32_t seed, uint32_t sampler, uint32_t pixel0, uint32_t npix,
                               uint32_t n1, uint32_t n2, float* d_u1, float* d_u2) {
    if (!ctx) return WPBR_ERR_INVALID_ARG;
    if (sampler > WPBR_SAMPLER_STRATIFIED_N) return set_err(ctx, WPBR_ERR_INVALID_ARG, "unknown sampler mode");
    if ((d_u1 && n1 == 0) || (d_u2 && n2 == 0))
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "empty sample array");
    if ((uint64_t)pixel0 + npix > 0x100000000ull) return set_err(ctx, WPBR_ERR_INVALID_ARG, "pixel range");
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    const SamplerDef S = make_sampler(seed, sampler, n1 ? n1 : 1, n2 ? n2 : 1);
    hipEventRecord(ctx->ev0, ctx->stream);
    wpbr_status s = sample_arrays_launch(S, pixel0, npix, d_u1, d_u2, ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("sample array launch failed: ") + hipGetErrorString(last_launch_error));
    s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sample arrays");
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

wpbr_status wpbr_film_rgb8(wpbr_ctx* ctx, uint32_t width, uint32_t height, const float* d_film,
                           uint8_t* d_rgb8) {
    if (!ctx || !d_film || !d_rgb8) return WPBR_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    (void)hipGetLastError();
    hipEventRecord(ctx->ev0, ctx->stream);
    wpbr_status s = film_rgb8_launch(d_film, width * height, d_rgb8, ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("film output launch failed: ") + hipGetErrorString(last_launch_error));
    s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "film output");
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

wpbr_status wpbr_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* rgb8) {
    if (!path || !rgb8 || width == 0 || height == 0) return WPBR_ERR_INVALID_ARG;
    // scanlines with filter type 0
    const size_t row = 3 * (size_t)width;
    std::vector<uint8_t> raw;
    raw.reserve((row + 1) * height);
    for (uint32_t y = 0; y < height; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb8 + y * row, rgb8 + (y + 1) * row);
    }
    // zlib stream of stored deflate blocks, Adler-32 at the end
    std::vector<uint8_t> z = {0x78, 0x01};
    for (size_t off = 0; off < raw.size() || off == 0; ) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)len);
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len);
        z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        off += len;
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, width);
    put32(ihdr, height);
    ihdr.push_back(8);  // bit depth
    ihdr.push_back(2);  // colour type: RGB
    ihdr.push_back(0);  // deflate
    ihdr.push_back(0);  // adaptive filtering
    ihdr.push_back(0);  // no interlace
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f) return WPBR_ERR_INVALID_ARG;
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    return (std::fclose(f) == 0 && ok) ? WPBR_OK : WPBR_ERR_INVALID_ARG;
}

}  // extern "C"
