// Scene ingestion, host side: the OBJ mesh reader standing in for
// MeshTools::loadMesh (MeshTools.h:14-99, assimp with aiProcess_Triangulate |
// aiProcess_ConvertToLeftHanded), the PNG decoder standing in for
// lodepng_decode24_file in readImagePNG (CSTextureProcess.h:462-496), and the
// C ABI of the device bake (STriangleMesh::create, CSTriangleMesh.cpp:11-39)
// and texel decode (SLoadTexture::load, CSTextureProcess.h:309-345).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "wpbr_ctx.hpp"

using namespace wpbr;

namespace wpbr {
wpbr_status bake_launch(const float* xform, uint32_t n_vertices, const float* pos, const float* nrm,
                        const float* tan, uint32_t n_triangles, const uint32_t* idx, float* pos_w,
                        float* nrm_w, float* tan_w, float* tri, uint32_t* bad, hipStream_t st);
wpbr_status srgb8_launch(size_t n, const uint8_t* in, float* out, int num_cus, hipStream_t st);
}  // namespace wpbr

namespace {

bool read_file(const char* path, std::vector<uint8_t>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + got);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

// ---------------------------------------------------------------- inflate --
// RFC 1951 (stored, fixed-Huffman and dynamic-Huffman blocks) inside the RFC
// 1950 zlib wrapper, Adler-32 checked.

struct BitReader {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t buf = 0;
    int cnt = 0;
    bool bad = false;
    uint32_t bits(int need) {
        uint32_t v = buf;
        while (cnt < need) {
            if (pos >= n) {
                bad = true;
                return 0;
            }
            v |= (uint32_t)p[pos++] << cnt;
            cnt += 8;
        }
        buf = need == 32 ? 0 : v >> need;
        cnt -= need;
        return need == 32 ? v : v & ((1u << need) - 1u);
    }
};

// Canonical Huffman code: number of codes per length, symbols ordered by
// (length, value).
struct Huffman {
    uint16_t count[16];
    uint16_t symbol[320];
};

// Returns false for an over-subscribed set of lengths.
bool huff_build(Huffman& h, const uint8_t* len, int n) {
    std::memset(h.count, 0, sizeof h.count);
    for (int s = 0; s < n; s++) h.count[len[s]]++;
    h.count[0] = 0;
    int left = 1;
    for (int l = 1; l < 16; l++) {
        left = 2 * left - h.count[l];
        if (left < 0) return false;
    }
    uint16_t offs[16];
    offs[1] = 0;
    for (int l = 1; l < 15; l++) offs[l + 1] = (uint16_t)(offs[l] + h.count[l]);
    for (int s = 0; s < n; s++)
        if (len[s]) h.symbol[offs[len[s]]++] = (uint16_t)s;
    return true;
}

// Codes are read most-significant bit first, one bit at a time: the first
// code of length l is `first`, the codes of that length are consecutive.
int huff_decode(BitReader& br, const Huffman& h) {
    int code = 0, first = 0, index = 0;
    for (int l = 1; l < 16; l++) {
        code |= (int)br.bits(1);
        if (br.bad) return -1;
        const int c = h.count[l];
        if (code - first < c) return h.symbol[index + code - first];
        index += c;
        first = (first + c) << 1;
        code <<= 1;
    }
    return -1;
}

// Length and distance bases and extra bits (RFC 1951 section 3.2.5).
const uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                               31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

bool inflate_codes(BitReader& br, const Huffman& lit, const Huffman& dist, std::vector<uint8_t>& out) {
    for (;;) {
        int s = huff_decode(br, lit);
        if (s < 0) return false;
        if (s < 256) {
            out.push_back((uint8_t)s);
        } else if (s == 256) {
            return true;
        } else {
            s -= 257;
            if (s >= 29) return false;
            const size_t len = kLenBase[s] + br.bits(kLenExtra[s]);
            const int d = huff_decode(br, dist);
            if (d < 0 || d >= 30) return false;
            const size_t back = kDistBase[d] + br.bits(kDistExtra[d]);
            if (br.bad || back > out.size()) return false;
            const size_t from = out.size() - back;
            for (size_t k = 0; k < len; k++) out.push_back(out[from + k]);
        }
    }
}

bool inflate(const uint8_t* src, size_t n, std::vector<uint8_t>& out) {
    BitReader br{src, n};
    int last;
    do {
        last = (int)br.bits(1);
        const uint32_t type = br.bits(2);
        if (br.bad) return false;
        if (type == 0) {
            br.buf = 0;
            br.cnt = 0;
            if (br.pos + 4 > n) return false;
            const uint32_t len = src[br.pos] | (uint32_t)src[br.pos + 1] << 8;
            const uint32_t nlen = src[br.pos + 2] | (uint32_t)src[br.pos + 3] << 8;
            br.pos += 4;
            if ((len ^ 0xFFFFu) != nlen || br.pos + len > n) return false;
            out.insert(out.end(), src + br.pos, src + br.pos + len);
            br.pos += len;
        } else if (type == 1) {
            // the fixed codes (RFC 1951 3.2.6), built once; a function-local static's
            // initialisation is thread-safe
            struct Fixed { Huffman lit, dist; };
            static const Fixed fixed = [] {
                Fixed f;
                uint8_t l[288];
                for (int s = 0; s < 288; s++) l[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
                huff_build(f.lit, l, 288);
                for (int s = 0; s < 30; s++) l[s] = 5;
                huff_build(f.dist, l, 30);
                return f;
            }();
            if (!inflate_codes(br, fixed.lit, fixed.dist, out)) return false;
        } else if (type == 2) {
            const int nlit = (int)br.bits(5) + 257, ndist = (int)br.bits(5) + 1, ncode = (int)br.bits(4) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t lens[320] = {0};
            for (int k = 0; k < ncode; k++) lens[order[k]] = (uint8_t)br.bits(3);
            if (br.bad || nlit > 286 || ndist > 30) return false;
            Huffman cl, lit, dist;
            if (!huff_build(cl, lens, 19)) return false;
            std::memset(lens, 0, sizeof lens);
            int k = 0;
            while (k < nlit + ndist) {
                const int s = huff_decode(br, cl);
                if (s < 0) return false;
                if (s < 16) {
                    lens[k++] = (uint8_t)s;
                    continue;
                }
                uint8_t v = 0;
                int rep;
                if (s == 16) {
                    if (k == 0) return false;
                    v = lens[k - 1];
                    rep = 3 + (int)br.bits(2);
                } else if (s == 17) {
                    rep = 3 + (int)br.bits(3);
                } else {
                    rep = 11 + (int)br.bits(7);
                }
                if (br.bad || k + rep > nlit + ndist) return false;
                while (rep--) lens[k++] = v;
            }
            if (lens[256] == 0) return false;
            if (!huff_build(lit, lens, nlit) || !huff_build(dist, lens + nlit, ndist)) return false;
            if (!inflate_codes(br, lit, dist, out)) return false;
        } else {
            return false;
        }
    } while (!last);
    return true;
}

bool zlib_decompress(const std::vector<uint8_t>& z, std::vector<uint8_t>& out) {
    if (z.size() < 6) return false;
    const uint32_t cmf = z[0], flg = z[1];
    if ((cmf & 15) != 8 || (cmf >> 4) > 7 || (cmf * 256 + flg) % 31 != 0 || (flg & 32)) return false;
    if (!inflate(z.data() + 2, z.size() - 6, out)) return false;
    uint32_t a = 1, b = 0;
    for (uint8_t c : out) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    const size_t e = z.size() - 4;
    const uint32_t want = (uint32_t)z[e] << 24 | (uint32_t)z[e + 1] << 16 | (uint32_t)z[e + 2] << 8 | z[e + 3];
    return want == ((b << 16) | a);
}

// -------------------------------------------------------------------- PNG --

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

uint32_t crc32_bytes(const uint8_t* p, size_t n) {
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
    uint32_t crc = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) crc = table.t[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

struct PngInfo {
    uint32_t w = 0, h = 0;
    int depth = 0, color = 0, interlace = 0, channels = 0;
    std::vector<uint8_t> palette;  // RGB triples
    std::vector<uint8_t> idat;
};

int png_channels(int color) {
    switch (color) {
        case 0: return 1;
        case 2: return 3;
        case 3: return 1;
        case 4: return 2;
        case 6: return 4;
        default: return 0;
    }
}

bool png_depth_ok(int color, int depth) {
    switch (color) {
        case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
        case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
        case 2: case 4: case 6: return depth == 8 || depth == 16;
        default: return false;
    }
}

// Chunk walk: signature, IHDR first, CRC of every chunk checked, IDAT
// concatenated, unknown critical chunks rejected, ancillary ones skipped.
bool png_parse(const std::vector<uint8_t>& f, PngInfo& pi, bool header_only) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 + 25 || std::memcmp(f.data(), sig, 8) != 0) return false;
    size_t p = 8;
    bool have_ihdr = false;
    while (p + 12 <= f.size()) {
        const uint32_t len = be32(&f[p]);
        if (len > f.size() - p - 12) return false;
        const uint8_t* type = &f[p + 4];
        const uint8_t* data = &f[p + 8];
        if (crc32_bytes(type, (size_t)len + 4) != be32(data + len)) return false;
        const std::string t((const char*)type, 4);
        if (!have_ihdr) {
            if (t != "IHDR" || len != 13) return false;
            pi.w = be32(data);
            pi.h = be32(data + 4);
            pi.depth = data[8];
            pi.color = data[9];
            pi.interlace = data[12];
            pi.channels = png_channels(pi.color);
            if (!pi.w || !pi.h || !png_depth_ok(pi.color, pi.depth) || data[10] != 0 || data[11] != 0 ||
                pi.interlace > 1)
                return false;
            have_ihdr = true;
            if (header_only) return true;
        } else if (t == "PLTE") {
            if (len % 3 || len == 0 || len > 768) return false;
            pi.palette.assign(data, data + len);
        } else if (t == "IDAT") {
            pi.idat.insert(pi.idat.end(), data, data + len);
        } else if (t == "IEND") {
            return !(pi.color == 3 && pi.palette.empty());
        } else if (!(type[0] & 32)) {
            return false;
        }
        p += (size_t)len + 12;
    }
    return false;
}

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    return (uint8_t)(pb <= pc ? b : c);
}

// Reverses the scanline filters of a w x h image starting at src[0] (each row
// one filter byte then `rowbytes` bytes) into dst; returns the bytes consumed.
bool unfilter(const uint8_t* src, size_t avail, uint32_t w, uint32_t h, const PngInfo& pi, std::vector<uint8_t>& dst,
              size_t& used) {
    const size_t bits = (size_t)pi.channels * pi.depth;
    const size_t rowbytes = (w * bits + 7) / 8;
    const size_t bpp = bits >= 8 ? bits / 8 : 1;
    used = (rowbytes + 1) * h;
    if (used > avail) return false;
    dst.assign(rowbytes * h, 0);
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t ft = src[y * (rowbytes + 1)];
        const uint8_t* in = src + y * (rowbytes + 1) + 1;
        uint8_t* cur = &dst[y * rowbytes];
        const uint8_t* prev = y ? &dst[(y - 1) * rowbytes] : nullptr;
        for (size_t i = 0; i < rowbytes; i++) {
            const int a = i >= bpp ? cur[i - bpp] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int v;
            switch (ft) {
                case 0: v = in[i]; break;
                case 1: v = in[i] + a; break;
                case 2: v = in[i] + b; break;
                case 3: v = in[i] + ((a + b) >> 1); break;
                case 4: v = in[i] + paeth(a, b, c); break;
                default: return false;
            }
            cur[i] = (uint8_t)v;
        }
    }
    return true;
}

// One channel of pixel x of an unfiltered row: the high byte of 16-bit
// samples, sub-byte samples most significant first.
uint32_t sample_at(const uint8_t* row, uint32_t x, int c, const PngInfo& pi) {
    const size_t k = (size_t)x * pi.channels + c;
    if (pi.depth == 16) return row[2 * k];
    if (pi.depth == 8) return row[k];
    const size_t bit = k * pi.depth;
    return (row[bit >> 3] >> (8 - pi.depth - (bit & 7))) & ((1u << pi.depth) - 1u);
}

// lodepng's conversion to 8-bit RGB: grey replicated (sub-byte grey scaled by
// 255 / (2^depth - 1)), palette looked up (an index beyond the palette gives
// black: lodepng fills the unused entries with black, lodepng.cpp:2607-2620, and
// does not check the index, :3378-3384), alpha dropped.
bool to_rgb(const uint8_t* row, uint32_t x, const PngInfo& pi, uint8_t* o) {
    switch (pi.color) {
        case 0: case 4: {
            uint32_t g = sample_at(row, x, 0, pi);
            if (pi.depth < 8) g = g * 255u / ((1u << pi.depth) - 1u);
            o[0] = o[1] = o[2] = (uint8_t)g;
            return true;
        }
        case 3: {
            const uint32_t i = sample_at(row, x, 0, pi);
            if (3 * (size_t)i + 2 >= pi.palette.size()) o[0] = o[1] = o[2] = 0;
            else std::memcpy(o, &pi.palette[3 * i], 3);
            return true;
        }
        default:
            for (int c = 0; c < 3; c++) o[c] = (uint8_t)sample_at(row, x, c, pi);
            return true;
    }
}

bool png_decode(const PngInfo& pi, uint8_t* rgb) {
    std::vector<uint8_t> raw;
    if (!zlib_decompress(pi.idat, raw)) return false;
    const size_t rowbytes_full = ((size_t)pi.w * pi.channels * pi.depth + 7) / 8;
    std::vector<uint8_t> img;
    size_t used = 0;
    if (!pi.interlace) {
        if (!unfilter(raw.data(), raw.size(), pi.w, pi.h, pi, img, used)) return false;
        for (uint32_t y = 0; y < pi.h; y++)
            for (uint32_t x = 0; x < pi.w; x++)
                if (!to_rgb(&img[y * rowbytes_full], x, pi, rgb + 3 * ((size_t)y * pi.w + x))) return false;
        return true;
    }
    // Adam7: seven reduced images, each filtered on its own, stored one after another.
    static const uint32_t x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const uint32_t dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
    size_t at = 0;
    for (int pass = 0; pass < 7; pass++) {
        const uint32_t pw = pi.w > x0[pass] ? (pi.w - x0[pass] + dx[pass] - 1) / dx[pass] : 0;
        const uint32_t ph = pi.h > y0[pass] ? (pi.h - y0[pass] + dy[pass] - 1) / dy[pass] : 0;
        if (!pw || !ph) continue;
        if (!unfilter(raw.data() + at, raw.size() - at, pw, ph, pi, img, used)) return false;
        at += used;
        const size_t rb = ((size_t)pw * pi.channels * pi.depth + 7) / 8;
        for (uint32_t y = 0; y < ph; y++)
            for (uint32_t x = 0; x < pw; x++) {
                const size_t X = x0[pass] + (size_t)x * dx[pass], Y = y0[pass] + (size_t)y * dy[pass];
                if (!to_rgb(&img[y * rb], x, pi, rgb + 3 * (Y * pi.w + X))) return false;
            }
    }
    return true;
}

// -------------------------------------------------------------------- OBJ --

struct ObjCorner {
    long v, t, n;  // 1-based after resolution, 0 = absent
};

// Resolves an OBJ index (1-based, negative = relative to the end) against the
// count read so far; 0 on an index out of range.
long obj_index(long i, size_t count) {
    if (i > 0) return (size_t)i <= count ? i : 0;
    if (i < 0) return (size_t)(-i) <= count ? (long)count + 1 + i : 0;
    return 0;
}

}  // namespace

struct wpbr_obj {
    std::vector<float> positions, normals, uvs;
    std::vector<uint32_t> indices;
    bool has_normals = false, has_uvs = false;
};

namespace {

// The first mesh of the file, as assimp hands MeshTools::loadMesh its
// mMeshes[0] (MeshTools.h:27): faces from the start up to the first `o`, `g` or
// `usemtl` line that follows a face.  Every face corner becomes its own vertex
// (no aiProcess_JoinIdenticalVertices).  aiProcess_ConvertToLeftHanded: z of
// positions and normals negated, v of texture coordinates replaced by 1 - v,
// and each face's corner order reversed; then aiProcess_Triangulate fans each
// face from its (reversed) first corner.  The mesh has normals (uvs) only when
// every corner of its faces names one.
bool obj_parse(const std::vector<uint8_t>& text, wpbr_obj& o, std::string& err) {
    std::vector<float> v, vt, vn;
    std::vector<std::vector<ObjCorner>> faces;
    bool all_n = true, all_t = true, ended = false;
    size_t p = 0, line_no = 0;
    std::string line;
    while (p < text.size() && !ended) {
        size_t e = p;
        while (e < text.size() && text[e] != '\n') e++;
        line.assign((const char*)&text[p], e - p);
        p = e + 1;
        line_no++;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const char* s = line.c_str();
        while (*s == ' ' || *s == '\t') s++;
        if (!*s || *s == '#') continue;
        const char* kw_end = s;
        while (*kw_end && *kw_end != ' ' && *kw_end != '\t') kw_end++;
        const std::string kw(s, kw_end - s);
        const char* a = kw_end;
        auto floats = [&](int need, int max, std::vector<float>& dst, float fill) -> bool {
            int got = 0;
            char* end;
            const char* q = a;
            float tmp[4];
            while (got < 4) {
                const float f = std::strtof(q, &end);
                if (end == q) break;
                tmp[got++] = f;
                q = end;
            }
            if (got < need) return false;
            for (int k = 0; k < max; k++) dst.push_back(k < got ? tmp[k] : fill);
            return true;
        };
        if (kw == "v") {
            if (!floats(3, 3, v, 0.0f)) { err = "bad vertex at line " + std::to_string(line_no); return false; }
        } else if (kw == "vn") {
            if (!floats(3, 3, vn, 0.0f)) { err = "bad normal at line " + std::to_string(line_no); return false; }
        } else if (kw == "vt") {
            if (!floats(1, 2, vt, 0.0f)) { err = "bad uv at line " + std::to_string(line_no); return false; }
        } else if (kw == "f") {
            std::vector<ObjCorner> f;
            const char* q = a;
            for (;;) {
                while (*q == ' ' || *q == '\t') q++;
                if (!*q) break;
                char* end;
                ObjCorner c{0, 0, 0};
                c.v = obj_index(std::strtol(q, &end, 10), v.size() / 3);
                if (end == q) { err = "bad face at line " + std::to_string(line_no); return false; }
                q = end;
                if (*q == '/') {
                    q++;
                    if (*q != '/') {
                        c.t = obj_index(std::strtol(q, &end, 10), vt.size() / 2);
                        if (end == q) { err = "bad face at line " + std::to_string(line_no); return false; }
                        q = end;
                    }
                    if (*q == '/') {
                        q++;
                        c.n = obj_index(std::strtol(q, &end, 10), vn.size() / 3);
                        if (end == q) { err = "bad face at line " + std::to_string(line_no); return false; }
                        q = end;
                    }
                }
                if (!c.v) { err = "face index out of range at line " + std::to_string(line_no); return false; }
                all_t = all_t && c.t;
                all_n = all_n && c.n;
                f.push_back(c);
            }
            if (f.size() < 3) continue;  // points and lines carry no triangles
            faces.push_back(std::move(f));
        } else if (kw == "o" || kw == "g" || kw == "usemtl") {
            if (!faces.empty()) ended = true;
        }
    }
    if (faces.empty()) { err = "no faces"; return false; }
    o.has_normals = all_n;
    o.has_uvs = all_t;
    uint32_t base = 0;
    for (const auto& f : faces) {
        const uint32_t m = (uint32_t)f.size();
        for (const auto& c : f) {
            const size_t vi = 3 * (size_t)(c.v - 1);
            o.positions.insert(o.positions.end(), {v[vi], v[vi + 1], -v[vi + 2]});
            if (o.has_normals) {
                const size_t ni = 3 * (size_t)(c.n - 1);
                o.normals.insert(o.normals.end(), {vn[ni], vn[ni + 1], -vn[ni + 2]});
            }
            if (o.has_uvs) {
                const size_t ti = 2 * (size_t)(c.t - 1);
                o.uvs.insert(o.uvs.end(), {vt[ti], 1.0f - vt[ti + 1]});
            }
        }
        // reversed corner order r_k = base + m - 1 - k, fanned from r_0
        for (uint32_t k = 1; k + 1 < m; k++) {
            o.indices.push_back(base + m - 1);
            o.indices.push_back(base + m - 1 - k);
            o.indices.push_back(base + m - 2 - k);
        }
        base += m;
    }
    return true;
}

}  // namespace

extern "C" {

wpbr_status wpbr_obj_load(const char* path, wpbr_obj** out) {
    if (!path || !out) return WPBR_ERR_INVALID_ARG;
    std::vector<uint8_t> text;
    if (!read_file(path, text)) return WPBR_ERR_INVALID_ARG;
    wpbr_obj* o = new wpbr_obj();
    std::string err;
    if (!obj_parse(text, *o, err)) {
        delete o;
        return WPBR_ERR_INVALID_ARG;
    }
    *out = o;
    return WPBR_OK;
}

wpbr_status wpbr_obj_counts(const wpbr_obj* obj, uint32_t counts[4]) {
    if (!obj || !counts) return WPBR_ERR_INVALID_ARG;
    counts[0] = (uint32_t)(obj->positions.size() / 3);
    counts[1] = (uint32_t)obj->indices.size();
    counts[2] = obj->has_normals;
    counts[3] = obj->has_uvs;
    return WPBR_OK;
}

wpbr_status wpbr_obj_copy(const wpbr_obj* obj, float* positions, float* normals, float* uvs, uint32_t* indices) {
    if (!obj) return WPBR_ERR_INVALID_ARG;
    if (positions) std::memcpy(positions, obj->positions.data(), obj->positions.size() * 4);
    if (normals && obj->has_normals) std::memcpy(normals, obj->normals.data(), obj->normals.size() * 4);
    if (uvs && obj->has_uvs) std::memcpy(uvs, obj->uvs.data(), obj->uvs.size() * 4);
    if (indices) std::memcpy(indices, obj->indices.data(), obj->indices.size() * 4);
    return WPBR_OK;
}

void wpbr_obj_free(wpbr_obj* obj) { delete obj; }

wpbr_status wpbr_read_png(const char* path, uint32_t* width, uint32_t* height, uint8_t* rgb8, size_t cap) {
    if (!path || !width || !height) return WPBR_ERR_INVALID_ARG;
    std::vector<uint8_t> f;
    if (!read_file(path, f)) return WPBR_ERR_INVALID_ARG;
    PngInfo pi;
    if (!png_parse(f, pi, rgb8 == nullptr)) return WPBR_ERR_INVALID_ARG;
    *width = pi.w;
    *height = pi.h;
    if (!rgb8) return WPBR_OK;
    if (cap < 3 * (size_t)pi.w * pi.h) return WPBR_ERR_INVALID_ARG;
    return png_decode(pi, rgb8) ? WPBR_OK : WPBR_ERR_INVALID_ARG;
}

wpbr_status wpbr_mesh_bake(wpbr_ctx* ctx, const float* xform, uint32_t n_vertices, const float* d_positions,
                           const float* d_normals, const float* d_tangents, uint32_t n_triangles,
                           const uint32_t* d_indices, float* d_positions_world, float* d_normals_world,
                           float* d_tangents_world, float* d_tri_verts) {
    if (!ctx || !xform || (n_vertices && !d_positions) || (n_triangles && (!d_indices || !d_tri_verts)))
        return WPBR_ERR_INVALID_ARG;
    if ((d_normals_world && !d_normals) || (d_tangents_world && !d_tangents))
        return set_err(ctx, WPBR_ERR_INVALID_ARG, "a world-space output needs its object-space input");
    (void)hipSetDevice(ctx->device);
    if (!ctx->err_flag) {
        wpbr_status s = hip_check(ctx, hipMalloc(&ctx->err_flag, 16), "hipMalloc");
        if (s) return s;
        s = hip_check(ctx, hipMemsetAsync(ctx->err_flag, 0, 16, ctx->stream), "memset");
        if (s) return s;
    }
    uint32_t* bad = ctx->err_flag + 3;  // word 3: this entry point's index check
    (void)hipGetLastError();
    hipEventRecord(ctx->ev0, ctx->stream);
    wpbr_status s = bake_launch(xform, n_vertices, d_positions, d_normals, d_tangents, n_triangles, d_indices,
                                d_positions_world, d_normals_world, d_tangents_world, d_tri_verts, bad,
                                ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("mesh bake launch failed: ") + hipGetErrorString(last_launch_error));
    uint32_t nbad = 0;
    s = hip_check(ctx, hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    if (!s) s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "mesh bake");
    if (s) return s;
    if (hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    if (nbad) {
        hipMemsetAsync(bad, 0, 4, ctx->stream);
        hipStreamSynchronize(ctx->stream);
        return set_err(ctx, WPBR_ERR_INVALID_ARG,
                       std::to_string(nbad) + " triangle indices at or beyond the vertex count");
    }
    return WPBR_OK;
}

wpbr_status wpbr_srgb8_to_linear(wpbr_ctx* ctx, size_t n, const uint8_t* d_rgb8, float* d_rgb) {
    if (!ctx || (n && (!d_rgb8 || !d_rgb))) return WPBR_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipGetLastError();
    hipEventRecord(ctx->ev0, ctx->stream);
    wpbr_status s = srgb8_launch(n, d_rgb8, d_rgb, ctx->num_cus, ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (s) return set_err(ctx, s, std::string("texel decode launch failed: ") + hipGetErrorString(last_launch_error));
    s = hip_check(ctx, hipStreamSynchronize(ctx->stream), "texel decode");
    if (!s && hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1) != hipSuccess) ctx->last_ms = 0.0f;
    return s;
}

}  // extern "C"
