// Oracle: CPU restatement of WoodenPBREngine's ray-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg may load liboracle.so, and only as the checker / the timed CPU
// baseline.  The product library (woodenpbrengine_amd/) never includes, links or
// calls anything here.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   * Morton codes     - closed-form bit interleave, brute-force recomputation.
//   * radix sort       - std::stable_sort on (code, index) (a library routine).
//   * treelets / emit  - closed-form invariants (each prim in one leaf, leaf n < 30
//                        unless all codes equal, bounds = union, preorder layout,
//                        split = first index with the bit set).
//   * upper SAH        - restatement only; std::partition order is
//                        implementation-defined, fixed here to the libstdc++
//                        two-pointer form and checked against std::partition.
//   * traversal        - brute force over all primitives (same intersectors) gives
//                        the same minimum t; ids agree unless t ties.
//   * triangle test    - closed-form hits (axis-aligned / edge-on cases).
//   * sphere test      - analytic closed form.
//   * shading          - closed forms listed in shade.cpp.
#pragma once
#include "common.h"
#include <array>
#include <vector>

namespace orc {

struct Scene {
    uint32_t n_spheres = 0;
    const float* sphere_xform = nullptr;   // per sphere: world 3x4 then inverse 3x4 (24 floats)
    const float* sphere_radius = nullptr;
    uint32_t n_triangles = 0;
    const float* tri_verts = nullptr;      // per triangle: v0, v1, v2 (9 floats)
};

struct LbvhResult {
    Bounds centroid_bounds;
    std::vector<uint32_t> morton;          // per shape, input order
    std::vector<uint32_t> sorted_code, sorted_index;
    std::vector<uint32_t> shape_entities;  // CSceneTreeLBVH::shapesEntities
    std::vector<uint32_t> treelet_start;
    std::vector<uint32_t> node_offset, node_nshapes;
    std::vector<uint8_t> node_axis;
    std::vector<float> node_bounds;        // 6 per node: min xyz, max xyz
    uint32_t n_nodes = 0, n_sah_nodes = 0, n_treelets = 0;
};

// libstdc++ std::partition for bidirectional iterators (the two-pointer form).
template <class It, class P>
inline
It partition_bidir(It first, It last, P pred) {
    while (true) {
        while (true) {
            if (first == last) return first;
            if (pred(*first)) ++first;
            else break;
        }
        --last;
        while (true) {
            if (first == last) return first;
            if (!pred(*last)) --last;
            else break;
        }
        std::iter_swap(first, last);
        ++first;
    }
}

void triangle_bounds(uint32_t n, const float* v, float* bounds, float* centroid);
void sphere_bounds(uint32_t n, const float* xform, const float* radius, float* bounds,
                   float* centroid);
Bounds centroid_bounds(uint32_t n, const float* centroid);
uint32_t morton_code(const Bounds& cb, const float* c);
// Key order of the LBVH build.  REFERENCE: what JobSortsShapesByMortonCode
// leaves in shapeMortonCodes (pass 3 of 5: bits 0..23, stable), so treelets are
// runs of equal top-12 bits in that order, up to n of them.  FULL: the complete
// 30-bit stable sort (at most 4096 treelets).
enum { LBVH_ORDER_REFERENCE = 0, LBVH_ORDER_FULL = 1 };
void lbvh_build(uint32_t n, const float* bounds, const float* centroid, const uint32_t* entities,
                LbvhResult& out, int order = LBVH_ORDER_REFERENCE);

bool box_hit(const float* bmn, const float* bmx, V3 o, V3 inv, const int* neg, float tMax,
             float* tNearOut);
float tri_intersect(const float* v, V3 o, V3 d, float* bary);
float sphere_intersect(const float* xform, float radius, V3 o, V3 d, V3* oL, V3* dL);
int64_t trace_ray(const Scene& sc, const LbvhResult& bvh, V3 o, V3 d, float* tOut,
                  uint64_t* nCandidates, int* stackOverflow);

// ---- spectrum (spectrum.cpp) ----------------------------------------------
constexpr int NSPEC = 64;
using Spec = std::array<float, NSPEC>;
struct Tables {
    float X[NSPEC], Y[NSPEC], Z[NSPEC];
    float refl[7][NSPEC], illum[7][NSPEC];  // white, cyan, magenta, yellow, red, green, blue
};
const Tables& tables();
float average_spectrum_samples(const float* lambda, const float* vals, int n, float l0, float l1);
Spec from_rgb(const float* rgb, bool illuminant);
void to_xyz(const Spec& s, float* xyz);
void xyz_to_rgb(const float* xyz, float* rgb);
// JobOutputFilmTitles (CSFilm.cpp:141-170) with gammaCorrect (CSFilm.cpp:129-140):
// per pixel the accumulated XYZ (3 floats) and filter weight sum -> 3 bytes.
void film_rgb8(uint32_t npix, const float* xyz, const float* w, uint8_t* out);

// ---- render (render.cpp) ----------------------------------------------------
enum MaterialType : uint32_t { MAT_CONDUCTOR = 0, MAT_DIELECTRIC = 1, MAT_SPECULAR = 2 };
struct Material {
    uint32_t type;
    float alpha;       // alphax == alphay
    float eta;         // dielectric etaT (etaI = 1)
    float eta_rgb[3];  // conductor etaT as RGB (fromRGB Reflectance)
    float k_rgb[3];    // conductor k as RGB (fromRGB Reflectance)
};

// Sample values of a render: mode 0 independent hash streams per (seed, pixel,
// index); modes 1, 2 the shuffled stratified arrays of
// JobSamplerStratifiedGenerateSampels1D/2D (CSSamplerStratified.cpp:8-90), n1 1D
// and n2 2D values per pixel (CSFilm.cpp:22-23 with spp passes: 16 + spp,
// 16 + 4 spp), 2D grid root x root scaled by 1/n2 (mode 1, as the reference) or
// 1/root (mode 2).
struct Sampler {
    uint32_t seed = 0, mode = 0;
    uint32_t n1 = 1, n2 = 1, root = 1;
    float inv1 = 1.0f, inv2 = 1.0f;
};
Sampler make_sampler(uint32_t seed, uint32_t mode, uint32_t n1, uint32_t n2);

// ---- textures (texture.cpp) ---------------------------------------------------
struct TexRGB {
    float v[3];
};
// CTexture2DRGB: the MIP levels (w[l] x h[l] texels, row-major).
struct Texture {
    std::vector<uint32_t> w, h;
    std::vector<std::vector<TexRGB>> mips;
    const TexRGB& texel(int x, int y, uint32_t l) const { return mips[l][(size_t)y * w[l] + x]; }
    TexRGB& texel(int x, int y, uint32_t l) { return mips[l][(size_t)y * w[l] + x]; }
};
Texture generate_mips(uint32_t w, uint32_t h, const float* rgb);
std::vector<float> texture_filter_table();
// EWA on one level; q = (s, t), dx = (ds/dx, dt/dx), dy = (ds/dy, dt/dy).
TexRGB texture_ewa(const float* q, const float* dx, const float* dy, const Texture& tex,
                   const std::vector<float>& filter, uint32_t mip);
// STexture2DSamplerAnisotropic::eval, maxAnisotropy 32; q = 6 floats as CTextureMappedPoint.
TexRGB texture_eval(const float* q, const Texture& tex, const std::vector<float>& filter);

struct RenderScene {
    Scene geom;
    // per-triangle shading data (optional, may be null)
    const uint8_t* tri_flags = nullptr;   // bit0: uvs, bit1: normals + tangents
    const float* tri_uv = nullptr;        // nt x 6
    const float* tri_normal = nullptr;    // nt x 9
    const float* tri_tangent = nullptr;   // nt x 9
    const uint32_t* shape_material = nullptr;  // per shape; 0xFFFFFFFF = none
    uint32_t n_materials = 0;
    const Material* materials = nullptr;
    // lights: sphere lights first, then the env light (if any)
    uint32_t n_sphere_lights = 0;
    const uint32_t* light_shape = nullptr;     // shape id of each sphere light
    const float* light_emit = nullptr;         // n_sphere_lights x 64
    int has_env = 0;
    const float* env_emit = nullptr;           // 64
    const float* env_rgb = nullptr;            // env_h x env_w x 3
    uint32_t env_w = 0, env_h = 0;
    const float* env_xform = nullptr;          // world 3x4 then inverse 3x4
    // camera
    const float* raster_camera = nullptr;      // 4x4 row-major
    const float* camera_world = nullptr;       // 3x4 row-major
    uint32_t width = 0, height = 0, spp = 0;
    uint32_t seed = 0;
    Sampler smp;
    // textured metals (CMaterialMetal.cpp:24-38): per material [metallic,
    // roughness] indices into textures (-1 none); null: no textures
    std::vector<Texture> textures;
    const int32_t* material_textures = nullptr;
    int32_t tex_index(uint32_t mid, int k) const { return material_textures ? material_textures[2 * mid + k] : -1; }
};

// The CTextureMappedPoint (s, t, ds/dx, dt/dx, ds/dy, dt/dy) of a camera ray's
// hit on `shape` at t (b: its barycentrics on a triangle), from the camera's ray
// differential at film position pFilm (interaction.cpp).
void camera_hit_texture_point(const RenderScene& sc, uint32_t shape, V3 o, V3 d, float t, const float* b,
                              const float* pFilm, float* q);

struct RenderOut {
    std::vector<float> sample_xyz;   // width*height*spp x 3, index (pixel*spp + s)
    std::vector<float> sample_pfilm; // x 2
    std::vector<float> film_rgb;     // width*height x 3 (toRGB(sum L w) / sum w)
    std::vector<float> film_xyz;     // width*height x 3 (sum of XYZ(L) w)
    std::vector<float> film_w;
    uint64_t counters[8] = {0};      // primary hits, misses, shadow rays, bsdf rays, specular hits,
                                     // clamped light picks, env picks, render ns
};

// ---- interactions at the collision boundary (interaction.cpp) ----------------
// CSurfaceInteraction (CSufraceInteraction.h:44-199) as SoA floats, field f of
// ray i at si[f * n + i] (the layout of wpbr.h's WPBR_SI_*).
enum {
    SI_P = 0, SI_WO = 3, SI_TIME = 6, SI_N = 7, SI_UV = 10, SI_DPDU = 12, SI_DPDV = 15, SI_DNDU = 18,
    SI_DNDV = 21, SI_SHADING_N = 24, SI_SHADING_DPDU = 27, SI_SHADING_DPDV = 30, SI_SHADING_DNDU = 33,
    SI_SHADING_DNDV = 36, SI_DPDX = 39, SI_DPDY = 42, SI_DUDX = 45, SI_DVDX = 46, SI_DUDY = 47, SI_DVDY = 48,
    SI_FLOATS = 49
};
// runCollisionSystem's results per ray cast: the closest hit (shape, t) and,
// when si is given, the interaction the reference attaches (a surface
// interaction for bSurfInteraction rays, surf == null meaning all; the plain
// CInteraction fields p, wo, time otherwise; zeros on a miss) with the
// differentials of the rays that carry a CRayDifferential (diff: 12 x n SoA,
// difXRay origin, dir, difYRay origin, dir; null: none).
void scene_collide(const RenderScene& sc, const LbvhResult& bvh, uint32_t n, const float* rays,
                   const uint8_t* surf, const float* diff, uint32_t* hit_shape, float* hit_t, float* si,
                   int nthreads);

// Texture-mapped points of camera rays through film positions (render.cpp).
void texture_points(const RenderScene& sc, const LbvhResult& bvh, uint32_t n, const float* pfilm, uint32_t* shape,
                    float* q);

// The BSDF Compute (sample = 0) / Sample (sample = 1) jobs for n requests (render.cpp).
void bsdf_eval(const RenderScene& sc, int sample, uint32_t n, const float* frame, const float* wo,
               const float* wi, const float* u, const uint32_t* material, float* value, float* pdf,
               float* wi_out);

// Every discrete decision of one sample (the finite-difference reference of
// diff.cpp re-evaluates the estimator with them frozen).
struct EnvCell {
    int32_t s0, t0, s0n, t0n;  // bilinear texels (clamped)
    int32_t rgb_case;          // fromRGB ordering case (0..5) of the looked-up rgb
    int32_t iu, iv;            // distribution cell of the pdf
};
struct SampleRec {
    uint8_t kind = 0;       // 0 miss, 1 hit without a contribution, 2 hit with a material
    uint8_t env_pick = 0;   // the env light was picked
    uint8_t light_on = 0;   // the light strategy contributes (reached, pdf != 0)
    uint8_t bsdf_on = 0;    // the BSDF ray reached the picked light (sample not rejected)
    uint32_t mid = 0, li = 0;
    float dir[3] = {0, 0, 0};  // primary ray direction (miss)
    float wo[3], ss[3], ts[3], ns[3];
    float wl[3];            // light direction (local frame)
    float pdfL = 0;         // light sample pdf
    float ub[2] = {0, 0};   // BSDF 2D sample
    float pdfL2s = 0;       // sphere light pdf in the BSDF strategy
    float pfilm[2];
    EnvCell env;            // env lookup of the miss or of the BSDF direction
};

float sample_1d(const Sampler& S, uint32_t pixel, uint32_t index);
void sample_2d(const Sampler& S, uint32_t pixel, uint32_t index, float* u);
// The arrays of pixel `pixel` as the reference's jobs build them: stratified
// elements in the reference's loop order, then shuffled (u1: n1, u2: n2 x 2).
void stratified_arrays(const Sampler& S, uint32_t pixel, float* u1, float* u2);
uint32_t permute_index(uint32_t c, uint32_t n, uint32_t key);
void render(const RenderScene& sc, const LbvhResult& bvh, RenderOut& out, int nthreads,
            const uint32_t* tile_subset, uint32_t n_tiles, SampleRec* recs = nullptr);
// Smits case index (0..5) of rgb, in the order of from_rgb's branches.
int rgb_case(const float* rgb);
float mitchell_weight(float x);  // the film's Mitchell filter (CSFilm.h:20-44)

// Finite-difference reference (diff.cpp).
struct DiffState;
DiffState* diff_record(const RenderScene& sc, const LbvhResult& bvh, int nthreads);
void diff_free(DiffState* d);
double diff_loss(const DiffState& d, const double* alpha, const double* etak, const double* env_rgb,
                 const float* target_rgb, float* rgb_out, int nthreads);
void shade_probe(int fn, const float* in, float* out);

}  // namespace orc
