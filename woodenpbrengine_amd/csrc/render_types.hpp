// Device-side scene and per-sample wavefront state of the renderer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "texture_device.hpp"
#include "wpbr_internal.hpp"

namespace wpbr {

constexpr int NSPEC = 64;
enum : uint32_t { MAT_CONDUCTOR = 0, MAT_DIELECTRIC = 1, MAT_SPECULAR = 2 };

struct DevMaterial {
    uint32_t type;
    float alpha;
    float eta;      // dielectric etaT (etaI = 1)
    uint32_t spec;  // conductor: index of its eta/k spectra in mat_spec (2 x 64 floats)
    int32_t tex_k;  // conductor: metallic texture (k = fromRGB of its value), -1 none
    int32_t tex_r;  // conductor: roughness texture (alpha from its red value), -1 none
};

struct SceneView {
    GeomView g;
    const uint8_t* tri_flags;
    const float* tri_uv;
    const float* tri_normal;
    const float* tri_tangent;
    const uint32_t* shape_material;
    const DevMaterial* materials;
    const float* mat_spec;       // per conductor: eta[64], k[64]
    uint32_t n_spec_floats;      // floats in mat_spec
    uint32_t n_sphere_lights;
    const uint32_t* light_shape;
    const float* light_emit;     // n_sphere_lights x 64
    uint32_t has_env;
    const float* env_emit;       // 64
    const float* env_rgb;        // h x w x 3
    uint32_t env_w, env_h;
    const float* env_xform;      // world 3x4, inverse 3x4
    const float* env_cond_func;  // h x w
    const float* env_cond_int;   // h
    const float* env_marg_int;   // 1
    const float* xyz;            // 3 x 64 (X, Y, Z)
    const float* illum;          // 7 x 64 Smits illuminant bases
    const float* refl;           // 7 x 64 Smits reflectance bases
    const TexView* textures;     // the scene's textures (textured metals)
    const float* raster_camera;  // 16
    const float* camera_world;   // 12
    uint32_t width, height;
};

// Requests of the BSDF Compute / Sample jobs (wpbr_bsdf_*), SoA device arrays.
struct BsdfArgs {
    const float* frame;       // 9 x n: ss, ts, ns
    const float* wo;          // 3 x n
    const float* wi;          // 3 x n (Compute)
    const float* u;           // 2 x n (Sample)
    const uint32_t* material; // n
    uint32_t n, n_materials;
    float* value;             // 64 x n
    float* pdf;               // n
    float* wi_out;            // 3 x n (Sample)
};

// Per-sample state between the wavefront stages (SoA, one entry per pixel).
struct WaveState {
    uint32_t* c1;         // 1D cursor
    uint32_t* c2;         // 2D cursor
    float* pfilm;         // 2 x N
    float* rays;          // primary rays, 6 x N
    uint32_t* hit;        // primary hit shape
    float* hit_t;
    float* bary;          // 3 x N
    // shade1 -> shade2
    uint32_t* flags;      // bit0 light ray, bit1 bsdf ray, bit2 env pick, bit3 done (miss/no bsdf)
    uint32_t* light;      // picked light index
    uint32_t* mat;        // material index
    float* l_cos;         // light strategy: dot(wi, wh) for the Fresnel term
    float* l_sc;          // scalar factor before |cos|
    float* l_ac;          // |cos theta_i|
    float* l_pdf;         // light pdf
    float* l_bpdf;        // BSDF pdf at the light direction
    uint32_t* l_zero;     // BSDF value is exactly 0 (degenerate geometry)
    float* b_cos;
    float* b_sc;
    float* b_ac;
    float* b_pdf;
    uint32_t* b_zero;
    float* b_wi;          // 3 x N world direction
    float* le;            // XYZ of the emitted light for primary misses (3 x N)
    float* tk;            // 3 x N: metallic texture value at a textured conductor hit (its k RGB)
    // compacted secondary ray queue
    float* qrays;         // 6 x 2N
    uint32_t* qhit;       // 2N
    float* qt;            // 2N
    uint32_t* qcount;     // 1
    uint32_t* l_ray;      // queue slot of the light ray per sample
    uint32_t* b_ray;      // queue slot of the BSDF ray per sample
    uint32_t* qtarget;    // 2N: light shape a queue ray must reach (0xFFFFFFFF: escape to the env)
    // shade1 record for the adjoint's path replay (written when RenderLaunch::record)
    float* r_wo;          // 3 x N: wo in the shading frame
    float* r_frame;       // 9 x N: ss, ts, ns (world)
    float* r_wl;          // 3 x N: light sample direction in the shading frame
    float* r_pdfl;        // N: light sample pdf
    float* r_ub;          // 2 x N: the BSDF sample's 2D sample
};

// Device tables of the adjoint (derivatives with respect to the parameters).
struct AdjView {
    uint32_t n_mat;
    const float* refl;      // 7 x 64 Smits reflectance bases
    const float* mat_rgb;   // n_mat x 6: conductor eta RGB, k RGB
    const int* mat_case;    // n_mat x 2: fromRGB case of eta and of k
    const int* env_src;     // env cells x 4: texels of the bilinear lookup behind a
                            // distribution cell (-1: the cell is never written)
    const float* env_srcw;  // env cells x 4: their weights times sin(theta) of the row
    const float* env_dI;    // texels: derivative of the marginal integral by the green channel
};

// Where a render's sample values come from (wpbr_scene_desc::sampler): mode 0
// independent hash streams; 1 and 2 the shuffled stratified arrays of n1 1D
// and n2 2D values per pixel, 2D strata on a root x root grid scaled by inv2
// (1/n2 for mode 1, 1/root for mode 2).
struct SamplerDef {
    uint32_t seed;
    uint32_t mode;
    uint32_t n1, n2, root;
    float inv1, inv2;
};

struct RenderLaunch {
    SceneView sv;
    WaveState ws;
    uint32_t n;           // pixels
    uint32_t spp_index;
    uint32_t spp;
    SamplerDef smp;
    float* film;          // W*H*4: sum X w, sum Y w, sum Z w, sum w
    float* sample_xyz;    // optional: (pixel*spp + j) x 3
    uint32_t record;      // shade1 writes the adjoint's replay record
};

// Arguments of the adjoint kernel.
struct AdjArgs {
    RenderLaunch L;
    AdjView A;
    const float4* film;  // forward film: sum XYZ w, sum w
    const float* dL;     // W*H*3: d loss / d rgb
    float* part;         // per-block rows of 7 n_mat + 1: alpha, eta rgb, k rgb, d/dI
    float* grad_env;     // env_h*env_w*3
};

}  // namespace wpbr
