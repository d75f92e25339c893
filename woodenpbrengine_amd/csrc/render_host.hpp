// Host-side scene and render state shared by the forward renderer
// (render_host.cpp) and the adjoint host code (adjoint_host.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "render_types.hpp"
#include "wpbr_ctx.hpp"

namespace wpbr {

wpbr_status render_pass_launch(RenderLaunch& L, int stage, hipStream_t st);
wpbr_status film_rgb8_launch(const float* film, uint32_t npix, uint8_t* out, hipStream_t st);
struct BsdfArgs;
wpbr_status bsdf_queues_launch(const SceneView& sv, const BsdfArgs& a, uint32_t* blk, uint32_t* counts,
                               uint32_t* queue, hipStream_t st);
wpbr_status bsdf_launch(const SceneView& sv, const BsdfArgs& a, int kind, bool sample, const uint32_t* queue,
                        uint32_t q, hipStream_t st);
wpbr_status scene_si_launch(const SceneView& sv, uint32_t n, const float* rays, const uint8_t* surf,
                            const float* diff, const uint32_t* hit_shape, const float* hit_t, float* si,
                            hipStream_t st);
wpbr_status sample_arrays_launch(const SamplerDef& S, uint32_t pixel0, uint32_t npix, float* u1, float* u2,
                                 hipStream_t st);
SamplerDef make_sampler(uint32_t seed, uint32_t mode, uint32_t n1, uint32_t n2);
// Batched stages over a device table of launch records: 0 camera, 1 shade1,
// 2 radiance + splat, 3 clear films and queue statistics, 4 radiance + splat
// with the spectral tables in LDS (every scene's tables fit: splat_tables_fit_h).
bool splat_tables_fit_h(const SceneView& sv);
wpbr_status render_batch_launch(const RenderLaunch* T, const SegMap& map, uint32_t grid, int stage, uint32_t j,
                                hipStream_t st);
wpbr_status scene_upload(wpbr_ctx* ctx, const wpbr_scene_desc* d, wpbr_scene** out);
// Carves one wave state for n pixels at p (p null: only the size); returns the bytes used.
size_t carve_wave(char* p, uint32_t n, WaveState& w);
wpbr_status ensure_wave(wpbr_ctx* ctx, uint32_t n);
wpbr_status ensure_stream2(wpbr_ctx* ctx);

struct HostTables {
    float xyz[3 * NSPEC];
    float refl[7 * NSPEC];
    float illum[7 * NSPEC];
};
const HostTables& host_tables();
// fromRGB(rgb, Reflectance) (CSpectrumSampled.h:110-158).
void from_rgb_refl(const float* rgb, float* s);
void build_materials(const float* h, uint32_t nm, const int32_t* mat_tex, std::vector<DevMaterial>& mats,
                     std::vector<float>& specs);
const TexView* texture_view(const wpbr_texture* t);
// Bilinear mip-0 lookup (CSTextureProcess.h:58-81), indices clamped to the texture.
void env_bilerp_h(const float* rgbTex, int W, int H, float s_, float t_, float* out);
// The env light's sampling distribution (CLightInfiniteArea.cpp:44-70).
void env_distribution(const float* rgbTex, uint32_t W, uint32_t H, std::vector<float>& func,
                      std::vector<float>& rowInt, float& margInt);

}  // namespace wpbr

struct wpbr_scene {
    wpbr_ctx* ctx = nullptr;
    std::vector<void*> bufs;
    wpbr::SceneView sv{};
    wpbr_geometry geom{};
    wpbr_lbvh* bvh = nullptr;
    uint32_t spp = 1, seed = 0;
    wpbr::SamplerDef smp{};
    // parameters (host copies) and the writable buffers behind sv's pointers
    std::vector<float> h_materials;  // n_materials x 9, the descriptor's layout
    std::vector<float> h_env;        // env_h x env_w x 3
    std::vector<int32_t> h_mat_tex;  // n_materials x 2: metallic, roughness texture (-1 none)
    bool textured = false;           // some material binds a texture
    wpbr::DevMaterial* d_materials = nullptr;
    float* d_mat_spec = nullptr;
    float* d_env_rgb = nullptr;
    float* d_env_func = nullptr;
    float* d_env_rowint = nullptr;
    float* d_env_margint = nullptr;
    wpbr::AdjView adj{};
    bool adj_ready = false;
};

struct wpbr_render_state {
    void* mem = nullptr;
    uint32_t cap = 0;
    // two wavefront states (sharing the sample-stream cursors): consecutive
    // sample passes alternate between them (and between the context's stream
    // and stream2), so that one pass's secondary trace and splat overlap the
    // next pass's camera, primary trace and shading
    wpbr::WaveState ws{}, ws2{};
    hipStream_t stream2 = nullptr;
    uint32_t* spill2 = nullptr;  // traversal stack overflow of stream2's launches
    size_t spill2_words = 0;
    uint32_t* claim2 = nullptr;  // persistent-kernel claim counters of stream2's launches
    std::vector<hipEvent_t> evs;
    double stats[8] = {0};
};

