/* wpbr.h - C ABI of libwpbr_hip.so, the MI355X (gfx950) implementation of
 * WoodenPBREngine's ray-tracing hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns a
 * wpbr_status, and never throws.  Pointers named d_* are device (HBM) pointers
 * obtained from wpbr_malloc (or any hipMalloc'd memory on the context's
 * device); all other pointers are host memory.  Calls are ordered on the
 * context's stream and return after the work completes unless the name ends
 * in _async.  A context is not re-entrant: use one per host thread.
 *
 * Reference interfaces each group replaces (WoodenPBREngine/...):
 *   wpbr_shape_bounds        JobUpdateBoundsAndCentroidTriangle / ...Sphere
 *                            (CSTriangle.cpp:116-142, CSSphere.h:126-153)
 *   wpbr_lbvh_build          MEngine::buildLBVH (MEngine.cpp:341-350): the jobs
 *                            JobGenerateShapesCentroidBound (SLBVH.cpp:164),
 *                            JobGenerateShapesMortonCode (:186),
 *                            JobSortsShapesByMortonCode (:231),
 *                            JobEmitLBVH (:328), JobBuildUpperSAH (:410)
 *   wpbr_collide             MEngine::runCollisionSystem (MEngine.cpp:352-361):
 *                            JobProcessRayCasts (SLBVH.cpp:85), the sphere and
 *                            triangle interaction-request jobs (CSSphere.cpp:16,
 *                            CSTriangle.cpp:9), JobProcessRayCastsResults
 *                            (SLBVH.cpp:9) and the full-interaction jobs
 *                            (CSSphere.cpp:53, CSTriangle.cpp:149)
 *   wpbr_render_*            MEngine::render (MEngine.cpp:380-448) job list
 */
#ifndef WPBR_H
#define WPBR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t wpbr_status;
enum {
    WPBR_OK = 0,
    WPBR_ERR_INVALID_ARG = 1,
    WPBR_ERR_OOM = 2,
    WPBR_ERR_HIP = 3,
    WPBR_ERR_STACK_OVERFLOW = 4, /* traversal stack exceeded (reference: UB past 256) */
    WPBR_ERR_TIMEOUT = 5,        /* a bounded device-side wait gave up */
    WPBR_ERR_NOT_BUILT = 6
};

typedef struct wpbr_ctx wpbr_ctx;
typedef struct wpbr_lbvh wpbr_lbvh;

/* ---- context and device memory --------------------------------------- */
wpbr_status wpbr_open(int device, wpbr_ctx** out);
void wpbr_close(wpbr_ctx* ctx);
const char* wpbr_status_string(wpbr_status s);
const char* wpbr_last_error(wpbr_ctx* ctx);
wpbr_status wpbr_malloc(wpbr_ctx* ctx, size_t bytes, void** d_ptr);
wpbr_status wpbr_free(wpbr_ctx* ctx, void* d_ptr);
wpbr_status wpbr_upload(wpbr_ctx* ctx, void* d_dst, const void* src, size_t bytes);
wpbr_status wpbr_download(wpbr_ctx* ctx, void* dst, const void* d_src, size_t bytes);
wpbr_status wpbr_sync(wpbr_ctx* ctx);
/* Device time in milliseconds of the last completed entry point (HIP events on
 * the context's stream). */
float wpbr_last_time_ms(wpbr_ctx* ctx);

/* ---- geometry ---------------------------------------------------------- *
 * Shape order (the order the reference's ECS iterates CBounds/CCentroid):
 * spheres first (created by SSphere::create while the scene is built), then
 * triangles (created afterwards by JobTriangleMeshGenerateTriangles).
 * Shape id s < n_spheres is sphere s; otherwise triangle s - n_spheres. */
typedef struct {
    uint32_t n_spheres;
    const float* d_sphere_xform;  /* n_spheres x 24: world 3x4 row-major, then its inverse */
    const float* d_sphere_radius; /* n_spheres */
    uint32_t n_triangles;
    const float* d_tri_verts;     /* n_triangles x 9: world-space v0, v1, v2 */
} wpbr_geometry;

/* a1: per-shape world bounds (6 floats: min xyz, max xyz) and centroid
 * (3 floats) into d_bounds / d_centroid, shape order as above. */
wpbr_status wpbr_shape_bounds(wpbr_ctx* ctx, const wpbr_geometry* geom, float* d_bounds,
                              float* d_centroid);

/* ---- LBVH --------------------------------------------------------------- */
/* Key order of a build.  REFERENCE (the default) is what the reference's
 * JobSortsShapesByMortonCode leaves in shapeMortonCodes for the rest of its
 * jobs: its 5 x 6-bit passes alternate between two buffers and the last pass's
 * output lands in the temporary one (SLBVH.cpp:245-291), so the order read by
 * the gather, the treelet split (SLBVH.cpp:297-319) and emitLBVH (:374-387) is
 * a stable sort on bits 0..23 only: treelets are runs of equal top-12 bits in
 * that order (up to one per shape) and the upper SAH (:457-563) runs over all
 * of them.  FULL: the complete 30-bit stable sort (at most 4096 treelets), the
 * order the job was written to produce; a different tree with the same hits. */
enum { WPBR_LBVH_ORDER_REFERENCE = 0, WPBR_LBVH_ORDER_FULL = 1 };
wpbr_status wpbr_lbvh_create(wpbr_ctx* ctx, wpbr_lbvh** out);
void wpbr_lbvh_destroy(wpbr_lbvh* bvh);
/* The key order of the next builds of bvh (created with the context's
 * default). */
wpbr_status wpbr_lbvh_set_order(wpbr_lbvh* bvh, int order);
int wpbr_lbvh_get_order(const wpbr_lbvh* bvh);
/* The key order of LBVH objects created afterwards on ctx, including the ones
 * wpbr_scene_create and wpbr_scenes_create make (default REFERENCE). */
wpbr_status wpbr_set_default_lbvh_order(wpbr_ctx* ctx, int order);
/* Builds the tree over the shapes of geom (bounds, centroid bounds, Morton
 * codes, radix sort, treelets, upper SAH, flatten, traversal layout).  Device
 * buffers are owned by the lbvh object and reused across builds. */
wpbr_status wpbr_lbvh_build(wpbr_ctx* ctx, const wpbr_geometry* geom, wpbr_lbvh* bvh);
/* Segmented build: the trees of n_scenes independent scenes (geoms[i] into
 * bvhs[i], each from wpbr_lbvh_create, all distinct, all of one key order),
 * one kernel launch per build kernel for all scenes together; each tree equals
 * what wpbr_lbvh_build gives for that scene alone.  Replaces MEngine::buildLBVH (MEngine.cpp:341-350)
 * run once per scene; the segmented form itself has no counterpart in the
 * reference.  At most 65535 scenes; wpbr_last_time_ms is the whole batch. */
wpbr_status wpbr_lbvh_build_segmented(wpbr_ctx* ctx, uint32_t n_scenes, const wpbr_geometry* geoms,
                                      wpbr_lbvh* const* bvhs);
/* The same build one job at a time, in order (MEngine::buildLBVH,
 * MEngine.cpp:341-350), so each of the reference's JOB_RUN calls maps to one
 * call: 0 JobUpdateBoundsAndCentroidSphere/Triangle (CSTriangle.cpp:116-142,
 * CSSphere.h:126-153), 1 JobGenerateShapesCentroidBound (SLBVH.cpp:164),
 * 2 JobGenerateShapesMortonCode (SLBVH.cpp:186), 3 JobSortsShapesByMortonCode
 * (SLBVH.cpp:231), 4 JobEmitLBVH (SLBVH.cpp:328), 5 JobBuildUpperSAH
 * (SLBVH.cpp:410).  Stages 0-4 return once queued on the context's stream;
 * stage 5 waits for the build.  A stage out of order, or a different geometry
 * than stage 0's, is WPBR_ERR_INVALID_ARG. */
wpbr_status wpbr_lbvh_build_stage(wpbr_ctx* ctx, const wpbr_geometry* geom, wpbr_lbvh* bvh, int stage);
/* Counts: [0] shapes, [1] nodes, [2] upper SAH nodes, [3] treelets. */
wpbr_status wpbr_lbvh_counts(wpbr_lbvh* bvh, uint32_t counts[4]);
/* Downloads one array of CSceneTreeLBVH (SLBVH.h:17-40) or build state to host
 * memory.  field: 0 Morton codes (u32, shape order), 1 sorted codes (u32),
 * 2 sorted shape index (u32) = shapesEntities, 3 node offset (u32:
 * shapeOffset for leaves, secondChildOffset for interior nodes), 4 node
 * nShapes (u32), 5 node axis (u8), 6 node bounds (6 x f32), 7 centroid bounds
 * (6 x f32), 8 treelet start (u32). */
wpbr_status wpbr_lbvh_download(wpbr_ctx* ctx, wpbr_lbvh* bvh, int field, void* dst, size_t bytes);
/* Per-kernel device times (ms) of the last build: [0] bounds, [1] morton+hist,
 * [2..5] sort passes, [6] treelet marks, [7] emit, [8] upper SAH, [9] flatten,
 * [10] traversal layout; returns the number written. */
int wpbr_lbvh_kernel_times(wpbr_lbvh* bvh, float* ms, int cap);

/* ---- collision ---------------------------------------------------------- *
 * Rays are SoA: d_rays = 6 x n floats (ox[n], oy[n], oz[n], dx[n], dy[n],
 * dz[n]); tmax = +inf as in the reference.  Output per ray: the hit shape id
 * (0xFFFFFFFF = INVALID_HANDLE, a miss) and t (+inf on a miss); optional
 * barycentrics (3 x n, triangles only).  Closest hit, ties to the first
 * candidate in the reference's depth-first order. */
wpbr_status wpbr_collide(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_lbvh* bvh,
                         uint32_t n, const float* d_rays, uint32_t* d_hit_shape, float* d_hit_t,
                         float* d_bary);
/* Same, without waiting; the caller synchronises with wpbr_sync. */
wpbr_status wpbr_collide_async(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_lbvh* bvh,
                               uint32_t n, const float* d_rays, uint32_t* d_hit_shape,
                               float* d_hit_t, float* d_bary);
/* Traversal launch configuration (no counterpart in the reference, whose
 * traversal is a CPU job).  Results do not depend on it.
 *   block        workgroup size: 64, 128, 256 or 512
 *   stack_depth  LDS stack entries per lane: 8, 16, 24, 32, 48 or 64 (deeper
 *                entries spill to global memory, 256 in all)
 *   mapping      1: ray i to workgroup i / block; 2: one contiguous part of the
 *                rays per XCD; 3: persistent lanes that fetch rays dynamically
 *   width        2: binary records; 4: records of four grandchildren
 *   refill       mapping 3: idle lanes of a wave that trigger a fetch (1..64)
 *   chunk        mapping 3: rays a wave claims per atomic (64..65536)
 * wpbr_set_trace_config: a 0 field keeps the current value. */
typedef struct {
    int32_t block;
    int32_t stack_depth;
    int32_t mapping;
    int32_t width;
    int32_t refill;
    int32_t chunk;
} wpbr_trace_config;
wpbr_status wpbr_set_trace_config(wpbr_ctx* ctx, const wpbr_trace_config* cfg);
wpbr_status wpbr_get_trace_config(wpbr_ctx* ctx, wpbr_trace_config* cfg);
/* Autotuner: times wpbr_collide over the given rays for each candidate launch
 * configuration (block, stack depth, mapping, width, refill), keeps the
 * fastest in the context and returns it in *best (optional).  The hit buffers
 * receive the (configuration-independent) result. */
wpbr_status wpbr_autotune_collide(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_lbvh* bvh,
                                  uint32_t n, const float* d_rays, uint32_t* d_hit_shape,
                                  float* d_hit_t, wpbr_trace_config* best);

typedef struct wpbr_texture wpbr_texture;  /* an RGB texture and its MIP pyramid (below) */

/* ---- rendering ---------------------------------------------------------- *
 * Scene description: host pointers, copied to the device by wpbr_scene_create.
 * Shape order as in wpbr_geometry.  Materials are 9 floats each: type (u32
 * bits: 0 conductor microfacet, 1 dielectric microfacet, 2 specular reflection
 * "perfect glass"), alpha, eta (dielectric etaT, etaI = 1), eta RGB[3] and k RGB[3]
 * (conductor; converted with fromRGB(Reflectance) as CMaterialMetal.cpp:46).
 * Lights: sphere lights (their shape ids and 64-bin emission), then the
 * infinite env light if has_env.  The camera is a raster->camera 4x4 and a
 * camera->world 3x4 (CSCameraProjective.h:25-52). */
typedef struct {
    uint32_t n_spheres;
    const float* sphere_xform;
    const float* sphere_radius;
    uint32_t n_triangles;
    const float* tri_verts;
    const uint8_t* tri_flags;     /* bit0 per-vertex uvs, bit1 normals + tangents */
    const float* tri_uv;          /* n_triangles x 6 */
    const float* tri_normal;      /* n_triangles x 9 */
    const float* tri_tangent;     /* n_triangles x 9 */
    const uint32_t* shape_material; /* per shape, 0xFFFFFFFF = none (light sphere) */
    uint32_t n_materials;
    const float* materials;
    uint32_t n_sphere_lights;
    const uint32_t* light_shape;
    const float* light_emit;      /* n_sphere_lights x 64 */
    uint32_t has_env;
    const float* env_emit;        /* 64 */
    const float* env_rgb;         /* env_h x env_w x 3 */
    uint32_t env_w, env_h;
    const float* env_xform;       /* world 3x4 then inverse 3x4 */
    const float* raster_camera;   /* 16 */
    const float* camera_world;    /* 12 */
    uint32_t width, height, spp, seed;
    /* Sample values (WPBR_SAMPLER_*).  The stratified modes give each pixel the
     * arrays of JobCreateCameraSamples (CSFilm.cpp:22-23, 57-64): 16 + spp 1D
     * and 16 + 4 spp 2D values, consumed in order over the spp passes. */
    uint32_t sampler;
    /* Textured metals (CMaterialMetal.cpp:24-38).  material_textures holds two
     * indices into textures per material (n_materials x 2, -1: none): [0] the
     * metallic texture, whose value gives the conductor's k (fromRGB,
     * Reflectance) while its eta is the constant spectrum of the material's
     * eta (the reference's Spectrum(1.5f) for 1.5); [1] the roughness texture,
     * whose red value r gives alpha = roughnessToAlpha(clamp(90 r, 0, 90))
     * (MicrofaceDistr.h:18-24).  Both are looked up (EWA, 32:1) at the camera
     * ray's hit with the camera's ray differentials (CSCameraPerspective.cpp:
     * 37-41), UV mapping on triangles and the sphere mapping on spheres
     * (CTextureMapping.h:39-90).  Only conductors bind textures.  The scene keeps
     * the texture pointers: they must outlive it.  n_textures 0 (null arrays):
     * no textures. */
    uint32_t n_textures;
    wpbr_texture* const* textures;
    const int32_t* material_textures;
} wpbr_scene_desc;

/* Sampler modes.  INDEPENDENT: a counter-based hash stream per (seed, pixel,
 * index).  STRATIFIED: the arrays of JobSamplerStratifiedGenerateSampels1D/2D
 * (CSSamplerStratified.cpp:8-90) with a defined shuffle; the 2D strata are
 * scaled by the array size as the reference does (CSSamplerStratified.cpp:79).
 * STRATIFIED_N: the same with the 2D strata scaled by n = sqrt(size), the
 * stratification the reference's outer invNSamples intends. */
enum { WPBR_SAMPLER_INDEPENDENT = 0, WPBR_SAMPLER_STRATIFIED = 1, WPBR_SAMPLER_STRATIFIED_N = 2 };

/* The per-pixel sample arrays of pixels pixel0 .. pixel0 + npix - 1 (film
 * row-major index) on the device: d_u1 receives npix x n1 floats, d_u2 npix x n2
 * float pairs, each pixel's array in consumption order.  For the stratified
 * modes these are the CSamples1D / CSamples2D data that
 * JobSamplerStratifiedGenerateSampels1D/2D (CSSamplerStratified.cpp:8-90)
 * write: 1D element k = min((k + xi) / n1, 1); 2D element (i, j) of the n x n
 * grid (n = (int)sqrt((float)n2)) = min((i + xi, j + xi') * s, 1) with s = 1/n2
 * (STRATIFIED) or 1/n (STRATIFIED_N), the n2 - n^2 tail (0, 0); then shuffled
 * (the reference's std::shuffle from a random_device seed is replaced by a
 * hash permutation of the array positions, defined per seed and pixel).  The
 * render with the same seed and mode reads exactly these values.  Either
 * output may be null. */
wpbr_status wpbr_sample_arrays(wpbr_ctx* ctx, uint32_t seed, uint32_t sampler, uint32_t pixel0,
                               uint32_t npix, uint32_t n1, uint32_t n2, float* d_u1, float* d_u2);

typedef struct wpbr_scene wpbr_scene;

/* Uploads the scene, runs the load-time preprocessing (spectral tables,
 * SampledSpectrum::init CSpectrumSampled.h:236-281; env-light distribution,
 * JobLightInfiniteAreaPreprocessSampling CLightInfiniteArea.cpp:44-70) and
 * builds its LBVH (wpbr_lbvh_build). */
wpbr_status wpbr_scene_create(wpbr_ctx* ctx, const wpbr_scene_desc* desc, wpbr_scene** out);
void wpbr_scene_destroy(wpbr_scene* scene);
/* Rebuilds the scene's LBVH from its device geometry. */
wpbr_status wpbr_scene_rebuild(wpbr_ctx* ctx, wpbr_scene* scene);
/* Renders spp samples per pixel (all film tiles per pass).  d_film receives
 * width*height*4 floats: the Mitchell-weighted sums of X, Y, Z and of the
 * weights (toRGB of the sum / weight is the reference's film value before its
 * tonemap).  d_sample_xyz (optional) receives width*height*spp*3 floats, the
 * XYZ of each sample's radiance at index (pixel * spp + j). */
wpbr_status wpbr_render(wpbr_ctx* ctx, wpbr_scene* scene, float* d_film, float* d_sample_xyz);
/* The render's JOB_RUN sequence (MEngine::render, MEngine.cpp:380-448) one
 * stage at a time, each stage one launch over every film tile on the context's
 * stream, finished when the call returns:
 *   WPBR_STAGE_BEGIN          film and queue counters cleared (once, before pass 0)
 *   WPBR_STAGE_CAMERA         JobSamplerUpdateCameraSamples, JobCameraPerspGenerateRaysDifferential (:391-392)
 *   WPBR_STAGE_PRIMARY_TRACE  runCollisionSystem of the camera rays (:393)
 *   WPBR_STAGE_SCATTER        JobScatteringRequestEmittedLight .. JobScatteringSampleLight,
 *                             JobScatteringSampleLightLI, the light sampling requests and
 *                             JobScatteringCastShadowRays; JobScatteringSampleBSDF, the BSDF
 *                             Sample jobs and JobScatteringCastShadowRaysWithInteraction (:394-433)
 *   WPBR_STAGE_SHADOW_TRACE   both shadow-ray runCollisionSystem calls (:412, :434) as one launch
 *                             over the compacted shadow-ray queue
 *   WPBR_STAGE_INTEGRATE      JobScatteringProcessShadowRay*, the BSDF Compute jobs,
 *                             JobSphereLightProcessComputeRequests, JobLightInfiniteAreaLiCompute,
 *                             JobScatteringIntegrateImportance*, JobScatteringFinish,
 *                             JobAccumalateLIFromSamples (:413-444)
 * BEGIN, then CAMERA .. INTEGRATE of pass 0, 1, .. spp - 1 leave d_film (and
 * d_sample_xyz) equal to wpbr_render's bit for bit; JobOutputFilmTitles /
 * JobOutputFilm are wpbr_film_rgb8 / wpbr_write_png on the film.  A stage out of
 * this order reads whatever the previous stages left. */
enum {
    WPBR_STAGE_BEGIN = 0, WPBR_STAGE_CAMERA = 1, WPBR_STAGE_PRIMARY_TRACE = 2, WPBR_STAGE_SCATTER = 3,
    WPBR_STAGE_SHADOW_TRACE = 4, WPBR_STAGE_INTEGRATE = 5
};
wpbr_status wpbr_render_stage(wpbr_ctx* ctx, wpbr_scene* scene, float* d_film, float* d_sample_xyz, uint32_t pass,
                              int stage);
/* overlap 1 (the default): consecutive sample passes of wpbr_render overlap on
 * two streams (pass j's camera, primary trace and shading run beside pass
 * j - 1's secondary trace and splat).  0: each pass starts after the previous
 * one has finished, so every kernel runs alone on the GPU (for per-kernel
 * timing).  Films and samples are the same bit for bit either way.  No
 * counterpart in the reference, which renders one tile at a time. */
wpbr_status wpbr_set_render_overlap(wpbr_ctx* ctx, int overlap);
/* runCollisionSystem on a scene's geometry (MEngine.cpp:352-361) with the
 * interactions it attaches to each CRayCast (CSufraceInteraction.h:36-199):
 *   d_rays      6 x n SoA floats (CRayCast::ray: origin xyz, dir xyz), tmax = inf
 *   d_surf      n bytes, CRayCast::bSurfInteraction (null: every ray 1)
 *   d_ray_diff  12 x n SoA floats, the rays' CRayDifferential (CRayDifferential.h:7-22:
 *               difXRay origin, dir, difYRay origin, dir), or null
 *   d_hit_shape, d_hit_t   closest hit as wpbr_collide (0xFFFFFFFF / +inf: a miss)
 *   d_si        WPBR_SI_FLOATS x n SoA floats (field f of ray i at d_si[f * n + i]),
 *               or null: for a hit with bSurfInteraction the CSurfaceInteraction of
 *               JobProcessTriangleFullInteractionRequests (CSTriangle.cpp:149-319) or
 *               JobProcessSphereFullInteractionRequests (CSSphere.cpp:53-186, in
 *               world space by CTransform.h:41-66), with the differentials of
 *               JobComputeDifferentialsForSurfInter (CSufraceInteraction.h:98-148,
 *               200-225) when d_ray_diff is given; for a hit without it the
 *               CInteraction of JobProcessRayCastsResults (SLBVH.cpp:55-59: p, wo,
 *               time); zeros on a miss (and in every field a job leaves default).
 * Replaces the collision jobs of wpbr_collide plus the full-interaction and
 * differential jobs. */
enum {
    WPBR_SI_P = 0, WPBR_SI_WO = 3, WPBR_SI_TIME = 6, WPBR_SI_N = 7, WPBR_SI_UV = 10,
    WPBR_SI_DPDU = 12, WPBR_SI_DPDV = 15, WPBR_SI_DNDU = 18, WPBR_SI_DNDV = 21,
    WPBR_SI_SHADING_N = 24, WPBR_SI_SHADING_DPDU = 27, WPBR_SI_SHADING_DPDV = 30,
    WPBR_SI_SHADING_DNDU = 33, WPBR_SI_SHADING_DNDV = 36, WPBR_SI_DPDX = 39, WPBR_SI_DPDY = 42,
    WPBR_SI_DUDX = 45, WPBR_SI_DVDX = 46, WPBR_SI_DUDY = 47, WPBR_SI_DVDY = 48,
    WPBR_SI_FLOATS = 49
};
wpbr_status wpbr_scene_collide(wpbr_ctx* ctx, wpbr_scene* scene, uint32_t n, const float* d_rays,
                               const uint8_t* d_surf, const float* d_ray_diff, uint32_t* d_hit_shape,
                               float* d_hit_t, float* d_si);
/* The same over bare geometry and its LBVH (as wpbr_collide), with the
 * triangles' mesh attributes as CTriangleMesh holds them after
 * STriangleMesh::create (device arrays, per triangle, triangle order):
 * d_tri_flags bit0 = the uvs are given, bit1 = the normals and tangents are
 * given (a null d_tri_flags or a null attrs: none, default uvs (0,0), (1,0),
 * (1,1) as CSTriangle.cpp:180-185). */
typedef struct {
    const uint8_t* d_tri_flags;   /* n_triangles */
    const float* d_tri_uv;        /* n_triangles x 6: uv of vertices 0, 1, 2 */
    const float* d_tri_normal;    /* n_triangles x 9 */
    const float* d_tri_tangent;   /* n_triangles x 9 */
} wpbr_tri_attributes;
wpbr_status wpbr_collide_interactions(wpbr_ctx* ctx, const wpbr_geometry* geom, const wpbr_tri_attributes* attrs,
                                      const wpbr_lbvh* bvh, uint32_t n, const float* d_rays,
                                      const uint8_t* d_surf, const float* d_ray_diff, uint32_t* d_hit_shape,
                                      float* d_hit_t, float* d_si);
/* ---- BSDF Compute / Sample jobs --------------------------------------------- *
 * The jobs the render runs on every hit's BSDF entity, over request queues:
 * JobGenerateBSDFRequests splits the hits into per-material queues (CMaterial.h:30-38)
 * and each BSDF family's Compute / Sample job runs over its own:
 *   WPBR_BSDF_CONDUCTOR_MF   JobBSDFConductorMicrofaceCompute / Sample (CBSDFConductorMicroface.cpp:84-115)
 *   WPBR_BSDF_DIELECTRIC_MF  JobBSDFDielectricMicrofaceCompute / Sample (CBSDFDielectricMicroface.cpp:84-115)
 *   WPBR_BSDF_SPECULAR_REFL  JobBSDFSpecularReflectionCompute / Sample (CBSDFSpecularReflectance.cpp:63-93)
 * A request (SoA device floats, n of them) is a BSDF entity's shading frame
 * (CBSDFTransform rows ss, ts, ns: d_frame 9 x n), wo (3 x n, world), wi
 * (Compute, 3 x n, world), the 2D sample (Sample, 2 x n) and the material index
 * into the scene's table (its type is the BSDF kind; 0xFFFFFFFF: none).
 * Results at each request's index: d_value n x 64 (16-byte aligned; a request's
 * spectrum contiguous, as the dense CSampledBSDFValue components, CBSDF.h:24-33),
 * f(wo, wi) |cos wi| per wavelength; d_pdf n, CSampledBSDFPDF; d_wi 3 x n
 * (Sample, world), CSampledWI.  A microfacet sample below the surface (the reference leaves its pdf
 * and wi unset) and a request without material give value 0, pdf 0, wi 0. */
enum { WPBR_BSDF_CONDUCTOR_MF = 0, WPBR_BSDF_DIELECTRIC_MF = 1, WPBR_BSDF_SPECULAR_REFL = 2, WPBR_BSDF_ANY = -1 };
typedef struct {
    const float* d_frame;
    const float* d_wo;
    const float* d_wi;
    const float* d_u;
    const uint32_t* d_material;
} wpbr_bsdf_requests;
/* The per-kind queues: d_queue (n entries) receives the indices of the
 * conductor requests, then the dielectric ones, then the specular ones, each in
 * request order; counts[k] their numbers (requests without material are left
 * out).  Compacted on the device with wave-ballot prefix sums. */
wpbr_status wpbr_bsdf_queues(wpbr_ctx* ctx, wpbr_scene* scene, const wpbr_bsdf_requests* req, uint32_t n,
                             uint32_t* d_queue, uint32_t counts[3]);
/* One family's Compute / Sample over the q requests listed in d_queue (all of
 * that kind), or with kind WPBR_BSDF_ANY every one of the n requests in one
 * launch, each by its own kind (d_queue and q ignored). */
wpbr_status wpbr_bsdf_compute(wpbr_ctx* ctx, wpbr_scene* scene, int kind, const wpbr_bsdf_requests* req,
                              uint32_t n, const uint32_t* d_queue, uint32_t q, float* d_value, float* d_pdf);
wpbr_status wpbr_bsdf_sample(wpbr_ctx* ctx, wpbr_scene* scene, int kind, const wpbr_bsdf_requests* req,
                             uint32_t n, const uint32_t* d_queue, uint32_t q, float* d_wi, float* d_value,
                             float* d_pdf);
/* Statistics of the last wpbr_render: [0] total ms, [1] trace ms (all trace
 * launches), [2] number of trace launches, [3] rays traced, [4] shade ms,
 * [5] camera ms, [6] secondary rays, [7] trace ms of the primary rays.
 * Returns the number written. */
int wpbr_render_stats(wpbr_ctx* ctx, double* out, int cap);
/* Autotuner for the render: renders the scene once per candidate traversal
 * configuration, keeps the one with the least traversal time in the context
 * and returns it in *best (optional).  d_film receives the last render. */
wpbr_status wpbr_autotune_render(wpbr_ctx* ctx, wpbr_scene* scene, float* d_film, wpbr_trace_config* best);

/* ---- batched many-scene render ------------------------------------------- *
 * The reference renders one scene, one 32x32 tile at a time (MEngine::render,
 * MEngine.cpp:380-448); these entry points render many scenes together.
 * wpbr_scenes_create uploads n_scenes descriptions and builds all their LBVHs
 * with one segmented build (each scene as wpbr_scene_create would make it).
 * A batch (wpbr_batch_create) holds the wavefront state of its scenes and
 * renders them with one launch per stage of every sample pass (camera,
 * primary trace, shade, secondary-ray trace, radiance + splat) over all
 * scenes' pixels, tiles and rays; the pass loop is captured in a HIP graph at
 * the first render (again after a traversal configuration change) and replayed.
 * Every film equals what wpbr_render gives for its scene alone.  All scenes of a
 * batch have the same spp; at most 65535 scenes; d_films[i] holds
 * width*height*4 floats of scene i (as wpbr_render). */
typedef struct wpbr_batch wpbr_batch;
wpbr_status wpbr_scenes_create(wpbr_ctx* ctx, uint32_t n_scenes, const wpbr_scene_desc* descs,
                               wpbr_scene** out);
wpbr_status wpbr_batch_create(wpbr_ctx* ctx, uint32_t n_scenes, wpbr_scene* const* scenes,
                              float* const* d_films, wpbr_batch** out);
void wpbr_batch_destroy(wpbr_batch* batch);
wpbr_status wpbr_batch_render(wpbr_ctx* ctx, wpbr_batch* batch);
/* [0] ms of the last render (graph replay), [1] samples, [2] scenes,
 * [3] launches per render, [4] secondary rays traced.  Returns the number written. */
int wpbr_batch_stats(wpbr_ctx* ctx, wpbr_batch* batch, double* out, int cap);
/* Autotuner for the batch: renders once per candidate traversal configuration
 * (block size, LDS stack depth, node width), keeps the fastest. */
wpbr_status wpbr_autotune_batch(wpbr_ctx* ctx, wpbr_batch* batch, wpbr_trace_config* best);

/* ---- differentiable render (no counterpart in the reference) -------------- *
 * An autograd.Function-style pair over the render: wpbr_render is the forward
 * (film of XYZ-weight sums), wpbr_render_backward the adjoint.
 * Parameters: per material its alpha (roughness) and its conductor eta and k
 * as RGB (fromRGB(Reflectance), CMaterialMetal.cpp:46), and the env map's RGB
 * texels (env_h x env_w x 3).  Gradient layout (floats): alpha[n_mat],
 * eta_k[n_mat][6] (eta RGB, k RGB), env[env_h * env_w * 3].
 * counts: [0] n_mat, [1] 6 n_mat, [2] env floats (0 without an env map). */
wpbr_status wpbr_scene_param_counts(wpbr_scene* scene, uint32_t counts[3]);
/* Replaces parameters (host arrays; a null array keeps its values). */
wpbr_status wpbr_scene_set_params(wpbr_ctx* ctx, wpbr_scene* scene, const float* alpha,
                                  const float* eta_k, const float* env_rgb);
/* The film's image, rgb = xyzToRGB(sum XYZ w) / sum w (CoefficientSpectrum.h:14-24,
 * CSFilm.cpp:141-170 before the tonemap), its L2 loss against d_target_rgb
 * (W*H*3) in *loss and the loss gradient 2 (rgb - target) in d_dL_drgb.  Any
 * output may be null. */
wpbr_status wpbr_film_loss(wpbr_ctx* ctx, wpbr_scene* scene, const float* d_film,
                           const float* d_target_rgb, float* d_dL_drgb, float* d_rgb, double* loss);
/* An RGB texture and its MIP pyramid on the device: CTexture2DRGB as
 * SLoadTexture::load builds it (CSTextureProcess.h:305-371) from texels already
 * linear (the reference applies gammaCorrectInv to its PNG's values first):
 * floor(log2(max(width, height))) - 1 levels, each max(1, half) of the previous
 * one per axis (2x2 means; the last column and row average the previous
 * level's last column / row; the last texel is copied), built on the device.
 * rgb: height x width x 3 floats, host, row-major; sides in [1, 16384], the
 * longer at least 4. */
wpbr_status wpbr_texture_create(wpbr_ctx* ctx, uint32_t width, uint32_t height, const float* rgb,
                                wpbr_texture** out);
void wpbr_texture_destroy(wpbr_texture* tex);
uint32_t wpbr_texture_levels(const wpbr_texture* tex);
/* The size of MIP level `level`. */
wpbr_status wpbr_texture_size(const wpbr_texture* tex, uint32_t level, uint32_t* width, uint32_t* height);
/* MIP level `level` to host memory (its height x width x 3 floats). */
wpbr_status wpbr_texture_level(wpbr_ctx* ctx, const wpbr_texture* tex, uint32_t level, float* rgb);
/* STexture2DSamplerAnisotropic::eval (CSTextureProcess.h:134-251) with
 * maxAnisotropy 32 (CMaterialMetal.cpp:28) and the 64-entry Gaussian table of
 * JobFilterTableGaussing (CSTextureProcess.cpp:16-32, MEngine.cpp:92-94) for n
 * texture-mapped points: d_query holds 6 floats each (s, t, ds/dx, dt/dx,
 * ds/dy, dt/dy as CTextureMappedPoint, CTextureMapping.h:17-30), d_rgb receives
 * 3 floats each. */
wpbr_status wpbr_texture_lookup(wpbr_ctx* ctx, const wpbr_texture* tex, uint32_t n, const float* d_query,
                                float* d_rgb);

/* wpbr_film_loss without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); no counterpart in the reference.  Same arguments and
 * results as wpbr_film_loss (the loss bit for bit).  Its per-block partials and
 * their sum live in a file-scope device array of the library, one copy per GPU,
 * held until the process ends and sized for films of at most 16 777 216 pixels
 * (larger films are WPBR_ERR_INVALID_ARG).  Calls that overlap on one GPU
 * share that array and race; each call overwrites what it reads. */
wpbr_status wpbr_film_loss_noalloc(wpbr_ctx* ctx, wpbr_scene* scene, const float* d_film,
                                   const float* d_target_rgb, float* d_dL_drgb, float* d_rgb, double* loss);
/* The film as an 8-bit RGB image (3 bytes per pixel, row-major, d_rgb8 on the
 * device): JobOutputFilmTitles (CSFilm.cpp:141-170) per pixel: black if the
 * accumulated radiance is black, else xyzToRGB(sum XYZ w) / (sum w * 800000),
 * clamped to [0, 1], gammaCorrect (CSFilm.cpp:129-140, slope 12.98 below
 * 0.0031308), times 255, truncated. */
wpbr_status wpbr_film_rgb8(wpbr_ctx* ctx, uint32_t width, uint32_t height, const float* d_film,
                           uint8_t* d_rgb8);
/* Writes an 8-bit RGB image (host memory, row-major) as a PNG file: the output
 * of JobOutputFilm (CSFilm.cpp:183-226, which merges the tiles and encodes with
 * lodepng).  Host code only; no context needed. */
wpbr_status wpbr_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* rgb8);

/* Gradient of the loss with respect to every parameter, given the forward's
 * film and d loss / d rgb per pixel (W*H*3): the sample passes are replayed from
 * their sample streams (nothing is kept from the forward but the film) and a
 * hand-written adjoint kernel differentiates each sample's radiance with the
 * forward's discrete decisions held fixed (visibility, light pick, env cells). */
wpbr_status wpbr_render_backward(wpbr_ctx* ctx, wpbr_scene* scene, const float* d_film,
                                 const float* d_dL_drgb, float* d_grad);

/* The backward pass's derivatives of the conductor microfacet BSDF
 * (CBSDFConductorMicroface.cpp:11-82 Compute/Sample, TR distribution
 * MicrofaceDistr.h:12-90, FrConductor CBXDF.h:200-238) for n requests of one
 * wavelength, evaluated by the same device code as wpbr_render_backward.
 * d_in: 11 x n floats SoA: wo xyz, wi xyz (shading frame), alpha, eta, k, u0, u1.
 * d_out: 13 x n floats SoA: D, dD/dalpha, pdf, dpdf/dalpha, f, df/dalpha
 * (f = Fr D / (4 |cos_i| |cos_o|) |cos_i|), Fr, dFr/deta, dFr/dk (Fr at wi.wh),
 * then for the half vector sampled from (u0, u1): wh_z, dwh_z/dalpha, and its
 * reflected direction: wi_z, dwi_z/dalpha.  No counterpart in the reference. */
wpbr_status wpbr_bsdf_derivatives(wpbr_ctx* ctx, uint32_t n, const float* d_in, float* d_out);

/* ---- scene ingestion (the formats before the path) ---------------------- */
/* An OBJ mesh as MeshTools::loadMesh reads it through assimp (MeshTools.h:14-99,
 * aiProcess_Triangulate | aiProcess_ConvertToLeftHanded): the first mesh of the
 * file (faces up to the first o / g / usemtl line after a face); every face
 * corner its own vertex; z of positions and normals negated, texture v
 * replaced by 1 - v, each face's corner order reversed, then fanned from its
 * first (reversed) corner.  Normals (uvs) are present only if every corner
 * names one.  Host code only; an unreadable or malformed file is
 * WPBR_ERR_INVALID_ARG. */
typedef struct wpbr_obj wpbr_obj;
wpbr_status wpbr_obj_load(const char* path, wpbr_obj** out);
/* counts: [0] vertices, [1] indices (3 per triangle), [2] 1 if normals, [3] 1 if uvs. */
wpbr_status wpbr_obj_counts(const wpbr_obj* obj, uint32_t counts[4]);
/* Copies the mesh to host arrays (a null array is skipped): positions and
 * normals 3 floats per vertex, uvs 2, indices u32 (CTriangleMesh::iVertices). */
wpbr_status wpbr_obj_copy(const wpbr_obj* obj, float* positions, float* normals, float* uvs, uint32_t* indices);
void wpbr_obj_free(wpbr_obj* obj);
/* A PNG file as 8-bit RGB, what lodepng_decode24_file returns to readImagePNG
 * (CSTextureProcess.h:462-496): every colour type and bit depth, Adam7 or not;
 * alpha dropped, 16-bit samples to their high byte, grey replicated (1, 2, 4-bit
 * grey scaled to 0..255), palette looked up; chunk CRCs and the zlib Adler-32
 * checked.  rgb8 null: only *width / *height; otherwise cap >= 3 w h bytes.
 * Host code only. */
wpbr_status wpbr_read_png(const char* path, uint32_t* width, uint32_t* height, uint8_t* rgb8, size_t cap);
/* STriangleMesh::create (CSTriangleMesh.cpp:11-39): the mesh to world space,
 * positions as points by the world 3x4, normals by the transpose of its
 * inverse, tangents as vectors, and the triangle vertex array wpbr_geometry
 * takes: d_tri_verts[9 i + 3 k ..] = world position of d_indices[3 i + k] (what
 * CSTriangle reads through CTriangleMesh::iVertices, CSTriangleMesh.h:24-29).
 * xform: host, 24 floats (world 3x4 row-major, then its inverse).  d_normals,
 * d_tangents and every world-space output may be null.  An index at or beyond
 * n_vertices is WPBR_ERR_INVALID_ARG (its corner is written as zeros). */
wpbr_status wpbr_mesh_bake(wpbr_ctx* ctx, const float* xform, uint32_t n_vertices, const float* d_positions,
                           const float* d_normals, const float* d_tangents, uint32_t n_triangles,
                           const uint32_t* d_indices, float* d_positions_world, float* d_normals_world,
                           float* d_tangents_world, float* d_tri_verts);
/* SLoadTexture::load's texel decode (CSTextureProcess.h:322-329 over
 * readImagePNG's values, :480-490): d_rgb[i] = gammaCorrectInv(d_rgb8[i] /
 * 255.0f) for n channel values: v / 12.92 up to 0.04045, else ((v + 0.055) /
 * 1.055)^2.4f with the power correctly rounded. */
wpbr_status wpbr_srgb8_to_linear(wpbr_ctx* ctx, size_t n, const uint8_t* d_rgb8, float* d_rgb);

#ifdef __cplusplus
}
#endif
#endif /* WPBR_H */
