// C++ host mirror of WoodenPBREngine's job-driven hot path over libwpbr_hip.so.
//
// The reference runs the path as fixed sequences of ECS jobs, each started by
// JOB_RUN (construct, set the temp allocator, run, reset; MEngine.cpp:44-50):
// MEngine::buildLBVH (MEngine.cpp:341-350), MEngine::runCollisionSystem
// (MEngine.cpp:352-361) and MEngine::render (MEngine.cpp:380-448).  This header
// keeps those job names, their order and the component layouts for a C++ host:
// every Job's run() issues the device work of the reference job through the C
// ABI in wpbr.h, on the engine's context.  Jobs whose work the GPU fuses into a
// neighbour (the traversal intersects candidates while it walks the tree; one
// render stage runs a run of scattering jobs) are kept as empty steps, so a
// caller's JOB_RUN sequence stays exactly the reference's.
//
// Header-only C++17; link with -lwpbr_hip.  Errors throw wpbr_host::Error
// (the C ABI itself never throws).
#pragma once
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "wpbr.h"

namespace wpbr_host {

// ---- component layouts (field order as in the reference) -------------------
struct DVector3f {
    float x, y, z;
};
using DPoint3f = DVector3f;
struct DBounds3f {  // WoodenMath bounds: pMin, pMax
    DPoint3f pMin, pMax;
};
struct DRayf {  // CRay.h:8-20 (origin, direction, t)
    DPoint3f origin;
    DVector3f dir;
    float t;
};
struct CRayCast {  // CSufraceInteraction.h:36-42
    DRayf ray;
    bool bSurfInteraction;
};
struct CInteraction {  // CSufraceInteraction.h:44-51
    DPoint3f p;
    DVector3f wo;
    uint32_t hCollision;  // shape entity, INVALID_HANDLE on a miss
    float time;
};
using DNormal3f = DVector3f;
struct DVector2f {
    float x, y;
};
// CSufraceInteraction.h:54-199 (its members in declaration order).
struct CSurfaceInteraction : CInteraction {
    DNormal3f n;
    DVector2f uv;
    DVector3f dpdu, dpdv;
    DNormal3f dndu, dndv;
    struct {
        DNormal3f n;
        DVector3f dpdu, dpdv;
        DVector3f dndu, dndv;
    } shading;
    DVector3f dpdx, dpdy;
    float dudx = 0, dvdx = 0, dudy = 0, dvdy = 0;
};
// CRayDifferential.h:7-22 over DRayDifferential.h:8-31: the ray and its x / y
// neighbours.
struct CRayDifferential : DRayf {
    DRayf difXRay, difYRay;
};
struct LinearBVHNode {  // SLBVH.h:17-27
    uint32_t offset;    // shapes offset (leaf) or second child offset (interior)
    uint32_t nShapes;   // 0 for interior nodes
    uint8_t axis;
};
struct CSceneTreeLBVH {  // SLBVH.h:28-40
    std::vector<LinearBVHNode> nodes;
    std::vector<DBounds3f> nodesBounds;
    std::vector<uint32_t> shapesEntities;
};
constexpr uint32_t INVALID_HANDLE = 0xFFFFFFFFu;

// CSTriangleMesh.h: per-vertex arrays (3 floats; uvs 2) and iVertices.
struct CTriangleMesh {
    std::vector<float> positions, normals, tangents, uvs;
    std::vector<uint32_t> iVertices;
    uint32_t nTriangles = 0, nVertices = 0;
};

// ---- render components (CBSDF.h, SScattering.h) ---------------------------------
constexpr int kSpectralSamples = 64;  // the device's SampledSpectrum bins
struct Spectrum {
    float c[kSpectralSamples];
};
struct CSampledBSDFValue : Spectrum {};  // CBSDF.h:24-33
struct CSampledBSDFPDF {                 // CBSDF.h:18-22
    float p;
};
struct CSampledWI : DVector3f {};  // SScattering.h:20-28
// CBSDF.h:36-43: the BSDF entity's world-to-shading transform; its rotation
// rows ss (normalised shading dpdu), ts = cross(ns, ss), ns (CBSDF.cpp:12-24).
struct CBSDFTransform {
    DVector3f ss, ts, ns;
};
// What one BSDF entity's Compute / Sample job reads (CBSDFConductorMicroface.cpp:84-115
// and its dielectric and specular siblings): its transform, wo, wi (Compute) or
// the 2D sample (Sample), and its material (the material's type is the family;
// INVALID_HANDLE: none).
struct CBSDFRequest {
    CBSDFTransform transform;
    DVector3f wo, wi;
    DVector2f u;
    uint32_t material;
};

class Error : public std::runtime_error {
public:
    Error(wpbr_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    wpbr_status status;
};

class MEngine;

// ---- MEngine::buildLBVH jobs (SLBVH.cpp, CSTriangle.cpp, CSSphere.h) -------
struct JobUpdateBoundsAndCentroidSphere { void run(MEngine* e); };    // CSSphere.h:126-153
struct JobUpdateBoundsAndCentroidTriangle { void run(MEngine* e); };  // CSTriangle.cpp:116-142
struct JobGenerateShapesCentroidBound { void run(MEngine* e); };      // SLBVH.cpp:164-181
struct JobGenerateShapesMortonCode { void run(MEngine* e); };         // SLBVH.cpp:186-227
struct JobSortsShapesByMortonCode { void run(MEngine* e); };          // SLBVH.cpp:231-324
struct JobEmitLBVH { void run(MEngine* e); };                         // SLBVH.cpp:328-406
struct JobBuildUpperSAH { void run(MEngine* e); };                    // SLBVH.cpp:410-563

// ---- MEngine::runCollisionSystem jobs ----------------------------------------
struct JobProcessRayCasts { void run(MEngine* e); };                       // SLBVH.cpp:85-162
struct JobProcessSphereSurfInteractionRequests { void run(MEngine*) {} };   // in the traversal
struct JobProcessTriangleInteractionRequests { void run(MEngine*) {} };     // in the traversal
struct JobProcessRayCastsResults { void run(MEngine* e); };                // SLBVH.cpp:9-66
struct JobProcessSphereFullInteractionRequests { void run(MEngine* e); };   // CSSphere.cpp:53-186
struct JobProcessTriangleFullInteractionRequests { void run(MEngine* e); }; // CSTriangle.cpp:149-319
struct JobCollisionFinish { void run(MEngine* e); };                       // SLBVH.cpp:72-82
// CSufraceInteraction.h:200-225 (the differentials come with the interactions
// from the device; this step moves them into surfaceInteractions)
struct JobComputeDifferentialsForSurfInter { void run(MEngine* e); };

// ---- MEngine::render jobs (MEngine.cpp:380-448) ----------------------------------
// The device renders every film tile at once, in five launches per sample pass
// (wpbr_render_stage); the job that starts a stage runs it, the jobs it covers
// are empty steps.
#define WPBR_FUSED_JOB(JobT) \
    struct JobT {            \
        void run(MEngine*) {} \
    };
struct JobGenerateFilmTiles { void run(MEngine* e); };                     // CSFilm.cpp: all tiles, film cleared
WPBR_FUSED_JOB(JobCreateCameraSamples)                                     // counter-based sample streams
WPBR_FUSED_JOB(JobSamplerStratifiedGenerateSampels1D)                      // values computed where consumed
WPBR_FUSED_JOB(JobSamplerStratifiedGenerateSampels2D)
WPBR_FUSED_JOB(JobSamplerUpdateCameraSamples)                              // in the camera stage
struct JobCameraPerspGenerateRaysDifferential { void run(MEngine* e); };   // CSCameraPerspective.cpp:9-52
struct JobScatteringRequestEmittedLight { void run(MEngine* e); };         // the scatter stage, SScattering.cpp:20-276
WPBR_FUSED_JOB(JobLightInfiniteAreaLeCompute)
WPBR_FUSED_JOB(JobScatteringAccumulateEmittedLight)
WPBR_FUSED_JOB(JobMapUVRequestsGenerate)
WPBR_FUSED_JOB(JobSphereProcessMapUVRequests)
WPBR_FUSED_JOB(JobTriangleProcessMapUVRequests)
struct JobGenerateBSDFRequests { void run(MEngine* e); };                  // CMaterial.h:30-38
WPBR_FUSED_JOB(JobGenerateBSDFMaterialMetal)
WPBR_FUSED_JOB(JobGenerateBSDFMaterialDielectric)
WPBR_FUSED_JOB(JobGenerateBSDFMaterialPerfectGlass)
WPBR_FUSED_JOB(JobBSDFComputeTransform)
WPBR_FUSED_JOB(JobScatteringSampleLight)
WPBR_FUSED_JOB(JobScatteringSampleLightLI)
WPBR_FUSED_JOB(JobSphereLightProcessSamplingRequests)
WPBR_FUSED_JOB(JobLightInfiniteAreaLiSample)
WPBR_FUSED_JOB(JobScatteringCastShadowRays)
struct JobScatteringProcessShadowRay { void run(MEngine* e); };            // the integrate stage
struct JobBSDFConductorMicrofaceCompute { void run(MEngine* e); };         // CBSDFConductorMicroface.cpp:84-99
struct JobBSDFSpecularReflectionCompute { void run(MEngine* e); };         // CBSDFSpecularReflectance.cpp:63-76
struct JobBSDFDielectricMicrofaceCompute { void run(MEngine* e); };        // CBSDFDielectricMicroface.cpp:84-99
WPBR_FUSED_JOB(JobScatteringIntegrateImportanceLight)
WPBR_FUSED_JOB(JobScatteringSampleBSDF)
struct JobBSDFConductorMicrofaceSample { void run(MEngine* e); };          // CBSDFConductorMicroface.cpp:100-115
struct JobBSDFSpecularReflectionSample { void run(MEngine* e); };          // CBSDFSpecularReflectance.cpp:77-93
struct JobBSDFDielectricMicrofaceSample { void run(MEngine* e); };         // CBSDFDielectricMicroface.cpp:100-115
WPBR_FUSED_JOB(JobScatteringCastShadowRaysWithInteraction)
WPBR_FUSED_JOB(JobScatteringProcessShadowRayWithInteraction)
WPBR_FUSED_JOB(JobSphereLightProcessComputeRequests)
WPBR_FUSED_JOB(JobLightInfiniteAreaLiCompute)
WPBR_FUSED_JOB(JobScatteringIntegrateImportanceBSDF)
WPBR_FUSED_JOB(JobScatteringFinish)
WPBR_FUSED_JOB(JobAccumalateLIFromSamples)
struct JobOutputFilmTitles { void run(MEngine* e); };                      // CSFilm.cpp:141-170
struct JobOutputFilm { void run(MEngine* e); };                            // CSFilm.cpp:183-226
#undef WPBR_FUSED_JOB

#define WPBR_JOB_RUN(JobT) \
    {                      \
        JobT j;            \
        j.run(this);       \
    }

class MEngine {
public:
    explicit MEngine(int device = 0) {
        check(wpbr_open(device, &ctx_), "wpbr_open");
        check(wpbr_lbvh_create(ctx_, &bvh_), "wpbr_lbvh_create");
    }
    ~MEngine() {
        release_shapes();
        release_rays();
        release_attributes();
        release_scene();
        release_bsdf();
        if (bvh_) wpbr_lbvh_destroy(bvh_);
        if (ctx_) wpbr_close(ctx_);
    }
    MEngine(const MEngine&) = delete;
    MEngine& operator=(const MEngine&) = delete;

    // Shapes in the reference's order: spheres (world 3x4 then inverse 3x4,
    // radius), then triangles (world-space vertices, 9 floats).
    void setShapes(const std::vector<float>& sphereXform24, const std::vector<float>& sphereRadius,
                   const std::vector<float>& triVerts9) {
        release_shapes();
        geom_.n_spheres = (uint32_t)sphereRadius.size();
        geom_.n_triangles = (uint32_t)(triVerts9.size() / 9);
        geom_.d_sphere_xform = upload(sphereXform24, &d_xform_);
        geom_.d_sphere_radius = upload(sphereRadius, &d_radius_);
        geom_.d_tri_verts = upload(triVerts9, &d_tris_);
    }

    // Per-triangle mesh attributes (triangle order, as setShapes' triangles):
    // flags bit0 uvs given, bit1 normals and tangents given; uv 6 floats,
    // normal and tangent 9 floats per triangle.  Empty vectors: none (default
    // uvs, geometric shading frame).
    void setTriangleAttributes(const std::vector<uint8_t>& flags, const std::vector<float>& uv6,
                               const std::vector<float>& normal9, const std::vector<float>& tangent9) {
        release_attributes();
        if (flags.empty()) return;
        attrs_.d_tri_flags = (const uint8_t*)upload_bytes(flags.data(), flags.size(), &d_flags_);
        attrs_.d_tri_uv = (const float*)upload_bytes(uv6.data(), uv6.size() * 4, &d_uv_);
        attrs_.d_tri_normal = (const float*)upload_bytes(normal9.data(), normal9.size() * 4, &d_nrm_);
        attrs_.d_tri_tangent = (const float*)upload_bytes(tangent9.data(), tangent9.size() * 4, &d_tan_);
    }

    // STriangleMesh::create (CSTriangleMesh.cpp:11-39) on the device: the mesh
    // to world space (world24 = world 3x4 then its inverse) and its triangles
    // appended to triVerts9, the array setShapes takes.
    void bakeTriangleMesh(const float* world24, CTriangleMesh& mesh, std::vector<float>& triVerts9) {
        void *d_pos = nullptr, *d_nrm = nullptr, *d_tan = nullptr, *d_idx = nullptr, *d_tri = nullptr;
        void *d_pw = nullptr, *d_nw = nullptr, *d_tw = nullptr;
        const bool hasN = !mesh.normals.empty(), hasT = !mesh.tangents.empty();
        const size_t vb = mesh.positions.size() * 4, tb = (size_t)mesh.nTriangles * 36;
        auto up = [&](const void* src, size_t bytes, void** d) {
            check(wpbr_malloc(ctx_, bytes, d), "wpbr_malloc");
            check(wpbr_upload(ctx_, *d, src, bytes), "wpbr_upload");
        };
        auto release = [&] {
            for (void* d : {d_pos, d_nrm, d_tan, d_idx, d_tri, d_pw, d_nw, d_tw})
                if (d) wpbr_free(ctx_, d);
        };
        try {
            up(mesh.positions.data(), vb, &d_pos);
            up(mesh.iVertices.data(), mesh.iVertices.size() * 4, &d_idx);
            if (hasN) up(mesh.normals.data(), vb, &d_nrm);
            if (hasT) up(mesh.tangents.data(), vb, &d_tan);
            check(wpbr_malloc(ctx_, vb, &d_pw), "wpbr_malloc");
            if (hasN) check(wpbr_malloc(ctx_, vb, &d_nw), "wpbr_malloc");
            if (hasT) check(wpbr_malloc(ctx_, vb, &d_tw), "wpbr_malloc");
            check(wpbr_malloc(ctx_, tb, &d_tri), "wpbr_malloc");
            check(wpbr_mesh_bake(ctx_, world24, mesh.nVertices, (const float*)d_pos, (const float*)d_nrm,
                                 (const float*)d_tan, mesh.nTriangles, (const uint32_t*)d_idx, (float*)d_pw,
                                 (float*)d_nw, (float*)d_tw, (float*)d_tri),
                  "wpbr_mesh_bake");
            check(wpbr_download(ctx_, mesh.positions.data(), d_pw, vb), "wpbr_download");
            if (hasN) check(wpbr_download(ctx_, mesh.normals.data(), d_nw, vb), "wpbr_download");
            if (hasT) check(wpbr_download(ctx_, mesh.tangents.data(), d_tw, vb), "wpbr_download");
            const size_t at = triVerts9.size();
            triVerts9.resize(at + (size_t)mesh.nTriangles * 9);
            check(wpbr_download(ctx_, triVerts9.data() + at, d_tri, tb), "wpbr_download");
        } catch (...) {
            release();
            throw;
        }
        release();
    }

    // MEngine.cpp:341-350
    void buildLBVH() {
        WPBR_JOB_RUN(JobUpdateBoundsAndCentroidSphere)
        WPBR_JOB_RUN(JobUpdateBoundsAndCentroidTriangle)
        WPBR_JOB_RUN(JobGenerateShapesCentroidBound)
        WPBR_JOB_RUN(JobGenerateShapesMortonCode)
        WPBR_JOB_RUN(JobSortsShapesByMortonCode)
        WPBR_JOB_RUN(JobEmitLBVH)
        WPBR_JOB_RUN(JobBuildUpperSAH)
    }

    // MEngine.cpp:352-361: rayCasts (and optionally rayDifferentials, one per
    // ray cast) in; interactions out (one per ray cast), and for every ray cast
    // with bSurfInteraction that hits, its surfaceInteractions entry.
    void runCollisionSystem() {
        WPBR_JOB_RUN(JobProcessRayCasts)
        WPBR_JOB_RUN(JobProcessSphereSurfInteractionRequests)
        WPBR_JOB_RUN(JobProcessTriangleInteractionRequests)
        WPBR_JOB_RUN(JobProcessRayCastsResults)
        WPBR_JOB_RUN(JobProcessSphereFullInteractionRequests)
        WPBR_JOB_RUN(JobProcessTriangleFullInteractionRequests)
        WPBR_JOB_RUN(JobCollisionFinish)
    }

    // A render scene (wpbr_scene_create: shapes, per-triangle attributes,
    // materials, sphere lights, env map, camera, film size, spp, sampler), with
    // its own LBVH built on upload.
    void setScene(const wpbr_scene_desc& desc) {
        release_scene();
        check(wpbr_scene_create(ctx_, &desc, &scene_), "wpbr_scene_create");
        width_ = desc.width;
        height_ = desc.height;
        spp_ = desc.spp;
        check(wpbr_malloc(ctx_, 16 * (size_t)width_ * height_, &d_film_), "wpbr_malloc");
        check(wpbr_malloc(ctx_, 3 * (size_t)width_ * height_, &d_rgb8_), "wpbr_malloc");
    }

    // MEngine.cpp:380-448.  The reference's outer loop renders one 32x32 tile
    // per iteration; here one iteration covers every tile.  Outputs: film
    // (width*height*4 floats: the Mitchell-weighted X, Y, Z sums and the weight
    // sum), filmRGB8 (JobOutputFilmTitles) and, when outputFilename is set, the
    // PNG file (JobOutputFilm).
    void render() {
        if (!scene_) throw Error(WPBR_ERR_INVALID_ARG, "render: no scene (setScene)");
        rendering_ = true;
        try {
            WPBR_JOB_RUN(JobGenerateFilmTiles)
            WPBR_JOB_RUN(JobCreateCameraSamples)
            WPBR_JOB_RUN(JobSamplerStratifiedGenerateSampels1D)
            WPBR_JOB_RUN(JobSamplerStratifiedGenerateSampels2D)
            for (pass_ = 0; pass_ < spp_; pass_++) {
                WPBR_JOB_RUN(JobSamplerUpdateCameraSamples)
                WPBR_JOB_RUN(JobCameraPerspGenerateRaysDifferential)
                renderCollisionSystem(0);
                WPBR_JOB_RUN(JobScatteringRequestEmittedLight)
                WPBR_JOB_RUN(JobLightInfiniteAreaLeCompute)
                WPBR_JOB_RUN(JobScatteringAccumulateEmittedLight)
                WPBR_JOB_RUN(JobComputeDifferentialsForSurfInter)
                WPBR_JOB_RUN(JobMapUVRequestsGenerate)
                WPBR_JOB_RUN(JobSphereProcessMapUVRequests)
                WPBR_JOB_RUN(JobTriangleProcessMapUVRequests)
                WPBR_JOB_RUN(JobGenerateBSDFRequests)
                WPBR_JOB_RUN(JobGenerateBSDFMaterialMetal)
                WPBR_JOB_RUN(JobGenerateBSDFMaterialDielectric)
                WPBR_JOB_RUN(JobGenerateBSDFMaterialPerfectGlass)
                WPBR_JOB_RUN(JobBSDFComputeTransform)
                WPBR_JOB_RUN(JobScatteringSampleLight)
                WPBR_JOB_RUN(JobScatteringSampleLightLI)
                WPBR_JOB_RUN(JobSphereLightProcessSamplingRequests)
                WPBR_JOB_RUN(JobLightInfiniteAreaLiSample)
                WPBR_JOB_RUN(JobScatteringCastShadowRays)
                renderCollisionSystem(1);
                WPBR_JOB_RUN(JobScatteringProcessShadowRay)
                WPBR_JOB_RUN(JobBSDFConductorMicrofaceCompute)
                WPBR_JOB_RUN(JobBSDFSpecularReflectionCompute)
                WPBR_JOB_RUN(JobBSDFDielectricMicrofaceCompute)
                WPBR_JOB_RUN(JobScatteringIntegrateImportanceLight)
                WPBR_JOB_RUN(JobScatteringSampleBSDF)
                WPBR_JOB_RUN(JobBSDFConductorMicrofaceSample)
                WPBR_JOB_RUN(JobBSDFSpecularReflectionSample)
                WPBR_JOB_RUN(JobBSDFDielectricMicrofaceSample)
                WPBR_JOB_RUN(JobScatteringCastShadowRaysWithInteraction)
                renderCollisionSystem(2);
                WPBR_JOB_RUN(JobScatteringProcessShadowRayWithInteraction)
                WPBR_JOB_RUN(JobSphereLightProcessComputeRequests)
                WPBR_JOB_RUN(JobLightInfiniteAreaLiCompute)
                WPBR_JOB_RUN(JobScatteringIntegrateImportanceBSDF)
                WPBR_JOB_RUN(JobScatteringFinish)
                WPBR_JOB_RUN(JobAccumalateLIFromSamples)
            }
            WPBR_JOB_RUN(JobOutputFilmTitles)
            WPBR_JOB_RUN(JobOutputFilm)
        } catch (...) {
            rendering_ = false;
            throw;
        }
        rendering_ = false;
    }
    std::vector<float> film;         // after render(): width*height*4
    std::vector<uint8_t> filmRGB8;   // after render(): width*height*3
    std::string outputFilename;      // JobOutputFilm's file (empty: none)

    // The BSDF jobs over host components, outside the render: the requests in
    // bsdfRequests, their queues (JobGenerateBSDFRequests), then each family's
    // Compute (MEngine.cpp:415-417 order) or Sample (:421-424) job.  Outputs per
    // request: sampledBSDFValues, sampledBSDFPDFs and (Sample) sampledWIs; a
    // request without material, or a rejected microfacet sample, gets zeros.
    void computeBSDFs() {
        WPBR_JOB_RUN(JobGenerateBSDFRequests)
        WPBR_JOB_RUN(JobBSDFConductorMicrofaceCompute)
        WPBR_JOB_RUN(JobBSDFSpecularReflectionCompute)
        WPBR_JOB_RUN(JobBSDFDielectricMicrofaceCompute)
        fetch_bsdf(false);
    }
    void sampleBSDFs() {
        WPBR_JOB_RUN(JobGenerateBSDFRequests)
        WPBR_JOB_RUN(JobBSDFConductorMicrofaceSample)
        WPBR_JOB_RUN(JobBSDFSpecularReflectionSample)
        WPBR_JOB_RUN(JobBSDFDielectricMicrofaceSample)
        fetch_bsdf(true);
    }
    std::vector<CBSDFRequest> bsdfRequests;
    std::vector<CSampledBSDFValue> sampledBSDFValues;
    std::vector<CSampledBSDFPDF> sampledBSDFPDFs;
    std::vector<CSampledWI> sampledWIs;
    uint32_t bsdfQueueCounts[3] = {0, 0, 0};  // conductor, dielectric, specular

    // The built tree as the reference's CSceneTreeLBVH.
    CSceneTreeLBVH tree() {
        uint32_t c[4];
        check(wpbr_lbvh_counts(bvh_, c), "wpbr_lbvh_counts");
        const uint32_t n = c[0], nn = c[1];
        std::vector<uint32_t> off(nn), ns(nn);
        std::vector<uint8_t> axis(nn);
        CSceneTreeLBVH t;
        t.nodesBounds.resize(nn);
        t.shapesEntities.resize(n);
        check(wpbr_lbvh_download(ctx_, bvh_, 3, off.data(), 4 * (size_t)nn), "download");
        check(wpbr_lbvh_download(ctx_, bvh_, 4, ns.data(), 4 * (size_t)nn), "download");
        check(wpbr_lbvh_download(ctx_, bvh_, 5, axis.data(), nn), "download");
        check(wpbr_lbvh_download(ctx_, bvh_, 6, t.nodesBounds.data(), 24 * (size_t)nn), "download");
        check(wpbr_lbvh_download(ctx_, bvh_, 2, t.shapesEntities.data(), 4 * (size_t)n), "download");
        t.nodes.resize(nn);
        for (uint32_t i = 0; i < nn; i++) t.nodes[i] = {off[i], ns[i], axis[i]};
        return t;
    }

    std::vector<CRayCast> rayCasts;                   // input of runCollisionSystem
    std::vector<CRayDifferential> rayDifferentials;   // optional input: empty, or one per ray cast
    std::vector<CInteraction> interactions;           // output, one per ray cast
    std::vector<CSurfaceInteraction> surfaceInteractions;  // output, one per ray cast (bSurfInteraction hits)

    wpbr_ctx* ctx() const { return ctx_; }
    wpbr_lbvh* lbvh() const { return bvh_; }
    const wpbr_geometry& geometry() const { return geom_; }

    // used by the jobs
    void build_stage(int stage) { check(wpbr_lbvh_build_stage(ctx_, &geom_, bvh_, stage), "wpbr_lbvh_build_stage"); }
    bool rendering() const { return rendering_; }
    void render_stage(int stage) {
        check(wpbr_render_stage(ctx_, scene_, (float*)d_film_, nullptr, pass_, stage), "wpbr_render_stage");
    }
    // The render's three runCollisionSystem calls (MEngine.cpp:393, 412, 434): the
    // camera rays, then both shadow-ray sets in one launch over their queue.
    void renderCollisionSystem(int which) {
        if (which == 0) render_stage(WPBR_STAGE_PRIMARY_TRACE);
        else if (which == 1) render_stage(WPBR_STAGE_SHADOW_TRACE);
    }
    void output_film_tiles() {
        const size_t npix = (size_t)width_ * height_;
        check(wpbr_film_rgb8(ctx_, width_, height_, (const float*)d_film_, (uint8_t*)d_rgb8_), "wpbr_film_rgb8");
        film.resize(4 * npix);
        filmRGB8.resize(3 * npix);
        check(wpbr_download(ctx_, film.data(), d_film_, 16 * npix), "download film");
        check(wpbr_download(ctx_, filmRGB8.data(), d_rgb8_, 3 * npix), "download 8-bit film");
    }
    void output_film() {
        if (!outputFilename.empty())
            check(wpbr_write_png(outputFilename.c_str(), width_, height_, filmRGB8.data()), "wpbr_write_png");
    }
    void bsdf_queues() {
        if (!scene_) throw Error(WPBR_ERR_INVALID_ARG, "BSDF jobs: no scene (setScene)");
        const uint32_t n = (uint32_t)bsdfRequests.size();
        release_bsdf();
        n_bsdf_ = n;
        if (!n) {
            bsdfQueueCounts[0] = bsdfQueueCounts[1] = bsdfQueueCounts[2] = 0;
            return;
        }
        std::vector<float> frame(9 * (size_t)n), wo(3 * (size_t)n), wi(3 * (size_t)n), u(2 * (size_t)n);
        std::vector<uint32_t> mat(n);
        for (uint32_t i = 0; i < n; i++) {
            const CBSDFRequest& r = bsdfRequests[i];
            const DVector3f rows[3] = {r.transform.ss, r.transform.ts, r.transform.ns};
            for (int k = 0; k < 3; k++) {
                frame[(size_t)(3 * k) * n + i] = rows[k].x;
                frame[(size_t)(3 * k + 1) * n + i] = rows[k].y;
                frame[(size_t)(3 * k + 2) * n + i] = rows[k].z;
            }
            wo[i] = r.wo.x; wo[n + i] = r.wo.y; wo[2 * (size_t)n + i] = r.wo.z;
            wi[i] = r.wi.x; wi[n + i] = r.wi.y; wi[2 * (size_t)n + i] = r.wi.z;
            u[i] = r.u.x; u[n + i] = r.u.y;
            mat[i] = r.material;
        }
        req_.d_frame = (const float*)upload_bytes(frame.data(), frame.size() * 4, &d_bframe_);
        req_.d_wo = (const float*)upload_bytes(wo.data(), wo.size() * 4, &d_bwo_);
        req_.d_wi = (const float*)upload_bytes(wi.data(), wi.size() * 4, &d_bwi_);
        req_.d_u = (const float*)upload_bytes(u.data(), u.size() * 4, &d_bu_);
        req_.d_material = (const uint32_t*)upload_bytes(mat.data(), mat.size() * 4, &d_bmat_);
        const std::vector<float> zeros((size_t)kSpectralSamples * n, 0.0f);
        upload_bytes(zeros.data(), zeros.size() * 4, &d_bvalue_);
        upload_bytes(zeros.data(), 4 * (size_t)n, &d_bpdf_);
        upload_bytes(zeros.data(), 12 * (size_t)n, &d_bwi_out_);
        check(wpbr_malloc(ctx_, 4 * (size_t)n, &d_bqueue_), "wpbr_malloc");
        check(wpbr_bsdf_queues(ctx_, scene_, &req_, n, (uint32_t*)d_bqueue_, bsdfQueueCounts), "wpbr_bsdf_queues");
    }
    void bsdf_job(int kind, bool sample) {
        const uint32_t q = bsdfQueueCounts[kind];
        if (!q) return;
        uint32_t first = 0;
        for (int k = 0; k < kind; k++) first += bsdfQueueCounts[k];
        const uint32_t* queue = (const uint32_t*)d_bqueue_ + first;
        if (sample)
            check(wpbr_bsdf_sample(ctx_, scene_, kind, &req_, n_bsdf_, queue, q, (float*)d_bwi_out_,
                                   (float*)d_bvalue_, (float*)d_bpdf_),
                  "wpbr_bsdf_sample");
        else
            check(wpbr_bsdf_compute(ctx_, scene_, kind, &req_, n_bsdf_, queue, q, (float*)d_bvalue_,
                                    (float*)d_bpdf_),
                  "wpbr_bsdf_compute");
    }
    void fetch_bsdf(bool sample) {
        const uint32_t n = n_bsdf_;
        sampledBSDFValues.resize(n);
        sampledBSDFPDFs.resize(n);
        sampledWIs.assign(sample ? n : 0, CSampledWI{});
        if (!n) return;
        check(wpbr_download(ctx_, sampledBSDFValues.data(), d_bvalue_, (size_t)kSpectralSamples * 4 * n), "download");
        check(wpbr_download(ctx_, sampledBSDFPDFs.data(), d_bpdf_, 4 * (size_t)n), "download");
        if (sample) {
            std::vector<float> w(3 * (size_t)n);
            check(wpbr_download(ctx_, w.data(), d_bwi_out_, w.size() * 4), "download");
            for (uint32_t i = 0; i < n; i++) {
                sampledWIs[i].x = w[i];
                sampledWIs[i].y = w[n + i];
                sampledWIs[i].z = w[2 * (size_t)n + i];
            }
        }
    }
    void trace_rays() {
        const uint32_t n = (uint32_t)rayCasts.size();
        const bool diff = !rayDifferentials.empty();
        if (diff && rayDifferentials.size() != n)
            throw Error(WPBR_ERR_INVALID_ARG, "rayDifferentials: none or one per ray cast");
        std::vector<float> soa(6 * (size_t)n), dsoa(diff ? 12 * (size_t)n : 0);
        std::vector<uint8_t> surf(n);
        for (uint32_t i = 0; i < n; i++) {
            const DRayf& r = rayCasts[i].ray;
            const float v[6] = {r.origin.x, r.origin.y, r.origin.z, r.dir.x, r.dir.y, r.dir.z};
            for (int k = 0; k < 6; k++) soa[(size_t)k * n + i] = v[k];
            surf[i] = rayCasts[i].bSurfInteraction ? 1 : 0;
            if (diff) {
                const DRayf& x = rayDifferentials[i].difXRay;
                const DRayf& y = rayDifferentials[i].difYRay;
                const float w[12] = {x.origin.x, x.origin.y, x.origin.z, x.dir.x, x.dir.y, x.dir.z,
                                     y.origin.x, y.origin.y, y.origin.z, y.dir.x, y.dir.y, y.dir.z};
                for (int k = 0; k < 12; k++) dsoa[(size_t)k * n + i] = w[k];
            }
        }
        ensure_rays(n);
        if (n) {
            check(wpbr_upload(ctx_, d_rays_, soa.data(), soa.size() * 4), "upload rays");
            check(wpbr_upload(ctx_, d_surf_, surf.data(), n), "upload bSurfInteraction");
            if (diff) check(wpbr_upload(ctx_, d_diff_, dsoa.data(), dsoa.size() * 4), "upload ray differentials");
        }
        check(wpbr_collide_interactions(ctx_, &geom_, attrs_.d_tri_flags ? &attrs_ : nullptr, bvh_, n,
                                        (const float*)d_rays_, (const uint8_t*)d_surf_,
                                        diff ? (const float*)d_diff_ : nullptr, (uint32_t*)d_hit_, (float*)d_t_,
                                        (float*)d_si_),
              "wpbr_collide_interactions");
        n_traced_ = n;
    }
    void fetch_hits() {
        hit_.resize(n_traced_);
        t_.resize(n_traced_);
        si_.resize((size_t)WPBR_SI_FLOATS * n_traced_);
        if (n_traced_) {
            check(wpbr_download(ctx_, hit_.data(), d_hit_, 4 * (size_t)n_traced_), "download hits");
            check(wpbr_download(ctx_, t_.data(), d_t_, 4 * (size_t)n_traced_), "download t");
            check(wpbr_download(ctx_, si_.data(), d_si_, si_.size() * 4), "download interactions");
        }
        surfaceInteractions.assign(n_traced_, CSurfaceInteraction{});
    }
    // The surface interactions of the ray casts with bSurfInteraction that hit
    // a sphere (spheres = true) or a triangle.
    void fill_surface_interactions(bool spheres) {
        const uint32_t n = n_traced_;
        auto f = [&](int field, uint32_t i) { return si_[(size_t)field * n + i]; };
        auto v3 = [&](int field, uint32_t i) { return DVector3f{f(field, i), f(field + 1, i), f(field + 2, i)}; };
        for (uint32_t i = 0; i < n; i++) {
            if (hit_[i] == INVALID_HANDLE || !rayCasts[i].bSurfInteraction) continue;
            if ((hit_[i] < geom_.n_spheres) != spheres) continue;
            CSurfaceInteraction& s = surfaceInteractions[i];
            s.p = v3(WPBR_SI_P, i);
            s.wo = v3(WPBR_SI_WO, i);
            s.hCollision = hit_[i];
            s.time = f(WPBR_SI_TIME, i);
            s.n = v3(WPBR_SI_N, i);
            s.uv = {f(WPBR_SI_UV, i), f(WPBR_SI_UV + 1, i)};
            s.dpdu = v3(WPBR_SI_DPDU, i);
            s.dpdv = v3(WPBR_SI_DPDV, i);
            s.dndu = v3(WPBR_SI_DNDU, i);
            s.dndv = v3(WPBR_SI_DNDV, i);
            s.shading.n = v3(WPBR_SI_SHADING_N, i);
            s.shading.dpdu = v3(WPBR_SI_SHADING_DPDU, i);
            s.shading.dpdv = v3(WPBR_SI_SHADING_DPDV, i);
            s.shading.dndu = v3(WPBR_SI_SHADING_DNDU, i);
            s.shading.dndv = v3(WPBR_SI_SHADING_DNDV, i);
        }
    }
    void fill_differentials() {
        const uint32_t n = n_traced_;
        if (rayDifferentials.empty()) return;
        auto f = [&](int field, uint32_t i) { return si_[(size_t)field * n + i]; };
        for (uint32_t i = 0; i < n; i++) {
            if (hit_[i] == INVALID_HANDLE || !rayCasts[i].bSurfInteraction) continue;
            CSurfaceInteraction& s = surfaceInteractions[i];
            s.dpdx = {f(WPBR_SI_DPDX, i), f(WPBR_SI_DPDX + 1, i), f(WPBR_SI_DPDX + 2, i)};
            s.dpdy = {f(WPBR_SI_DPDY, i), f(WPBR_SI_DPDY + 1, i), f(WPBR_SI_DPDY + 2, i)};
            s.dudx = f(WPBR_SI_DUDX, i);
            s.dvdx = f(WPBR_SI_DVDX, i);
            s.dudy = f(WPBR_SI_DUDY, i);
            s.dvdy = f(WPBR_SI_DVDY, i);
        }
    }
    // SLBVH.cpp:40-62: a miss keeps INVALID_HANDLE; a hit records p = o + d t,
    // wo = -d, the shape and t (the device wrote them for rays without
    // bSurfInteraction; the surface ones take the same fields).
    void finish_interactions() {
        interactions.resize(n_traced_);
        for (uint32_t i = 0; i < n_traced_; i++) {
            const DRayf& r = rayCasts[i].ray;
            CInteraction& it = interactions[i];
            it.hCollision = hit_[i];
            it.time = t_[i];
            if (hit_[i] == INVALID_HANDLE) {
                it.p = {0, 0, 0};
                it.wo = {0, 0, 0};
                continue;
            }
            it.p = {r.origin.x + r.dir.x * t_[i], r.origin.y + r.dir.y * t_[i], r.origin.z + r.dir.z * t_[i]};
            it.wo = {-r.dir.x, -r.dir.y, -r.dir.z};
        }
    }

private:
    void check(wpbr_status s, const char* what) {
        if (s != WPBR_OK)
            throw Error(s, std::string(what) + ": " + wpbr_status_string(s) + " (" +
                               (ctx_ ? wpbr_last_error(ctx_) : "") + ")");
    }
    const void* upload_bytes(const void* src, size_t bytes, void** d) {
        if (!bytes) return nullptr;
        check(wpbr_malloc(ctx_, bytes, d), "wpbr_malloc");
        check(wpbr_upload(ctx_, *d, src, bytes), "wpbr_upload");
        return *d;
    }
    void release_attributes() {
        for (void** p : {&d_flags_, &d_uv_, &d_nrm_, &d_tan_})
            if (*p) { wpbr_free(ctx_, *p); *p = nullptr; }
        attrs_ = wpbr_tri_attributes{};
    }
    const float* upload(const std::vector<float>& v, void** d) {
        if (v.empty()) return nullptr;
        check(wpbr_malloc(ctx_, v.size() * 4, d), "wpbr_malloc");
        check(wpbr_upload(ctx_, *d, v.data(), v.size() * 4), "wpbr_upload");
        return (const float*)*d;
    }
    void release_shapes() {
        for (void** p : {&d_xform_, &d_radius_, &d_tris_})
            if (*p) { wpbr_free(ctx_, *p); *p = nullptr; }
    }
    void release_scene() {
        for (void** p : {&d_film_, &d_rgb8_})
            if (*p) { wpbr_free(ctx_, *p); *p = nullptr; }
        if (scene_) wpbr_scene_destroy(scene_);
        scene_ = nullptr;
    }
    void release_bsdf() {
        for (void** p : {&d_bframe_, &d_bwo_, &d_bwi_, &d_bu_, &d_bmat_, &d_bvalue_, &d_bpdf_, &d_bwi_out_, &d_bqueue_})
            if (*p) { wpbr_free(ctx_, *p); *p = nullptr; }
        req_ = wpbr_bsdf_requests{};
    }
    void ensure_rays(uint32_t n) {
        if (n <= ray_cap_) return;
        release_rays();
        check(wpbr_malloc(ctx_, 24 * (size_t)n, &d_rays_), "wpbr_malloc");
        check(wpbr_malloc(ctx_, (size_t)n, &d_surf_), "wpbr_malloc");
        check(wpbr_malloc(ctx_, 48 * (size_t)n, &d_diff_), "wpbr_malloc");
        check(wpbr_malloc(ctx_, 4 * (size_t)n, &d_hit_), "wpbr_malloc");
        check(wpbr_malloc(ctx_, 4 * (size_t)n, &d_t_), "wpbr_malloc");
        check(wpbr_malloc(ctx_, 4 * (size_t)WPBR_SI_FLOATS * n, &d_si_), "wpbr_malloc");
        ray_cap_ = n;
    }
    void release_rays() {
        for (void** p : {&d_rays_, &d_surf_, &d_diff_, &d_hit_, &d_t_, &d_si_})
            if (*p) { wpbr_free(ctx_, *p); *p = nullptr; }
        ray_cap_ = 0;
    }

    wpbr_ctx* ctx_ = nullptr;
    wpbr_lbvh* bvh_ = nullptr;
    wpbr_geometry geom_{};
    void *d_xform_ = nullptr, *d_radius_ = nullptr, *d_tris_ = nullptr;
    void *d_rays_ = nullptr, *d_surf_ = nullptr, *d_diff_ = nullptr, *d_hit_ = nullptr, *d_t_ = nullptr,
         *d_si_ = nullptr;
    void *d_flags_ = nullptr, *d_uv_ = nullptr, *d_nrm_ = nullptr, *d_tan_ = nullptr;
    wpbr_tri_attributes attrs_{};
    uint32_t ray_cap_ = 0, n_traced_ = 0;
    wpbr_scene* scene_ = nullptr;
    uint32_t width_ = 0, height_ = 0, spp_ = 0, pass_ = 0;
    bool rendering_ = false;
    void *d_film_ = nullptr, *d_rgb8_ = nullptr;
    wpbr_bsdf_requests req_{};
    uint32_t n_bsdf_ = 0;
    void *d_bframe_ = nullptr, *d_bwo_ = nullptr, *d_bwi_ = nullptr, *d_bu_ = nullptr, *d_bmat_ = nullptr;
    void *d_bvalue_ = nullptr, *d_bpdf_ = nullptr, *d_bwi_out_ = nullptr, *d_bqueue_ = nullptr;
    std::vector<uint32_t> hit_;
    std::vector<float> t_, si_;
};

// Spheres and triangles are bounded by one kernel (stage 0); the sphere job
// is the empty step before it.
inline void JobUpdateBoundsAndCentroidSphere::run(MEngine*) {}
inline void JobUpdateBoundsAndCentroidTriangle::run(MEngine* e) { e->build_stage(0); }
inline void JobGenerateShapesCentroidBound::run(MEngine* e) { e->build_stage(1); }
inline void JobGenerateShapesMortonCode::run(MEngine* e) { e->build_stage(2); }
inline void JobSortsShapesByMortonCode::run(MEngine* e) { e->build_stage(3); }
inline void JobEmitLBVH::run(MEngine* e) { e->build_stage(4); }
inline void JobBuildUpperSAH::run(MEngine* e) { e->build_stage(5); }
inline void JobProcessRayCasts::run(MEngine* e) { e->trace_rays(); }
inline void JobProcessRayCastsResults::run(MEngine* e) { e->fetch_hits(); }
inline void JobProcessSphereFullInteractionRequests::run(MEngine* e) { e->fill_surface_interactions(true); }
inline void JobProcessTriangleFullInteractionRequests::run(MEngine* e) { e->fill_surface_interactions(false); }
inline void JobCollisionFinish::run(MEngine* e) { e->finish_interactions(); }
inline void JobComputeDifferentialsForSurfInter::run(MEngine* e) {
    if (!e->rendering()) e->fill_differentials();  // in a render: inside the scatter stage
}
inline void JobGenerateFilmTiles::run(MEngine* e) { e->render_stage(WPBR_STAGE_BEGIN); }
inline void JobCameraPerspGenerateRaysDifferential::run(MEngine* e) { e->render_stage(WPBR_STAGE_CAMERA); }
inline void JobScatteringRequestEmittedLight::run(MEngine* e) { e->render_stage(WPBR_STAGE_SCATTER); }
inline void JobScatteringProcessShadowRay::run(MEngine* e) { e->render_stage(WPBR_STAGE_INTEGRATE); }
inline void JobOutputFilmTitles::run(MEngine* e) { e->output_film_tiles(); }
inline void JobOutputFilm::run(MEngine* e) { e->output_film(); }
// In a render the BSDF jobs run inside the scatter (Sample) and integrate
// (Compute) stages; outside it they run over bsdfRequests.
inline void JobGenerateBSDFRequests::run(MEngine* e) {
    if (!e->rendering()) e->bsdf_queues();
}
inline void JobBSDFConductorMicrofaceCompute::run(MEngine* e) {
    if (!e->rendering()) e->bsdf_job(WPBR_BSDF_CONDUCTOR_MF, false);
}
inline void JobBSDFSpecularReflectionCompute::run(MEngine* e) {
    if (!e->rendering()) e->bsdf_job(WPBR_BSDF_SPECULAR_REFL, false);
}
inline void JobBSDFDielectricMicrofaceCompute::run(MEngine* e) {
    if (!e->rendering()) e->bsdf_job(WPBR_BSDF_DIELECTRIC_MF, false);
}
inline void JobBSDFConductorMicrofaceSample::run(MEngine* e) {
    if (!e->rendering()) e->bsdf_job(WPBR_BSDF_CONDUCTOR_MF, true);
}
inline void JobBSDFSpecularReflectionSample::run(MEngine* e) {
    if (!e->rendering()) e->bsdf_job(WPBR_BSDF_SPECULAR_REFL, true);
}
inline void JobBSDFDielectricMicrofaceSample::run(MEngine* e) {
    if (!e->rendering()) e->bsdf_job(WPBR_BSDF_DIELECTRIC_MF, true);
}

#undef WPBR_JOB_RUN

// MeshTools::loadMesh (MeshTools.h:14-99): the OBJ file's first mesh,
// triangulated and left-handed as assimp hands it over.
namespace MeshTools {
inline CTriangleMesh loadMesh(const std::string& filename) {
    wpbr_obj* obj = nullptr;
    if (wpbr_obj_load(filename.c_str(), &obj) != WPBR_OK)
        throw Error(WPBR_ERR_INVALID_ARG, "The file " + filename + " can't be opened");
    uint32_t c[4];
    wpbr_obj_counts(obj, c);
    CTriangleMesh mesh;
    mesh.nVertices = c[0];
    mesh.nTriangles = c[1] / 3;
    mesh.positions.resize(3 * (size_t)c[0]);
    if (c[2]) mesh.normals.resize(3 * (size_t)c[0]);
    if (c[3]) mesh.uvs.resize(2 * (size_t)c[0]);
    mesh.iVertices.resize(c[1]);
    wpbr_obj_copy(obj, mesh.positions.data(), c[2] ? mesh.normals.data() : nullptr,
                  c[3] ? mesh.uvs.data() : nullptr, mesh.iVertices.data());
    wpbr_obj_free(obj);
    return mesh;
}
}  // namespace MeshTools

}  // namespace wpbr_host
