/* phg.h — C ABI of the MI355X-native BNEEPT hot path of Photon-v2.
 *
 * This is the thin host->device boundary that the reference's renderer seam would call.
 * The reference has no device layer; each entry point below names the reference
 * interface whose work it takes over (paths relative to the Photon-v2 source tree,
 * Engine/Engine/Source/...).
 *
 * Conventions
 *   - Plain C types only; no exceptions cross the ABI; every call returns phg_status.
 *   - A context owns all device memory it allocates; host arrays passed in are borrowed
 *     for the duration of the call only.
 *   - One host thread drives a context (mirrors the blocking phRender model,
 *     Engine/CEngine/Include/ph_c_core.h:7-21).  phg_get_stats may be called from another
 *     thread while no other call on the same context runs.
 *   - Raster (0,0) is the lower-left image corner; films are stored row by row from the
 *     bottom row up (HdrRgbFilm sensor order, Core/Filmic/HdrRgbFilm.cpp:103-121).
 */
#ifndef PHG_H
#define PHG_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum phg_status {
    PHG_OK = 0,
    PHG_ERR_INVALID_ARG = 1,
    PHG_ERR_HIP = 2,          /* a HIP runtime call failed; see phg_last_error */
    PHG_ERR_OUT_OF_MEMORY = 3,
    PHG_ERR_NO_SCENE = 4,     /* render/intersect before phg_upload_scene */
    PHG_ERR_NO_DEVICE = 5,    /* no usable gfx950 device */
    PHG_ERR_UNSUPPORTED = 6   /* e.g. BVH too deep for the traversal stack */
} phg_status;

/* Surface optics (Core/SurfaceBehavior/SurfaceOptics): one record per material. */
enum {
    PHG_MAT_ABSORBER = 0,   /* IdealAbsorber: area-light surfaces (Actor/Light/AAreaLight.cpp:30-34) */
    PHG_MAT_LAMBERT = 1,    /* LambertianReflector via MatteOpaque */
    PHG_MAT_CONDUCTOR = 2,  /* OpaqueMicrofacet + Schlick conductor Fresnel via AbradedOpaque */
    PHG_MAT_DIELECTRIC = 3  /* TranslucentMicrofacet + exact dielectric Fresnel via AbradedTranslucent */
};

/* Conductor Fresnel (ConductiveInterfaceInfo::genFresnelEffect,
 * Actor/Material/Component/ConductiveInterfaceInfo.cpp): Schlick's approximation from F0
 * (SchlickApproxConductorFresnel.cpp:35-43) or the exact conductor-dielectric formula from the
 * complex index of refraction n + ik inside and the real index outside
 * (ExactConductorFresnel.cpp:66-101). */
enum { PHG_FRESNEL_SCHLICK = 0, PHG_FRESNEL_EXACT = 1 };

typedef struct phg_material {
    uint32_t type;
    float rgb[3];        /* albedo (Lambert) or F0 (conductor, Schlick) */
    float roughness;     /* GGX alpha = max(roughness^2, 0.001) */
    float ior_outer;     /* dielectric; conductor with PHG_FRESNEL_EXACT: the outer index */
    float ior_inner;     /* dielectric only */
    uint32_t fresnel;    /* conductor: PHG_FRESNEL_SCHLICK (F0 = rgb) or PHG_FRESNEL_EXACT (n, k) */
    float n[3];          /* conductor with PHG_FRESNEL_EXACT: inner index, real part per channel */
    float k[3];          /* ... imaginary part per channel */
    float pad[2];
} phg_material;          /* 64 bytes */

/* Area-light emitter group (DiffuseSurfaceEmitter / GroupedDiffuseSurfaceEmitter):
 * constant front-face radiance over triangles emitter_tris[tri_begin, tri_begin+tri_count). */
typedef struct phg_emitter {
    float radiance[3];
    uint32_t tri_begin;
    uint32_t tri_count;
    uint32_t pad[3];
} phg_emitter;           /* 32 bytes */

/* Pinhole camera (Core/Receiver/PinholeCamera.cpp:26-69), horizontal FOV
 * (EngineEnv/Observer/SingleLensObserver.cpp:70-81).  right/up/forward orthonormal. */
typedef struct phg_camera {
    float pos[3];
    float right[3];
    float up[3];
    float forward[3];
    float tan_half_fov;
    float pad[3];
} phg_camera;            /* 64 bytes */

/* Flattened, world-space scene (what a renderer seam derives from the cooked VisualWorld,
 * World/VisualWorld.cpp:150-198).  Triangles carry face normals only (PTriangle,
 * Core/Intersection/PTriangle.cpp:19-35). */
typedef struct phg_scene_desc {
    uint32_t num_tris;
    uint32_t num_materials;
    uint32_t num_emitters;
    uint32_t num_emitter_tris;
    const float* verts;            /* 9 floats per triangle: A, B, C */
    const uint32_t* tri_material;  /* per triangle */
    const int32_t* tri_emitter;    /* per triangle; -1 if not emissive */
    const phg_material* materials;
    const phg_emitter* emitters;
    const uint32_t* emitter_tris;  /* triangle ids of all emitter groups, concatenated */
    phg_camera camera;
} phg_scene_desc;

typedef struct phg_device_opts {
    int32_t device;        /* HIP device ordinal */
    uint32_t flags;        /* reserved, 0 */
} phg_device_opts;

typedef struct phg_bvh_opts {
    uint32_t max_leaf_tris;   /* 1..8, default 4 */
    uint32_t sah_bins;        /* 8..64, default 16 */
    uint32_t nthreads;        /* host build threads, 0 = all cores */
    uint32_t width;           /* children per node, 2..4 (0 = 4); the node format stays 4-wide */
} phg_bvh_opts;

/* One ray: origin, minT, direction, maxT (TLineSegment, Math/Geometry/TLineSegment.ipp). */
typedef struct phg_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} phg_ray;                /* 32 bytes */

/* Closest-hit record.  prim = input triangle id (-1 on a miss); b, c are the barycentric
 * weights of vertices B and C as computed by the watertight test. */
typedef struct phg_hit {
    float t;
    int32_t prim;
    float b;
    float c;
} phg_hit;                /* 16 bytes */

enum { PHG_PTR_HOST = 0, PHG_PTR_DEVICE = 1 };
enum { PHG_QUERY_CLOSEST = 0, PHG_QUERY_ANY = 1 };

/* Estimators (Core/Estimator/): BNEEPT (BNEEPTEstimator.cpp:38-262), BVPT — paths find light
 * only by hitting it (BVPTEstimator.cpp:25-160), BVPTDL — emission seen directly or after one
 * BSDF sample (BVPTDLEstimator.cpp:25-98). */
enum { PHG_ESTIMATOR_BNEEPT = 0, PHG_ESTIMATOR_BVPT = 1, PHG_ESTIMATOR_BVPTDL = 2 };
/* Film reconstruction filters (SampleFilter.cpp:8-56): box of 1 pixel; Gaussian (sigma 0.5,
 * submerged), Mitchell-Netravali (b = c = 1/3) and Blackman-Harris, each 4 x 4 pixels.  With a
 * filter the film's sums are weighted: sum w r, g, b, sum w, sum w r^2, g^2, b^2. */
enum { PHG_FILTER_BOX = 0, PHG_FILTER_GAUSSIAN = 1, PHG_FILTER_MN = 2, PHG_FILTER_BH = 3 };

/* Sample source of a job (EngineEnv/SampleSource: sample-source(uniform-random | stratified |
 * halton)).
 * INDEPENDENT: every path draws its dimensions from its own PCG32 stream
 * (SGUniformRandom). STRATIFIED: SGStratified (Core/SampleGenerator/SGStratified.cpp:16-154):
 * sample s of every pixel forms one batch, and the five dimensions the estimator reads first at
 * the first surface are stratified across the batch's pixels (2-D, 2-D and 1-D jittered strata
 * in keyed shuffled order); the raster offset and all later dimensions are as INDEPENDENT.
 * Per-pixel distributions are the same; the pixels of one batch are negatively correlated.
 * HALTON: SGHalton (Core/SampleGenerator/SGHalton.cpp:22-206): the sample of pixel p in batch s
 * is Halton point s * W * H + p; dimensions 0, 1 (bases 2, 3) place it over the whole sample
 * window (it lands in any pixel, and the film splats it there), dimensions 2..6 (bases 5..17)
 * are the first surface's five; later dimensions are as INDEPENDENT.  OR in a permutation
 * (HaltonSampleSource "permutation"; default fixed) and a sequence ("sequence"; default
 * original).  The reference draws the fixed and per-digit permutations, the random leap and
 * the random start from its random source; here they are keyed by the job's seed.  Owen
 * scrambling is the reference's own (keyed by the dimension only). */
enum { PHG_SAMPLER_INDEPENDENT = 0, PHG_SAMPLER_STRATIFIED = 1, PHG_SAMPLER_HALTON = 2 };
enum {
    PHG_HALTON_PERM_FIXED = 0 << 8,      /* one permutation of the digit values per dimension */
    PHG_HALTON_PERM_NONE = 1 << 8,       /* plain radical inverse */
    PHG_HALTON_PERM_PER_DIGIT = 2 << 8,  /* one permutation per dimension and digit position */
    PHG_HALTON_PERM_OWEN = 3 << 8,       /* each digit permuted by a hash of the digits before it */
    PHG_HALTON_SEQ_ORIGINAL = 0 << 12,   /* points 0, 1, 2, ... */
    PHG_HALTON_SEQ_LEAP = 1 << 12,       /* every 19961st point (the 257th good prime) */
    PHG_HALTON_SEQ_RANDOM_LEAP = 2 << 12,  /* a leap drawn from the good primes 152..512 */
    PHG_HALTON_SEQ_RANDOM_START = 3 << 12  /* a 64-bit start offset per dimension */
};

typedef struct phg_render_opts {
    uint32_t width, height;
    uint32_t spp;              /* samples per pixel rendered by this call */
    uint32_t spp_begin;        /* index of the first sample (stream key), for progressive passes */
    uint32_t max_bounces;      /* bounce-loop cap; 10000 reproduces the reference's MAX_RAY_BOUNCES */
    uint32_t spp_per_pass;     /* samples per pixel per wavefront pass; 0 = automatic */
    uint64_t seed;
    uint32_t flags;            /* PHG_RENDER_* */
    float guide_prob;          /* PHG_RENDER_GUIDED: probability of sampling the learned lobes (one-sample MIS);
                                  < 0: the network's own per-query probability (output 40, in [0.05, 0.95]) */
    float guide_lr;            /* PHG_RENDER_TRAIN: Adam learning rate */
    uint32_t guide_batch;      /* PHG_RENDER_TRAIN: records per training step (0 = 2^18) */
    uint32_t guide_steps;      /* PHG_RENDER_TRAIN: Adam steps after each pass */
    uint32_t estimator;        /* PHG_ESTIMATOR_*; guiding needs BNEEPT */
    uint32_t filter;           /* PHG_FILTER_*: film reconstruction filter (SampleFilter.cpp:8-56) */
    uint32_t guide_depth;      /* PHG_RENDER_GUIDED: path vertices (from the first hit) that sample the lobes; 0 = all.
                                  PHG_RENDER_TRAIN keeps training records at those vertices only */
    uint32_t sampler;          /* PHG_SAMPLER_* (ReceiverSamplingWork.cpp:76-84 with the scene's sample source) */
} phg_render_opts;

enum {
    PHG_RENDER_COUNT_TRAVERSAL = 1,  /* count node visits / triangle tests (slower) */
    PHG_RENDER_GUIDED = 2,           /* sample directions from the guiding network too (one-sample MIS with the BSDF) */
    PHG_RENDER_TRAIN = 4,            /* train the guiding network online on this render's radiance samples */
    PHG_RENDER_GRAPH = 8             /* captured HIP graphs, same results as the eager loop: an unguided render of
                                        max_bounces < 62 replays one graph of the whole pass; any other (unbounded,
                                        guided, training) replays graphs of 8 bounces -- extend, guide inference,
                                        shade, shadow -- until the path queue is empty (one host round trip per 8
                                        bounces instead of per bounce) */
};

typedef struct phg_stats {
    uint64_t paths;             /* completed camera paths (samples) */
    uint64_t rays_primary;
    uint64_t rays_extension;
    uint64_t rays_shadow;
    uint64_t node_visits;       /* only with PHG_RENDER_COUNT_TRAVERSAL / counted intersect */
    uint64_t tri_tests;
    double render_ms;           /* wall time of the last phg_render call */
    double trace_ms;            /* GPU time in closest-hit traversal kernels (last call) */
    double shadow_ms;           /* GPU time in any-hit traversal kernels (last call) */
    double shade_ms;            /* GPU time in generate/shade/film kernels (last call) */
    uint32_t bvh_nodes;
    uint32_t bvh_depth;
    uint32_t bvh_stack_need;
    uint32_t pad;
    double bvh_build_ms;
    uint64_t trace_node_visits;   /* closest-hit kernels, with PHG_RENDER_COUNT_TRAVERSAL */
    uint64_t trace_tri_tests;
    uint64_t shadow_node_visits;  /* any-hit kernels, with PHG_RENDER_COUNT_TRAVERSAL */
    uint64_t shadow_tri_tests;
    uint32_t trace_launches;      /* closest-hit kernel launches in the last render */
    uint32_t shadow_launches;
    double guide_infer_ms;        /* GPU time in guiding-network inference (last render) */
    double guide_train_ms;        /* GPU time in guiding-network training (last render) */
    uint64_t guide_queries;       /* network evaluations (last render) */
    uint64_t guide_records;       /* records processed by training steps (last render; a record per step counts once) */
    uint32_t guide_steps;         /* Adam steps taken so far on this context */
    uint32_t batch_passes;        /* wavefront passes of the last phg_render_batch */
    double graph_ms;              /* GPU time of the replayed pass graphs (PHG_RENDER_GRAPH / PHG_BATCH_GRAPH);
                                     trace/shadow/shade times then cover only eager work */
} phg_stats;

/* Context lifetime.  Replaces nothing in the reference: the CPU engine needs no device. */
phg_status phg_create(void** out_ctx, const phg_device_opts* opts);
phg_status phg_destroy(void* ctx);
const char* phg_last_error(void* ctx);

/* Scene upload + acceleration structure build.
 * Replaces VisualWorld's top-level accelerator build, TBinaryBvhIntersector::update ->
 * TBvhBuilder<2>::buildInformativeBvh -> TLinearDepthFirstBinaryBvh::build
 * (Core/Intersection/BVH/TBinaryBvhIntersector.ipp:66-97, Math/Algorithm/BVH/TBvhBuilder.ipp:34-86),
 * and ESPowerFavoring::update (Core/Emitter/Sampler/ESPowerFavoring.cpp:21-49). */
phg_status phg_upload_scene(void* ctx, const phg_scene_desc* scene, const phg_bvh_opts* bvh);

/* Batched closest-hit / any-hit ray queries.
 * Replaces Scene::isIntersecting / Scene::isOccluding (World/Scene.cpp:36-68) over
 * TBinaryBvhIntersector::isIntersecting (TBinaryBvhIntersector.ipp:25-51) and
 * PTriangle::isIntersecting -> TWatertightTriangle::isIntersecting
 * (PTriangle.cpp:37-49, Math/Geometry/TWatertightTriangle.ipp:16-149).
 * For PHG_QUERY_ANY, hits[i].prim is 1 if occluded and -1 otherwise.
 * ptr_kind selects whether rays/hits are host or device pointers. */
phg_status phg_intersect(void* ctx, const phg_ray* rays, uint64_t n, phg_hit* hits,
                         int query, int ptr_kind);

/* Allocation-free variant of phg_intersect for device pointers.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or allocating library routine: its scratch (error flag and traversal-stack spill)
 * is a file-scope device array of 16 MiB per GPU, held until the process ends, and its grid is
 * sized to fit that array (a BVH whose per-block spill exceeds it is PHG_ERR_UNSUPPORTED).
 * Results equal phg_intersect's.  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (from several contexts or threads) race; calls are synchronous. */
phg_status phg_intersect_noalloc(void* ctx, const phg_ray* rays, uint64_t n, phg_hit* hits, int query);

/* Same as phg_intersect(PHG_QUERY_CLOSEST) but also returns per-ray node visits and
 * triangle tests (for the algorithmic-byte roofline). */
phg_status phg_intersect_counted(void* ctx, const phg_ray* rays, uint64_t n, phg_hit* hits,
                                 uint32_t* node_visits, uint32_t* tri_tests, int ptr_kind);

/* Film.  Replaces HdrRgbFilm's FP64 RadianceSensor storage (Core/Filmic/HdrRgbFilm.cpp:62-122,
 * Core/Receiver/RadianceSensor.h:10-21) with a box filter: 8 doubles per pixel =
 * sum r, g, b, sample weight, sum r^2, g^2, b^2, 0 (the squares feed per-pixel CIs). */
phg_status phg_film_clear(void* ctx, uint32_t width, uint32_t height);
phg_status phg_film_read(void* ctx, double* out, uint64_t out_len);
/* Host film output (no device needed).  phg_film_develop: RGB float per pixel = weighted sum /
 * weight (HdrRgbFilm::developRegion, HdrRgbFilm.cpp:167-207; 0 where no weight).
 * phg_film_write_pfm: the developed frame as a .pfm file, as PfmFile::save writes it
 * (DataIO/PfmFile.cpp:109-145): header "PF\n<w> <h>\n-1\n", rows bottom to top. */
phg_status phg_film_develop(const double* film, uint32_t width, uint32_t height, float* out_rgb);
phg_status phg_film_write_pfm(const double* film, uint32_t width, uint32_t height, const char* path);

/* Render spp samples per pixel with BNEEPT into the film.
 * Replaces EqualSamplingRenderer::doRender (Core/Renderer/Sampling/EqualSamplingRenderer.cpp:119-203)
 * -> ReceiverSamplingWork::doWork (ReceiverSamplingWork.cpp:61-138)
 * -> BNEEPTEstimator::estimate (Core/Estimator/BNEEPTEstimator.cpp:38-262)
 * -> HdrRgbFilm::addSample (HdrRgbFilm.cpp:62-122). */
phg_status phg_render(void* ctx, const phg_render_opts* opts);

/* ---- Batched rendering (no counterpart in the reference, which renders one scene per
 * engine): several scenes uploaded together, then many render jobs over them whose paths
 * share every wavefront pass, so small jobs fill the device alongside large ones.
 *
 * phg_upload_scenes: like phg_upload_scene for n scenes at once (BVHs built in parallel);
 * scene i is addressed by index i in the jobs.  phg_render (and phg_intersect) use scene 0. */
phg_status phg_upload_scenes(void* ctx, const phg_scene_desc* scenes, uint32_t n, const phg_bvh_opts* bvh);

typedef struct phg_batch_job {
    uint32_t scene;            /* index into the scenes of phg_upload_scenes */
    uint32_t width, height;
    uint32_t spp;
    uint32_t max_bounces;
    uint32_t seed;             /* < 2^24; sample streams as in phg_render */
    uint32_t flags;            /* PHG_JOB_CAMERA: render through `camera` instead of the scene's */
    uint32_t estimator;        /* PHG_ESTIMATOR_* */
    phg_camera camera;
    uint32_t filter;           /* PHG_FILTER_* */
    uint32_t sampler;          /* PHG_SAMPLER_* */
    uint32_t pad2[2];
} phg_batch_job;

enum { PHG_JOB_CAMERA = 1 };
enum { PHG_BATCH_ORDER_GIVEN = 0, PHG_BATCH_ORDER_SHORTEST_FIRST = 1 };
enum {
    PHG_BATCH_COUNT_TRAVERSAL = 1,
    PHG_BATCH_GRAPH = 2,       /* as PHG_RENDER_GRAPH (guided and training batches too) */
    PHG_BATCH_GUIDED = 4,      /* as PHG_RENDER_GUIDED; all jobs must render the same scene (one network) */
    PHG_BATCH_TRAIN = 8        /* as PHG_RENDER_TRAIN, on the records of every job in the pass */
};

typedef struct phg_batch_opts {
    uint32_t paths_per_pass;   /* paths in flight per pass (0 = the context's tuning); passes carry whole samples */
    uint32_t order;            /* PHG_BATCH_ORDER_*: which jobs the passes take first */
    uint32_t flags;            /* PHG_BATCH_* */
    uint32_t guide_steps;      /* PHG_BATCH_TRAIN: Adam steps after each pass (0 = 4) */
    float guide_prob;          /* PHG_BATCH_GUIDED: probability of sampling the learned lobes; < 0: learned per query */
    float guide_lr;            /* PHG_BATCH_TRAIN: Adam learning rate (0 = 1e-3) */
    uint32_t guide_batch;      /* PHG_BATCH_TRAIN: records per step (0 = 2^18) */
    uint32_t guide_depth;      /* PHG_BATCH_GUIDED: path vertices that sample the lobes (0 = all); training
                                  records are kept at those vertices only */
} phg_batch_opts;

typedef struct phg_batch_result {
    double done_ms;            /* from the start of the batch to the end of the job's last pass */
    uint64_t paths;            /* camera paths of the job */
} phg_batch_result;

/* Render all jobs (each into its own FP64 film, cleared first).  Every job's film equals
 * what phg_render of that job alone produces: the sample streams depend only on (seed,
 * pixel, sample) and each pixel accumulates its samples in order. */
phg_status phg_render_batch(void* ctx, const phg_batch_job* jobs, uint32_t n, const phg_batch_opts* opts,
                            phg_batch_result* results);
/* The film of job `job` of the last batch: 8 doubles per pixel as phg_film_read. */
phg_status phg_batch_film_read(void* ctx, uint32_t job, double* out, uint64_t out_len);

/* Launch tuning of the path kernels (no counterpart in the reference).  trace_block: threads
 * per block of the traversal kernels (64, 128 or 256; 0 = 128); paths_per_pass: paths in
 * flight per wavefront pass (2^16..2^26; 0 = 32M), the default for phg_render passes and
 * batches that do not set their own (renders and batches that train the guiding network use
 * at most 16M unless they set their own).  Results do not depend on either.  phg_get_tuning also
 * reports the resident traversal blocks per compute unit that size the persistent grid. */
typedef struct phg_tuning {
    uint32_t trace_block;
    uint32_t paths_per_pass;
    uint32_t trace_blocks_per_cu;   /* out only */
    uint32_t pad;
} phg_tuning;
phg_status phg_set_tuning(void* ctx, const phg_tuning* t);
phg_status phg_get_tuning(void* ctx, phg_tuning* t);

/* Autotuner for batched rendering (no counterpart in the reference).  Renders a sample of
 * the jobs (up to sample_jobs of them, spread over their sizes, at most sample_spp samples
 * each) over their scenes and times it: first every traversal block size x paths-per-pass
 * candidate with the default BVH, then every BVH leaf size x width candidate with the best
 * launch parameters.  On return the context holds the sample scenes (upload the real ones
 * again with best.max_leaf_tris / best.width) and the best launch tuning.  trials (optional,
 * up to max_trials entries) receives every measurement. */
typedef struct phg_tune_trial {
    uint32_t max_leaf_tris, width;      /* BVH */
    uint32_t trace_block, paths_per_pass;
    double msamples_per_s;              /* sample batch, captured-graph passes */
} phg_tune_trial;
typedef struct phg_autotune_opts {
    uint32_t sample_jobs;      /* 0 = 8 */
    uint32_t sample_spp;       /* 0 = 4 */
    uint32_t repeats;          /* timed renders per candidate, best kept; 0 = 2 */
    uint32_t pad;
} phg_autotune_opts;
phg_status phg_autotune_batch(void* ctx, const phg_scene_desc* scenes, uint32_t nscenes, const phg_batch_job* jobs,
                              uint32_t njobs, const phg_autotune_opts* opts, phg_tune_trial* best,
                              phg_tune_trial* trials, uint32_t max_trials, uint32_t* ntrials);

/* Statistics of the last render/intersect (Renderer::asyncQueryRenderStats,
 * Core/Renderer/Sampling/EqualSamplingRenderer.cpp:258-280, reports paths/second). */
phg_status phg_get_stats(void* ctx, phg_stats* out);

/* ---- Path-guiding network (no counterpart in the reference, which has no guiding).
 * A 64-64-64-64-64-41 ReLU MLP over a frequency encoding of (position, normal, outgoing
 * direction) producing 8 von Mises-Fisher lobes over the incident direction and the logit z of
 * the lobes' selection probability a = 0.05 + 0.9 sigmoid(z); BF16 on the matrix cores with
 * FP32 accumulation.  Parameters: one FP32 blob of PHG_GUIDE_PARAMS values, W0 b0 W1 b1 W2 b2
 * W3 b3 W4 b4 with each W stored [in][out].
 * Training record: 16 floats = 9 features in [0,1] (position in the scene box, normal and
 * outgoing direction mapped from [-1,1]), incident direction xyz, weight (target / sampling
 * density), BSDF density b of the direction, the selection probability the direction was
 * sampled with (0: the BSDF alone), and, where that probability is to be learned, the density
 * p_s the direction was drawn with (0: a fixed probability).  The lobe outputs minimise
 * -weight * log(a q + (1 - a) b) with q the lobes' density and a the record's fixed
 * probability or the network's own: the lobes are fitted as the one-sample MIS mixture that
 * samples the direction; with b = 0 (or a = 0) it is the plain weighted negative
 * log-likelihood of q.  The selection logit (a head on the last hidden layer whose gradient
 * does not reach the hidden layers) minimises weight^2 p_s / (a q + (1 - a) b), the
 * one-sample estimate of the mixture estimator's second moment. */
enum { PHG_GUIDE_PARAMS = 19305, PHG_GUIDE_OUTPUTS = 41, PHG_GUIDE_RECORD = 16 };

phg_status phg_guide_set_params(void* ctx, const float* params, uint32_t n);
phg_status phg_guide_get_params(void* ctx, float* params, uint32_t n);
/* raw network outputs (41 per query: 8 lobes x (mean direction 3, log kappa, weight logit), selection
 * logit) for 9 features per query */
phg_status phg_guide_forward(void* ctx, const float* features, uint64_t n, float* out);
/* decoded mixture (41 per query: 8 lobes x (unit mean direction 3, kappa, normalised weight), then
 * the lobes' selection probability 0.05 + 0.9 sigmoid(logit)), as the guided shade stage reads it */
phg_status phg_guide_lobes(void* ctx, const float* features, uint64_t n, float* out);
/* gradient of the mean of the records' losses (above) */
phg_status phg_guide_gradients(void* ctx, const float* records, uint64_t n, float* grads);
/* `steps` Adam steps on the same records (learning rate lr) */
phg_status phg_guide_train(void* ctx, const float* records, uint64_t n, float lr, uint32_t steps);

/* ---- Procedural scene generation (host).  Not a reference interface: produces the
 * benchmark workloads as flattened scenes.  The returned arrays are owned by the library
 * until phg_scene_free. */
enum {
    PHG_GEN_CORNELL = 1,        /* unit-cube Cornell box, 34 + 2 light triangles */
    PHG_GEN_PROCEDURAL = 2,     /* displaced icospheres + ground + area quads */
    PHG_GEN_OCCLUDED_ROOM = 3,  /* room lit through a slit by a hidden quad light */
    PHG_GEN_TERRAIN = 4         /* displaced icospheres over an FBM terrain grid (20M triangles), 64 area quads */
};

typedef struct phg_gen_opts {
    uint32_t kind;
    uint32_t seed;
    uint32_t target_tris;      /* PROCEDURAL/ROOM: total triangle count target */
    uint32_t num_instances;    /* PROCEDURAL: icosphere instances */
    uint32_t num_lights;       /* PROCEDURAL: area-light quads */
    uint32_t subdiv;           /* PROCEDURAL: icosphere subdivision level */
    uint32_t width, height;    /* for the camera aspect */
    float fov_deg;             /* horizontal */
    uint32_t pad;
} phg_gen_opts;

typedef struct phg_scene_host {
    phg_scene_desc desc;       /* pointers into library-owned arrays */
} phg_scene_host;

phg_status phg_scene_generate(const phg_gen_opts* opts, phg_scene_host** out);
void phg_scene_free(phg_scene_host* scene);

/* ---- .p2 scene ingestion (host).
 * Replaces, for scenes that render on this path, SdlSceneFileReader / SdlCommandParser
 * (SDL/SdlCommandParser.cpp:100-700) followed by the actor cooking that flattens models,
 * materials and area lights into the visual world (Actor/AModel.cpp:69-84,
 * Actor/Light/AAreaLight.cpp:37-87, EngineEnv/Visualizer/PathTracingVisualizer.cpp:18-36).
 * Understood: option(single-frame-render-session), observer(single-lens),
 * sample-source(*), visualizer(path-tracing) with estimator bneept | bvpt | bvptdl,
 * geometry(rectangle | cuboid | sphere | triangle | triangle-mesh), image(constant),
 * material(matte-opaque | abraded-opaque | abraded-translucent), actor(model),
 * actor(rectangle-light), and the translate / rotate / scale executors.  Spheres are
 * tessellated; unsupported commands and fields are skipped and counted in `warnings`.
 * The returned scene is freed with phg_scene_free.  err (optional) receives a message. */

typedef struct phg_p2_settings {
    uint32_t width, height;    /* frame-size */
    uint32_t spp;              /* sample-source samples */
    uint32_t estimator;        /* PHG_ESTIMATOR_* */
    uint32_t max_bounces;      /* visualizer max-path-length if given, else 10000 (no cap) */
    uint32_t filter;           /* PHG_FILTER_*; the GPU film accumulates with a box filter */
    uint32_t warnings;         /* skipped commands / fields / approximations */
    uint32_t sampler;          /* PHG_SAMPLER_* of the sample-source */
} phg_p2_settings;

phg_status phg_scene_parse_p2(const char* text, phg_p2_settings* settings, phg_scene_host** out, char* err,
                              uint32_t err_len);
phg_status phg_scene_load_p2(const char* path, phg_p2_settings* settings, phg_scene_host** out, char* err,
                             uint32_t err_len);

#ifdef __cplusplus
}
#endif
#endif /* PHG_H */
