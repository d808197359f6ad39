/* petrichor_hip.h — C-ABI of the MI355X (gfx950) rendering hot path.
 *
 * Plain pointers, sizes and status codes; no C++ or torch types cross this
 * boundary. Each entry point names the reference interface it replaces
 * (paths relative to the reference's LibPetrichor/).
 *
 * Status codes: 0 = ok, < 0 = error (see PT_ERR_*); pt_last_error() returns a
 * description of the last failure on a context. No entry point aborts the process.
 * Calls on one context are externally serialised (one HIP stream per context).
 */
#ifndef PETRICHOR_HIP_H
#define PETRICHOR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_OK 0
#define PT_ERR_ARG (-1)         /* bad argument or descriptor */
#define PT_ERR_OOM (-2)         /* device allocation failed */
#define PT_ERR_BUDGET (-3)      /* request exceeds the memory budget (serving) */
#define PT_ERR_UNSUPPORTED (-4) /* feature outside the supported scene model */
#define PT_ERR_HIP (-5)         /* HIP runtime error */
#define PT_ERR_STATE (-6)       /* wrong object state (e.g. no BVH built) */

/* Material types in the order of Core/Material/MaterialBase.h:15-22. */
#define PT_MAT_LAMBERT 0
#define PT_MAT_GLASS 1
#define PT_MAT_EMISSION 2
#define PT_MAT_GLOSSY 3 /* GGX */
#define PT_MAT_MIX 4

#define PT_INTEGRATOR_MIS 0    /* Core/Integrator/PathTracing.cpp */
#define PT_INTEGRATOR_SIMPLE 1 /* Core/Integrator/SimplePathTracing.cpp */

typedef struct pt_context pt_context;
typedef struct pt_scene pt_scene;
typedef struct pt_bvh pt_bvh;

/* Flattened Core::Scene (Core/Scene.h:60-218). Triangles are given as three
 * positions each (the reference's Triangle holds three Vertex pointers,
 * Core/Geometry/Triangle.h:59); a hit identifies a triangle by its index in this
 * array instead of a GeometryBase pointer. */
typedef struct pt_scene_desc {
    int32_t num_triangles;
    const float* positions;      /* 9 floats per triangle: p0, p1, p2 */
    const float* normals;        /* 9 floats per triangle (vertex normals) or NULL */
    const int32_t* material_ids; /* one per triangle */
    const int32_t* smooth;       /* 1 = ShadingTypes::Smooth per triangle, or NULL (all flat) */
    int32_t num_materials;
    const int32_t* material_types; /* PT_MAT_* */
    /* 8 floats per material: color[3] (Lambert kd, Glass colour, Emission Le, GGX f0),
     * p3 (GGX roughness, Glass ior, Mix weight), p4, p5 (Mix sub-material indices),
     * p6 (importance sampling enabled: 1 or 0), p7 unused */
    const float* material_params;
    int32_t num_lights;
    const int32_t* light_triangles; /* Scene::AppendLight order (Core/Scene.h:86-92); primitive indices */
    /* Camera state of Core/Camera.h:121-130: pos[3] right[3] up[3] forward[3]
     * focus_dist h_perf aperture, then one unused float. */
    float camera[16];
    int32_t env_width, env_height;
    const float* env_rgb; /* env_width*env_height RGB floats, row-major, or NULL */
    float env_base_color[3];
    float env_z_rotation;
    int32_t env_importance_sampling; /* Environment::useEnvImportanceSampling */
    /* Texture coordinates (Vertex::uv, Core/Geometry/Vertex.h:21): 6 floats per triangle
     * (u, v of p0, p1, p2), or NULL for all zero. Read by the UV AOV and the material maps. */
    const float* uvs;
    /* Material maps (Lambert::SetTexAlbedo, GGX::SetF0Texture, GGX::SetRoughnessMap):
     * num_textures textures, texture t of texture_sizes[2t] x texture_sizes[2t+1] RGB floats
     * at textures[t] (row-major, as Texture2D holds them); material_maps, 2 per material:
     * the albedo map of a Lambert or the reflectance map of a GGX, then the roughness map of
     * a GGX; -1 for none. NULL material_maps: no maps. */
    int32_t num_textures;
    const int32_t* texture_sizes;
    const float* const* textures;
    const int32_t* material_maps;
    /* GGX normal maps (GGX::SetNormalMap, Core/Material/GGX.h:50-54,96-119): one texture
     * index per material, read by GGX materials only, -1 for none; NULL for none. */
    const int32_t* material_normal_maps;
    /* Map strengths, 2 per material: GGX::SetNormalMapStrength (GGX.h:84-87; the setter
     * clamps it to [0, 1], values outside are rejected here) and SetRoughnessMapStrength
     * (GGX.h:89-93); NULL for 1 and 1, the reference's defaults. */
    const float* material_map_strengths;
    /* Spheres (Core/Geometry/Sphere.cpp:17-118), primitives num_triangles ..
     * num_triangles + num_spheres - 1 after the triangles: centre and radius (4 floats) and
     * a material each. light_triangles may name them by that primitive index (Scene::AppendLight
     * of a Sphere); primitive indices in hits count the same way. */
    int32_t num_spheres;
    const float* spheres;
    const int32_t* sphere_material_ids;
} pt_scene_desc;

typedef struct pt_bvh_desc {
    int32_t num_bins; /* 16 in the reference (Core/Accel/BinnedSAHBVH.h:174); 32 for the microbenchmark */
    int32_t leaf_max; /* 4 in the reference (BinnedSAHBVH.cpp:79) */
    int32_t width;    /* traversal node width: 2 walks the binary tree in the reference's
                       * order; 4 and 8 add wide nodes (64 B / 128 B, quantised child boxes,
                       * nearest child first), which reach the same nearest hit except on rays
                       * that graze a box within rounding of their hit */
    int32_t builder;  /* PT_BUILD_* below */
} pt_bvh_desc;

/* pt_bvh_desc.builder. The device builder takes 2..64 bins, the host builders 2..256.
 * PT_BUILD_DEVICE and PT_BUILD_HOST give the same tree: the reference's nodes, with the
 * primitives of a leaf ordered by the reference's sort key and, among equal keys, by
 * primitive index. PT_BUILD_HOST_REFERENCE_ORDER sorts every split node's range as
 * BinnedSAHBVH::Build does (std::sort by the centroid on the widest axis,
 * Core/Accel/BinnedSAHBVH.cpp:104-107), so the leaf order among equal keys is the
 * reference's too (one host sort per node: several seconds for millions of triangles). */
#define PT_BUILD_DEVICE 0
#define PT_BUILD_HOST 1
#define PT_BUILD_HOST_REFERENCE_ORDER 2

typedef struct pt_render_desc {
    int32_t width, height; /* Core/RenderSetting.h:12-13 */
    int32_t spp;           /* numSamplesPerPixel */
    int32_t max_bounces;   /* numMaxBounces */
    int32_t integrator;    /* PT_INTEGRATOR_* */
    uint32_t seed;         /* job seed of the per-sample sampler streams */
    int32_t flags;         /* PT_RENDER_* */
    /* The 2D sampler of the integrators: 0 for RandomSampler2D (what Core::Petrichor::Render
     * uses, Core/Petrichor.cpp:76-77); 1..PT_MAX_SAMPLER_LENGTH for MicroJitteredSampler
     * (Core/Sampler/MicroJitteredSampler.cpp:10-48) initialised with that sequence length.
     * Sample s of a pixel starts at Halton point s mod length with its own offsets. The
     * denoising AOVs always use the random sampler. */
    int32_t sampler_length;
} pt_render_desc;
#define PT_MAX_SAMPLER_LENGTH 65536

/* Every launch on its own, timed with HIP events (same image). */
#define PT_RENDER_TIME_KERNELS 1
/* Plain launches: one host synchronisation per wavefront iteration, no graph (same
 * image; the per-job baseline of the batched and served forms). */
#define PT_RENDER_PLAIN_LAUNCHES 2
/* With PT_RENDER_TIME_KERNELS: count node visits and triangle tests per trace kind
 * (binary tree: child boxes tested, 32 B each; 8-wide tree: 128 B node records). */
#define PT_RENDER_TRACE_STATS 4
/* An image of at least 2^19 pixels renders as two halves of its pixels on two streams,
 * whose launches the GPU overlaps (same image). It pays when the two streams get
 * hardware queues of their own: with more streams alive in the process than the device
 * has queues (4 by default), the halves can share one and run slower than one piece. */
#define PT_RENDER_TWO_STREAMS 8
/* The production schedule (graph replays, and two streams with PT_RENDER_TWO_STREAMS) with a
 * launch log: every trace launch records the span of its blocks on the device clock and
 * its rays (with PT_RENDER_TRACE_STATS also its node visits and triangle tests), and every
 * path-completion launch is timed with events; the sums fill the log_* and finish_* fields
 * of pt_render_stats (same image). Its only extra work is one single-thread kernel after
 * each trace launch. */
#define PT_RENDER_LAUNCH_LOG 16
/* With PT_RENDER_TWO_STREAMS: three parts of the pixels on three streams instead of two
 * (same image). The process then needs three hardware queues for them. */
#define PT_RENDER_THREE_STREAMS 32

/* Per-kernel figures: index 0 extension (closest hit), 1 environment shadow rays (any
 * hit), 2 area-light shadow rays (closest hit), 3 shading. Times are filled only with
 * PT_RENDER_TIME_KERNELS. */
typedef struct pt_render_stats {
    int64_t queries;         /* BVH queries issued (extension + shadow) */
    int64_t iterations;      /* wavefront iterations */
    double elapsed_ms;       /* device time of the render loop */
    int64_t kernel_items[4]; /* rays traced / paths shaded per kernel kind */
    int64_t kernel_launches[4];
    double kernel_ms[4];     /* summed launch durations */
    int64_t kernel_nodes[4]; /* node records visited by the trace kinds (PT_RENDER_TRACE_STATS) */
    int64_t kernel_tris[4];  /* triangle tests of the trace kinds (PT_RENDER_TRACE_STATS) */
    /* PT_RENDER_LAUNCH_LOG: trace launches of the graphs, their rays, summed spans (ms), node
     * visits and triangle tests (with PT_RENDER_TRACE_STATS); path-completion launches, the rays
     * they traced and their summed durations (ms) */
    int64_t log_launches, log_items;
    double log_trace_ms;
    int64_t log_nodes, log_tris;
    int64_t finish_launches, finish_rays;
    double finish_ms;
} pt_render_stats;

/* Context = one device + one HIP stream. */
int pt_ctx_create(int device, pt_context** out);
int pt_ctx_destroy(pt_context* ctx);
const char* pt_last_error(pt_context* ctx);

/* Uploads a scene (replaces building Core::Scene + Environment::Load/PreCalc,
 * Core/Scene.h:60-218, Core/Environment.cpp:14-128). Host pointers. */
int pt_scene_create(pt_context* ctx, const pt_scene_desc* desc, pt_scene** out);
int pt_scene_destroy(pt_scene* scene);

/* Binned-SAH BVH over the scene (replaces BinnedSAHBVH::Build,
 * Core/Accel/BinnedSAHBVH.cpp:15-225). */
int pt_bvh_build(pt_scene* scene, const pt_bvh_desc* desc, pt_bvh** out);
int pt_bvh_destroy(pt_bvh* bvh);
/* Leaf ordering (the reference's m_primitiveIDs, BinnedSAHBVH.h:178): prim_order[k]
 * is the primitive at leaf slot k (triangles, then spheres). Host pointer of
 * num_triangles + num_spheres ints. */
int pt_bvh_export_order(pt_bvh* bvh, int32_t* prim_order);
/* Node table of the binary tree in the reference's node order (BinnedSAHBVH::m_nodes,
 * BinnedSAHBVH.h:22-52): per node box[6] (lower, upper), child[2] (-1 for leaves),
 * range[2] (first and end leaf slot) and leaf flag. Host pointers sized by pt_bvh_info's
 * num_nodes; any may be NULL. */
int pt_bvh_export_nodes(pt_bvh* bvh, float* boxes, int32_t* child, int32_t* range, int32_t* leaf);
/* Node and depth counts of the binary tree. */
int pt_bvh_info(pt_bvh* bvh, int32_t* num_nodes, int32_t* max_depth, double* build_ms);

/* Builds the binary tree on the host only, without a device (for validation and for
 * callers that build offline): tri_pos 9 floats per triangle. Outputs, each sized by the
 * caller: prim_order[num_triangles], and per node (at most 2*num_triangles - 1, or 1)
 * boxes[6], child[2], range[2], leaf[1]. Returns the node count (>= 1) or a PT_ERR_*. */
int pt_bvh_build_host(const float* tri_pos, int32_t num_triangles, const pt_bvh_desc* desc,
                      int32_t* prim_order, float* boxes, int32_t* child, int32_t* range, int32_t* leaf,
                      int32_t* max_depth);

/* Closest hit for n rays (replaces BinnedSAHBVH::Intersect, BinnedSAHBVH.cpp:272-374).
 * Device pointers. rays: 8 floats per ray (ox oy oz dist_min dx dy dz dist_max);
 * hits: t_out (+inf on a miss) and prim_out (triangle index, -1 on a miss).
 * stream: a hipStream_t or NULL for the context's stream. A launch of at least 2^20 rays
 * traces them in coherence order (sorted by the Morton code of the origin) and writes
 * every hit at its ray's index; the hits are those of any other order. The BVH keeps the
 * sort's buffers (16 bytes per ray of its largest such launch, plus the sort's scratch)
 * until pt_bvh_destroy; calls on one BVH must not overlap. Same for pt_trace_any. */
int pt_trace_closest(pt_bvh* bvh, int64_t n, const float* rays, float* t_out, int32_t* prim_out,
                     void* stream);
/* pt_trace_closest over host memory: rays, t_out and prim_out are host pointers; the rays
 * go through device buffers the BVH keeps (40 bytes per ray of its largest such call) and
 * the call returns when the hits are in t_out / prim_out. The BVH query of
 * AccelBase::Intersect (Core/Accel/AccelBase.h:37-56) for host code that holds its rays on
 * the host (the C++ BinnedSAHBVH class of petrichor/Petrichor.h). */
int pt_trace_closest_host(pt_bvh* bvh, int64_t n, const float* rays, float* t_out, int32_t* prim_out);
/* Any hit (occlusion) for n rays: hit_out[i] = 1 if some triangle is hit within
 * [dist_min, dist_max], else 0. No reference counterpart (the reference's shadow
 * rays call the closest-hit query, Core/Integrator/PathTracing.cpp:293). */
int pt_trace_any(pt_bvh* bvh, int64_t n, const float* rays, int32_t* hit_out, void* stream);

/* The same queries over structure-of-arrays buffers (BinnedSAHBVH::Intersect over many
 * rays, Core/Accel/BinnedSAHBVH.cpp:272-374; the Ray fields of Core/Ray.h:10-49 and the
 * HitInfo of Core/HitInfo.h:14-27). Device pointers of n floats (ints) each. dist_min and
 * dist_max may be NULL: 0 and FLT_MAX, as AccelBase::Intersect(ray, scene)
 * (Core/Accel/AccelBase.h:44-48). Hits: t (+inf on a miss), prim (triangle index, then
 * spheres; -1 on a miss) and, when not NULL, u and v: the barycentric coordinates of the
 * hit as Triangle::Intersect forms them (Core/Geometry/Triangle.cpp:57-72; 0 for spheres
 * and misses). The BVH keeps a staging copy of the rays (32 bytes per ray of its largest
 * such call) until pt_bvh_destroy; calls on one BVH must not overlap. Hits equal
 * pt_trace_closest's / pt_trace_any's for the same rays. */
typedef struct pt_rays_soa {
    const float *ox, *oy, *oz;
    const float *dx, *dy, *dz;
    const float *dist_min, *dist_max;
} pt_rays_soa;
typedef struct pt_hits_soa {
    float* t;
    int32_t* prim;
    float *u, *v;
} pt_hits_soa;
int pt_trace_closest_soa(pt_bvh* bvh, int64_t n, const pt_rays_soa* rays, const pt_hits_soa* hits, void* stream);
int pt_trace_any_soa(pt_bvh* bvh, int64_t n, const pt_rays_soa* rays, int32_t* hit_out, void* stream);

/* Closest hit as pt_trace_closest, allocating nothing: no hipMalloc, hipFree or allocating
 * library call, for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation). No counterpart in the reference. Its
 * work cursors and, for launches of at least 2^20 rays, the coherence order's keys,
 * indices and sort storage live in a fixed scratch array in each GPU's memory (about
 * 36 MiB, held from the library's load on that GPU to the end of the process). n may be at
 * most PT_NOALLOC_MAX_RAYS (PT_ERR_ARG above). Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams, other contexts or threads)
 * race and must be serialised by the caller. Hits equal pt_trace_closest's. */
#define PT_NOALLOC_MAX_RAYS (1 << 21)
int pt_trace_closest_noalloc(pt_bvh* bvh, int64_t n, const float* rays, float* t_out, int32_t* prim_out,
                             void* stream);

/* Closest hit as pt_trace_closest, synchronous, also returning the traversal work:
 * host_stats[0] = node slab tests, host_stats[1] = triangle tests, summed over rays. */
int pt_trace_stats(pt_bvh* bvh, int64_t n, const float* rays, float* t_out, int32_t* prim_out,
                   int64_t* host_stats);

/* Renders one image (replaces Core::Petrichor::Render, Core/Petrichor.cpp:34-102, with
 * per-pixel sampler seeds). rgb_out: width*height*3 floats, host pointer when
 * out_on_device == 0, device pointer otherwise. stats may be NULL. */
int pt_render(pt_scene* scene, pt_bvh* bvh, const pt_render_desc* desc, float* rgb_out,
              int32_t out_on_device, pt_render_stats* stats);

/* Denoising AOVs (the DenoisingAlbedo and DenoisingNormal passes of
 * Core::Petrichor::Render, Core/Petrichor.cpp:104-261, with Core/AOV/AOVDenoisingAlbedo.cpp
 * and AOVDenoisingNormal.cpp): each pixel averages desc->spp camera rays (the reference's
 * RenderSetting::numSppForDenoising), followed through glass and zero-roughness GGX to the
 * first other hit, for at most desc->max_bounces hits. Albedo: throughput x (1 for
 * Lambert, F0 for GGX, the environment on a miss); normal: throughput x the shading normal
 * (-direction on a miss). UV: the mean over desc->spp camera rays of the interpolated
 * texture coordinates (u, v, 0) of the first hit (misses add zero). Seeds per pixel as
 * pt_render; desc->integrator and flags are not read. rgb_out as pt_render; synchronous. */
#define PT_AOV_DENOISING_ALBEDO 0
#define PT_AOV_DENOISING_NORMAL 1
#define PT_AOV_UV 2 /* Core/AOV/AOVUVCoordinate.cpp: mean uv of the first hit, spp = numSamplesPerPixel */
int pt_render_aov(pt_scene* scene, pt_bvh* bvh, int32_t aov, const pt_render_desc* desc, float* rgb_out,
                  int32_t out_on_device, pt_render_stats* stats);

/* ---- batched rendering ---------------------------------------------------------
 * An engine holds a pool of path slots (one per pixel of the jobs in flight) in device
 * memory. Jobs of different scenes, sizes and settings share every launch: one
 * persistent trace launch and one shading launch per wavefront iteration, replayed as
 * a captured hipGraph. Jobs are admitted as slots free up and retired when done. No
 * reference counterpart: the reference renders one scene per Petrichor::Render call
 * (Core/Petrichor.cpp:34-102). */
typedef struct pt_engine pt_engine;

typedef struct pt_job_desc {
    pt_scene* scene;
    pt_bvh* bvh;
    pt_render_desc render;
    float* rgb_out;        /* width*height*3 floats */
    int32_t out_on_device; /* 1: rgb_out is a device pointer */
} pt_job_desc;

typedef struct pt_batch_stats {
    int64_t queries;    /* BVH queries of every job */
    int64_t iterations; /* wavefront iterations (graph replays) */
    double elapsed_ms;  /* device time from the first admission to the last retirement */
    double wall_ms;     /* host time of the same span */
    int64_t syncs;      /* host synchronisations */
} pt_batch_stats;

/* max_paths: path slots (pixels in flight); max_jobs: jobs in flight; multi_job = 1
 * for engines that mix jobs (pt_render_batch), 0 for one job at a time. */
int pt_engine_create(pt_context* ctx, int64_t max_paths, int32_t max_jobs, int32_t multi_job,
                     pt_engine** out);
int pt_engine_destroy(pt_engine* engine);

/* Launch shapes of an engine, chosen by petrichor_amd/autotune.py. No reference
 * counterpart (the reference's ThreadPool takes a thread count, ThreadPool.cpp:19-26). */
typedef struct pt_tuning {
    int32_t iterations_per_sync; /* graph replays between host synchronisations, 1..256 */
    int32_t trace_blocks_per_cu; /* persistent trace blocks of 256 threads per CU, 1..8 (0 = as many
                                    as the LDS stacks allow; larger values are clamped to that) */
    int32_t trace_chunk;         /* rays a wave takes from the shared cursor at a time, 64..4096 */
    int32_t shade_blocks_per_cu; /* shading blocks of 256 threads per CU, 1..32 */
    int32_t admission;           /* PT_ADMIT_*: when pt_render_batch admits the next job */
    int32_t solo_log2_samples;   /* 0, or k in 10..40: pt_render_batch renders each job of at least
                                    2^k samples on its own with the single-job kernels (one job at a
                                    time, on the context's single-job engine and stream) while the
                                    other jobs share this engine; the two streams' launches overlap.
                                    Images do not depend on it. */
    int32_t threads_per_ray;     /* 1 (0 reads as 1), or 2: the two lanes of a pair walk one ray of a
                                    binary BVH of a plain tree, each testing one child box or one
                                    triangle per step (the same walk and hits). Wide walks, deep
                                    trees and scenes with spheres keep one lane per ray. */
    int32_t leaf_batch;          /* k in 1..4 (0 reads as 3), or -1: the one-lane binary walk of a
                                    plain tree runs in wave-wide phases, node steps until at least
                                    1/2^(k-1) of a wave's rays are in a leaf, then triangle steps (the
                                    same walk and hits per ray; a wave runs one kind of step at a
                                    time); -1: every step runs both kinds. */
} pt_tuning;
/* pt_tuning.admission. PT_ADMIT_FIT: any pending job whose path slots are free (largest
 * first). PT_ADMIT_TAILS: a job only once every running job has handed out all its samples,
 * so that large jobs run one after another with up to half the pool each and the next job
 * starts in the slots the previous one's last paths leave free. Images do not depend on it. */
#define PT_ADMIT_FIT 0
#define PT_ADMIT_TAILS 1
int pt_engine_set_tuning(pt_engine* engine, const pt_tuning* tuning);
int pt_engine_get_tuning(pt_engine* engine, pt_tuning* tuning);
/* Launch shapes of this BVH's traces (pt_trace_* and pt_render of this BVH): its
 * trace_blocks_per_cu, trace_chunk, threads_per_ray and leaf_batch (the other fields are ignored). Hits
 * and images do not depend on them. */
int pt_bvh_set_trace_tuning(pt_bvh* bvh, const pt_tuning* tuning);
/* Renders n jobs to completion, admitting them in order as slots free up;
 * iterations_per_sync graph replays run between host synchronisations (0 = the
 * engine's tuning). */
int pt_render_batch(pt_engine* engine, int32_t n, const pt_job_desc* jobs, int32_t iterations_per_sync,
                    pt_batch_stats* stats);

/* ---- Serving ----------------------------------------------------------------------------
 * A render server keeps scenes and their BVHs resident on the device and runs requests
 * with continuous batching: a host thread admits queued requests into the shared path
 * pool whenever path slots and budget allow, so requests that arrive while others run
 * join the same graph-replayed iterations. No reference counterpart: the reference
 * renders one scene per Petrichor::Render call (Core/Petrichor.cpp:34-102), with a
 * progress callback (Petrichor.h:33-38) that pt_server_preview stands in for.
 * Every function is thread-safe. */
typedef struct pt_server pt_server;

typedef struct pt_server_desc {
    int32_t device;
    int32_t max_jobs;            /* requests running at once (0 = 64) */
    int64_t hbm_budget_bytes;    /* device memory for scenes, BVHs, the path pool and request images */
    int64_t max_paths;           /* path slots shared by running requests (0 = 4M) */
    int32_t iterations_per_sync; /* graph replays between host synchronisations (0 = 8) */
    int32_t solo_log2_samples;   /* 0, or k in 10..40: a request of at least 2^k samples renders alone on
                                    a second, single-job engine (the single-job kernels, its own stream)
                                    beside the shared pool, one such request at a time */
    int64_t solo_paths;          /* path slots of that engine (0 = 8M), charged to the budget */
} pt_server_desc;

#define PT_REQ_QUEUED 0
#define PT_REQ_RUNNING 1
#define PT_REQ_DONE 2
#define PT_REQ_FAILED 3

typedef struct pt_request_info {
    int32_t state;     /* PT_REQ_* */
    float progress;    /* completed samples / (pixels * spp) as of the last preview; 1 when done */
    double queue_ms;   /* submission to admission (running or done) */
    double latency_ms; /* submission to the final image on the host (done) */
} pt_request_info;

typedef struct pt_server_stats {
    int64_t submitted, completed, running, queued;
    int64_t resident_scenes, evicted_scenes;
    int64_t bytes_in_use, budget_bytes;
    int64_t queries;    /* BVH queries of every request so far */
    int64_t iterations; /* wavefront iterations */
    int64_t syncs;      /* host synchronisations of the request loop */
    double busy_ms;     /* loop time with requests on the device */
} pt_server_stats;

int pt_server_create(const pt_server_desc* desc, pt_server** out);
/* Stops the request loop (running requests are abandoned) and frees everything. */
int pt_server_destroy(pt_server* server);
/* Error text of the calling thread's last failed pt_server_* call. */
const char* pt_server_last_error(void);
/* Uploads a scene and builds its BVH (desc NULL = 16 bins, leaf <= 4, device builder).
 * When the budget is short, idle scenes (no queued or running request) are evicted,
 * least recently used first; PT_ERR_BUDGET if that is not enough. */
int pt_server_add_scene(pt_server* server, const pt_scene_desc* scene, const pt_bvh_desc* bvh, int64_t* scene_id);
/* PT_ERR_STATE if the scene was evicted or removed. A scene with requests is removed
 * when they finish. */
int pt_server_remove_scene(pt_server* server, int64_t scene_id);
int pt_server_scene_resident(pt_server* server, int64_t scene_id);
/* Queues a request; PT_ERR_STATE for an unknown or evicted scene, PT_ERR_BUDGET for a
 * request that can never fit the path pool or the budget. */
int pt_server_submit(pt_server* server, int64_t scene_id, const pt_render_desc* render, int64_t* request_id);
int pt_server_poll(pt_server* server, int64_t request_id, pt_request_info* info);
/* Waits until the request is done or failed, at most timeout_ms (< 0: no limit);
 * returns PT_OK with info filled, whichever state it is in. */
int pt_server_wait(pt_server* server, int64_t request_id, int32_t timeout_ms, pt_request_info* info);
/* Progressive result: the current estimate (finished pixels final, others the mean of
 * their completed samples, black before the first) into a host buffer of W*H*3 floats.
 * Blocks until the loop's next synchronisation while the request runs. */
int pt_server_preview(pt_server* server, int64_t request_id, float* rgb_out, float* progress);
/* Final image of a done request into a host buffer of W*H*3 floats; the request is
 * then released. PT_ERR_STATE before it is done. */
int pt_server_fetch(pt_server* server, int64_t request_id, float* rgb_out);
int pt_server_get_stats(pt_server* server, pt_server_stats* stats);

/* ---- output --------------------------------------------------------------------
 * 8-bit RGB of an image as the reference writes it to a PNG (the pixel loop of
 * Texture2D::Save, Core/Texture2D.cpp:199-222: ACES filmic, the RTCamp7 curve, degamma
 * and truncation, Core/Texture2D.cpp:18-91): rgb8_out[3 i + c] is the byte of
 * rgb[3 i + c], for any float input, NaN and infinities included. num_pixels * 3 values
 * each: device pointers when on_device != 0 (the call returns once the kernel is queued
 * on stream, a hipStream_t, or the context's stream when NULL), else host pointers (the
 * call stages them through device memory and returns when the bytes are on the host).
 * Writing the PNG file (stbi_write_png, Core/Texture2D.cpp:217-222) stays with the host. */
int pt_tonemap_rgb8(pt_context* ctx, int64_t num_pixels, const float* rgb, uint8_t* rgb8_out, int32_t on_device,
                    void* stream);

/* ---- input -----------------------------------------------------------------------
 * Triangles of a Wavefront OBJ text (len bytes) as Mesh::Load receives them from Assimp
 * 4.1.0 read with no post-processing (Core/Geometry/Mesh.cpp:18-125): one triangle per
 * face (its first three corners), in file order; positions[9 t..] the corners' positions,
 * normals[9 t..] their normals normalised (+Z for a face without normal indices), uvs[6 t..]
 * their texture coordinates (zero without; uvs may be NULL).
 * Numbers are read as Assimp's fast_atof reads them. Fills at most max_triangles (pass 0
 * to count) and sets *num_triangles to the file's count. Host pointers, no device.
 * Errors (with the line in err, err_len bytes): PT_ERR_ARG for malformed text or an index
 * out of range, PT_ERR_UNSUPPORTED for point and line elements. */
int pt_obj_parse(const char* text, int64_t len, int64_t max_triangles, float* positions, float* normals,
                 float* uvs, int64_t* num_triangles, char* err, int32_t err_len);

/* A PNG file's pixels as Texture2D::Load reads them (Core/Texture2D.cpp:107-160): 8-bit
 * RGB as stbi_load(..., 3) gives it (grey replicated, alpha dropped, palettes expanded,
 * 16-bit samples reduced to their high byte), each byte / 255, then powf(x, 2.2) per
 * channel when color != 0 (ApplyGamma, Core/Texture2D.cpp:44-56). Sets *width, *height;
 * with rgb_out == NULL only those, else fills width*height*3 floats (row-major, top row
 * first). Non-interlaced images of bit depth 8 or 16. Host pointers, no device. Errors:
 * PT_ERR_ARG with the reason in err. */
int pt_png_decode(const uint8_t* data, int64_t len, int32_t color, int32_t* width, int32_t* height, float* rgb_out,
                  char* err, int32_t err_len);

/* An image file's pixels as Texture2D::Load reads a non-HDR texture with stbi_load(..., 3)
 * (Core/Texture2D.cpp:107-160), for the formats whose pixels are fixed by the file: PNG
 * (as pt_png_decode), BMP (1/4/8-bit palettes, 16/24/32-bit, bit-field masks; not
 * run-length encoded), TGA (colour-mapped, true-colour and grey, raw or run-length) and
 * binary PNM (P5, P6, maximum value <= 255), told apart by their headers. 8-bit RGB as stb
 * gives it, each byte / 255, then powf(x, 2.2) when color != 0. Outputs as pt_png_decode.
 * JPEG, GIF and PSD files: PT_ERR_UNSUPPORTED; malformed files: PT_ERR_ARG (reason in err). */
int pt_image_decode(const uint8_t* data, int64_t len, int32_t color, int32_t* width, int32_t* height, float* rgb_out,
                    char* err, int32_t err_len);

#ifdef __cplusplus
}
#endif
#endif
