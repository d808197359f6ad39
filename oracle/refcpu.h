/* refcpu.h — C API of the CPU oracle for Petrichor's rendering hot path.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a clean-room CPU restatement of the
 * reference renderer (LibPetrichor/Core, see refcpu.cpp for per-function
 * citations). Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it, and only as the checker or as the timed CPU baseline; the
 * product library (petrichor_amd) never calls it.
 *
 * Scene layout shared with the product's C-ABI (include/petrichor_hip.h):
 *   tri_pos      9 floats per triangle (p0, p1, p2)
 *   tri_nrm      9 floats per triangle (vertex normals) or NULL
 *   tri_mat      material index per triangle
 *   tri_smooth   0 = flat, 1 = smooth shading per triangle, or NULL (all flat)
 *   mat_type     0 Lambert, 1 Glass, 2 Emission, 3 GGX ("Glossy"), 4 Mix
 *                (the order of Core/Material/MaterialBase.h:15-22)
 *   mat_param    8 floats per material: color[3] (kd / ior-colour / Le / f0),
 *                p3 (GGX roughness, Glass ior, Mix weight), p4, p5 (Mix
 *                sub-material indices), p6 (importance sampling enabled), p7 unused
 *   light_tris   triangle indices registered with Scene::AppendLight, in order
 *   cam          16 floats: pos[3] right[3] up[3] forward[3] focus_dist h_perf aperture 0
 *   env_rgb      env_w*env_h*3 floats (row-major, RGB) or NULL (no texture)
 *   env_param    5 floats: base_color[3], z_rotation, use_importance_sampling
 */
#ifndef PETRICHOR_ORACLE_REFCPU_H
#define PETRICHOR_ORACLE_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void* oracle_scene_create(int num_tris, const float* tri_pos, const float* tri_nrm,
                          const int32_t* tri_mat, const int32_t* tri_smooth,
                          int num_mats, const int32_t* mat_type, const float* mat_param,
                          int num_lights, const int32_t* light_tris, const float* cam,
                          int env_w, int env_h, const float* env_rgb, const float* env_param);
void oracle_scene_destroy(void* scene);
/* Vertex texture coordinates, 6 floats per triangle (u, v of p0, p1, p2); zero without. */
void oracle_scene_set_uv(void* scene, const float* uv);
/* Material maps: num_textures RGB textures (sizes: w, h each) and 2 map indices per
 * material (albedo or reflectance map, roughness map; -1 none). */
void oracle_scene_set_maps(void* scene, int num_textures, const int32_t* sizes, const float* const* rgb,
                           const int32_t* mat_maps);
/* Spheres (centre, radius; a material each) as primitives after the triangles. */
void oracle_scene_set_spheres(void* scene, int n, const float* spheres, const int32_t* mats);
/* Sphere::Intersect, Interpolate and SampleSurface of one sphere (4 floats): out12 = hit,
 * t, the normal at the hit, the sampled position and its pdf from point p (RandomSampler2D
 * seeds), the sampled normal. */
void oracle_sphere(const float* sphere, const float* ray6, const float* p, uint32_t seed0, uint32_t seed1,
                   float* out12);
/* GGX normal maps (one texture index per material, -1 none; or NULL) and map strengths
 * (2 per material: normal, roughness; or NULL), over the textures set above. */
void oracle_scene_set_normal_maps(void* scene, const int32_t* nmaps, const float* strengths);
/* GGX::GetNormal of a normal-mapped GGX (GGX.h:96-119) for a shading normal, tangent, texel
 * and strength. */
void oracle_ggx_map_normal(const float* normal, const float* tangent, const float* texel, float strength,
                           float* out);
/* Triangle::Interpolate's normal and tangent (Triangle.cpp:90-147) at distance t along a ray
 * (6 floats: o, dir) on one triangle (9 position floats; vertex normals or NULL for flat;
 * 6 texture coordinates or NULL). */
void oracle_interpolate_frame(const float* pos, const float* nrm, const float* uv, const float* ray, float t,
                              float* normal_out, float* tangent_out);

/* Binary binned-SAH BVH, Core/Accel/BinnedSAHBVH.cpp:15-225 (num_bins = 16 there). */
void* oracle_bvh_build(void* scene, int num_bins);
void oracle_bvh_destroy(void* bvh);
/* A tree given as node tables (as oracle_bvh_export writes them), e.g. one built by the
 * product, so that traversal results can be compared on an identical leaf order. */
void* oracle_bvh_from_tables(void* scene, int num_nodes, const float* boxes, const int32_t* child,
                             const int32_t* range, const int32_t* leaf, const int32_t* prim_ids);
int oracle_bvh_num_nodes(void* bvh);
int oracle_bvh_max_depth(void* bvh);
/* Node table in the reference's order: box[6] (lower, upper), child[2], range[2], leaf. */
void oracle_bvh_export(void* bvh, float* boxes, int32_t* child, int32_t* range, int32_t* leaf,
                       int32_t* prim_ids);

/* Closest-hit queries. rays: 8 floats each (o[3], dist_min, d[3], dist_max), the layout
 * of the product's trace entry points.
 * Outputs t (+inf on a miss) and prim (-1 on a miss). stats (optional, 2 int64):
 * node visits (slab tests) and triangle tests summed over the rays. */
void oracle_trace_closest(void* scene, void* bvh, int64_t n, const float* rays, float* t_out,
                          int32_t* prim_out, int64_t* stats);
/* Core/Accel/BruteForce.cpp:21-50: layout-independent closest hit. */
void oracle_trace_brute(void* scene, int64_t n, const float* rays, float* t_out, int32_t* prim_out);

/* Render with the MIS integrator (integrator 0, Core/Integrator/PathTracing.cpp), the
 * simple path tracer (integrator 1, Core/Integrator/SimplePathTracing.cpp) or a denoising
 * AOV (2: albedo, Core/AOV/AOVDenoisingAlbedo.cpp; 3: normal, AOVDenoisingNormal.cpp;
 * spp is then the denoising sample count; 4: the UV AOV, Core/AOV/AOVUVCoordinate.cpp), with
 * samplers seeded per sample (oracle_sample_seed), 16x16 tiles on num_threads workers
 * (0 = all hardware threads).
 * If pixels != NULL only those num_pixels pixel indices are rendered (rgb_out then holds
 * 3 floats per listed pixel). Returns the number of BVH queries issued. work (optional,
 * 9 int64): per query kind (extension, area-light shadow, environment shadow) the
 * number of queries, node visits and triangle tests. */
int64_t oracle_render(void* scene, void* bvh, int width, int height, int spp, int max_bounces,
                      int integrator, uint32_t seed, int num_threads, const int32_t* pixels,
                      int64_t num_pixels, float* rgb_out, int64_t* work);
/* oracle_render with the renders' 2D sampler chosen: sampler_length 0 for the random
 * sampler, > 0 for MicroJitteredSampler over Halton sequences of that length (the
 * integrators 0 and 1 only). */
int64_t oracle_render_sampler(void* scene, void* bvh, int width, int height, int spp, int max_bounces,
                              int integrator, uint32_t seed, int num_threads, const int32_t* pixels,
                              int64_t num_pixels, float* rgb_out, int64_t* work, int sampler_length);
/* Math::Halton::Initialize's table (length values in `base`). */
void oracle_halton(int length, int base, float* out);
/* n draws (x, y pairs) of MicroJitteredSampler(seed) after Initialize(length). */
void oracle_microjittered(uint32_t seed, int length, int n, float* out);

/* Building blocks, exposed for unit tests against the reference translation units. */
uint32_t oracle_pixel_seed(uint32_t job_seed, uint32_t pixel, uint32_t stream);
/* The seed of stream `stream` (0: 1D, 1: 2D x, 2: 2D y) of sample `sample` of a pixel; sample
 * 0's seeds are oracle_pixel_seed's. */
uint32_t oracle_sample_seed(uint32_t job_seed, uint32_t pixel, uint32_t sample, uint32_t stream);
void oracle_xorshift(uint32_t seed, int n, uint32_t* u32_out, float* f_out);
/* Camera::GenerateRay with a RandomSampler2D(seed0, seed1): out 10 floats
 * (o[3], dir[3], throughput[3], samples consumed). */
void oracle_camera_ray(const float* cam, int i, int j, int w, int h, uint32_t seed0, uint32_t seed1,
                       float* out);
/* The camera ray of sample `sample` of pixel (i, j) as the renderers draw it: the sample's
 * 2D sampler (seeded per sample, after the stream warm-up). out as oracle_camera_ray. */
void oracle_sample_camera_ray(const float* cam, int i, int j, int w, int h, uint32_t job_seed, uint32_t sample,
                              float* out);
/* Material::CreateNextRay for material `mat` (mat_type/mat_param as above) on a shading
 * point {pos[3], normal[3]} for an incoming ray {o[3], dir[3], throughput[3], bounce, prob,
 * ior}; RandomSampler2D(seed0, seed1). out: o[3] dir[3] throughput[3] bounce prob ior. */
void oracle_next_ray(int mat_type, const float* mat_param, const float* shading, const float* ray_in,
                     uint32_t seed0, uint32_t seed1, float* out);
/* BxDF (3 floats) and PDF (1 float) for an incoming/outgoing direction pair. */
void oracle_bsdf_eval(int mat_type, const float* mat_param, const float* normal, const float* dir_in,
                      const float* dir_out, float* out4);
/* Triangle::Intersect on one triangle: returns 1 on a hit and writes t. */
int oracle_triangle_intersect(const float* tri9, const float* ray6, float* t);
/* Triangle::SampleSurface: out pos[3] normal[3] pdf_area. */
void oracle_triangle_sample(const float* tri9, uint32_t seed0, uint32_t seed1, float* out7);
/* Environment: GetColor (3), ImportanceSampling (dir[3], pdf), GetImportanceSamplingPDF. */
void oracle_env_color(void* scene, const float* dir, float* rgb);
void oracle_env_sample(void* scene, uint32_t seed0, uint32_t seed1, float* out4);
float oracle_env_pdf(void* scene, const float* dir);
/* The double-precision libm routines the oracle and the device share (pt_libm.h):
 * fn 0 sin(a), 1 cos(a), 2 acos(a), 3 atan2(a, b), n values. */
void oracle_libm(int fn, int64_t n, const double* a, const double* b, double* out);
/* PathTracing::SampleLight (PathTracing.cpp:329-386) with RandomSampler2D(seed0, seed1)
 * and the random value rnd; want_env != 0 for the call of CalcLightContribution (an extra
 * environment slot), 0 for the MIS weight of an emitter hit. Returns the slot picked (a
 * light index) or -1 for the environment slot; out: pos[3], normal[3], pdf_area (-1 when
 * not written), then the number of 2D draws consumed. */
int oracle_sample_light(void* scene, float rnd, uint32_t seed0, uint32_t seed1, int want_env, float* out8);
/* OrthonormalBasis::Build(w) + GetDir(theta, phi): out u[3] v[3] w[3] dir[3]. */
void oracle_onb(const float* w, float theta, float phi, float* out12);
/* Texture2D::Save's PNG bytes of num_pixels RGB pixels (Core/Texture2D.cpp:199-215), on
 * num_threads host threads (0: all cores). */
void oracle_tonemap_rgb8(int64_t num_pixels, const float* rgb, uint8_t* out, int num_threads);

#ifdef __cplusplus
}
#endif
#endif
