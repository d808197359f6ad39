// pt_kernels.h — device data views and kernel launchers of the render path.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/petrichor_hip.h"
#include "pt_device.h"

// Traversal stack entries per ray. A walk holds at most one pending sibling per
// tree level plus the current node, so trees up to PT_MAX_STACK - 2 levels deep
// are accepted (checked when the BVH is built).
#define PT_MAX_STACK 64
// Wide trees: a level leaves at most 7 (8-wide) or 3 (4-wide) children pending, so those
// stacks take trees of up to 27 and 31 wide levels (checked when the wide nodes are made).
#define PT_STACK8 192
#define PT_STACK4 96

namespace pt {

// Binary BVH in device memory: inner nodes as 4 float4 (left box, right box, refs),
// triangles in leaf order as 3 float4 (p0, e1, e2, n).
struct BvhView {
    const float4* nodes;
    const float4* tris;
    const int32_t* prim_ids;  // leaf slot -> triangle index
    const int2* big_leaves;   // (first slot, count) of leaves over PT_REF_MAX_SMALL_LEAF triangles
    V3 root_lo, root_hi;
    uint32_t root_ref;        // PT_REF_* encoding, PT_REF_EMPTY for an empty scene
    const int2* inner_range = nullptr; // leaf-slot range of every binary inner node (ties, wide walks)
    const uint4* nodes8 = nullptr;  // wide nodes (pt_wide.hip), null for the binary traversal
    uint32_t root8 = PT_REF_EMPTY;  // root of the wide tree (wide node 0, or a leaf)
    int width = 2;                  // node width of the walk: 2, or 4 / 8 with nodes8
    // parent of every binary inner node (-1 for the root), set only for trees deeper than
    // the traversal stack, which are then walked without one
    const int32_t* parents = nullptr;
    // per leaf slot: 1 for a sphere (its record holds centre and radius); null without spheres
    const int8_t* kinds = nullptr;
    // 1 when every child box of the binary nodes has lo <= hi on every axis (no NaN): rays
    // with a finite origin and finite, nonzero 1/dir then take the branch-free slab test
    int boxes_ordered = 0;
};

struct SceneView {
    BvhView bvh;
    const int32_t* tri_mat;         // material per leaf slot (may be a Mix)
    const int32_t* tri_mat_shading; // GetMaterial(0.0f) per leaf slot (Triangle.cpp:151)
    const int32_t* tri_smooth;      // per leaf slot, or null
    const float* tri_nrm;           // 9 floats per leaf slot, or null
    const float* tri_uv = nullptr;  // 6 floats per leaf slot (texture coordinates), or null
    const TexDev* textures = nullptr; // material map textures
    const int2* mat_maps = nullptr;   // per material: albedo / reflectance map, roughness map (-1: none); or null
    const int* mat_nmaps = nullptr;   // per material: the normal map of a GGX (-1: none); or null
    const float2* map_strength = nullptr; // per material: normal and roughness map strengths; or null (1, 1)
    const MaterialDev* mats;
    const float4* lights;           // lights in AppendLight order, 3 float4 each: a triangle record, or a sphere's
    const int8_t* light_kinds = nullptr; // per light: 1 for a sphere; null without sphere lights
    int num_lights;
    CameraDev cam;
    EnvDev env;
};

// A job's work is num_pixels * spp samples, handed out to its path slots one sample at a
// time (sample-major: sample s of the job's pixel i is item s * num_pixels + i). A finished
// sample's colour goes to `samples` at the same index; once every sample is done, the
// pixel is the sum of its samples in sample order over spp (k_resolve).
struct JobView {
    int width, height, spp, max_bounces, integrator;
    uint32_t seed;
    float* out;                 // width*height*3 floats of the image
    unsigned long long* done;   // samples finished so far
    unsigned long long* cursor; // next item to hand out
    float* samples;             // 3 floats per item
    unsigned long long items;   // num_pixels * spp
    int path_base;   // first path slot of this job
    int num_slots;   // path slots of this job
    int job_index;   // index in the engine's job table
    int pixel_base;  // the job renders pixels [pixel_base, pixel_base + num_pixels) of the image
    int num_pixels;
    const float2* mj = nullptr; // micro-jittered 2D sampler's Halton table (Samplers), or null
};

// A job as the kernels see it in the engine's job table.
struct JobDev {
    SceneView sc;
    JobView job;
};

// Per-path state, structure of arrays indexed by path slot.
struct PathView {
    uint4* rng;    // 3 per path: 1D stream, 2D x stream, 2D y stream
    float4* color; // sample colour, w = sin of the last BSDF direction to the normal
    float4* ray_o; // origin, w = prob
    float4* ray_d; // direction, w = ior
    float4* thr;   // throughput, w = BSDF pdf of the last sampled direction
    int4* meta;    // sample index, ray bounce, phase | pending << 8, loop bounce
    int32_t* pixel; // the job's pixel index of the current sample
    int32_t* job;   // engine job index of each path slot
    float4* pend_a; // pending shadow data (see k_shade)
    float4* pend_b;
    float4* pend_c;
    float4* pend_d;
    float4* pend_o;
    float4* pend_dir;
};

struct Queue {
    float4* rays; // 2 per entry: (o, dist_min), (d, dist_max)
    int32_t* path;
    int* count;
};

// One segment of a trace launch: `n` rays (8 floats each) traced closest-hit or any-hit.
// Results go to out_t / out_k indexed by path[i] (or by i when path is null); with
// map_prim the leaf slot is replaced by the triangle index, with hit_flag out_k gets 1/0.
struct TraceSegment {
    const float4* rays;
    const int32_t* path;
    const int* count; // device-side ray count
    int any;
    int map_prim;
    int hit_flag;
    float* out_t;
    int32_t* out_k;
    int rays_by_path = 0; // ray i is read at rays[path[i]] (the API's coherence-ordered launches)
};
// One record of a launch log: the device-clock span of a trace launch (first block's start
// to last block's end, in wall-clock ticks), the rays of its three segments and, when the
// launch counts its work, its node visits and triangle tests.
struct LaunchRec {
    unsigned long long t0, t1;
    unsigned long long nodes, tris;
    int items[3];
    int pad;
};
#define PT_LAUNCH_LOG_CAP 8192
struct TraceJob {
    TraceSegment seg[3];
    int num_seg;
    // Multi-job launches: the BVH of a ray is views[job_of_path[path]]; otherwise `bvh`.
    const BvhView* views;
    const int32_t* job_of_path;
    unsigned long long* queries; // optional: += rays traced by this launch
    int chunk = 64;              // rays a wave takes from the cursor at a time (>= 64; set by the launcher)
    int width = 2;               // node width of the walk (every BVH of a multi-job launch has the same)
    int ext = 0;                 // multi-job launches: some BVH has a deep tree or spheres (the extended walk)
    int leaf_shift = 0;          // phased binary walk: a triangle step runs once (leaf rays << leaf_shift) >= rays
    // optional launch log: the launch writes its span (and with log_counts its node visits
    // and triangle tests) into launch_log[*launch_index]; launch_log_tick then records the
    // ray counts and advances the index
    LaunchRec* launch_log = nullptr;
    const int* launch_index = nullptr;
    int log_counts = 0;
};

// Launch shapes chosen by the autotuner (defaults fit every workload reasonably).
struct LaunchTuning {
    int trace_blocks_per_cu = 0; // 256-thread blocks per CU (0: as many as their LDS stacks allow)
    int trace_chunk = 64;        // >= 64, the wave width
    int shade_blocks_per_cu = 8;
    int admission = 0; // PT_ADMIT_* (host-side scheduling of pt_render_batch)
    int solo_log2_samples = 0; // jobs of at least 2^k samples render alone on the single-job engine (0: none)
    int threads_per_ray = 1;   // 1, or 2 lanes per ray for the binary walk of plain trees (trace_step_pair)
    int leaf_batch = 3;        // -1 (off), or k in 1..4: phased binary walk (trace_step_phase), triangle
                               // steps once 1/2^(k-1) of a wave's rays are in a leaf
};
// The rays of a trace launch are handed out in PT_TRACE_PARTS parts, one per XCD: part x
// holds the x-th eighth of every segment, and the blocks that the hardware places on XCD
// x (workgroups go to the XCDs in turn) take their chunks from part x first, then from the
// other parts. A part's rays are neighbours in the queues (the same job, nearby pixels),
// so each XCD's L2 holds the BVH nodes of fewer jobs and screen regions, and the chunk
// atomics, which every wave issues on the same cursor otherwise, are spread over eight.
#define PT_TRACE_PARTS 8
#define PT_TRACE_STEAL 1             // other parts a wave takes rays from once its own part is empty
#define PT_TRACE_STEAL_ALL_RAYS (1 << 24) // ... and in launches of more rays, every other part
// A part's cursor is cursor[x * PT_CURSOR_STRIDE]: each on a 256-byte line of its own, so
// that the cursors' atomics are spread over memory channels.
#define PT_CURSOR_STRIDE 64
#define PT_CURSOR_INTS (PT_TRACE_PARTS * PT_CURSOR_STRIDE)

// Persistent trace launch over every segment of a job (rays handed out to waves in
// chunks from the PT_TRACE_PARTS cursors, which must be zero when the launch starts). stats (optional,
// 2 counters) accumulates node visits and triangle tests. The grid does not depend
// on the ray counts, so the launch can be captured in a graph.
hipError_t launch_trace_job(const BvhView& b, const TraceJob& job, int* cursor, unsigned long long* stats,
                            hipStream_t st, const LaunchTuning& tune = LaunchTuning());
// Zeroes counters[i * stride] for i < n (n <= 64).
hipError_t launch_zero(int* counters, int n, hipStream_t st, int stride = 1);
// Structure-of-arrays rays to the trace kernel's ray records, and its per-ray (t, leaf
// slot) results to structure-of-arrays hits (primitive index, +inf / -1 on a miss, and
// the barycentric u, v when u / v are not null), in place (pt_trace.hip).
hipError_t launch_pack_rays_soa(int64_t n, const float* ox, const float* oy, const float* oz, const float* dx,
                                const float* dy, const float* dz, const float* dmin, const float* dmax, float4* out,
                                hipStream_t st);
hipError_t launch_hits_soa(int64_t n, const BvhView& b, const float4* rays, float* t, int32_t* prim, float* u, float* v,
                           hipStream_t st);
// *p = value, by a kernel (no host copy).
// Zeroes n trace cursors (at the given stride) and the four queue counts at `counts`.
hipError_t launch_zero_iteration(int* cursors, int n, int stride, int* counts, hipStream_t st);
hipError_t launch_set(int* p, int value, hipStream_t st);
// After a logged trace launch: log[*index].items = the three segment counts, ++*index
// (while *index < PT_LAUNCH_LOG_CAP - 1).
hipError_t launch_log_tick(const int* seg_counts, LaunchRec* log, int* index, hipStream_t st);
// dst[i] = src[0] (copies a device counter into a log slot in stream order).
hipError_t launch_copy_u64(const unsigned long long* src, unsigned long long* dst, hipStream_t st);
// Exclusive prefix sum of n integers (rocPRIM); call with temp == nullptr to size temp.
template <typename T>
hipError_t exclusive_sum(void* temp, size_t& temp_bytes, const T* in, T* out, size_t n, hipStream_t st);
// Compute units of the current device (queried once, then cached; called before any
// graph capture so that no device query happens while capturing).
int cu_count();

// Progressive image of a running job into out (3 floats per pixel): the mean of each
// pixel's completed samples (the sample buffer must have been cleared when the job
// started); adds the job's completed samples to *samples.
hipError_t launch_preview(const JobView& job, const PathView& ps, float* out, unsigned long long* samples,
                          hipStream_t st);
// Sets up every path slot of a job with the job's first items (seeds, first camera ray),
// sets the job's item cursor and sample count, and appends the paths to the live list and
// their camera rays to the extension queue.
hipError_t launch_paths_init(const SceneView& sc, const JobView& job, const PathView& ps, const Queue& ext,
                             int* active, int* active_count, int* base, hipStream_t st);
// The image of a job whose samples are all done: each pixel is the sum of its samples in
// sample order, divided by spp (PathTracing.cpp:194-198).
hipError_t launch_resolve(const JobView& job, hipStream_t st);
// Fills a job's sample buffer with the "not done" pattern read by the progressive preview.
hipError_t launch_samples_clear(const JobView& job, hipStream_t st);
// Shading of every live path (count read on the device). With `jobs` non-null each
// path takes its scene and settings from jobs[ps.job[p]], else from sc / job.
struct ShadeArgs {
    PathView ps;
    const int* active_in;
    const int* active_count;
    int* active_out;
    int* active_out_count;
    Queue ext, sh_any, sh_closest;
    const float* hit_t;
    const int* hit_k;
    const float* sh_t;
    const int* sh_k;
};
hipError_t launch_shade(int integrator, const SceneView& sc, const JobView& job, const JobDev* jobs,
                        const ShadeArgs& a, hipStream_t st, const LaunchTuning& tune = LaunchTuning());

// The hit buffers of ShadeArgs, writable.
struct HitBuffers {
    float* hit_t;
    int* hit_k;
    float* sh_t;
    int* sh_k;
};
// Runs the live paths a.active_in[0 .. *a.active_count) to the end of their pixels, one
// thread per path (shading events and the rays they ask for, in the wavefront's order,
// with the same results); adds the rays traced to *queries. jobs: the multi-job table (the
// job of each path, then sc and job are unused), or null for a single-job render. width:
// the BVHs' node width (2, 4 or 8). The grid holds PT_FINISH_BLOCKS_PER_CU blocks of 256 per CU, looping over
// the paths.
#define PT_FINISH_BLOCKS_PER_CU 4
hipError_t launch_finish_paths(int integrator, const SceneView& sc, const JobView& job, const JobDev* jobs,
                               int width, const ShadeArgs& a, const HitBuffers& h, unsigned long long* queries,
                               hipStream_t st);

// An AOV of a w x h image, one thread per pixel (pt_kernels.hip): mode 0 the denoising
// albedo, 1 the denoising normal, 2 the first hit's texture coordinates; queries
// (optional) receives the rays traced.
hipError_t launch_render_aov(const SceneView& sc, int w, int h, int spp, int max_bounces, uint32_t seed,
                             int mode, int width, float* out, unsigned long long* queries, hipStream_t st);

// Coherence order of n rays (pt_order.hip): the ray indices sorted by the Morton code of
// the origin in the box (stable) end in *sorted, one of the two halves of idx; keys and
// idx hold 2n entries, temp ray_order_temp_bytes(n) bytes.
size_t ray_order_temp_bytes(int n);
// dst[i * row + c] = src[idx[i] * row + c] for i < n (pt_order.hip): per-triangle rows put
// in leaf-slot order.
hipError_t launch_gather_rows(const float* src, const int32_t* idx, int n, int row, float* dst, hipStream_t st);
hipError_t launch_ray_order(const float4* rays, int n, V3 box_lo, V3 box_hi, uint32_t* keys, int32_t* idx,
                            void* temp, size_t temp_bytes, hipStream_t st, int32_t** sorted);

// The fixed scratch of the allocation-free trace entry point (pt_order.hip): a static
// device array per GPU, addressed once per device by noalloc_scratch_init (with
// hipGetSymbolAddress; the array's name is not a device address in host code).
#define PT_MAX_DEVICES 64
struct NoallocScratch {
    int* cursor;   // PT_CURSOR_INTS work cursors, then the ray count
    uint32_t* keys; // 2 x PT_NOALLOC_MAX_RAYS
    int32_t* idx;   // 2 x PT_NOALLOC_MAX_RAYS
    void* temp;     // the radix sort's temporary storage
    size_t temp_bytes;
};
hipError_t noalloc_scratch_init(int device);
// Null members until noalloc_scratch_init succeeded for the device.
NoallocScratch noalloc_scratch(int device);

// The reference's 8-bit PNG bytes of num_pixels RGB float pixels (pt_post.hip).
hipError_t launch_tonemap_rgb8(const float* rgb, uint8_t* out, int64_t num_pixels, hipStream_t st);

} // namespace pt
