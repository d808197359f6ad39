// pt_kernels.hip — gfx950 kernels of the wavefront path tracer: path set-up and
// camera rays, the shading stage (NEE + MIS, BSDF sampling, accumulation) and queue
// compaction. BVH traversal lives in pt_trace.hip.
#include <hip/hip_runtime.h>

#include "pt_device.h"
#include "pt_kernels.h"
#include "pt_traverse.h"

namespace pt {

// ---------------------------------------------------------------------------------
// Wavefront path tracer
// ---------------------------------------------------------------------------------

enum Phase { kPhCamera = 0, kPhBounce = 1, kPhWait = 2, kPhDone = 3 };
enum Pending { kPendNone = 0, kPendEnv = 1, kPendArea = 2 };

__device__ __forceinline__ V3 ld3(const float4 v) { return v3(v.x, v.y, v.z); }
__device__ __forceinline__ float4 st3(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__device__ __forceinline__ uint4 pack_xs(const XorShift& s) { return make_uint4(s.x, s.y, s.z, s.w); }
__device__ __forceinline__ XorShift unpack_xs(uint4 v)
{
    XorShift s;
    s.x = v.x;
    s.y = v.y;
    s.z = v.z;
    s.w = v.w;
    return s;
}

__device__ __forceinline__ TriDev load_tri(const float4* tris, int k)
{
    const float4 r0 = tris[3 * k], r1 = tris[3 * k + 1], r2 = tris[3 * k + 2];
    TriDev t;
    t.p0 = v3(r0.x, r0.y, r0.z);
    t.e1 = v3(r0.w, r1.x, r1.y);
    t.e2 = v3(r1.z, r1.w, r2.x);
    t.n = v3(r2.y, r2.z, r2.w);
    return t;
}

__device__ __forceinline__ MaterialDev load_mat(const SceneView& s, int m) { return s.mats[m]; }

// GeometryBase::GetMaterial (GeometryBase.h:59-69): Mix picks one of two.
__device__ __forceinline__ int material_of(const SceneView& s, int slot, float rnd)
{
    const int m = s.tri_mat[slot];
    const MaterialDev& md = s.mats[m];
    if (md.type == kMatMix) return md.mix <= rnd ? md.mat0 : md.mat1;
    return m;
}

__device__ __forceinline__ bool is_sphere(const SceneView& s, int k) { return s.bvh.kinds != nullptr && s.bvh.kinds[k] != 0; }

// Shading point of a hit (Triangle::Interpolate, Triangle.cpp:90-153, or
// Sphere::Interpolate, Sphere.cpp:61-88) at leaf slot k, of the record tr loaded from that
// slot.
__device__ __forceinline__ void shade_point_tr(const SceneView& s, const TriDev& tr, int k, V3 o, V3 d, float t,
                                               V3* pos, V3* nrm)
{
    if (is_sphere(s, k)) {
        sphere_shading(tr.p0, tr.e1.x, o, d, t, pos, nrm);
        return;
    }
    const bool smooth = s.tri_smooth != nullptr && s.tri_smooth[k] != 0;
    V3 n0 = v3(0.0f, 0.0f, 1.0f), n1 = n0, n2 = n0;
    if (smooth) {
        const float* q = s.tri_nrm + 9 * (size_t)k;
        n0 = v3(q[0], q[1], q[2]);
        n1 = v3(q[3], q[4], q[5]);
        n2 = v3(q[6], q[7], q[8]);
    }
    tri_shading(tr, smooth, n0, n1, n2, o, d, t, pos, nrm);
}
__device__ __forceinline__ void shade_point(const SceneView& s, int k, V3 o, V3 d, float t, V3* pos, V3* nrm)
{
    shade_point_tr(s, load_tri(s.bvh.tris, k), k, o, d, t, pos, nrm);
}

// Texture coordinates of a hit at leaf slot k (Triangle::Interpolate, Triangle.cpp:90-108):
// the barycentric weights of the hit, uv = w1 (uv1 - uv0) + w2 (uv2 - uv0) + uv0, with the
// third component zero; all zero when the scene has no texture coordinates.
__device__ __forceinline__ V3 hit_uv(const SceneView& s, int k, V3 o, V3 d)
{
    if (s.tri_uv == nullptr || is_sphere(s, k)) return v3(0.0f, 0.0f, 0.0f); // a sphere has none (Sphere.cpp:81)
    const TriDev tr = load_tri(s.bvh.tris, k);
    const float inv_det = 1.0f / dot(-d, tr.n);
    const V3 vec = o - tr.p0;
    const float w1 = dot(vec, cross(d, tr.e2)) * inv_det;
    const float w2 = dot(vec, cross(tr.e1, d)) * inv_det;
    const float* q = s.tri_uv + 6 * (size_t)k;
    const V3 uv0 = v3(q[0], q[1], 0.0f), uv1 = v3(q[2], q[3], 0.0f), uv2 = v3(q[4], q[5], 0.0f);
    return w1 * (uv1 - uv0) + w2 * (uv2 - uv0) + uv0;
}

// Material m at a hit with texture coordinates uv, its maps applied: Lambert::GetAlbedo =
// kd x the albedo map, GGX::GetF0's reflectance x the reflectance map, GGX::GetAlpha =
// (strength x luminance of the roughness map)^2, the strength 1 unless set (Lambert.cpp:
// 117-126, GGX.h:57-75,89-93,126-139). Without maps the material as stored.
__device__ __forceinline__ MaterialDev material_at(const SceneView& s, int m, V3 uv)
{
    MaterialDev md = load_mat(s, m);
    if (s.mat_maps == nullptr) return md;
    const int2 mm = s.mat_maps[m];
    if (md.type == kMatLambert && mm.x >= 0) {
        md.albedo = md.color * tex_by_uv(s.textures[mm.x], uv.x, uv.y);
    } else if (md.type == kMatGlossy) {
        if (mm.x >= 0) md.color = md.color * tex_by_uv(s.textures[mm.x], uv.x, uv.y);
        if (mm.y >= 0) {
            const float k = s.map_strength != nullptr ? s.map_strength[m].y : 1.0f;
            const float r = k * luminance(tex_by_uv(s.textures[mm.y], uv.x, uv.y));
            md.alpha = r * r;
        }
    }
    return md;
}

// The normal the GGX functions take at leaf slot k for material m (of type `type`) with
// shading normal n and texture coordinates uv: GGX::GetNormal (GGX.h:96-119) when the
// material has a normal map, else n. The tangent is Triangle::Interpolate's
// (Triangle.cpp:110-147): diffUV2.y e1 - diffUV1.y e2, made orthogonal to a smooth
// triangle's normal, UnitX when its squared length is 0. The frame is
// OrthonormalBasis::Build(n, tangent) (OrthonormalBasis.cpp:29-35), the map's texel t
// gives l = normalize(2 t - 1), and the normal is (1 - s) l.x U + l.y V + (1 - s) l.z W
// with the strength s, not normalised again.
__device__ __forceinline__ V3 ggx_map_normal_at(const SceneView& s, int m, int k, V3 n, V3 uv)
{
    V3 tg = v3(0.0f, 0.0f, 0.0f); // Sphere::Interpolate leaves the tangent zero
    if (!is_sphere(s, k)) {
        const TriDev tr = load_tri(s.bvh.tris, k);
        float du1 = 0.0f, du2 = 0.0f;
        if (s.tri_uv != nullptr) {
            const float* q = s.tri_uv + 6 * (size_t)k;
            du1 = q[3] - q[1]; // diffUV1.y = uv1.v - uv0.v
            du2 = q[5] - q[1]; // diffUV2.y = uv2.v - uv0.v
        }
        tg = du2 * tr.e1 - du1 * tr.e2;
        if (s.tri_smooth != nullptr && s.tri_smooth[k] != 0) tg = tg - dot(n, tg) * n;
        if (sq_length(tg) == 0.0f) tg = v3(1.0f, 0.0f, 0.0f);
    }
    const V3 w = normalized(n), u = normalized(tg), v = cross(w, u);
    const V3 t = tex_by_uv(s.textures[s.mat_nmaps[m]], uv.x, uv.y);
    const V3 l = normalized(2.0f * t - v3(1.0f, 1.0f, 1.0f));
    const float st = s.map_strength != nullptr ? s.map_strength[m].x : 1.0f;
    return ((1.0f - st) * l.x) * u + l.y * v + ((1.0f - st) * l.z) * w;
}
__device__ __forceinline__ V3 ggx_map_normal(const SceneView& s, int m, int type, int k, V3 n, V3 uv)
{
    if (s.mat_nmaps == nullptr || type != kMatGlossy || s.mat_nmaps[m] < 0) return n;
    return ggx_map_normal_at(s, m, k, n, uv);
}
// Texture coordinates of the hit, read only when the scene has maps.
__device__ __forceinline__ V3 maps_uv(const SceneView& s, int k, V3 o, V3 d)
{
    return s.mat_maps != nullptr ? hit_uv(s, k, o, d) : v3(0.0f, 0.0f, 0.0f);
}

// PathTracing::SampleLight (PathTracing.cpp:329-386) for the shadow-ray origin p (which only
// sphere lights read). Returns false for the env slot.
__device__ __forceinline__ bool sample_light(const SceneView& s, V3 p0, float rnd, Samplers& smp, bool want_env,
                                             float* pdf_area, V3* pl_pos)
{
    const int L = s.num_lights;
    int index;
    if (!want_env) {
        index = (int)(rnd * (float)(uint64_t)L);
        index = min(index, (int)((uint64_t)L - 1));
    } else {
        index = (int)(rnd * (float)((uint64_t)L + 1));
        index = min(index, L);
        if (index >= L) return false;
    }
    float sum_area = 0.0f;
    *pl_pos = v3(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < L; ++i) {
        const TriDev tr = load_tri(s.lights, i);
        V3 p, n;
        float pa;
        if (s.light_kinds != nullptr && s.light_kinds[i] != 0) sphere_sample(tr.p0, tr.e1.x, p0, smp, &p, &n, &pa);
        else tri_sample(tr, smp, &p, &n, &pa);
        sum_area += 1.0f / pa;
        if (i == index) *pl_pos = p;
    }
    *pdf_area = 1.0f / sum_area;
    return true;
}

// Wave-aggregated queue append: one atomic per wavefront, lanes get consecutive slots.
__device__ __forceinline__ int queue_slot(int* counter, bool want)
{
    const unsigned long long mask = __ballot(want);
    if (!want) return -1;
    const int leader = __ffsll((long long)mask) - 1;
    const int rank = __popcll(mask & ((1ull << (threadIdx.x & 63)) - 1ull));
    int base = 0;
    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader, 64);
    return base + rank;
}

// Appends to four queues at once. Lanes 0-3 each own one queue's wave-level atomic, so
// the four atomics leave as one instruction and the wave waits for one round trip to
// L2 instead of four. Every lane of the wave must call it.
__device__ __forceinline__ void queue_slots4(int* const counter[4], const bool want[4], int slot[4])
{
    const int lane = threadIdx.x & 63;
    const unsigned long long lane_lt = (1ull << lane) - 1ull;
    unsigned long long m[4];
    for (int q = 0; q < 4; ++q) m[q] = __ballot(want[q]);
    int* const c = lane == 0 ? counter[0] : (lane == 1 ? counter[1] : (lane == 2 ? counter[2] : counter[3]));
    const unsigned long long mine = lane == 0 ? m[0] : (lane == 1 ? m[1] : (lane == 2 ? m[2] : m[3]));
    int base = 0;
    if (lane < 4 && mine != 0ull) base = atomicAdd(c, __popcll(mine));
    for (int q = 0; q < 4; ++q) {
        const int b = __shfl(base, q, 64);
        slot[q] = want[q] ? b + __popcll(m[q] & lane_lt) : -1;
    }
}

// Adds one per lane with `want` to that lane's counter: one atomic per distinct counter in
// the wave (the lanes of a wave usually share one job). Every lane of the wave must call it.
__device__ __forceinline__ void wave_count(int* counter, bool want)
{
    const int lane = threadIdx.x & 63;
    unsigned long long left = __ballot(want);
    while (left) {
        const int leader = __ffsll((long long)left) - 1;
        const unsigned long long lp = (unsigned long long)__shfl((long long)(uintptr_t)counter, leader, 64);
        const unsigned long long same = __ballot(want && (unsigned long long)(uintptr_t)counter == lp);
        if (lane == leader) atomicAdd(counter, __popcll(same));
        left &= ~same;
    }
}

// 64-bit form of wave_count (the samples a job has finished).
__device__ __forceinline__ void wave_count64(unsigned long long* counter, bool want)
{
    const int lane = threadIdx.x & 63;
    unsigned long long left = __ballot(want);
    while (left) {
        const int leader = __ffsll((long long)left) - 1;
        const unsigned long long lp = (unsigned long long)__shfl((long long)(uintptr_t)counter, leader, 64);
        const unsigned long long same = __ballot(want && (unsigned long long)(uintptr_t)counter == lp);
        if (lane == leader) atomicAdd(counter, (unsigned long long)__popcll(same));
        left &= ~same;
    }
}

// Takes one item per lane with `want` from that lane's cursor, one atomic per distinct
// cursor in the wave; lanes without `want` get ~0. Every lane of the wave must call it.
__device__ __forceinline__ unsigned long long wave_take(unsigned long long* cursor, bool want)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long lane_lt = (1ull << lane) - 1ull;
    unsigned long long left = __ballot(want);
    unsigned long long mine = ~0ull;
    while (left) {
        const int leader = __ffsll((long long)left) - 1;
        const unsigned long long lp = (unsigned long long)__shfl((long long)(uintptr_t)cursor, leader, 64);
        const unsigned long long same = __ballot(want && (unsigned long long)(uintptr_t)cursor == lp);
        long long base = 0;
        if (lane == leader) base = (long long)atomicAdd(cursor, (unsigned long long)__popcll(same));
        base = __shfl(base, leader, 64);
        if ((same >> lane) & 1ull) mine = (unsigned long long)base + (unsigned long long)__popcll(same & lane_lt);
        left &= ~same;
    }
    return mine;
}

// Block-level forms for the shading kernels (blocks of PT_SHADE_WAVES waves): one atomic
// per counter per block instead of per wave. Every thread of the block must call them, in
// the same order; each holds two barriers. Many atomics on one address are served one
// after another at the memory side (about 11 ns each on MI355X, measured), so a launch
// that appends millions of paths spends most of its time on them unless they are few.
#define PT_SHADE_WAVES 4
struct BlockScratch {
    int q_cnt[PT_SHADE_WAVES][4];  // per wave: entries of the four queues
    int q_base[4];                 // the block's first entry per queue
    int e_cnt[PT_SHADE_WAVES];     // per wave: lanes that take an item
    unsigned long long e_base;     // the block's first item
};

__device__ __forceinline__ void block_queue_slots4(int* const counter[4], const bool want[4], int slot[4],
                                                   BlockScratch& sh)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long lane_lt = (1ull << lane) - 1ull;
    unsigned long long m[4];
    for (int q = 0; q < 4; ++q) m[q] = __ballot(want[q]);
    if (lane < 4) sh.q_cnt[wave][lane] = __popcll(lane == 0 ? m[0] : (lane == 1 ? m[1] : (lane == 2 ? m[2] : m[3])));
    __syncthreads();
    if (threadIdx.x < 4) {
        int tot = 0;
        for (int w = 0; w < PT_SHADE_WAVES; ++w) tot += sh.q_cnt[w][threadIdx.x];
        sh.q_base[threadIdx.x] = tot != 0 ? atomicAdd(counter[threadIdx.x], tot) : 0;
    }
    __syncthreads();
    for (int q = 0; q < 4; ++q) {
        int b = sh.q_base[q];
        for (int w = 0; w < wave; ++w) b += sh.q_cnt[w][q];
        slot[q] = want[q] ? b + __popcll(m[q] & lane_lt) : -1;
    }
}

// Takes one item per lane with `want` from the block's cursor and adds as many finished
// samples to `done`; lanes without `want` get ~0.
__device__ __forceinline__ unsigned long long block_take(unsigned long long* cursor, unsigned long long* done,
                                                         bool want, BlockScratch& sh)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(want);
    if (lane == 0) sh.e_cnt[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < PT_SHADE_WAVES; ++w) tot += sh.e_cnt[w];
        unsigned long long b = 0;
        if (tot != 0) {
            b = atomicAdd(cursor, (unsigned long long)tot);
            atomicAdd(done, (unsigned long long)tot);
        }
        sh.e_base = b;
    }
    __syncthreads();
    if (!want) return ~0ull;
    unsigned long long b = sh.e_base;
    for (int w = 0; w < wave; ++w) b += (unsigned long long)sh.e_cnt[w];
    return b + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ void push_ray(const Queue& q, int slot, int path, V3 o, float dmin, V3 d, float dmax)
{
    q.rays[2 * slot] = make_float4(o.x, o.y, o.z, dmin);
    q.rays[2 * slot + 1] = make_float4(d.x, d.y, d.z, dmax);
    q.path[slot] = path;
}

// Start sample `s` of a path: Camera::GenerateRay (Camera.cpp:52-91).
__device__ __forceinline__ Ray camera_ray(const SceneView& sc, const JobView& job, int pixel, Samplers& smp)
{
    const int x = pixel % job.width, y = pixel / job.width;
    return generate_ray(sc.cam, x, y, job.width, job.height, smp);
}

// Starts item `item` (< job.items) on path p: sample s = item / num_pixels of the job's
// pixel item % num_pixels, with that sample's own sampler streams and its camera ray.
__device__ __forceinline__ Ray start_sample(const SceneView& sc, const JobView& job, const PathView& ps, int p,
                                            unsigned long long item)
{
    const unsigned long long np = (unsigned long long)job.num_pixels;
    const int i = (int)(item % np), s = (int)(item / np);
    const int pixel = job.pixel_base + i;
    Samplers smp = samplers_for_sample(job.seed, (uint32_t)pixel, (uint32_t)s, job.mj);
    const Ray r = camera_ray(sc, job, pixel, smp);
    ps.color[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ps.ray_o[p] = st3(r.o, r.prob);
    ps.ray_d[p] = st3(r.dir, r.ior);
    ps.thr[p] = st3(r.throughput, 0.0f);
    ps.meta[p] = make_int4(s, r.bounce, kPhCamera, 0);
    ps.rng[3 * p] = pack_xs(smp.s1);
    ps.rng[3 * p + 1] = pack_xs(smp.s2x);
    ps.rng[3 * p + 2] = pack_xs(smp.s2y);
    ps.pixel[p] = pixel;
    ps.job[p] = job.job_index;
    return r;
}

// The extension-ray interval of a camera ray: [0, FLT_MAX] for MIS (AccelBase.h:44-48),
// [kEps, kInfinity] for the simple integrator (SimplePathTracing.cpp:47).
__device__ __forceinline__ float camera_dmin(int integrator) { return integrator == 0 ? 0.0f : PT_EPS; }
__device__ __forceinline__ float camera_dmax(int integrator) { return integrator == 0 ? FLT_MAX : PT_INFINITY; }

// Reserves a job's first n paths in the live list and its camera rays in the extension
// queue: base[0] and base[1] receive the two counts before they grow by n.
__global__ void k_reserve_init(int* active_count, int* ext_count, int n, int* base)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    base[0] = *active_count;
    *active_count += n;
    base[1] = *ext_count;
    *ext_count += n;
}

// One-time set-up of every path slot of a job with the job's first items; the path joins
// the live list and its camera ray the extension queue, at the entries k_reserve_init
// reserved (the live slots are a prefix, so slot i takes entry base + i, without atomics
// on the two counts). Thread 0 sets the job's cursor past the items handed out here and
// zeroes its sample count.
__global__ void __launch_bounds__(256) k_paths_init(SceneView sc, JobView job, PathView ps, Queue ext, int* active,
                                                    const int* base)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < job.num_slots && (unsigned long long)i < job.items;
    const int p = job.path_base + i;
    if (i == 0) {
        *job.cursor = (unsigned long long)job.num_slots < job.items ? (unsigned long long)job.num_slots : job.items;
        *job.done = 0ull;
    }
    if (!live) return;
    const Ray r = start_sample(sc, job, ps, p, (unsigned long long)i);
    active[base[0] + i] = p;
    push_ray(ext, base[1] + i, p, r.o, camera_dmin(job.integrator), r.dir, camera_dmax(job.integrator));
}

// The shading stage of the MIS integrator (PathTracing::Render, PathTracing.cpp:20-200,
// CalcLightContribution :202-327) and of the simple integrator
// (SimplePathTracing::Render, SimplePathTracing.cpp:16-101). One thread per live path:
//   1. resolve the previous bounce's shadow ray,
//   2. consume the extension hit,
//   3. at a surface, sample a light (issue a shadow ray) and the BSDF (issue the next
//      extension ray); at the end of a sample, accumulate and start the next sample.
// What shading one path asks of the queues.
struct ShadeOut {
    bool keep, want_ext, want_any, want_closest;
    bool sample_end; // the path finished its sample (its colour is in the job's sample buffer)
    V3 ext_o, ext_d, sh_o, sh_d;
    float ext_dmin, ext_dmax;
};

template <int INTEGRATOR>
__device__ __forceinline__ void shade_one(const SceneView& sc, const JobView& job, const ShadeArgs& A, int p,
                                          ShadeOut& out)
{
    const PathView& ps = A.ps;
    const float* hit_t = A.hit_t;
    const int* hit_k = A.hit_k;
    const float* sh_t = A.sh_t;
    const int* sh_k = A.sh_k;
    bool keep = false, want_ext = false, want_any = false, want_closest = false, sample_end = false;
    V3 ext_o = v3(0.0f, 0.0f, 0.0f), ext_d = ext_o, sh_o = ext_o, sh_d = ext_o;
    float ext_dmin = 0.0f, ext_dmax = FLT_MAX;
    {
        Samplers smp;
        smp.s1 = unpack_xs(ps.rng[3 * p]);
        smp.s2x = unpack_xs(ps.rng[3 * p + 1]);
        smp.s2y = unpack_xs(ps.rng[3 * p + 2]);
        smp.mj = job.mj;
        const float4 o4 = ps.ray_o[p], d4 = ps.ray_d[p], t4 = ps.thr[p];
        Ray ray;
        ray.o = ld3(o4);
        ray.prob = o4.w;
        ray.dir = ld3(d4);
        ray.ior = d4.w;
        ray.throughput = ld3(t4);
        float pdf_bsdf = t4.w;
        int4 meta = ps.meta[p]; // x sample, y ray.bounce, z phase | pending << 8, w loop bounce
        ray.bounce = meta.y;
        int phase = meta.z & 0xff;
        const int pend = (meta.z >> 8) & 0xff;
        V3 color = ld3(ps.color[p]);
        float sin_n = ps.color[p].w;
        V3 direct = v3(0.0f, 0.0f, 0.0f); // a camera ray's miss or emitter: added unconditionally
        const int pixel = ps.pixel[p];
        bool sample_done = false, add_color = true;

        if (INTEGRATOR == 0) {
            // 1. shadow ray of the previous bounce
            if (pend == kPendEnv) {
                if (sh_k[p] < 0) color = color + ld3(ps.pend_a[p]);
            } else if (pend == kPendArea) {
                const int k = sh_k[p];
                if (k >= 0 && sc.mats[material_of(sc, k, next1(smp))].type == kMatEmission) {
                    const float4 pa = ps.pend_a[p]; // pl.xyz, pdf_area
                    const float4 pb = ps.pend_b[p]; // si.pos.xyz, cs
                    const float4 pc = ps.pend_c[p]; // T_in.xyz, pdf_mat
                    const float4 pd = ps.pend_d[p]; // f.xyz
                    const float4 so = ps.pend_o[p], sd = ps.pend_dir[p];
                    V3 lpos, lnrm;
                    shade_point(sc, k, ld3(so), ld3(sd), sh_t[p], &lpos, &lnrm);
                    const float dist = sq_length(ld3(pa) - lpos);
                    if (fabsf_(dist - 0.0f) < PT_EPS) {
                        const float l2 = sq_length(lpos - ld3(pb));
                        const float cos_p = fabsf_(dot(ld3(sd), lnrm));
                        float mis = 0.0f;
                        const float pdf_area = pa.w;
                        if (l2 > 0.0f) {
                            const float pl = pdf_area;
                            const float pbs = pc.w * cos_p / l2;
                            mis = pl * pl / (pl * pl + pbs * pbs);
                        }
                        const V3 li = sc.mats[sc.tri_mat_shading[k]].color;
                        const V3 kk = li * ld3(pd) * pb.w * cos_p / (pdf_area * l2);
                        color = color + mis * ld3(pc) * kk;
                    }
                }
            }
        }

        // 2. the extension hit
        bool at_surface = false;
        V3 si_pos = v3(0.0f, 0.0f, 0.0f), si_n = v3(0.0f, 0.0f, 1.0f), si_uv = v3(0.0f, 0.0f, 0.0f);
        int si_mat = 0;
        int loop_b = meta.w;
        if (phase == kPhWait) {
            sample_done = true;
        } else if (INTEGRATOR == 0) {
            const int k = hit_k[p];
            const float t = hit_t[p];
            // the hit triangle and its material are read before the branches that use
            // them (slot 0 on a miss), so that these loads overlap the path-state loads;
            // an empty scene (no slot 0, no triangle buffers) reads nothing
            TriDev pre_tr;
            pre_tr.p0 = pre_tr.e1 = pre_tr.e2 = pre_tr.n = v3(0.0f, 0.0f, 0.0f);
            int pre_mat = 0;
            if (sc.bvh.root_ref != PT_REF_EMPTY) {
                const int kk = k < 0 ? 0 : k;
                pre_tr = load_tri(sc.bvh.tris, kk);
                pre_mat = sc.tri_mat_shading[kk];
            }
            if (phase == kPhCamera) {
                if (k < 0) {
                    direct = ray.throughput * env_color(sc.env, ray.dir);
                    sample_done = true;
                    add_color = false;
                } else {
                    shade_point_tr(sc, pre_tr, k, ray.o, ray.dir, t, &si_pos, &si_n);
                    si_mat = pre_mat;
                    si_uv = maps_uv(sc, k, ray.o, ray.dir);
                    if (sc.mats[si_mat].type == kMatEmission) {
                        direct = ray.throughput * sc.mats[si_mat].color;
                        sample_done = true;
                        add_color = false;
                    } else {
                        at_surface = true;
                        loop_b = 0;
                    }
                }
            } else { // kPhBounce
                if (k >= 0) {
                    V3 npos, nn;
                    shade_point_tr(sc, pre_tr, k, ray.o, ray.dir, t, &npos, &nn);
                    const int nmat = pre_mat;
                    if (sc.mats[nmat].type == kMatEmission) {
                        const float l2 = sq_length(npos - ray.o);
                        const float cos_p = fabsf_(dot(ray.dir, nn));
                        float pdf_area = 0.0f;
                        V3 dummy;
                        sample_light(sc, ray.o, next1(smp), smp, false, &pdf_area, &dummy);
                        const float pb = pdf_bsdf;
                        const float pl = l2 / cos_p * pdf_area;
                        const float mis = pb * pb / (pl * pl + pb * pb);
                        color = color + mis * ray.throughput * sc.mats[nmat].color;
                        sample_done = true;
                    } else {
                        loop_b += 1;
                        if (loop_b < job.max_bounces) {
                            at_surface = true;
                            si_pos = npos;
                            si_n = nn;
                            si_mat = nmat;
                            si_uv = maps_uv(sc, k, ray.o, ray.dir);
                        } else {
                            sample_done = true;
                        }
                    }
                } else {
                    // the environment's colour and pdf share the direction's angles
                    if (env_use_is(sc.env)) {
                        const EnvAngles ang = env_angles(ray.dir);
                        float mis = 0.0;
                        if (ray.dir.x != 0 && ray.dir.y != 0 && sin_n > 0) {
                            const float pe = env_pdf_at(sc.env, ang) / (2.0f * PT_PI * PT_PI * sin_n);
                            const float pb = pdf_bsdf;
                            mis = pb * pb / (pe * pe + pb * pb);
                        } else {
                            mis = 1.0f;
                        }
                        color = color + mis * ray.throughput * env_color_at(sc.env, ang);
                    } else {
                        color = color + ray.throughput * env_color(sc.env, ray.dir);
                    }
                    sample_done = true;
                }
            }
        } else {
            // simple integrator: one hit, material chosen with a 1D sample
            const int k = hit_k[p];
            const float t = hit_t[p];
            if (k < 0) {
                color = color + ray.throughput * env_color(sc.env, ray.dir);
                sample_done = true;
            } else {
                const int m = material_of(sc, k, next1(smp));
                const V3 uv = maps_uv(sc, k, ray.o, ray.dir);
                const MaterialDev mat = material_at(sc, m, uv);
                if (mat.type == kMatEmission) {
                    color = color + ray.throughput * mat.color;
                    sample_done = true;
                } else {
                    shade_point(sc, k, ray.o, ray.dir, t, &si_pos, &si_n);
                    ray = next_ray(mat, ray, si_pos, ggx_map_normal(sc, m, mat.type, k, si_n, uv), smp);
                    bool cont = true;
                    if (ray.bounce > job.max_bounces) {
                        ray.prob *= 0.9f;
                        ray.prob = smax(0.1f, ray.prob);
                        if (next1(smp) < ray.prob) ray.throughput = ray.throughput / ray.prob;
                        else cont = false;
                    }
                    if (cont) {
                        want_ext = true;
                        ext_o = ray.o;
                        ext_d = ray.dir;
                        ext_dmin = PT_EPS;
                        ext_dmax = PT_INFINITY;
                        phase = kPhBounce;
                    } else {
                        sample_done = true;
                    }
                }
            }
        }

        // 3. a surface of the MIS loop: NEE then BSDF sampling
        int new_pend = kPendNone;
        if (INTEGRATOR == 0 && at_surface) {
            const MaterialDev mat = material_at(sc, si_mat, si_uv);
            // the GGX functions take the normal-mapped normal; the integrator's own cosines
            // and offsets keep the shading normal (PathTracing.cpp:206-300)
            const V3 bn = ggx_map_normal(sc, si_mat, mat.type, hit_k[p], si_n, si_uv);
            const float dd = -dot(ray.dir, si_n);
            const V3 pnee = si_pos + PT_EPS * copysignf_(1.0f, dd) * si_n;
            float pdf_area = 1.0f;
            V3 pl;
            const bool is_area = sample_light(sc, pnee, next1(smp), smp, true, &pdf_area, &pl);
            if (is_area) {
                if (sc.num_lights > 0) {
                    const V3 ldir = normalized(pl - pnee);
                    want_closest = true;
                    sh_o = pnee;
                    sh_d = ldir;
                    new_pend = kPendArea;
                    const V3 f = bxdf(mat, ray.dir, ldir, bn);
                    ps.pend_a[p] = st3(pl, pdf_area);
                    ps.pend_b[p] = st3(si_pos, fabsf_(dot(ldir, si_n)));
                    ps.pend_c[p] = st3(ray.throughput, bsdf_pdf(mat, ray.dir, ldir, bn));
                    ps.pend_d[p] = st3(f, 0.0f);
                    ps.pend_o[p] = st3(pnee, 0.0f);
                    ps.pend_dir[p] = st3(ldir, 0.0f);
                }
            } else if (env_use_is(sc.env)) {
                float pdf_uv = 0.0f;
                const V3 edir = env_sample(sc.env, smp, &pdf_uv);
                const V3 eo = si_pos + PT_EPS * copysignf_(1.0f, dd) * si_n;
                want_any = true;
                sh_o = eo;
                sh_d = edir;
                new_pend = kPendEnv;
                const float cs = fabsf_(dot(edir, si_n));
                const float sn = __builtin_sqrtf(smax(0.0f, 1.0f - cs * cs));
                const float pb = 2.0f * PT_PI * PT_PI * sn * bsdf_pdf(mat, ray.dir, edir, bn);
                const float mis = pdf_uv * pdf_uv / (pdf_uv * pdf_uv + pb * pb);
                const V3 li = env_color(sc.env, edir);
                const V3 f = bxdf(mat, ray.dir, edir, bn);
                const V3 c = mis * ray.throughput * (li * f * sn * cs * 2.0f * PT_PI * PT_PI) / pdf_uv;
                ps.pend_a[p] = st3(c, 0.0f);
            }
            const V3 prev_dir = ray.dir;
            ray = next_ray(mat, ray, si_pos, bn, smp);
            if (fabsf_(sq_length(ray.throughput) - 0.0f) < PT_EPS) {
                if (new_pend != kPendNone) phase = kPhWait;
                else sample_done = true;
            } else {
                pdf_bsdf = bsdf_pdf(mat, prev_dir, ray.dir, bn);
                const float cs = fabsf_(dot(si_n, ray.dir));
                sin_n = __builtin_sqrtf(smax(0.0f, 1.0f - cs * cs));
                want_ext = true;
                ext_o = ray.o;
                ext_d = ray.dir;
                ext_dmin = 0.0f;
                ext_dmax = FLT_MAX;
                phase = kPhBounce;
            }
        }

        // end of a sample: its colour goes to the job's sample buffer (the MIS integrator
        // drops a sample with a negative component, PathTracing.cpp:191-195: it adds zero)
        const int sample = meta.x;
        if (sample_done) {
            V3 v;
            if (INTEGRATOR == 0) v = add_color ? (min_elem(color) >= 0.0f ? color : v3(0.0f, 0.0f, 0.0f)) : direct;
            else v = color;
            const size_t item = (size_t)sample * (size_t)job.num_pixels + (size_t)(pixel - job.pixel_base);
            float* o = job.samples + 3 * item;
            o[0] = v.x;
            o[1] = v.y;
            o[2] = v.z;
            sample_end = true;
            phase = kPhDone;
        }
        keep = phase != kPhDone;
        if (keep) {
            ps.rng[3 * p] = pack_xs(smp.s1);
            ps.rng[3 * p + 1] = pack_xs(smp.s2x);
            ps.rng[3 * p + 2] = pack_xs(smp.s2y);
            ps.ray_o[p] = st3(ray.o, ray.prob);
            ps.ray_d[p] = st3(ray.dir, ray.ior);
            ps.thr[p] = st3(ray.throughput, pdf_bsdf);
            ps.meta[p] = make_int4(sample, ray.bounce, phase | (new_pend << 8), loop_b);
            ps.color[p] = st3(color, sin_n);
        }
    }
    out.keep = keep;
    out.sample_end = sample_end;
    out.want_ext = want_ext;
    out.want_any = want_any;
    out.want_closest = want_closest;
    out.ext_o = ext_o;
    out.ext_d = ext_d;
    out.sh_o = sh_o;
    out.sh_d = sh_d;
    out.ext_dmin = ext_dmin;
    out.ext_dmax = ext_dmax;
}

// Grid-stride over the live paths (count read on the device, so the grid is fixed).
// Per-wave ballots compact the new rays into the queues and the surviving paths into
// the next live list.
template <int INTEGRATOR, bool MULTI>
__device__ __forceinline__ void shade_live_paths(const SceneView& sc1, const JobView& job1,
                                                 const JobDev* __restrict__ jobs, const ShadeArgs& A)
{
    static_assert(PT_SHADE_WAVES * 64 == 256, "shading blocks of 256 threads");
    __shared__ BlockScratch sh;
    const int n = *A.active_count;
    const int stride = gridDim.x * blockDim.x;
    // the loop is uniform over the block (its block-level appends hold barriers)
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const int i = base + threadIdx.x;
        const bool in_range = i < n;
        const int p = in_range ? A.active_in[i] : 0;
        ShadeOut o;
        o.keep = o.want_ext = o.want_any = o.want_closest = o.sample_end = false;
        const JobDev* J = nullptr;
        if (MULTI) {
            // A wave's paths usually belong to one job (a job's slots are contiguous and its
            // paths stay together in the live list). Then the job's scene and settings are
            // read at a wave-uniform address, with scalar loads; otherwise per lane.
            const int my_job = in_range ? A.ps.job[p] : -1;
            const int j0 = __builtin_amdgcn_readfirstlane(my_job);
            const bool uniform = __ballot(in_range && my_job != j0) == 0ull;
            if (uniform && j0 >= 0) {
                const JobDev* Ju = jobs + j0;
                J = Ju;
                if (Ju->job.integrator == 0) {
                    if (in_range) shade_one<0>(Ju->sc, Ju->job, A, p, o);
                } else {
                    if (in_range) shade_one<1>(Ju->sc, Ju->job, A, p, o);
                }
            } else if (in_range) {
                // the integrator is a per-job setting here
                J = jobs + my_job;
                if (J->job.integrator == 0) shade_one<0>(J->sc, J->job, A, p, o);
                else shade_one<1>(J->sc, J->job, A, p, o);
            }
        } else if (in_range) {
            shade_one<INTEGRATOR>(sc1, job1, A, p, o);
        }
        // a finished sample is counted and the path takes the job's next item, if any (one
        // job: one take per block; many jobs: per wave and job)
        {
            unsigned long long item;
            if (MULTI) {
                unsigned long long* done = nullptr;
                unsigned long long* cursor = nullptr;
                if (o.sample_end) {
                    done = J->job.done;
                    cursor = J->job.cursor;
                }
                wave_count64(done, o.sample_end);
                item = wave_take(cursor, o.sample_end);
            } else {
                item = block_take(job1.cursor, job1.done, o.sample_end, sh);
            }
            if (o.sample_end && item < (MULTI ? J->job.items : job1.items)) {
                const JobView& jv = MULTI ? J->job : job1;
                const Ray r = start_sample(MULTI ? J->sc : sc1, jv, A.ps, p, item);
                o.keep = true;
                o.want_ext = true;
                o.ext_o = r.o;
                o.ext_d = r.dir;
                o.ext_dmin = camera_dmin(jv.integrator);
                o.ext_dmax = camera_dmax(jv.integrator);
            }
        }
        int* const counters[4] = {A.ext.count, A.sh_any.count, A.sh_closest.count, A.active_out_count};
        const bool want[4] = {o.want_ext, o.want_any, o.want_closest, o.keep};
        int slot[4];
        block_queue_slots4(counters, want, slot, sh);
        if (o.want_ext) push_ray(A.ext, slot[0], p, o.ext_o, o.ext_dmin, o.ext_d, o.ext_dmax);
        if (o.want_any) push_ray(A.sh_any, slot[1], p, o.sh_o, 0.0f, o.sh_d, FLT_MAX);
        if (o.want_closest) push_ray(A.sh_closest, slot[2], p, o.sh_o, 0.0f, o.sh_d, FLT_MAX);
        if (o.keep) A.active_out[slot[3]] = p;
    }
}

template <int INTEGRATOR>
__global__ void __launch_bounds__(256) k_shade(SceneView sc, JobView job, ShadeArgs A)
{
    shade_live_paths<INTEGRATOR, false>(sc, job, nullptr, A);
}

// The multi-job form: the integrator is a per-job setting.
__global__ void __launch_bounds__(256) k_shade_multi(const JobDev* __restrict__ jobs, ShadeArgs A)
{
    shade_live_paths<0, true>(SceneView{}, JobView{}, jobs, A);
}

// One ray traced by one thread with the walk of the trace kernel; t and k as the trace
// kernel writes them for the path queues (leaf slot, -1 on a miss).
template <int W>
__device__ __forceinline__ void trace_single(const BvhView& b, V3 o, float dmin, V3 d, float dmax, bool any,
                                             uint2* lds, StackEntry* spill, float* t, int* k)
{
    const float4 r0 = make_float4(o.x, o.y, o.z, dmin), r1 = make_float4(d.x, d.y, d.z, dmax);
    TraceRay r;
    TraceState s;
    if (trace_begin(b, r0, r1, any, r, s)) {
        while (!trace_any_width_step<false, true, W>(b, r, s, lds, spill)) {
        }
    }
    *t = s.best_t;
    *k = s.best;
}

// Denoising AOVs (AOVDenoisingAlbedo / AOVDenoisingNormal::Render and
// CalcPathContribution, Core/AOV/AOVDenoisingAlbedo.cpp:19-133,
// Core/AOV/AOVDenoisingNormal.cpp:19-128), one thread per pixel: spp camera rays, each
// followed through glass and zero-roughness GGX to its first other hit. Albedo: the
// throughput times GetAlbedo (one for an untextured Lambert), times F0 (GGX), or times
// the environment on a miss; normal: the throughput times the shading normal, or times
// -dir on a miss. An Emission hit gives zero (the reference's default branch), and so
// does a path still going after max_bounces hits. The paths are short (most end at the
// first hit), so they are not run through the wavefront queues.
// The UV AOV (AOVUVCoordinate::Render, Core/AOV/AOVUVCoordinate.cpp:11-51) is mode 2: the
// mean over spp camera rays of the first hit's texture coordinates (a miss adds zero).
template <int W>
__global__ void __launch_bounds__(256, W == 8 ? 1 : 4) k_render_aov(SceneView sc, int w, int h, int spp, int max_bounces,
                                                    uint32_t seed, int mode, float* out,
                                                    unsigned long long* queries)
{
    const bool normal_aov = mode == 1;
    __shared__ uint2 lds_stack[lds_stack_entries<W>() * 256];
    uint2* lds = lds_stack + threadIdx.x;
    StackEntry spill[stack_entries<W>() - lds_stack_entries<W>()];
    const int n = w * h;
    unsigned long long q = 0;
    for (int pix = blockIdx.x * blockDim.x + threadIdx.x; pix < n; pix += gridDim.x * blockDim.x) {
        const int px = pix % w, py = pix / w;
        V3 sum = v3(0.0f, 0.0f, 0.0f);
        for (int s = 0; s < spp; ++s) {
            Samplers smp = samplers_for_sample(seed, (uint32_t)pix, (uint32_t)s);
            Ray ray = generate_ray(sc.cam, px, py, w, h, smp);
            V3 contrib = v3(0.0f, 0.0f, 0.0f);
            if (mode == 2) {
                float t;
                int k;
                trace_single<W>(sc.bvh, ray.o, PT_EPS, ray.dir, PT_INFINITY, false, lds, spill, &t, &k);
                ++q;
                if (k >= 0) sum = sum + hit_uv(sc, k, ray.o, ray.dir);
                continue;
            }
            for (int bounce = 0; bounce < max_bounces; ++bounce) {
                float t;
                int k;
                trace_single<W>(sc.bvh, ray.o, PT_EPS, ray.dir, PT_INFINITY, false, lds, spill, &t, &k);
                ++q;
                if (k < 0) {
                    contrib = ray.throughput * (normal_aov ? -ray.dir : env_color(sc.env, ray.dir));
                    break;
                }
                const int m = material_of(sc, k, next1(smp));
                V3 pos, sn;
                shade_point(sc, k, ray.o, ray.dir, t, &pos, &sn);
                const V3 uv = maps_uv(sc, k, ray.o, ray.dir);
                const MaterialDev mat = material_at(sc, m, uv);
                if (mat.type == kMatLambert) {
                    contrib = ray.throughput * (normal_aov ? sn : mat.albedo);
                    break;
                }
                if (mat.type == kMatGlossy && mat.alpha != 0.0f) {
                    contrib = ray.throughput * (normal_aov ? sn : ggx_f0(mat));
                    break;
                }
                if (mat.type != kMatGlass && mat.type != kMatGlossy) break; // Emission: zero
                ray = next_ray(mat, ray, pos, ggx_map_normal(sc, m, mat.type, k, sn, uv), smp); // glass, smooth GGX
            }
            sum = sum + contrib;
        }
        const V3 r = sum / (float)spp;
        out[3 * (size_t)pix] = r.x;
        out[3 * (size_t)pix + 1] = r.y;
        out[3 * (size_t)pix + 2] = r.z;
    }
    if (queries && q) atomicAdd(queries, q);
}

hipError_t launch_render_aov(const SceneView& sc, int w, int h, int spp, int max_bounces, uint32_t seed,
                             int mode, int width, float* out, unsigned long long* queries, hipStream_t st)
{
    const int64_t n = (int64_t)w * h;
    if (n <= 0) return hipSuccess;
    const int64_t want = (n + 255) / 256, cap = (int64_t)cu_count() * PT_FINISH_BLOCKS_PER_CU;
    const dim3 grid((unsigned)(want < cap ? want : cap)), block(256);
    if (width == 8)
        hipLaunchKernelGGL(k_render_aov<8>, grid, block, 0, st, sc, w, h, spp, max_bounces, seed, mode, out, queries);
    else if (width == 4)
        hipLaunchKernelGGL(k_render_aov<4>, grid, block, 0, st, sc, w, h, spp, max_bounces, seed, mode, out, queries);
    else
        hipLaunchKernelGGL(k_render_aov<2>, grid, block, 0, st, sc, w, h, spp, max_bounces, seed, mode, out, queries);
    return hipGetLastError();
}

// Events of path p from its current hits to the end of its sample, then of the job's next
// items while there are any. Returns the rays traced.
template <int INTEGRATOR, int W>
__device__ __forceinline__ unsigned long long finish_path(const SceneView& sc, const JobView& job,
                                                          const ShadeArgs& A, const HitBuffers& H, int p,
                                                          uint2* lds, StackEntry* spill)
{
    unsigned long long q = 0;
    for (;;) {
        ShadeOut o;
        o.keep = o.want_ext = o.want_any = o.want_closest = o.sample_end = false;
        shade_one<INTEGRATOR>(sc, job, A, p, o);
        if (o.sample_end) {
            atomicAdd(job.done, 1ull);
            const unsigned long long item = atomicAdd(job.cursor, 1ull);
            if (item >= job.items) break;
            const Ray r = start_sample(sc, job, A.ps, p, item);
            o.keep = true;
            o.want_ext = true;
            o.ext_o = r.o;
            o.ext_d = r.dir;
            o.ext_dmin = camera_dmin(job.integrator);
            o.ext_dmax = camera_dmax(job.integrator);
        }
        if (!o.keep) break;
        float t;
        int k;
        if (o.want_any) {
            trace_single<W>(sc.bvh, o.sh_o, 0.0f, o.sh_d, FLT_MAX, true, lds, spill, &t, &k);
            H.sh_k[p] = k;
            ++q;
        }
        if (o.want_closest) {
            trace_single<W>(sc.bvh, o.sh_o, 0.0f, o.sh_d, FLT_MAX, false, lds, spill, &t, &k);
            H.sh_t[p] = t;
            H.sh_k[p] = k;
            ++q;
        }
        if (o.want_ext) {
            trace_single<W>(sc.bvh, o.ext_o, o.ext_dmin, o.ext_d, o.ext_dmax, false, lds, spill, &t, &k);
            H.hit_t[p] = t;
            H.hit_k[p] = k;
            ++q;
        }
    }
    return q;
}

// Completion of the live paths of a render, one thread per path: every shading event is
// followed at once by tracing its rays in the same thread, until the job has no item left. The events, their order and every result are those of the wavefront
// iterations; what changes is that a path no longer waits for the slowest ray of each
// iteration, which near the end of a render (few live paths, each iteration as long as
// its longest ray) is where the time goes. Multi-job launches read each path's job.
template <int INTEGRATOR, int W, bool MULTI>
__global__ void __launch_bounds__(256) k_finish_paths(SceneView sc1, JobView job1, const JobDev* jobs, ShadeArgs A,
                                                      HitBuffers H, unsigned long long* queries)
{
    __shared__ uint2 lds_stack[lds_stack_entries<W>() * 256];
    uint2* lds = lds_stack + threadIdx.x;
    StackEntry spill[stack_entries<W>() - lds_stack_entries<W>()];
    const int n = *A.active_count;
    unsigned long long q = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int p = A.active_in[i];
        if (MULTI) {
            const JobDev* J = jobs + A.ps.job[p];
            if (J->job.integrator == 0) q += finish_path<0, W>(J->sc, J->job, A, H, p, lds, spill);
            else q += finish_path<1, W>(J->sc, J->job, A, H, p, lds, spill);
        } else {
            q += finish_path<INTEGRATOR, W>(sc1, job1, A, H, p, lds, spill);
        }
    }
    if (q) atomicAdd(queries, q);
}

__global__ void k_zero_counters(int* c, int n, int stride)
{
    const int i = threadIdx.x;
    if (i < n) c[i * stride] = 0;
}

// One launch for an iteration's two resets: the trace cursors (n at the given stride) and
// the four counts of the queue set that shading writes (the trace only reads the other set).
__global__ void k_zero_iteration(int* cursors, int n, int stride, int* counts)
{
    const int i = threadIdx.x;
    if (i < n) cursors[i * stride] = 0;
    if (i >= 32 && i < 36) counts[i - 32] = 0;
}

__global__ void k_set_int(int* p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

__global__ void k_log_tick(const int* seg_counts, LaunchRec* log, int* index)
{
    if (threadIdx.x != 0) return;
    const int i = *index;
    for (int k = 0; k < 3; ++k) log[i].items[k] = seg_counts[k];
    if (i < PT_LAUNCH_LOG_CAP - 1) *index = i + 1;
}

__global__ void k_copy_u64(const unsigned long long* src, unsigned long long* dst)
{
    if (threadIdx.x == 0) *dst = *src;
}

// ---------------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------------

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

hipError_t launch_zero(int* counters, int n, hipStream_t st, int stride)
{
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, counters, n, stride);
    return hipGetLastError();
}

hipError_t launch_zero_iteration(int* cursors, int n, int stride, int* counts, hipStream_t st)
{
    if (n < 0 || n > 32) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_iteration, dim3(1), dim3(64), 0, st, cursors, n, stride, counts);
    return hipGetLastError();
}

hipError_t launch_set(int* p, int value, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, p, value);
    return hipGetLastError();
}

hipError_t launch_log_tick(const int* seg_counts, LaunchRec* log, int* index, hipStream_t st)
{
    hipLaunchKernelGGL(k_log_tick, dim3(1), dim3(64), 0, st, seg_counts, log, index);
    return hipGetLastError();
}

hipError_t launch_copy_u64(const unsigned long long* src, unsigned long long* dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(64), 0, st, src, dst);
    return hipGetLastError();
}

hipError_t launch_paths_init(const SceneView& sc, const JobView& job, const PathView& ps, const Queue& ext,
                             int* active, int* active_count, int* base, hipStream_t st)
{
    const int64_t n = job.num_slots;
    if (n <= 0) return hipSuccess;
    const int live = (int)((unsigned long long)n < job.items ? (unsigned long long)n : job.items);
    hipLaunchKernelGGL(k_reserve_init, dim3(1), dim3(64), 0, st, active_count, ext.count, live, base);
    hipLaunchKernelGGL(k_paths_init, dim3(grid_for(n, 256)), dim3(256), 0, st, sc, job, ps, ext, active,
                       (const int*)base);
    return hipGetLastError();
}

// Progressive image of a running job: per pixel the mean of its completed samples (a
// sample not done yet holds the all-ones pattern written by k_samples_clear), black before
// the first; adds the completed samples to *samples.
__global__ void __launch_bounds__(256) k_preview(JobView job, float* out, unsigned long long* samples)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = job.num_pixels;
    unsigned long long c = 0;
    if (i < n) {
        float r = 0.0f, g = 0.0f, b = 0.0f;
        int k = 0;
        for (int s = 0; s < job.spp; ++s) {
            const float* v = job.samples + 3 * ((size_t)s * n + i);
            if (__float_as_uint(v[0]) == 0xffffffffu) continue;
            r += v[0];
            g += v[1];
            b += v[2];
            ++k;
        }
        const size_t px = (size_t)job.pixel_base + i;
        out[3 * px] = k ? r / (float)k : 0.0f;
        out[3 * px + 1] = k ? g / (float)k : 0.0f;
        out[3 * px + 2] = k ? b / (float)k : 0.0f;
        c = (unsigned long long)k;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(samples, c);
}

hipError_t launch_preview(const JobView& job, const PathView& ps, float* out, unsigned long long* samples,
                          hipStream_t st)
{
    (void)ps;
    const int64_t n = job.num_pixels;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_preview, dim3(grid_for(n, 256)), dim3(256), 0, st, job, out, samples);
    return hipGetLastError();
}

// PathTracing::Render's average (PathTracing.cpp:194-198): the samples of a pixel summed in
// sample order, then divided by spp.
__global__ void __launch_bounds__(256) k_resolve(JobView job)
{
    const int n = job.num_pixels;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        V3 sum = v3(0.0f, 0.0f, 0.0f);
        for (int s = 0; s < job.spp; ++s) {
            const float* v = job.samples + 3 * ((size_t)s * n + i);
            sum = sum + v3(v[0], v[1], v[2]);
        }
        const V3 avg = sum / (float)job.spp;
        float* o = job.out + 3 * ((size_t)job.pixel_base + i);
        o[0] = avg.x;
        o[1] = avg.y;
        o[2] = avg.z;
    }
}

hipError_t launch_resolve(const JobView& job, hipStream_t st)
{
    const int64_t n = job.num_pixels;
    if (n <= 0) return hipSuccess;
    const int64_t want = (n + 255) / 256, cap = (int64_t)cu_count() * 16;
    hipLaunchKernelGGL(k_resolve, dim3((unsigned)(want < cap ? want : cap)), dim3(256), 0, st, job);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_fill_ones(uint4* p, size_t n16, uint32_t* tail, size_t n_tail)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
    if (blockIdx.x == 0 && threadIdx.x < n_tail) tail[threadIdx.x] = 0xffffffffu;
}

hipError_t launch_samples_clear(const JobView& job, hipStream_t st)
{
    const size_t words = 3 * (size_t)job.items;
    if (words == 0) return hipSuccess;
    // the buffer starts 256-byte aligned (hipMalloc); the last words that do not fill a
    // 16-byte store are written one by one
    const size_t n16 = words / 4, n_tail = words - 4 * n16;
    const int64_t want = (int64_t)((n16 + 255) / 256), cap = (int64_t)cu_count() * 16;
    hipLaunchKernelGGL(k_fill_ones, dim3((unsigned)(want < 1 ? 1 : (want < cap ? want : cap))), dim3(256), 0, st,
                       (uint4*)job.samples, n16, (uint32_t*)job.samples + 4 * n16, n_tail);
    return hipGetLastError();
}

int cu_count()
{
    static int cached = 0;
    if (cached == 0) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached = cus;
    }
    return cached;
}

hipError_t launch_shade(int integrator, const SceneView& sc, const JobView& job, const JobDev* jobs,
                        const ShadeArgs& a, hipStream_t st, const LaunchTuning& tune)
{
    const int bpc = tune.shade_blocks_per_cu < 1 ? 1 : (tune.shade_blocks_per_cu > 32 ? 32 : tune.shade_blocks_per_cu);
    const dim3 grid((unsigned)(cu_count() * bpc)), block(256);
    if (jobs) hipLaunchKernelGGL(k_shade_multi, grid, block, 0, st, jobs, a);
    else if (integrator == 0) hipLaunchKernelGGL((k_shade<0>), grid, block, 0, st, sc, job, a);
    else hipLaunchKernelGGL((k_shade<1>), grid, block, 0, st, sc, job, a);
    return hipGetLastError();
}

hipError_t launch_finish_paths(int integrator, const SceneView& sc, const JobView& job, const JobDev* jobs,
                               int width, const ShadeArgs& a, const HitBuffers& h, unsigned long long* queries,
                               hipStream_t st)
{
    const dim3 grid((unsigned)(cu_count() * PT_FINISH_BLOCKS_PER_CU)), block(256);
#define PT_LAUNCH_FINISH(I, W, M) \
    hipLaunchKernelGGL((k_finish_paths<I, W, M>), grid, block, 0, st, sc, job, jobs, a, h, queries)
#define PT_LAUNCH_FINISH_W(I, M)                  \
    do {                                          \
        if (width == 8) PT_LAUNCH_FINISH(I, 8, M);      \
        else if (width == 4) PT_LAUNCH_FINISH(I, 4, M); \
        else PT_LAUNCH_FINISH(I, 2, M);                 \
    } while (0)
    if (jobs) PT_LAUNCH_FINISH_W(0, true);
    else if (integrator == 0) PT_LAUNCH_FINISH_W(0, false);
    else PT_LAUNCH_FINISH_W(1, false);
#undef PT_LAUNCH_FINISH_W
#undef PT_LAUNCH_FINISH
    return hipGetLastError();
}

} // namespace pt
