// pt_trace.hip — the persistent trace kernel over the ray queues (closest hit and any
// hit; the walk itself is in pt_traverse.h).
//
// Execution: persistent waves. A wave takes rays in chunks from the cursor of its XCD's
// part of the queues (then from the other parts) and refills its lanes as rays finish,
// so lanes do not wait for the longest ray of the wave. The LDS size per block (the
// traversal stacks) sets the blocks per CU.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "pt_device.h"
#include "pt_kernels.h"
#include "pt_traverse.h"

namespace pt {

// Part x of a segment of n rays: [min(x q, n), min(x q + q, n)) with q = ceil(n / parts).
__device__ __forceinline__ void part_range(int n, int x, int* lo, int* hi)
{
    const int q = (n + PT_TRACE_PARTS - 1) / PT_TRACE_PARTS;
    *lo = min(x * q, n);
    *hi = min(*lo + q, n);
}

// Rays in part x over all segments.
__device__ __forceinline__ int part_total(const int* seg_n, int num_seg, int x)
{
    int t = 0;
    for (int k = 0; k < num_seg; ++k) {
        int lo, hi;
        part_range(seg_n[k], x, &lo, &hi);
        t += hi - lo;
    }
    return t;
}

// The g-th ray of part x (g < part_total): its segment and index in the segment.
__device__ __forceinline__ void part_item(const int* seg_n, int num_seg, int x, int g, int* seg_id, int* seg_i)
{
    for (int k = 0; k < num_seg; ++k) {
        int lo, hi;
        part_range(seg_n[k], x, &lo, &hi);
        if (g < hi - lo || k == num_seg - 1) {
            *seg_id = k;
            *seg_i = lo + g;
            return;
        }
        g -= hi - lo;
    }
}

// TPR: lanes per ray, 1, or 2 for the binary walk of plain trees (trace_step_pair): the two
// lanes of a pair hold the same ray and take it from the pool together.
template <bool STATS, bool MULTI, int W, bool EXT, int TPR, bool PH>
__device__ __forceinline__ void trace_body(BvhView b1, const TraceJob& job, int* cursor, unsigned long long* stats)
{
    static_assert(TPR == 1 || (TPR == 2 && W == 2 && !EXT), "two lanes per ray: binary walk of plain trees");
    static_assert(!PH || (TPR == 1 && W == 2 && !EXT), "phased steps: binary walk of plain trees");
    // segment sizes come from device counters (written by the previous stage)
    int seg_n[3] = {0, 0, 0};
    int total = 0;
    for (int k = 0; k < job.num_seg; ++k) {
        seg_n[k] = *job.seg[k].count;
        total += seg_n[k];
    }
    if (job.queries && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(job.queries, (unsigned long long)total);
    if (total == 0) return;
    // launch log: the index stays put during the launch (the tick after it advances it)
    LaunchRec* rec = job.launch_log ? job.launch_log + *job.launch_index : nullptr;
    if (rec && threadIdx.x == 0) atomicMin(&rec->t0, (unsigned long long)wall_clock64());
    __shared__ uint2 lds_stack[lds_stack_entries<W>() * PT_TRACE_BLOCK];
    uint2* lds = lds_stack + threadIdx.x;
    StackEntry spill[stack_entries<W>() - lds_stack_entries<W>()];
    const int lane = threadIdx.x & 63;
    // rays are ranked by their first lane: every lane (one lane per ray), or the even lanes
    const unsigned long long ray_lanes = TPR == 2 ? 0x5555555555555555ull : ~0ull;
    const unsigned long long lane_lt = ((1ull << (lane & ~(TPR - 1))) - 1ull) & ray_lanes;
    const int sub = TPR == 2 ? (lane & 1) : 0;
    // wave-uniform: this wave's share of rays [pool_next, pool_end) in part pool_part, and
    // how many parts it has found empty
    int pool_next = 0, pool_end = 0, pool_part = (int)(blockIdx.x % PT_TRACE_PARTS), parts_empty = 0;
    bool exhausted = false; // wave-uniform
    // Parts a wave helps with after its own. Every part has waves of its own (the grid
    // has at least PT_TRACE_PARTS blocks), which drain it before they stop, so helping is
    // only for balance: one more part in launches of up to PT_TRACE_STEAL_ALL_RAYS rays,
    // where each probe of an empty part is an atomic on a cursor that every wave probes
    // at the launch's end; all parts in larger launches.
    const int steal_parts =
        gridDim.x >= PT_TRACE_PARTS && total <= PT_TRACE_STEAL_ALL_RAYS ? PT_TRACE_STEAL : PT_TRACE_PARTS - 1;
    int item = -1, seg_id = 0, seg_i = 0;
    TraceRay r;
    TraceState s;
    BvhView b = b1; // per lane in multi-job launches
    uint64_t nv = 0, nt = 0;
    for (;;) {
        // refill lanes without a ray, first from the wave's pool, then from a new chunk,
        // once a quarter of the wave is idle (the refill's bookkeeping then runs every few
        // steps instead of whenever one ray ends)
        const bool need = item < 0;
        const unsigned long long need_mask = __ballot(need) & ray_lanes;
        if (__popcll(need_mask) >= 16 / TPR && !exhausted) {
            const int n_need = __popcll(need_mask);
            const int rank = __popcll(need_mask & lane_lt);
            const int leader = __ffsll((long long)need_mask) - 1;
            const int avail = pool_end - pool_next;
            int part2 = pool_part, base2 = 0, end2 = 0;
            if (avail < n_need) {
                // a new chunk: from the current part, else from the next parts in turn
                for (;;) {
                    const int len = part_total(seg_n, job.num_seg, part2);
                    int got = len;
                    if (lane == leader) got = atomicAdd(cursor + part2 * PT_CURSOR_STRIDE, job.chunk);
                    got = __shfl(got, leader, 64);
                    if (got < len) {
                        base2 = got;
                        end2 = min(got + job.chunk, len);
                        break;
                    }
                    if (++parts_empty > steal_parts) {
                        exhausted = true;
                        break;
                    }
                    part2 = (part2 + 1) % PT_TRACE_PARTS;
                }
            }
            if (need) {
                int g = -1, gp = pool_part;
                if (rank < avail) {
                    g = pool_next + rank;
                } else if (base2 + (rank - avail) < end2) {
                    g = base2 + (rank - avail);
                    gp = part2;
                }
                item = g;
                if (g >= 0) {
                    part_item(seg_n, job.num_seg, gp, g, &seg_id, &seg_i);
                    const TraceSegment& sg = job.seg[seg_id];
                    if (MULTI) b = job.views[job.job_of_path[sg.path[seg_i]]];
                    const size_t ri = sg.rays_by_path ? (size_t)sg.path[seg_i] : (size_t)seg_i;
                    const float4 r0 = sg.rays[2 * ri], r1 = sg.rays[2 * ri + 1];
                    if (!trace_begin(b, r0, r1, sg.any != 0, r, s)) {
                        if (STATS && sub == 0) nv += s.nv;
                        if (sub == 0) trace_finish(b, sg, seg_i, s);
                        item = -1;
                    }
                }
            }
            if (avail >= n_need) {
                pool_next += n_need;
            } else {
                pool_part = part2;
                pool_next = min(base2 + (n_need - avail), end2);
                pool_end = end2;
            }
        }
        const unsigned long long busy = __ballot(item >= 0);
        if (busy == 0ull) {
            if (exhausted) break; // no ray left anywhere for this wave
            continue;
        }
        // phased walk: a triangle step once enough of the wave's rays are in a leaf
        bool tri_phase = false;
        if constexpr (PH) {
            const unsigned long long in_leaf = __ballot(item >= 0 && s.leaf_k < s.leaf_end);
            tri_phase = (__popcll(in_leaf) << job.leaf_shift) >= __popcll(busy);
        }
        if (item >= 0) {
            bool done;
            if constexpr (TPR == 2) done = trace_step_pair<STATS>(b, r, s, lds, spill, sub);
            else if constexpr (PH) done = trace_step_phase<STATS>(b, r, s, lds, spill, tri_phase);
            else done = trace_any_width_step<STATS, EXT, W>(b, r, s, lds, spill);
            if (done) {
                if (STATS && sub == 0) {
                    nv += s.nv;
                    nt += s.nt;
                }
                if (sub == 0) trace_finish(b, job.seg[seg_id], seg_i, s);
                item = -1;
            }
        }
    }
    if (STATS) {
        atomicAdd(rec ? &rec->nodes : &stats[0], (unsigned long long)nv);
        atomicAdd(rec ? &rec->tris : &stats[1], (unsigned long long)nt);
    }
    if (rec) {
        __syncthreads(); // the block's last wave is done
        if (threadIdx.x == 0) atomicMax(&rec->t1, (unsigned long long)wall_clock64());
    }
}

// The multi-job binary walk is held to 64 VGPRs, so that its 8 LDS-bound blocks per CU
// are all resident (8 waves per SIMD; unbounded it takes 67 and fits 7, 2.8 % slower on
// the 64-job batch). The single-job binary walk needs 61 by itself, and the 4-wide walk
// is held to the same 8 blocks; the 8-wide walk is LDS-bound at 5 blocks per CU.
template <bool STATS, bool MULTI, int W, bool EXT, int TPR, bool PH>
__global__ void __launch_bounds__(PT_TRACE_BLOCK, W == 8 ? 1 : (W == 4 ? 6 : 8)) k_trace_persistent(BvhView b1, TraceJob job,
                                                                                    int* cursor,
                                                                                    unsigned long long* stats)
{
    trace_body<STATS, MULTI, W, EXT, TPR, PH>(b1, job, cursor, stats);
}

// Ray records (o, dist_min), (d, dist_max) from structure-of-arrays inputs; a null
// dist_min / dist_max array reads as 0 / FLT_MAX (AccelBase::Intersect(ray, scene),
// Core/Accel/AccelBase.h:44-48).
__global__ void k_pack_rays_soa(int64_t n, const float* __restrict__ ox, const float* __restrict__ oy,
                                const float* __restrict__ oz, const float* __restrict__ dx,
                                const float* __restrict__ dy, const float* __restrict__ dz,
                                const float* __restrict__ dmin, const float* __restrict__ dmax, float4* out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        out[2 * i] = make_float4(ox[i], oy[i], oz[i], dmin ? dmin[i] : 0.0f);
        out[2 * i + 1] = make_float4(dx[i], dy[i], dz[i], dmax ? dmax[i] : FLT_MAX);
    }
}

// Structure-of-arrays hits from a trace that wrote t and the leaf slot (-1 on a miss) per
// ray: the slot becomes the primitive index, a miss gets t = +inf, and u, v (when asked
// for) are the hit triangle's barycentric coordinates as Triangle::Intersect forms them
// (Core/Geometry/Triangle.cpp:57-72; 0 for spheres and misses).
__global__ void k_hits_soa(int64_t n, BvhView b, const float4* __restrict__ rays, float* t, int32_t* prim, float* u,
                           float* v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int k = prim[i];
        float bu = 0.0f, bv = 0.0f;
        if (k < 0) {
            t[i] = __builtin_inff();
            prim[i] = -1;
        } else {
            prim[i] = b.prim_ids[k];
            if (u || v) {
                if (b.kinds == nullptr || b.kinds[k] == 0) {
                    const float4 t0 = b.tris[3 * (size_t)k], t1 = b.tris[3 * (size_t)k + 1],
                                 t2 = b.tris[3 * (size_t)k + 2];
                    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
                    const V3 o = v3(r0.x, r0.y, r0.z), d = v3(r1.x, r1.y, r1.z);
                    const V3 p0 = v3(t0.x, t0.y, t0.z), e1 = v3(t0.w, t1.x, t1.y);
                    const V3 e2 = v3(t1.z, t1.w, t2.x), nn = v3(t2.y, t2.z, t2.w);
                    const float inv_det = 1.0f / dot(-d, nn);
                    const V3 vec = o - p0;
                    bu = dot(vec, cross(d, e2)) * inv_det;
                    bv = dot(vec, cross(e1, d)) * inv_det;
                }
            }
        }
        if (u) u[i] = bu;
        if (v) v[i] = bv;
    }
}

hipError_t launch_pack_rays_soa(int64_t n, const float* ox, const float* oy, const float* oz, const float* dx,
                                const float* dy, const float* dz, const float* dmin, const float* dmax, float4* out,
                                hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t want = (n + 255) / 256, cap = (int64_t)cu_count() * 16;
    hipLaunchKernelGGL(k_pack_rays_soa, dim3((unsigned)(want < cap ? want : cap)), dim3(256), 0, st, n, ox, oy, oz,
                       dx, dy, dz, dmin, dmax, out);
    return hipGetLastError();
}

hipError_t launch_hits_soa(int64_t n, const BvhView& b, const float4* rays, float* t, int32_t* prim, float* u, float* v,
                           hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t want = (n + 255) / 256, cap = (int64_t)cu_count() * 16;
    hipLaunchKernelGGL(k_hits_soa, dim3((unsigned)(want < cap ? want : cap)), dim3(256), 0, st, n, b, rays, t, prim,
                       u, v);
    return hipGetLastError();
}

hipError_t launch_trace_job(const BvhView& b, const TraceJob& job_in, int* cursor, unsigned long long* stats,
                            hipStream_t st, const LaunchTuning& tune)
{
    if (job_in.num_seg <= 0) return hipSuccess;
    TraceJob job = job_in;
    // a wave refills at most 64 lanes from one chunk, so chunks hold at least 64 rays
    job.chunk = tune.trace_chunk < 64 ? 64 : tune.trace_chunk;
    const bool multi = job.views != nullptr;
    const int width = multi ? job.width : (b.nodes8 != nullptr ? b.width : 2);
    // persistent grid: blocks of 4 waves; the LDS stacks of a block bound the blocks per CU
    // (160 KB of LDS, at most 8 waves per SIMD)
    const int lds_block =
        (width == 8 ? lds_stack_entries<8>() : (width == 4 ? lds_stack_entries<4>() : lds_stack_entries<2>())) *
        PT_TRACE_BLOCK * (int)sizeof(uint2);
    const int max_bpc = (160 * 1024) / lds_block < 8 ? (160 * 1024) / lds_block : 8;
    const int bpc = tune.trace_blocks_per_cu < 1 || tune.trace_blocks_per_cu > max_bpc ? max_bpc
                                                                                     : tune.trace_blocks_per_cu;
    const dim3 grid((unsigned)(cu_count() * bpc)), block(PT_TRACE_BLOCK);
    // the extended walk (deep trees, sphere leaves) only where a BVH needs it
    const bool ext = multi ? job.ext != 0 : (b.parents != nullptr || b.kinds != nullptr);
    const bool counted = stats || (job.launch_log && job.log_counts);
    // two lanes per ray on request, for the binary walk of plain trees
    const bool pair = tune.threads_per_ray == 2 && width == 2 && !ext;
    // phased steps on request, for the one-lane binary walk of plain trees
    const bool phased = !pair && tune.leaf_batch >= 1 && width == 2 && !ext;
    job.leaf_shift = phased ? (tune.leaf_batch > 4 ? 3 : tune.leaf_batch - 1) : 0;
#define PT_LAUNCH_TRACE(S, M, W, E) \
    hipLaunchKernelGGL((k_trace_persistent<S, M, W, E, 1, false>), grid, block, 0, st, b, job, cursor, stats)
#define PT_LAUNCH_TRACE_E(S, M, W) \
    do {                          \
        if (ext) PT_LAUNCH_TRACE(S, M, W, true); \
        else PT_LAUNCH_TRACE(S, M, W, false);    \
    } while (0)
#define PT_LAUNCH_TRACE_W(S, M) \
    do {                       \
        if (width == 8) PT_LAUNCH_TRACE_E(S, M, 8);      \
        else if (width == 4) PT_LAUNCH_TRACE_E(S, M, 4); \
        else PT_LAUNCH_TRACE_E(S, M, 2);                 \
    } while (0)
#define PT_LAUNCH_PAIR(S, M) \
    hipLaunchKernelGGL((k_trace_persistent<S, M, 2, false, 2, false>), grid, block, 0, st, b, job, cursor, stats)
#define PT_LAUNCH_PHASED(S, M) \
    hipLaunchKernelGGL((k_trace_persistent<S, M, 2, false, 1, true>), grid, block, 0, st, b, job, cursor, stats)
    if (pair) {
        if (counted) {
            if (multi) PT_LAUNCH_PAIR(true, true);
            else PT_LAUNCH_PAIR(true, false);
        } else {
            if (multi) PT_LAUNCH_PAIR(false, true);
            else PT_LAUNCH_PAIR(false, false);
        }
    } else if (phased) {
        if (counted) {
            if (multi) PT_LAUNCH_PHASED(true, true);
            else PT_LAUNCH_PHASED(true, false);
        } else {
            if (multi) PT_LAUNCH_PHASED(false, true);
            else PT_LAUNCH_PHASED(false, false);
        }
    } else if (counted) {
        if (multi) PT_LAUNCH_TRACE_W(true, true);
        else PT_LAUNCH_TRACE_W(true, false);
    } else {
        if (multi) PT_LAUNCH_TRACE_W(false, true);
        else PT_LAUNCH_TRACE_W(false, false);
    }
#undef PT_LAUNCH_PHASED
#undef PT_LAUNCH_PAIR
#undef PT_LAUNCH_TRACE_W
#undef PT_LAUNCH_TRACE_E
#undef PT_LAUNCH_TRACE
    return hipGetLastError();
}

} // namespace pt
