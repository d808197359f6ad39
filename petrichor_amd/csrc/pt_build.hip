// pt_build.hip — binned-SAH BVH construction on gfx950, in the reference's topology
// (BinnedSAHBVH::Build, Core/Accel/BinnedSAHBVH.cpp:15-225), and packing into the
// traversal layout.
//
// Level-synchronous: every level processes all splittable nodes at once.
//  1. centroid bounds per node (AABB::Merge of centroids, BinnedSAHBVH.cpp:92-100),
//  2. widest axis (AABB::GetWidestAxis, AABB.cpp:35-52) and the bin of every primitive,
//     int(nextafter(bins * l / len, 0)) (BinnedSAHBVH.cpp:110-125),
//  3. per-bin boxes and counts, the SAH sweep nL*SA(L) + nR*SA(R) with the first minimum
//     kept (GetSAHCost, BinnedSAHBVH.cpp:128-151,227-270), leaf on k = 0,
//  4. partition of the node's primitives into the two children.
// Box merges are min/max, exact and order-free, so they are done with atomics on
// order-preserving integer keys; the node set, boxes and ranges equal the reference's.
// Only the order inside a leaf depends on the reference's std::sort; every leaf is
// sorted here at the end by the key of the last sort the reference applied to it
// (parent's widest axis, or its own for an SAH leaf), ties by triangle index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <vector>

#include "pt_build.h"
#include "pt_device.h"
#include "pt_kernels.h"

namespace pt {

namespace {

// Order-preserving float <-> unsigned map: a < b as floats iff key(a) < key(b).
__device__ __forceinline__ uint32_t fkey(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k)
{
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}
constexpr uint32_t kKeyMax = 0xffffffffu; // +inf and above: "empty" lower bound
constexpr uint32_t kKeyMin = 0u;          // "empty" upper bound

// std::nextafter(x, 0.0f) for the finite, non-negative bin positions (and NaN).
__device__ __forceinline__ float next_toward_zero(float x)
{
    if (!(x == x) || x == 0.0f) return x;
    return __uint_as_float(__float_as_uint(x) - 1u);
}


// Per-primitive centroid and box, and the root box. A wave handles PT_PRIMS_SPAN groups of
// 64 triangles and adds its running bounds to the root keys once (with one atomic per
// wave and group, every wave of the grid would queue on the same six keys).
#define PT_PRIMS_SPAN 16
__global__ void k_prims(const float* __restrict__ tri_pos, int ntri, const float4* __restrict__ spheres, int n,
                        float4* cen, float4* blo, float4* bhi, uint32_t* root_keys)
{
    const int lane = (int)(threadIdx.x & 63);
    const int first = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64 * PT_PRIMS_SPAN;
    uint32_t k[6] = {kKeyMax, kKeyMax, kKeyMax, kKeyMin, kKeyMin, kKeyMin};
    for (int it = 0; it < PT_PRIMS_SPAN; ++it) {
        const int i = first + 64 * it + lane;
        if (i >= n) break;
        V3 lo, hi, c;
        if (i < ntri) {
            const float* p = tri_pos + 9 * (size_t)i;
            const V3 p0 = v3(p[0], p[1], p[2]), p1 = v3(p[3], p[4], p[5]), p2 = v3(p[6], p[7], p[8]);
            lo = v3(smin(smin(FLT_MAX, p0.x), p1.x), smin(smin(FLT_MAX, p0.y), p1.y), smin(smin(FLT_MAX, p0.z), p1.z));
            lo = v3(smin(lo.x, p2.x), smin(lo.y, p2.y), smin(lo.z, p2.z));
            hi = v3(smax(smax(-FLT_MAX, p0.x), p1.x), smax(smax(-FLT_MAX, p0.y), p1.y), smax(smax(-FLT_MAX, p0.z), p1.z));
            hi = v3(smax(hi.x, p2.x), smax(hi.y, p2.y), smax(hi.z, p2.z));
            c = (p0 + p1 + p2) / 3.0f; // Triangle::GetCentroid (Triangle.h:45-50)
        } else {
            // Sphere::CalcBoundary (Sphere.cpp:24-30) and GetCentroid (Sphere.h): origin
            const float4 q = spheres[i - ntri];
            c = v3(q.x, q.y, q.z);
            const V3 rr = v3(1.0f, 1.0f, 1.0f) * q.w;
            lo = c - rr;
            hi = c + rr;
        }
        cen[i] = make_float4(c.x, c.y, c.z, 0.0f);
        blo[i] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        bhi[i] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        k[0] = min(k[0], fkey(lo.x));
        k[1] = min(k[1], fkey(lo.y));
        k[2] = min(k[2], fkey(lo.z));
        k[3] = max(k[3], fkey(hi.x));
        k[4] = max(k[4], fkey(hi.y));
        k[5] = max(k[5], fkey(hi.z));
    }
    // root box: wave reduction, then one atomic per wave and bound
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 6; ++a) {
            const uint32_t o = __shfl_xor(k[a], off, 64);
            k[a] = a < 3 ? min(k[a], o) : max(k[a], o);
        }
    if (lane == 0 && first < n) {
        for (int a = 0; a < 3; ++a) atomicMin(&root_keys[a], k[a]);
        for (int a = 3; a < 6; ++a) atomicMax(&root_keys[a], k[a]);
    }
}

struct LevelArgs {
    const float4* cen;
    const float4* blo;
    const float4* bhi;
    const int* ids;
    const int* pos_node;  // active index of the node owning each position, -1 = in a leaf
    int n;
    int num_bins;
    int bin_nodes;        // nodes of k_bin's many-node LDS table
    // per active node
    const int* act;       // global node id
    uint32_t* cbk;        // centroid bounds keys, 6 per node
    float* len;
    int* axis;
    int* bad;             // 1 when every bin position is NaN (len == 0)
    uint32_t* bin_keys;   // 6 per (node, bin)
    int* bin_cnt;         // per (node, bin)
    int* pos_bin;         // bin of each position
};

// Wave-uniform test: do all live lanes of the wave see the same value?
__device__ __forceinline__ bool wave_uniform(int v, bool live)
{
    const int first = __shfl(v, __ffsll((long long)__ballot(live)) - 1, 64);
    return __ballot(live && v != first) == 0ull;
}

// Centroid bounds per node. A wave walks PT_CB_SPAN consecutive groups of 64 positions
// (a node's positions are contiguous) and keeps a running min/max while its lanes share
// one node, so it adds to a node's keys once per node met instead of once per group;
// groups that straddle nodes add once per node run.
#define PT_CB_SPAN 16
__device__ __forceinline__ void cb_flush(uint32_t* cbk, int node, uint32_t key[6])
{
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 6; ++q) {
            const uint32_t o = __shfl_xor(key[q], off, 64);
            key[q] = q < 3 ? min(key[q], o) : max(key[q], o);
        }
    if ((threadIdx.x & 63) == 0) {
        for (int q = 0; q < 3; ++q) atomicMin(&cbk[6 * node + q], key[q]);
        for (int q = 3; q < 6; ++q) atomicMax(&cbk[6 * node + q], key[q]);
    }
}

__global__ void k_centroid_bounds(LevelArgs a)
{
    const int lane = (int)(threadIdx.x & 63);
    const int first = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64 * PT_CB_SPAN;
    uint32_t acc[6] = {kKeyMax, kKeyMax, kKeyMax, kKeyMin, kKeyMin, kKeyMin};
    int acc_node = -1; // wave-uniform: the node acc belongs to
    for (int it = 0; it < PT_CB_SPAN; ++it) {
        const int k = first + 64 * it + lane;
        if (first + 64 * it >= a.n) break;
        const int pn = k < a.n ? a.pos_node[k] : -1;
        const bool live = pn >= 0;
        if (__ballot(live) == 0ull) continue;
        float4 c = make_float4(0, 0, 0, 0);
        if (live) c = a.cen[a.ids[k]];
        const uint32_t key[6] = {live ? fkey(c.x) : kKeyMax, live ? fkey(c.y) : kKeyMax,
                                 live ? fkey(c.z) : kKeyMax, live ? fkey(c.x) : kKeyMin,
                                 live ? fkey(c.y) : kKeyMin, live ? fkey(c.z) : kKeyMin};
        if (wave_uniform(pn, live)) {
            const int node = __shfl(pn, __ffsll((long long)__ballot(live)) - 1, 64);
            if (node != acc_node) {
                if (acc_node >= 0) cb_flush(a.cbk, acc_node, acc);
                for (int q = 0; q < 3; ++q) acc[q] = kKeyMax;
                for (int q = 3; q < 6; ++q) acc[q] = kKeyMin;
                acc_node = node;
            }
            for (int q = 0; q < 3; ++q) acc[q] = min(acc[q], key[q]);
            for (int q = 3; q < 6; ++q) acc[q] = max(acc[q], key[q]);
        } else {
            // The group straddles nodes. A node's positions are contiguous, so its lanes form
            // one run: a segmented scan leaves each run's bounds in its last lane, which adds
            // them once.
            uint32_t kk[6];
            for (int q = 0; q < 6; ++q) kk[q] = key[q];
            for (int off = 1; off < 64; off <<= 1) {
                const int opn = __shfl_up(pn, off, 64);
                for (int q = 0; q < 6; ++q) {
                    const uint32_t o = __shfl_up(kk[q], off, 64);
                    if (lane >= off && opn == pn) kk[q] = q < 3 ? min(kk[q], o) : max(kk[q], o);
                }
            }
            const int npn = __shfl_down(pn, 1, 64);
            if (live && (lane == 63 || npn != pn)) {
                for (int q = 0; q < 3; ++q) atomicMin(&a.cbk[6 * pn + q], kk[q]);
                for (int q = 3; q < 6; ++q) atomicMax(&a.cbk[6 * pn + q], kk[q]);
            }
        }
    }
    if (acc_node >= 0) cb_flush(a.cbk, acc_node, acc);
}

__global__ void k_axis(LevelArgs a, int num_active)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_active) return;
    const uint32_t* k = a.cbk + 6 * i;
    const V3 lo = v3(funkey(k[0]), funkey(k[1]), funkey(k[2]));
    const V3 hi = v3(funkey(k[3]), funkey(k[4]), funkey(k[5]));
    // AABB::GetWidestAxis: strict '>' from 0, lowest axis on ties
    float w = 0.0f;
    int ax = 0;
    for (int q = 0; q < 3; ++q) {
        const float x = comp(hi, q) - comp(lo, q);
        if (x > w) {
            w = x;
            ax = q;
        }
    }
    a.axis[i] = ax;
    a.len[i] = comp(hi, ax) - comp(lo, ax);
    a.bad[i] = 0;
}

// Bin of every primitive and the per-bin boxes and counts. A block walks PT_BIN_SPAN
// consecutive groups of 256 positions. While a group's positions all belong to one node
// (the common case near the root) they are reduced in LDS, across groups of the same
// node, and the block adds one atomic per bin and value when the node changes or at the
// end. A group over at most bin_table_nodes nodes (deeper levels; a node's positions are
// contiguous and the active nodes are numbered in position order) is reduced in an LDS
// table of those nodes' bins; other groups add per position.
#define PT_BIN_SPAN 8
// Nodes of the many-node table for nb bins: up to 32, within 16 KB of LDS.
inline int bin_table_nodes(int nb) { return std::max(1, std::min(32, 16384 / (28 * nb))); }
// LDS words of k_bin for nb bins: the one-node table, the many-node table and the group's
// lowest and highest node
inline size_t bin_lds_words(int nb) { return (size_t)7 * nb * (1 + bin_table_nodes(nb)) + 2; }

// Adds the bin tables of nodes [node0, node0 + count) held in lds (7 * nb words per node:
// 6 keys per bin, then a count per bin) to the global bins.
__device__ __forceinline__ void bin_flush(const LevelArgs& a, const uint32_t* lds, int node0, int count)
{
    const int nb = a.num_bins;
    const int per = 7 * nb;
    for (int i = threadIdx.x; i < count * per; i += blockDim.x) {
        const int off = i / per, j = i - off * per;
        const uint32_t* t = lds + off * per;
        const int bin = j < 6 * nb ? j / 6 : j - 6 * nb;
        if (t[6 * nb + bin] == 0u) continue;
        const size_t s = (size_t)(node0 + off) * nb + bin;
        if (j < 6 * nb) {
            const int q = j % 6;
            if (q < 3) atomicMin(&a.bin_keys[6 * s + q], t[j]);
            else atomicMax(&a.bin_keys[6 * s + q], t[j]);
        } else {
            atomicAdd(&a.bin_cnt[s], (int)t[j]);
        }
    }
}

__device__ __forceinline__ void bin_clear(uint32_t* t, int words, int nb)
{
    for (int i = threadIdx.x; i < words; i += blockDim.x) {
        const int j = i % (7 * nb);
        t[i] = j < 6 * nb ? ((j % 6) < 3 ? kKeyMax : kKeyMin) : 0u;
    }
}

__device__ __forceinline__ void bin_add(uint32_t* t, int nb, int b, float4 blo, float4 bhi)
{
    uint32_t* bk = t + 6 * b;
    atomicMin(&bk[0], fkey(blo.x));
    atomicMin(&bk[1], fkey(blo.y));
    atomicMin(&bk[2], fkey(blo.z));
    atomicMax(&bk[3], fkey(bhi.x));
    atomicMax(&bk[4], fkey(bhi.y));
    atomicMax(&bk[5], fkey(bhi.z));
    atomicAdd(&t[6 * nb + b], 1u);
}

__global__ __launch_bounds__(256) void k_bin(LevelArgs a)
{
    extern __shared__ uint32_t lds[];
    const int nb = a.num_bins;
    uint32_t* one = lds;                                   // the one-node table
    uint32_t* many = lds + 7 * nb;                          // the many-node table
    int* grp = (int*)(lds + 7 * nb * (1 + a.bin_nodes)); // lowest, highest node of a group
    const int first = (int)blockIdx.x * (int)blockDim.x * PT_BIN_SPAN;
    int one_node = -1; // block-uniform: the node whose bins the one-node table holds
    for (int it = 0; it < PT_BIN_SPAN; ++it) {
        const int base = first + it * (int)blockDim.x;
        if (base >= a.n) break;
        const int k = base + threadIdx.x;
        const int pn = k < a.n ? a.pos_node[k] : -1;
        const int pn0 = a.pos_node[base];
        const bool uni = __syncthreads_and(pn < 0 || pn == pn0) && pn0 >= 0;
        int b = -1;
        float4 blo = make_float4(0, 0, 0, 0), bhi = blo;
        if (pn >= 0) {
            const int id = a.ids[k];
            const float4 c = a.cen[id];
            const int ax = a.axis[pn];
            const float len = a.len[pn];
            const float lo = funkey(a.cbk[6 * pn + ax]);
            const float cc = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
            const float l = cc - lo;
            const float x = next_toward_zero((float)nb * l / len);
            if (!(x == x)) {
                a.bad[pn] = 1;
                a.pos_bin[k] = 0;
            } else {
                // bins * l / len can round up to bins when bins is not a power of two:
                // the last bin takes it, as in the host builder
                b = min(max((int)x, 0), nb - 1);
                a.pos_bin[k] = b;
                blo = a.blo[id];
                bhi = a.bhi[id];
            }
        }
        if (uni) {
            if (pn0 != one_node) {
                __syncthreads();
                if (one_node >= 0) bin_flush(a, one, one_node, 1);
                __syncthreads();
                bin_clear(one, 7 * nb, nb);
                __syncthreads();
                one_node = pn0;
            }
            if (b >= 0) bin_add(one, nb, b, blo, bhi);
            continue;
        }
        // the group's node range
        if (threadIdx.x == 0) {
            grp[0] = 0x7fffffff;
            grp[1] = -1;
        }
        __syncthreads();
        if (pn >= 0) {
            atomicMin(&grp[0], pn);
            atomicMax(&grp[1], pn);
        }
        __syncthreads();
        const int lo_node = grp[0], count = grp[1] - grp[0] + 1;
        if (count >= 1 && count <= a.bin_nodes) {
            bin_clear(many, count * 7 * nb, nb);
            __syncthreads();
            if (b >= 0) bin_add(many + (pn - lo_node) * 7 * nb, nb, b, blo, bhi);
            __syncthreads();
            bin_flush(a, many, lo_node, count);
        } else if (b >= 0) {
            const size_t s = (size_t)pn * nb + b;
            uint32_t* bk = a.bin_keys + 6 * s;
            atomicMin(&bk[0], fkey(blo.x));
            atomicMin(&bk[1], fkey(blo.y));
            atomicMin(&bk[2], fkey(blo.z));
            atomicMax(&bk[3], fkey(bhi.x));
            atomicMax(&bk[4], fkey(bhi.y));
            atomicMax(&bk[5], fkey(bhi.z));
            atomicAdd(&a.bin_cnt[s], 1);
        }
        __syncthreads(); // grp and the many-node table are reused by the next group
    }
    __syncthreads();
    if (one_node >= 0) bin_flush(a, one, one_node, 1);
}

__device__ __forceinline__ float key_area(const uint32_t* lo, const uint32_t* hi)
{
    const V3 d = v3(funkey(hi[0]), funkey(hi[1]), funkey(hi[2])) - v3(funkey(lo[0]), funkey(lo[1]), funkey(lo[2]));
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x); // AABB::GetSurfaceArea (AABB.h:76-81)
}

struct SahArgs {
    int num_active;
    int num_bins; // <= 64: one lane per bin
    int leaf_max;
    const int* act;
    const uint32_t* bin_keys;
    const int* bin_cnt;
    const int* bad;
    const int* axis;
    BuildNode* nodes;
    int* best_k;
    int* n_left;
    uint32_t* child_keys;         // 12 per active node: left lo/hi, right lo/hi
    unsigned long long* scan_val; // (splits << 32) | non-leaf children, per active node
};

// SAH sweep (GetSAHCost, BinnedSAHBVH.cpp:227-270), one wave per node, lane q = bin q.
// Left of split k holds bins [0, k) (exclusive prefix), right holds [k, bins) (inclusive
// suffix). Empty bins carry the identity keys, as the reference's empty AABBs do. The
// sequential loop keeps the first strict minimum below FLT_MAX: here the lowest cost wins,
// ties go to the lowest k, NaN and >= FLT_MAX costs never win, and k = 0 when none does.
__global__ __launch_bounds__(256) void k_sah(SahArgs s)
{
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= s.num_active) return;
    const int nb = s.num_bins;
    const bool inb = lane < nb;
    const size_t bi = (size_t)i * nb + lane;
    uint32_t lo[3], hi[3];
    int cnt = 0;
    for (int q = 0; q < 3; ++q) {
        lo[q] = inb ? s.bin_keys[6 * bi + q] : kKeyMax;
        hi[q] = inb ? s.bin_keys[6 * bi + 3 + q] : kKeyMin;
    }
    if (inb) cnt = s.bin_cnt[bi];
    // inclusive prefix (left) and suffix (right) scans
    uint32_t plo[3] = {lo[0], lo[1], lo[2]}, phi[3] = {hi[0], hi[1], hi[2]};
    uint32_t slo[3] = {lo[0], lo[1], lo[2]}, shi[3] = {hi[0], hi[1], hi[2]};
    int pc = cnt, sc = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        const bool up = lane >= off, dn = lane + off < 64;
        for (int q = 0; q < 3; ++q) {
            const uint32_t a0 = __shfl_up(plo[q], off, 64), a1 = __shfl_up(phi[q], off, 64);
            const uint32_t b0 = __shfl_down(slo[q], off, 64), b1 = __shfl_down(shi[q], off, 64);
            if (up) {
                plo[q] = min(plo[q], a0);
                phi[q] = max(phi[q], a1);
            }
            if (dn) {
                slo[q] = min(slo[q], b0);
                shi[q] = max(shi[q], b1);
            }
        }
        const int ac = __shfl_up(pc, off, 64), bc = __shfl_down(sc, off, 64);
        if (up) pc += ac;
        if (dn) sc += bc;
    }
    // exclusive prefix: the left side of split k
    uint32_t llo[3], lhi[3];
    for (int q = 0; q < 3; ++q) {
        const uint32_t a0 = __shfl_up(plo[q], 1, 64), a1 = __shfl_up(phi[q], 1, 64);
        llo[q] = lane ? a0 : kKeyMax;
        lhi[q] = lane ? a1 : kKeyMin;
    }
    const int lc_ = __shfl_up(pc, 1, 64);
    const int lcnt = lane ? lc_ : 0;
    const bool usable = inb && !s.bad[i];
    float cost = 0.0f;
    if (usable) {
        const float sal = lcnt ? key_area(llo, lhi) : 0;
        const float sar = sc ? key_area(slo, shi) : 0;
        cost = lcnt * sal + sc * sar;
    }
    const bool valid = usable && cost < FLT_MAX;
    // first minimum: (cost, k) lexicographic, invalid lanes lose
    float bc = valid ? cost : FLT_MAX;
    int bk = valid ? lane : 64;
    for (int off = 32; off > 0; off >>= 1) {
        const float oc = __shfl_xor(bc, off, 64);
        const int ok = __shfl_xor(bk, off, 64);
        if (oc < bc || (oc == bc && ok < bk)) {
            bc = oc;
            bk = ok;
        }
    }
    const int best = bk == 64 ? 0 : bk;
    const int src = best;
    const int best_left = __shfl(lcnt, src, 64);
    uint32_t ck[12];
    for (int q = 0; q < 3; ++q) {
        ck[q] = __shfl(llo[q], src, 64);
        ck[3 + q] = __shfl(lhi[q], src, 64);
        ck[6 + q] = __shfl(slo[q], src, 64);
        ck[9 + q] = __shfl(shi[q], src, 64);
    }
    if (lane != 0) return;
    const int g = s.act[i];
    BuildNode& nd = s.nodes[g];
    const int total = nd.end - nd.begin;
    s.best_k[i] = best;
    s.n_left[i] = best_left;
    unsigned long long v = 0;
    if (best == 0) {
        nd.leaf = 1;
        nd.sort_axis = s.axis[i]; // the node's own sort decides its order
    } else {
        const int nl = best_left, nr = total - best_left;
        v = (1ull << 32) | (unsigned long long)((nl > s.leaf_max) + (nr > s.leaf_max));
        for (int q = 0; q < 12; ++q) s.child_keys[12 * (size_t)i + q] = ck[q];
    }
    s.scan_val[i] = v;
}

struct ChildArgs {
    int num_active;
    int leaf_max;
    int node_base; // id of the first child created at this level
    const int* act;
    const int* axis;
    const int* best_k;
    const int* n_left;
    const uint32_t* child_keys;
    const unsigned long long* scan_excl;
    BuildNode* nodes;
    int* next_act;
    int* child_next; // per active node: next-level active index of the left / right child (-1 = leaf)
    int* cursor;     // per active node: fill positions of left / right
};

__global__ void k_children(ChildArgs c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.num_active) return;
    c.cursor[2 * i] = 0;
    c.cursor[2 * i + 1] = 0;
    c.child_next[2 * i] = -1;
    c.child_next[2 * i + 1] = -1;
    if (c.best_k[i] == 0) return;
    const int g = c.act[i];
    BuildNode& nd = c.nodes[g];
    const unsigned long long e = c.scan_excl[i];
    const int split_rank = (int)(e >> 32);
    const int act_base = (int)(e & 0xffffffffull);
    const int lid = c.node_base + 2 * split_rank, rid = lid + 1;
    const int nl = c.n_left[i];
    const uint32_t* ck = c.child_keys + 12 * (size_t)i;
    BuildNode L, R;
    for (int q = 0; q < 3; ++q) {
        L.lo[q] = funkey(ck[q]);
        L.hi[q] = funkey(ck[3 + q]);
        R.lo[q] = funkey(ck[6 + q]);
        R.hi[q] = funkey(ck[9 + q]);
    }
    L.begin = nd.begin;
    L.end = nd.begin + nl;
    R.begin = nd.begin + nl;
    R.end = nd.end;
    L.child = R.child = -1;
    L.leaf = (L.end - L.begin) <= c.leaf_max;
    R.leaf = (R.end - R.begin) <= c.leaf_max;
    L.sort_axis = R.sort_axis = c.axis[i]; // a child leaf keeps its parent's order
    L.depth = R.depth = nd.depth + 1;
    c.nodes[lid] = L;
    c.nodes[rid] = R;
    nd.child = lid;
    int a = act_base;
    if (!L.leaf) {
        c.next_act[a] = lid;
        c.child_next[2 * i] = a++;
    }
    if (!R.leaf) {
        c.next_act[a] = rid;
        c.child_next[2 * i + 1] = a;
    }
}

struct PartArgs {
    int n;
    const int* ids;
    const int* pos_node;
    const int* pos_bin;
    const int* act;
    const int* best_k;
    const int* n_left;
    const int* child_next;
    const BuildNode* nodes;
    int* cursor;
    int* ids_out;
    int* pos_node_out;
};

// Partition of one group of 64 positions (every lane of the wave calls it). Cursor
// increments are aggregated per (node, side) key over the wave: one atomic for each of the
// first four keys present, then one per remaining lane.
__device__ __forceinline__ void partition_group(const PartArgs& p, int k, int lane)
{
    const bool inr = k < p.n;
    const int pn = inr ? p.pos_node[k] : -1;
    int side = -1; // -1: stays, 0 left, 1 right
    const int id = inr ? p.ids[k] : 0;
    if (pn >= 0 && p.best_k[pn] != 0) side = p.pos_bin[k] < p.best_k[pn] ? 0 : 1;
    const bool moving = side >= 0;
    const int key = moving ? 2 * pn + side : -1;
    int slot = -1;
    bool pending = moving;
    for (int it = 0; it < 4; ++it) {
        const unsigned long long m = __ballot(pending);
        if (m == 0ull) break;
        const int leader = __ffsll((long long)m) - 1;
        const int k0 = __shfl(key, leader, 64);
        const bool mine = pending && key == k0;
        const unsigned long long same = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&p.cursor[k0], __popcll(same));
        base = __shfl(base, leader, 64);
        if (mine) {
            slot = base + __popcll(same & ((1ull << lane) - 1ull));
            pending = false;
        }
    }
    if (pending) slot = atomicAdd(&p.cursor[key], 1);
    if (!inr) return;
    if (!moving) {
        p.ids_out[k] = id;
        p.pos_node_out[k] = -1;
        return;
    }
    const BuildNode& nd = p.nodes[p.act[pn]];
    const int dst = side == 0 ? nd.begin + slot : nd.begin + p.n_left[pn] + slot;
    p.ids_out[dst] = id;
    p.pos_node_out[dst] = p.child_next[2 * pn + side];
}

// Partition of the positions into the children. A wave handles PT_PART_SPAN groups of 64
// positions. When all of them belong to one node (near the root), it counts both sides
// over the span first and takes its slots with one atomic per side, then places the
// positions in order; otherwise it partitions group by group.
#define PT_PART_SPAN 16
__global__ void k_partition(PartArgs p)
{
    const int lane = (int)(threadIdx.x & 63);
    const unsigned long long lane_lt = (1ull << lane) - 1ull;
    const int first = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64 * PT_PART_SPAN;
    if (first >= p.n) return;
    int lo = 0x7fffffff, hi = -1;
    for (int it = 0; it < PT_PART_SPAN; ++it) {
        const int k = first + 64 * it + lane;
        const int pn = k < p.n ? p.pos_node[k] : -1;
        if (pn >= 0) {
            lo = min(lo, pn);
            hi = max(hi, pn);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off, 64));
        hi = max(hi, __shfl_xor(hi, off, 64));
    }
    if (hi < 0 || lo != hi) {
        for (int it = 0; it < PT_PART_SPAN; ++it) partition_group(p, first + 64 * it + lane, lane);
        return;
    }
    const int node = lo;
    const int bk = p.best_k[node];
    if (bk == 0) { // the node becomes a leaf: every position stays
        for (int it = 0; it < PT_PART_SPAN; ++it) {
            const int k = first + 64 * it + lane;
            if (k < p.n) {
                p.ids_out[k] = p.ids[k];
                p.pos_node_out[k] = -1;
            }
        }
        return;
    }
    int nl = 0, nr = 0;
    for (int it = 0; it < PT_PART_SPAN; ++it) {
        const int k = first + 64 * it + lane;
        const bool live = k < p.n && p.pos_node[k] >= 0;
        const bool left = live && p.pos_bin[k] < bk;
        nl += __popcll(__ballot(left));
        nr += __popcll(__ballot(live && !left));
    }
    int bl = 0, br = 0;
    if (lane == 0) {
        if (nl) bl = atomicAdd(&p.cursor[2 * node], nl);
        if (nr) br = atomicAdd(&p.cursor[2 * node + 1], nr);
    }
    bl = __shfl(bl, 0, 64);
    br = __shfl(br, 0, 64);
    const BuildNode& nd = p.nodes[p.act[node]];
    const int right0 = nd.begin + p.n_left[node];
    const int child_l = p.child_next[2 * node], child_r = p.child_next[2 * node + 1];
    for (int it = 0; it < PT_PART_SPAN; ++it) {
        const int k = first + 64 * it + lane;
        const bool inr = k < p.n;
        const bool live = inr && p.pos_node[k] >= 0;
        const bool left = live && p.pos_bin[k] < bk;
        const bool right = live && !left;
        const unsigned long long ml = __ballot(left), mr = __ballot(right);
        if (inr) {
            const int id = p.ids[k];
            if (left) {
                const int dst = nd.begin + bl + __popcll(ml & lane_lt);
                p.ids_out[dst] = id;
                p.pos_node_out[dst] = child_l;
            } else if (right) {
                const int dst = right0 + br + __popcll(mr & lane_lt);
                p.ids_out[dst] = id;
                p.pos_node_out[dst] = child_r;
            } else {
                p.ids_out[k] = id;
                p.pos_node_out[k] = -1;
            }
        }
        bl += __popcll(ml);
        br += __popcll(mr);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int* p, int v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_iota(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}
__global__ void k_fill_keyboxes(uint32_t* p, size_t boxes)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6 * boxes) p[i] = (i % 6) < 3 ? kKeyMax : kKeyMin;
}

// Leaves of up to kSmallLeaf triangles are sorted by (centroid on the sort axis, index)
// with one thread each; larger ones are listed for the host.
constexpr int kSmallLeaf = 32;

__global__ void k_sort_leaves(const BuildNode* nodes, int num_nodes, const float4* cen, int* ids, int* big,
                              int* num_big)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= num_nodes) return;
    const BuildNode nd = nodes[g];
    if (!nd.leaf || nd.sort_axis < 0) return;
    const int n = nd.end - nd.begin;
    if (n < 2) return;
    if (n > kSmallLeaf) {
        big[atomicAdd(num_big, 1)] = g;
        return;
    }
    const int ax = nd.sort_axis;
    auto key = [&](int id) {
        const float4 c = cen[id];
        return ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
    };
    for (int i = nd.begin + 1; i < nd.end; ++i) {
        const int v = ids[i];
        const float kv = key(v);
        int j = i - 1;
        while (j >= nd.begin) {
            const int u = ids[j];
            const float ku = key(u);
            if (ku < kv || (!(kv < ku) && u < v)) break;
            ids[j + 1] = u;
            --j;
        }
        ids[j + 1] = v;
    }
}

// ---- packing into the traversal layout (the layout of pack_bvh2) ---------------------

__global__ void k_subtree(const BuildNode* nodes, int g0, int g1, int* inner_count)
{
    const int g = g0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= g1) return;
    const BuildNode& nd = nodes[g];
    inner_count[g] = nd.leaf ? 0 : 1 + inner_count[nd.child] + inner_count[nd.child + 1];
}

// Depth-first preorder index of every inner node, left subtree first.
__global__ void k_preorder(const BuildNode* nodes, int g0, int g1, const int* inner_count, int* pre)
{
    const int g = g0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= g1) return;
    const BuildNode& nd = nodes[g];
    if (nd.leaf) return;
    const int p = pre[g];
    const int l = nd.child, r = nd.child + 1;
    if (!nodes[l].leaf) pre[l] = p + 1;
    if (!nodes[r].leaf) pre[r] = p + 1 + inner_count[l];
}

__device__ __forceinline__ bool is_big_leaf(const BuildNode& x)
{
    return x.leaf && x.end - x.begin > PT_REF_MAX_SMALL_LEAF;
}

// Big leaves are numbered in leaf-slot order: mark their first slot, then scan.
__global__ void k_mark_big(const BuildNode* nodes, int num_nodes, int* mark)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < num_nodes && is_big_leaf(nodes[g])) mark[nodes[g].begin] = 1;
}

__global__ void k_fill_big(const BuildNode* nodes, int num_nodes, const int* big_index, int2* big_leaves)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= num_nodes) return;
    const BuildNode& x = nodes[g];
    if (is_big_leaf(x)) big_leaves[big_index[x.begin]] = make_int2(x.begin, x.end - x.begin);
}

__device__ __forceinline__ uint32_t child_ref(const BuildNode& c, int c_pre, const int* big_index)
{
    if (!c.leaf) return pt_ref_inner((uint32_t)c_pre);
    const int cnt = c.end - c.begin;
    if (cnt <= PT_REF_MAX_SMALL_LEAF) return pt_ref_leaf((uint32_t)c.begin, (uint32_t)cnt);
    return pt_ref_big_leaf((uint32_t)big_index[c.begin]);
}

__global__ void k_pack_nodes(const BuildNode* nodes, int num_nodes, const int* pre, const int* big_index,
                             float4* out, int2* inner_range)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= num_nodes) return;
    const BuildNode& x = nodes[g];
    if (x.leaf) return;
    inner_range[pre[g]] = make_int2(x.begin, x.end);
    const BuildNode& l = nodes[x.child];
    const BuildNode& r = nodes[x.child + 1];
    float4* o = out + 4 * (size_t)pre[g];
    o[0] = make_float4(l.lo[0], l.lo[1], l.lo[2], l.hi[0]);
    o[1] = make_float4(l.hi[1], l.hi[2], r.lo[0], r.lo[1]);
    o[2] = make_float4(r.lo[2], r.hi[0], r.hi[1], r.hi[2]);
    const uint32_t rl = child_ref(l, pre[x.child], big_index);
    const uint32_t rr = child_ref(r, pre[x.child + 1], big_index);
    o[3] = make_float4(__uint_as_float(rl), __uint_as_float(rr), 0.0f, 0.0f);
}

__global__ void k_boxes_ordered(const float4* nodes, int num_inner, int* bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_inner) return;
    // left box (a.xyz, (a.w, b.x, b.y)), right box ((b.z, b.w, c.x), c.yzw)
    const float4 a = nodes[4 * (size_t)i], b = nodes[4 * (size_t)i + 1], c = nodes[4 * (size_t)i + 2];
    const bool ok = a.x <= a.w && a.y <= b.x && a.z <= b.y && b.z <= c.y && b.w <= c.z && c.x <= c.w;
    if (!ok) bad[0] = 1;
}

struct GatherArgs {
    const float* tri_pos;
    const float4* spheres; // primitives ntri.. are spheres (centre, radius)
    int ntri;
    int8_t* kinds;         // per leaf slot: 1 for a sphere; null without spheres
    const int* ids;
    int n;
    const int* tri_mat;
    const int* mat_shading_of;
    const int* smooth;
    const float* nrm;
    float4* tris;
    int* prim_ids;
    int* mat;
    int* mat_sh;
    int* smooth_out;
    float* nrm_out;
};

// Leaf-ordered triangle records and per-slot data.
__global__ void k_gather(GatherArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const int id = a.ids[k];
    if (id < a.ntri) {
        const float* p = a.tri_pos + 9 * (size_t)id;
        const TriDev t = tri_from_points(v3(p[0], p[1], p[2]), v3(p[3], p[4], p[5]), v3(p[6], p[7], p[8]));
        a.tris[3 * (size_t)k] = make_float4(t.p0.x, t.p0.y, t.p0.z, t.e1.x);
        a.tris[3 * (size_t)k + 1] = make_float4(t.e1.y, t.e1.z, t.e2.x, t.e2.y);
        a.tris[3 * (size_t)k + 2] = make_float4(t.e2.z, t.n.x, t.n.y, t.n.z);
    } else {
        // a sphere's slot: (centre, radius), then zeros
        a.tris[3 * (size_t)k] = a.spheres[id - a.ntri];
        a.tris[3 * (size_t)k + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.tris[3 * (size_t)k + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (a.kinds) a.kinds[k] = id < a.ntri ? 0 : 1;
    a.prim_ids[k] = id;
    const int m = a.tri_mat[id];
    a.mat[k] = m;
    a.mat_sh[k] = a.mat_shading_of[m];
    if (a.smooth) a.smooth_out[k] = a.smooth[id];
    if (a.nrm)
        for (int q = 0; q < 9; ++q) a.nrm_out[9 * (size_t)k + q] = a.nrm[9 * (size_t)id + q];
}

inline unsigned blocks(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Temporary device buffers, freed when the build returns.
struct Scratch {
    std::vector<void*> ptrs;
    template <typename T>
    hipError_t alloc(T** p, size_t count)
    {
        *p = nullptr;
        hipError_t e = hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
    ~Scratch()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

#define B_CHECK(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return e_;                                                               \
    } while (0)

float host_unkey(uint32_t k)
{
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

} // namespace

void gpu_bvh_free(GpuBvhResult* r)
{
    if (r->nodes_d) (void)hipFree(r->nodes_d);
    if (r->ids_d) (void)hipFree(r->ids_d);
    r->nodes_d = nullptr;
    r->ids_d = nullptr;
}

hipError_t gpu_build_bvh(const float* tri_pos_d, int ntri, const float4* spheres_d, int num_spheres, int num_bins,
                         int leaf_max, hipStream_t st, GpuBvhResult* out)
{
    const int n = ntri + num_spheres;
    *out = GpuBvhResult();
    if (n <= 0 || num_bins < 2 || num_bins > 64 || leaf_max < 1) return hipErrorInvalidValue;
    Scratch tmp;
    float4 *cen, *blo, *bhi;
    int *ids[2], *pos_node[2], *pos_bin;
    uint32_t* root_keys;
    B_CHECK(tmp.alloc(&cen, n));
    B_CHECK(tmp.alloc(&blo, n));
    B_CHECK(tmp.alloc(&bhi, n));
    B_CHECK(hipMalloc((void**)&out->ids_d, sizeof(int) * (size_t)n));
    ids[0] = out->ids_d;
    B_CHECK(tmp.alloc(&ids[1], n));
    B_CHECK(tmp.alloc(&pos_node[0], n));
    B_CHECK(tmp.alloc(&pos_node[1], n));
    B_CHECK(tmp.alloc(&pos_bin, n));
    B_CHECK(tmp.alloc(&root_keys, 6));
    const size_t max_nodes = 2 * (size_t)n + 1;
    B_CHECK(hipMalloc((void**)&out->nodes_d, sizeof(BuildNode) * max_nodes));
    BuildNode* nodes = out->nodes_d;
    // a level's splittable nodes hold more than leaf_max triangles each, disjointly
    const size_t max_active = (size_t)n / (leaf_max + 1) + 2;
    int *act[2], *axis, *bad, *bin_cnt, *best_k, *n_left, *child_next, *cursor;
    uint32_t *cbk, *bin_keys, *child_keys;
    float* len;
    unsigned long long *scan_val, *scan_excl;
    B_CHECK(tmp.alloc(&act[0], max_active));
    B_CHECK(tmp.alloc(&act[1], max_active));
    B_CHECK(tmp.alloc(&cbk, 6 * max_active));
    B_CHECK(tmp.alloc(&len, max_active));
    B_CHECK(tmp.alloc(&axis, max_active));
    B_CHECK(tmp.alloc(&bad, max_active));
    B_CHECK(tmp.alloc(&bin_keys, 6 * max_active * num_bins));
    B_CHECK(tmp.alloc(&bin_cnt, max_active * num_bins));
    B_CHECK(tmp.alloc(&child_keys, 12 * max_active));
    B_CHECK(tmp.alloc(&best_k, max_active));
    B_CHECK(tmp.alloc(&n_left, max_active));
    B_CHECK(tmp.alloc(&child_next, 2 * max_active));
    B_CHECK(tmp.alloc(&cursor, 2 * max_active));
    B_CHECK(tmp.alloc(&scan_val, max_active + 1));
    B_CHECK(tmp.alloc(&scan_excl, max_active + 1));
    size_t scan_bytes = 0;
    B_CHECK(exclusive_sum(nullptr, scan_bytes, scan_val, scan_excl, (size_t)((int)max_active + 1), st));
    char* scan_tmp;
    B_CHECK(tmp.alloc(&scan_tmp, scan_bytes));

    // primitives, the root box and the identity order
    hipLaunchKernelGGL(k_fill_keyboxes, dim3(1), dim3(64), 0, st, root_keys, (size_t)1);
    hipLaunchKernelGGL(k_prims, dim3(blocks(((size_t)n + PT_PRIMS_SPAN - 1) / PT_PRIMS_SPAN)), dim3(256), 0, st,
                       tri_pos_d, ntri, spheres_d, n, cen, blo, bhi, root_keys);
    hipLaunchKernelGGL(k_iota, dim3(blocks(n)), dim3(256), 0, st, ids[0], n);
    B_CHECK(hipGetLastError());
    uint32_t rk[6];
    B_CHECK(hipMemcpyAsync(rk, root_keys, sizeof rk, hipMemcpyDeviceToHost, st));
    B_CHECK(hipStreamSynchronize(st));
    BuildNode root;
    for (int q = 0; q < 3; ++q) {
        root.lo[q] = host_unkey(rk[q]);
        root.hi[q] = host_unkey(rk[3 + q]);
    }
    root.begin = 0;
    root.end = n;
    root.child = -1;
    root.leaf = n <= leaf_max;
    root.sort_axis = -1; // a root leaf keeps the input order
    root.depth = 1;
    B_CHECK(hipMemcpyAsync(nodes, &root, sizeof root, hipMemcpyHostToDevice, st));
    int num_nodes = 1, max_depth = 1;
    int num_active = root.leaf ? 0 : 1;
    if (num_active) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, act[0], 0, (size_t)1);
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks(n)), dim3(256), 0, st, pos_node[0], 0, (size_t)n);
    }
    std::vector<int> level_begin = {0, 1};
    int cur = 0;
    while (num_active > 0) {
        LevelArgs la;
        la.cen = cen;
        la.blo = blo;
        la.bhi = bhi;
        la.ids = ids[cur];
        la.pos_node = pos_node[cur];
        la.n = n;
        la.num_bins = num_bins;
        la.bin_nodes = bin_table_nodes(num_bins);
        la.act = act[cur];
        la.cbk = cbk;
        la.len = len;
        la.axis = axis;
        la.bad = bad;
        la.bin_keys = bin_keys;
        la.bin_cnt = bin_cnt;
        la.pos_bin = pos_bin;
        const size_t nbins = (size_t)num_active * num_bins;
        hipLaunchKernelGGL(k_fill_keyboxes, dim3(blocks(6 * (size_t)num_active)), dim3(256), 0, st, cbk,
                           (size_t)num_active);
        hipLaunchKernelGGL(k_fill_keyboxes, dim3(blocks(6 * nbins)), dim3(256), 0, st, bin_keys, nbins);
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks(nbins)), dim3(256), 0, st, bin_cnt, 0, nbins);
        hipLaunchKernelGGL(k_centroid_bounds, dim3(blocks(((size_t)n + PT_CB_SPAN - 1) / PT_CB_SPAN)), dim3(256), 0,
                           st, la);
        hipLaunchKernelGGL(k_axis, dim3(blocks(num_active)), dim3(256), 0, st, la, num_active);
        hipLaunchKernelGGL(k_bin, dim3(blocks(((size_t)n + PT_BIN_SPAN - 1) / PT_BIN_SPAN)), dim3(256),
                           bin_lds_words(num_bins) * sizeof(uint32_t), st, la);
        SahArgs sa;
        sa.num_active = num_active;
        sa.num_bins = num_bins;
        sa.leaf_max = leaf_max;
        sa.act = act[cur];
        sa.bin_keys = bin_keys;
        sa.bin_cnt = bin_cnt;
        sa.bad = bad;
        sa.axis = axis;
        sa.nodes = nodes;
        sa.best_k = best_k;
        sa.n_left = n_left;
        sa.child_keys = child_keys;
        sa.scan_val = scan_val;
        hipLaunchKernelGGL(k_sah, dim3(blocks(64 * (size_t)num_active)), dim3(256), 0, st, sa);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t*)(scan_val + num_active), 0u, (size_t)2);
        B_CHECK(exclusive_sum(scan_tmp, scan_bytes, scan_val, scan_excl, (size_t)(num_active + 1), st));
        unsigned long long tot = 0;
        B_CHECK(hipMemcpyAsync(&tot, scan_excl + num_active, sizeof tot, hipMemcpyDeviceToHost, st));
        ChildArgs ca;
        ca.num_active = num_active;
        ca.leaf_max = leaf_max;
        ca.node_base = num_nodes;
        ca.act = act[cur];
        ca.axis = axis;
        ca.best_k = best_k;
        ca.n_left = n_left;
        ca.child_keys = child_keys;
        ca.scan_excl = scan_excl;
        ca.nodes = nodes;
        ca.next_act = act[1 - cur];
        ca.child_next = child_next;
        ca.cursor = cursor;
        hipLaunchKernelGGL(k_children, dim3(blocks(num_active)), dim3(256), 0, st, ca);
        PartArgs pa;
        pa.n = n;
        pa.ids = ids[cur];
        pa.pos_node = pos_node[cur];
        pa.pos_bin = pos_bin;
        pa.act = act[cur];
        pa.best_k = best_k;
        pa.n_left = n_left;
        pa.child_next = child_next;
        pa.nodes = nodes;
        pa.cursor = cursor;
        pa.ids_out = ids[1 - cur];
        pa.pos_node_out = pos_node[1 - cur];
        hipLaunchKernelGGL(k_partition, dim3(blocks(((size_t)n + PT_PART_SPAN - 1) / PT_PART_SPAN)), dim3(256), 0,
                           st, pa);
        B_CHECK(hipGetLastError());
        B_CHECK(hipStreamSynchronize(st));
        const int splits = (int)(tot >> 32);
        cur = 1 - cur;
        num_nodes += 2 * splits;
        if (splits > 0) {
            max_depth += 1;
            level_begin.push_back(num_nodes);
        }
        num_active = (int)(tot & 0xffffffffull);
    }
    if (cur == 1) B_CHECK(hipMemcpyAsync(ids[0], ids[1], sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, st));
    // leaf order
    int *big, *num_big;
    B_CHECK(tmp.alloc(&big, (size_t)num_nodes));
    B_CHECK(tmp.alloc(&num_big, 1));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, num_big, 0, (size_t)1);
    hipLaunchKernelGGL(k_sort_leaves, dim3(blocks(num_nodes)), dim3(256), 0, st, nodes, num_nodes, cen, ids[0], big,
                       num_big);
    B_CHECK(hipGetLastError());
    int nbig = 0;
    B_CHECK(hipMemcpyAsync(&nbig, num_big, sizeof(int), hipMemcpyDeviceToHost, st));
    B_CHECK(hipStreamSynchronize(st));
    if (nbig > 0) { // leaves of more than kSmallLeaf triangles (coincident centroids), sorted on the host
        std::vector<int> bl(nbig);
        std::vector<float4> c(n);
        B_CHECK(hipMemcpy(bl.data(), big, sizeof(int) * (size_t)nbig, hipMemcpyDeviceToHost));
        B_CHECK(hipMemcpy(c.data(), cen, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost));
        for (int g : bl) {
            BuildNode nd;
            B_CHECK(hipMemcpy(&nd, nodes + g, sizeof nd, hipMemcpyDeviceToHost));
            const int m = nd.end - nd.begin;
            std::vector<int> sub(m);
            B_CHECK(hipMemcpy(sub.data(), ids[0] + nd.begin, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost));
            const int ax = nd.sort_axis;
            auto key = [&](int id) { return ax == 0 ? c[id].x : (ax == 1 ? c[id].y : c[id].z); };
            std::sort(sub.begin(), sub.end(), [&](int a, int b) {
                const float ka = key(a), kb = key(b);
                return ka < kb || (!(kb < ka) && a < b);
            });
            B_CHECK(hipMemcpy(ids[0] + nd.begin, sub.data(), sizeof(int) * (size_t)m, hipMemcpyHostToDevice));
        }
    }
    out->num_nodes = num_nodes;
    out->max_depth = max_depth;
    out->level_begin = level_begin;
    return hipSuccess;
}

hipError_t gpu_pack_bvh(const GpuBvhResult& r, int ntri, const float* tri_pos_d, const float4* spheres_d,
                        int num_spheres, const int32_t* tri_mat_d, const int32_t* mat_shading_of_d,
                        const int32_t* smooth_d, const float* nrm_d, hipStream_t st, GpuPacked* out)
{
    const int num_tris = ntri + num_spheres; // leaf slots
    *out = GpuPacked();
    Scratch tmp;
    const int nn = r.num_nodes;
    int *inner_count, *pre, *mark, *big_index;
    B_CHECK(tmp.alloc(&inner_count, (size_t)nn));
    B_CHECK(tmp.alloc(&pre, (size_t)nn));
    B_CHECK(tmp.alloc(&mark, (size_t)num_tris + 1));
    B_CHECK(tmp.alloc(&big_index, (size_t)num_tris + 1));
    const std::vector<int>& lb = r.level_begin;
    for (int l = (int)lb.size() - 2; l >= 0; --l) // bottom-up subtree sizes
        hipLaunchKernelGGL(k_subtree, dim3(blocks(lb[l + 1] - lb[l])), dim3(256), 0, st, r.nodes_d, lb[l], lb[l + 1],
                           inner_count);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, pre, 0, (size_t)1);
    for (int l = 0; l + 1 < (int)lb.size(); ++l) // top-down preorder indices
        hipLaunchKernelGGL(k_preorder, dim3(blocks(lb[l + 1] - lb[l])), dim3(256), 0, st, r.nodes_d, lb[l], lb[l + 1],
                           inner_count, pre);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks((size_t)num_tris + 1)), dim3(256), 0, st, mark, 0,
                       (size_t)num_tris + 1);
    hipLaunchKernelGGL(k_mark_big, dim3(blocks(nn)), dim3(256), 0, st, r.nodes_d, nn, mark);
    size_t scan_bytes = 0;
    B_CHECK(exclusive_sum(nullptr, scan_bytes, mark, big_index, (size_t)(num_tris + 1), st));
    char* scan_tmp;
    B_CHECK(tmp.alloc(&scan_tmp, scan_bytes));
    B_CHECK(exclusive_sum(scan_tmp, scan_bytes, mark, big_index, (size_t)(num_tris + 1), st));
    int counts[2] = {0, 0};
    B_CHECK(hipMemcpyAsync(&counts[0], inner_count, sizeof(int), hipMemcpyDeviceToHost, st));
    B_CHECK(hipMemcpyAsync(&counts[1], big_index + num_tris, sizeof(int), hipMemcpyDeviceToHost, st));
    BuildNode root;
    B_CHECK(hipMemcpyAsync(&root, r.nodes_d, sizeof root, hipMemcpyDeviceToHost, st));
    B_CHECK(hipStreamSynchronize(st));
    out->num_inner = counts[0];
    out->num_big = counts[1];
    B_CHECK(hipMalloc((void**)&out->nodes_d, sizeof(float4) * 4 * (size_t)std::max(out->num_inner, 1)));
    B_CHECK(hipMalloc((void**)&out->big_leaves_d, sizeof(int2) * (size_t)std::max(out->num_big, 1)));
    B_CHECK(hipMalloc((void**)&out->inner_range_d, sizeof(int2) * (size_t)std::max(out->num_inner, 1)));
    hipLaunchKernelGGL(k_pack_nodes, dim3(blocks(nn)), dim3(256), 0, st, r.nodes_d, nn, pre, big_index, out->nodes_d,
                       out->inner_range_d);
    hipLaunchKernelGGL(k_fill_big, dim3(blocks(nn)), dim3(256), 0, st, r.nodes_d, nn, big_index, out->big_leaves_d);
    B_CHECK(hipMalloc((void**)&out->tris_d, sizeof(float4) * 3 * (size_t)num_tris));
    B_CHECK(hipMalloc((void**)&out->prim_ids_d, sizeof(int) * (size_t)num_tris));
    B_CHECK(hipMalloc((void**)&out->mat_d, sizeof(int) * (size_t)num_tris));
    B_CHECK(hipMalloc((void**)&out->mat_sh_d, sizeof(int) * (size_t)num_tris));
    if (smooth_d) B_CHECK(hipMalloc((void**)&out->smooth_d, sizeof(int) * (size_t)num_tris));
    if (nrm_d) B_CHECK(hipMalloc((void**)&out->nrm_d, sizeof(float) * 9 * (size_t)num_tris));
    if (num_spheres > 0) B_CHECK(hipMalloc((void**)&out->kinds_d, (size_t)num_tris));
    GatherArgs ga;
    ga.tri_pos = tri_pos_d;
    ga.spheres = spheres_d;
    ga.ntri = ntri;
    ga.kinds = out->kinds_d;
    ga.ids = r.ids_d;
    ga.n = num_tris;
    ga.tri_mat = tri_mat_d;
    ga.mat_shading_of = mat_shading_of_d;
    ga.smooth = smooth_d;
    ga.nrm = nrm_d;
    ga.tris = out->tris_d;
    ga.prim_ids = out->prim_ids_d;
    ga.mat = out->mat_d;
    ga.mat_sh = out->mat_sh_d;
    ga.smooth_out = out->smooth_d;
    ga.nrm_out = out->nrm_d;
    hipLaunchKernelGGL(k_gather, dim3(blocks(num_tris)), dim3(256), 0, st, ga);
    B_CHECK(hipGetLastError());
    B_CHECK(hipStreamSynchronize(st));
    out->root_lo = v3(root.lo[0], root.lo[1], root.lo[2]);
    out->root_hi = v3(root.hi[0], root.hi[1], root.hi[2]);
    if (!root.leaf) out->root_ref = pt_ref_inner(0);
    else if (num_tris <= PT_REF_MAX_SMALL_LEAF) out->root_ref = pt_ref_leaf(0, (uint32_t)num_tris);
    else out->root_ref = pt_ref_big_leaf(0);
    return hipSuccess;
}

void gpu_packed_free(GpuPacked* p)
{
    for (void* q : {(void*)p->nodes_d, (void*)p->big_leaves_d, (void*)p->inner_range_d, (void*)p->tris_d, (void*)p->prim_ids_d,
                    (void*)p->mat_d, (void*)p->mat_sh_d, (void*)p->smooth_d, (void*)p->nrm_d, (void*)p->kinds_d})
        if (q) (void)hipFree(q);
    *p = GpuPacked();
}

hipError_t gpu_boxes_ordered(const float4* nodes_d, int num_inner, hipStream_t st, bool* ordered)
{
    *ordered = true;
    if (num_inner <= 0) return hipSuccess;
    int* bad_d = nullptr;
    int bad = 0;
    hipError_t e = hipMalloc((void**)&bad_d, sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(bad_d, 0, sizeof(int), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_boxes_ordered, dim3((unsigned)((num_inner + 255) / 256)), dim3(256), 0, st, nodes_d,
                           num_inner, bad_d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, bad_d, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (bad_d) (void)hipFree(bad_d);
    *ordered = e == hipSuccess && bad == 0;
    return e;
}

} // namespace pt
