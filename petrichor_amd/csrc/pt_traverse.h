// pt_traverse.h — single-ray BVH traversal for gfx950, shared by the persistent trace
// kernel (pt_trace.hip) and the path-completion kernel (pt_kernels.hip): closest hit and
// any hit over the binary tree in the reference's topology, or over the 8-wide tree.
//
// Walk order and culling are the reference's (BinnedSAHBVH::Intersect,
// Core/Accel/BinnedSAHBVH.cpp:272-374): from an inner node the child whose box centroid
// is nearer the ray origin (squared distance, AABB::CalcCentroid, AABB.h:70-74) is
// visited first, the other is pushed with its slab distance; a popped child is skipped
// when a hit exists, the origin is not strictly inside the child's box and the hit is
// nearer than the child's slab distance; leaf triangles are tested in leaf order and a
// hit replaces the best one only when strictly nearer. The reference pushes children
// whose box the ray misses and rejects them when popped; they are not pushed here.
//
// The wide traversal (BVHs built with width 4 or 8, nodes in pt_wide.hip) tests the up
// to 4 or 8 child boxes of a node against [max(dist_min, 0), min(nearest hit, dist_max)], visits
// the nearest hit child next and pushes the others farthest first with their entry
// distance; a popped entry is skipped when the nearest hit is closer than its entry
// distance. Leaves, triangle tests and the hit rule are the reference's, so it finds the
// same nearest hit, and of several hits at exactly the same distance it keeps the one
// the reference's walk reaches first (reference_first); only rays that graze a box
// within rounding of their hit can differ.
//
// A step handles one inner node or one triangle (wide: one node or one whole leaf).
// Each lane keeps its traversal stack in LDS (the first lds_stack_entries entries) with
// deeper entries in private memory.
//
// Binary trees deeper than that stack (BvhView::parents set) are walked without one, in
// the same order: the stack only ever holds the far children of the ancestors of the
// current entry whose near child was entered while the far one was hit, deepest on top.
// So when an entry is done, the next one is found by walking up the parent links to the
// nearest ancestor entered through its near child whose far child the ray hits; its
// distance and inside bit are the same slab and containment tests the push would have
// stored. The lane's first LDS stack slot holds the current entry's parent.
#pragma once

#include <hip/hip_runtime.h>

#include "pt_device.h"
#include "pt_kernels.h"

namespace pt {

// LDS stack entries per lane, by node width W. Binary: 10 (20 KB per 256-thread block, 8
// blocks or 32 waves per CU). 4-wide: 13 (26 KB, the 6 blocks per CU its registers
// allow). 8-wide: 16 (32 KB, 5 blocks per CU, which its registers allow anyway), as its
// stacks grow by up to 7 entries per node.
template <int W>
constexpr int lds_stack_entries()
{
    return W == 8 ? 16 : (W == 4 ? 13 : 10);
}
#define PT_TRACE_BLOCK 256
// Whole traversal stack per lane (LDS entries first, the rest in private memory).
template <int W>
constexpr int stack_entries()
{
    return W == 8 ? PT_STACK8 : (W == 4 ? PT_STACK4 : PT_MAX_STACK);
}

struct StackEntry {
    uint32_t ref; // PT_REF_* word with the inside bit
    float d;      // slab distance of that child's box
};

struct TraceRay {
    V3 o, d, inv;
    float dmin, dmax;
    float tnear; // max(dist_min, 0): where child boxes start to matter (8-wide traversal)
    int sx, sy, sz;
    bool any;
    bool ordered; // slab_ordered applies: finite origin, finite nonzero inv, ordered boxes
};

struct TraceState {
    int best;     // leaf slot of the nearest hit so far, -1 = none
    float best_t;
    uint32_t cur; // entry being processed
    float cur_d;
    int sp;
    uint32_t nv, nt; // node visits (slab tests) and triangle tests
    uint32_t leaf_k, leaf_end; // triangles of the current leaf still to test
};

template <int W>
__device__ __forceinline__ void stack_push(uint2* lds, StackEntry* spill, int sp, uint32_t ref, float d)
{
    constexpr int NL = lds_stack_entries<W>();
    if (sp < NL) {
        lds[sp * PT_TRACE_BLOCK] = make_uint2(ref, __float_as_uint(d));
    } else {
        spill[sp - NL].ref = ref;
        spill[sp - NL].d = d;
    }
}

template <int W>
__device__ __forceinline__ void stack_pop(const uint2* lds, const StackEntry* spill, int sp, uint32_t* ref,
                                          float* d)
{
    constexpr int NL = lds_stack_entries<W>();
    if (sp < NL) {
        const uint2 e = lds[sp * PT_TRACE_BLOCK];
        *ref = e.x;
        *d = __uint_as_float(e.y);
    } else {
        *ref = spill[sp - NL].ref;
        *d = spill[sp - NL].d;
    }
}

// Starts a ray at the root: returns false when the root box is missed (or the scene is
// empty), which finishes the ray with no hit.
__device__ __forceinline__ bool trace_begin(const BvhView& b, const float4 r0, const float4 r1, bool any,
                                            TraceRay& r, TraceState& s)
{
    r.o = v3(r0.x, r0.y, r0.z);
    r.d = v3(r1.x, r1.y, r1.z);
    r.dmin = r0.w;
    r.dmax = r1.w;
    r.tnear = fmaxf(r0.w, 0.0f);
    r.any = any;
    r.inv = v3(1.0f, 1.0f, 1.0f) / r.d; // Vector3f::One() / ray.dir (BinnedSAHBVH.cpp:280)
    r.sx = r.inv.x < 0.0f;
    r.sy = r.inv.y < 0.0f;
    r.sz = r.inv.z < 0.0f;
    const float big = __builtin_huge_valf();
    r.ordered = b.boxes_ordered != 0 && fabsf(r.o.x) < big && fabsf(r.o.y) < big && fabsf(r.o.z) < big &&
                fabsf(r.inv.x) < big && fabsf(r.inv.y) < big && fabsf(r.inv.z) < big && r.inv.x != 0.0f &&
                r.inv.y != 0.0f && r.inv.z != 0.0f;
    s.best = -1;
    s.best_t = 0.0f;
    s.sp = 0;
    s.nv = 0;
    s.nt = 0;
    s.leaf_k = s.leaf_end = 0;
    if (b.root_ref == PT_REF_EMPTY) return false;
    s.nv = 1;
    s.cur = b.nodes8 ? b.root8 : b.root_ref;
    return slab(r.o, r.inv, r.sx, r.sy, r.sz, b.root_lo, b.root_hi, &s.cur_d);
}

// Whether the reference's walk reaches leaf slot a before leaf slot c, both hit by the
// ray: below their lowest common ancestor in the binary tree it descends first into the
// child whose box centroid is nearer the ray origin (and finishes that subtree before
// the other); inside one leaf the slots are tested in order. Used by the 8-wide
// traversal to break exact distance ties the way the reference does.
__device__ __noinline__ bool reference_first(const BvhView& b, const TraceRay& r, int a, int c)
{
    uint32_t ref = b.root_ref;
    // down the binary tree: ends at a leaf or where a and c part, within its depth
    for (int depth = 0; depth <= (int)PT_REF_INDEX_MASK; ++depth) {
        if (pt_ref_cnt(ref) != 0) return a < c;
        const float4* nd = b.nodes + 4 * (size_t)pt_ref_index(ref);
        const float4 q0 = nd[0], q1 = nd[1], q2 = nd[2], m = nd[3];
        const uint32_t ml = __float_as_uint(m.x), mr = __float_as_uint(m.y);
        int lb, le;
        if (pt_ref_cnt(ml) == 0) {
            const int2 rg = b.inner_range[pt_ref_index(ml)];
            lb = rg.x;
            le = rg.y;
        } else if (pt_ref_cnt(ml) == PT_REF_BIG) {
            const int2 bl = b.big_leaves[pt_ref_index(ml)];
            lb = bl.x;
            le = bl.x + bl.y;
        } else {
            lb = (int)pt_ref_index(ml);
            le = lb + (int)pt_ref_cnt(ml);
        }
        const bool a_left = a >= lb && a < le, c_left = c >= lb && c < le;
        if (a_left == c_left) {
            ref = a_left ? ml : mr;
            continue;
        }
        const V3 llo = v3(q0.x, q0.y, q0.z), lhi = v3(q0.w, q1.x, q1.y);
        const V3 rlo = v3(q1.z, q1.w, q2.x), rhi = v3(q2.y, q2.z, q2.w);
        const float dl = sq_length(0.5f * (llo + lhi) - r.o);
        const float dr = sq_length(0.5f * (rlo + rhi) - r.o);
        return a_left == (dl < dr);
    }
    return a < c;
}

// Tests the primitive of leaf slot k, a triangle or a sphere (the reference's hit rule);
// true when an any-hit ray has found its hit. The 8-wide traversal meets leaves out of the reference's order, so an exactly
// equal distance goes to the triangle the reference would have reached first.
template <bool STATS, int W, bool EXT>
__device__ __forceinline__ bool tri_test(const BvhView& b, const TraceRay& r, TraceState& s, uint32_t k)
{
    if (STATS) ++s.nt;
    const float4 t0 = b.tris[3 * (size_t)k];
    const float4 t1 = b.tris[3 * (size_t)k + 1];
    const float4 t2 = b.tris[3 * (size_t)k + 2];
    TriDev tr;
    tr.p0 = v3(t0.x, t0.y, t0.z);
    tr.e1 = v3(t0.w, t1.x, t1.y);
    tr.e2 = v3(t1.z, t1.w, t2.x);
    tr.n = v3(t2.y, t2.z, t2.w);
    float t;
    bool hit = tri_intersect(tr, r.o, r.d, &t);
    // a sphere's slot (only in scenes with spheres): its record read again, so that the
    // triangle path keeps its registers
    if (EXT && b.kinds != nullptr && b.kinds[k] != 0) {
        const float4 q = b.tris[3 * (size_t)k];
        hit = sphere_intersect(v3(q.x, q.y, q.z), q.w, r.o, r.d, &t);
    }
    if (hit && !(t < r.dmin || t > r.dmax)) {
        if (s.best < 0 || t < s.best_t) {
            s.best = (int)k;
            s.best_t = t;
            if (r.any) return true;
        } else if (W != 2 && t == s.best_t && reference_first(b, r, (int)k, s.best)) {
            s.best = (int)k;
        }
    }
    return false;
}

// Enters the leaf s.cur: its triangles are tested in leaf order from the next call of
// leaf_or_pop on.
__device__ __forceinline__ void enter_leaf(const BvhView& b, TraceState& s)
{
    uint32_t first = pt_ref_index(s.cur), cnt = pt_ref_cnt(s.cur);
    if (cnt == PT_REF_BIG) {
        const int2 bl = b.big_leaves[first];
        first = (uint32_t)bl.x;
        cnt = (uint32_t)bl.y;
    }
    s.leaf_k = first;
    s.leaf_end = first + cnt;
}

// End of a step: one triangle of the current leaf (so that lanes in leaves and lanes in
// inner nodes do similar work per step), then, when the step did not move to a child
// (`next`) and no triangle is left, the next entry from the stack. Returns true when the
// ray is finished.
// Deep binary trees: moves to the entry the stack would pop next (see the header) and
// returns true, or returns false when there is none (the ray is done).
__device__ __forceinline__ bool deep_next(const BvhView& b, const TraceRay& r, TraceState& s, uint2* lds)
{
    uint32_t child = s.cur & ~PT_REF_INSIDE;
    if (child == b.root_ref) return false;
    uint32_t p = lds[0].x; // parent of the entry just done
    // the parent links form a tree rooted at root_ref (checked when they are made), so
    // the walk ends at the root; the bound only caps a walk over links that do not
    for (int guard = 0; guard <= (int)PT_REF_INDEX_MASK; ++guard) {
        const float4* nd = b.nodes + 4 * (size_t)p;
        const float4 a = nd[0], bb = nd[1], c = nd[2], mf = nd[3];
        const uint32_t ml = __float_as_uint(mf.x), mr = __float_as_uint(mf.y);
        const V3 llo = v3(a.x, a.y, a.z), lhi = v3(a.w, bb.x, bb.y);
        const V3 rlo = v3(bb.z, bb.w, c.x), rhi = v3(c.y, c.z, c.w);
        const float dl = sq_length(0.5f * (llo + lhi) - r.o);
        const float dr = sq_length(0.5f * (rlo + rhi) - r.o);
        const bool left_first = dl < dr;
        if (child == (left_first ? ml : mr)) {
            // entered through the near child: the far one is next if the ray hits it
            const V3 flo = left_first ? rlo : llo, fhi = left_first ? rhi : lhi;
            float d2;
            if (slab(r.o, r.inv, r.sx, r.sy, r.sz, flo, fhi, &d2)) {
                s.cur = (left_first ? mr : ml) | (box_contains(flo, fhi, r.o) ? PT_REF_INSIDE : 0u);
                s.cur_d = d2;
                lds[0].x = p;
                return true;
            }
        }
        if (p == b.root_ref) return false;
        child = p; // an inner node's reference is its index
        p = (uint32_t)b.parents[p];
    }
    return false;
}

template <bool STATS, int W, bool EXT>
__device__ __forceinline__ bool leaf_or_pop(const BvhView& b, const TraceRay& r, TraceState& s, uint2* lds,
                                            StackEntry* spill, bool next)
{
    if (s.leaf_k < s.leaf_end) {
        if (W != 2) { // a wide node's work per step is larger: whole leaves keep steps alike
            for (; s.leaf_k < s.leaf_end; ++s.leaf_k)
                if (tri_test<STATS, W, EXT>(b, r, s, s.leaf_k)) return true;
        } else {
            if (tri_test<STATS, W, EXT>(b, r, s, s.leaf_k++)) return true;
            if (s.leaf_k < s.leaf_end) return false;
        }
        next = false;
    }
    if (!next) {
        if (W == 2 && EXT && b.parents) return !deep_next(b, r, s, lds);
        if (s.sp == 0) return true;
        --s.sp;
        stack_pop<W>(lds, spill, s.sp, &s.cur, &s.cur_d);
    }
    return false;
}

// The inner node s.cur of the binary walk: its two child boxes are tested and the ray
// moves to the child the reference visits first (the other one, when hit, is pushed).
// Returns true when it moved to a child; false when it hit neither (the entry is done).
template <bool STATS, bool EXT>
__device__ __forceinline__ bool node_step(const BvhView& b, const TraceRay& r, TraceState& s, uint2* lds,
                                          StackEntry* spill)
{
    const float4* nd = b.nodes + 4 * (size_t)pt_ref_index(s.cur);
    const float4 a = nd[0], bb = nd[1], c = nd[2], mf = nd[3];
    const uint32_t ml = __float_as_uint(mf.x), mr = __float_as_uint(mf.y);
    const V3 llo = v3(a.x, a.y, a.z), lhi = v3(a.w, bb.x, bb.y);
    const V3 rlo = v3(bb.z, bb.w, c.x), rhi = v3(c.y, c.z, c.w);
    if (STATS) s.nv += 2;
    const float dl = sq_length(0.5f * (llo + lhi) - r.o);
    const float dr = sq_length(0.5f * (rlo + rhi) - r.o);
    float tl, tr_;
    bool hl, hr;
    if (__ballot(!r.ordered) == 0ull) { // every ray stepping here: the branch-free test
        hl = slab_ordered(r.o, r.inv, llo, lhi, &tl);
        hr = slab_ordered(r.o, r.inv, rlo, rhi, &tr_);
    } else {
        hl = slab(r.o, r.inv, r.sx, r.sy, r.sz, llo, lhi, &tl);
        hr = slab(r.o, r.inv, r.sx, r.sy, r.sz, rlo, rhi, &tr_);
    }
    const uint32_t el = ml | (box_contains(llo, lhi, r.o) ? PT_REF_INSIDE : 0u);
    const uint32_t er = mr | (box_contains(rlo, rhi, r.o) ? PT_REF_INSIDE : 0u);
    // the child the reference pops first: left if dl < dr, else right
    const bool left_first = dl < dr;
    const bool h1 = left_first ? hl : hr, h2 = left_first ? hr : hl;
    const uint32_t e1 = left_first ? el : er, e2 = left_first ? er : el;
    const float d1 = left_first ? tl : tr_, d2 = left_first ? tr_ : tl;
    if (EXT && b.parents) {
        if (h1 || h2) {
            lds[0].x = pt_ref_index(s.cur); // deep tree: the parent of the child entered
            s.cur = h1 ? e1 : e2;
            s.cur_d = h1 ? d1 : d2;
            return true;
        }
    } else if (h1) {
        if (h2) {
            stack_push<2>(lds, spill, s.sp, e2, d2);
            ++s.sp;
        }
        s.cur = e1;
        s.cur_d = d1;
        return true;
    } else if (h2) {
        s.cur = e2;
        s.cur_d = d2;
        return true;
    }
    return false;
}

// Whether the entry s.cur is culled: a nearer hit exists and the origin is outside its box.
__device__ __forceinline__ bool entry_culled(const TraceState& s)
{
    return s.best >= 0 && !(s.cur & PT_REF_INSIDE) && s.best_t < s.cur_d;
}

// One traversal step: an inner node or triangles of a leaf, then the next entry from the
// stack when the current one is done. Returns true when the ray is finished. EXT: the walk
// handles trees deeper than the stack and sphere leaves (BvhView::parents, ::kinds); the
// kernels take it only for BVHs that have them, so ordinary trees keep the plain walk.
template <bool STATS, bool EXT>
__device__ __forceinline__ bool trace_step(const BvhView& b, const TraceRay& r, TraceState& s, uint2* lds,
                                           StackEntry* spill)
{
    bool next = false;
    if (s.leaf_k < s.leaf_end) {
        // in the middle of a leaf
    } else if (entry_culled(s)) {
        // culled
    } else if (pt_ref_cnt(s.cur) != 0) {
        enter_leaf(b, s);
    } else {
        next = node_step<STATS, EXT>(b, r, s, lds, spill);
    }
    return leaf_or_pop<STATS, 2, EXT>(b, r, s, lds, spill, next);
}

// One step of the binary walk of a plain tree in one of two wave-wide phases. In a node
// phase (tri_phase false) the rays that are not in a leaf take their next entry: an inner
// node is tested, a culled entry is dropped, a leaf is entered (its triangles wait for a
// triangle phase); rays in a leaf wait. In a triangle phase the rays in a leaf test their
// next triangle and the others wait. Each ray still makes the moves of trace_step in the
// same order, so its hits are the same; a wave only runs the code of one kind of move per
// step, where trace_step runs both whenever its rays are at both. Returns true when the
// ray is finished.
template <bool STATS>
__device__ __forceinline__ bool trace_step_phase(const BvhView& b, const TraceRay& r, TraceState& s, uint2* lds,
                                                 StackEntry* spill, bool tri_phase)
{
    const bool in_leaf = s.leaf_k < s.leaf_end;
    if (tri_phase) {
        if (!in_leaf) return false;
        if (tri_test<STATS, 2, false>(b, r, s, s.leaf_k++)) return true;
        if (s.leaf_k < s.leaf_end) return false;
    } else {
        if (in_leaf) return false;
        if (entry_culled(s)) {
            // culled
        } else if (pt_ref_cnt(s.cur) != 0) {
            enter_leaf(b, s);
            if (s.leaf_k < s.leaf_end) return false;
        } else if (node_step<STATS, false>(b, r, s, lds, spill)) {
            return false;
        }
    }
    // the entry is done: the next one from the stack
    if (s.sp == 0) return true;
    --s.sp;
    stack_pop<2>(lds, spill, s.sp, &s.cur, &s.cur_d);
    return false;
}

__device__ __forceinline__ void sort_pair(float& da, uint32_t& ra, float& db, uint32_t& rb)
{
    const bool sw = db < da;
    const float lo_d = sw ? db : da, hi_d = sw ? da : db;
    const uint32_t lo_r = sw ? rb : ra, hi_r = sw ? ra : rb;
    da = lo_d;
    ra = lo_r;
    db = hi_d;
    rb = hi_r;
}

__device__ __forceinline__ float byte_f(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xffu); }

// One step of the wide traversal (W = 4 or 8 children per node): the current entry (a
// wide node or a leaf), then the next entry from the stack. Returns true when the ray is
// finished.
template <bool STATS, bool EXT, int W>
__device__ __forceinline__ bool trace_wide_step(const BvhView& b, const TraceRay& r, TraceState& s, uint2* lds,
                                                StackEntry* spill)
{
    static_assert(W == 4 || W == 8, "wide nodes hold 4 or 8 children");
    bool next = false;
    if (s.leaf_k < s.leaf_end) {
        // in the middle of a leaf
    } else if (s.best >= 0 && s.best_t < s.cur_d) {
        // culled: the box starts beyond the nearest hit
    } else if (pt_ref_cnt(s.cur) != 0) {
        enter_leaf(b, s);
    } else {
        // quantised near / far plane bytes by axis (q[0..2] the near planes, q[3..5] the
        // far ones for this ray's octant) and the child references, per 4 children
        uint32_t nq[3][W / 4], fq[3][W / 4], ref[W];
        uint4 w0;
        if constexpr (W == 8) {
            const uint4* nd = b.nodes8 + 8 * (size_t)pt_ref_index(s.cur);
            w0 = nd[0];
            const uint4 w1 = nd[1], w2 = nd[2], w3 = nd[3], w4 = nd[4], w5 = nd[5];
            // qlo: w1.xy, w1.zw, w2.xy; qhi: w2.zw, w3.xy, w3.zw
            const uint32_t lo[3][2] = {{w1.x, w1.y}, {w1.z, w1.w}, {w2.x, w2.y}};
            const uint32_t hi[3][2] = {{w2.z, w2.w}, {w3.x, w3.y}, {w3.z, w3.w}};
            const int sg[3] = {r.sx, r.sy, r.sz};
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    nq[a][h] = sg[a] ? hi[a][h] : lo[a][h];
                    fq[a][h] = sg[a] ? lo[a][h] : hi[a][h];
                }
            ref[0] = w4.x; ref[1] = w4.y; ref[2] = w4.z; ref[3] = w4.w;
            ref[4] = w5.x; ref[5] = w5.y; ref[6] = w5.z; ref[7] = w5.w;
        } else {
            const uint4* nd = b.nodes8 + 4 * (size_t)pt_ref_index(s.cur);
            w0 = nd[0];
            const uint4 w1 = nd[1], w2 = nd[2], w3 = nd[3];
            // qlo: w1.x, w1.y, w1.z; qhi: w1.w, w2.x, w2.y; references w2.zw, w3.xy
            nq[0][0] = r.sx ? w1.w : w1.x;
            fq[0][0] = r.sx ? w1.x : w1.w;
            nq[1][0] = r.sy ? w2.x : w1.y;
            fq[1][0] = r.sy ? w1.y : w2.x;
            nq[2][0] = r.sz ? w2.y : w1.z;
            fq[2][0] = r.sz ? w1.z : w2.y;
            ref[0] = w2.z; ref[1] = w2.w; ref[2] = w3.x; ref[3] = w3.y;
        }
        const int cnt = (int)(w0.w >> 24);
        // the ray in the node's quantised frame: t = q * (inv * step) + (p - o) * inv
        const float ix = r.inv.x * __uint_as_float((w0.w & 0xffu) << 23);
        const float iy = r.inv.y * __uint_as_float(((w0.w >> 8) & 0xffu) << 23);
        const float iz = r.inv.z * __uint_as_float(((w0.w >> 16) & 0xffu) << 23);
        const float ox = (__uint_as_float(w0.x) - r.o.x) * r.inv.x;
        const float oy = (__uint_as_float(w0.y) - r.o.y) * r.inv.y;
        const float oz = (__uint_as_float(w0.z) - r.o.z) * r.inv.z;
        const float tfar = s.best >= 0 ? fminf(s.best_t, r.dmax) : r.dmax;
        float d[W];
        int hits = 0;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const int h = i >> 2, k = i & 3;
            const float tx0 = fmaf(byte_f(nq[0][h], k), ix, ox);
            const float ty0 = fmaf(byte_f(nq[1][h], k), iy, oy);
            const float tz0 = fmaf(byte_f(nq[2][h], k), iz, oz);
            const float tx1 = fmaf(byte_f(fq[0][h], k), ix, ox);
            const float ty1 = fmaf(byte_f(fq[1][h], k), iy, oy);
            const float tz1 = fmaf(byte_f(fq[2][h], k), iz, oz);
            const float tmin = fmaxf(fmaxf(tx0, ty0), fmaxf(tz0, r.tnear));
            const float tmax = fminf(fminf(tx1, ty1), fminf(tz1, tfar));
            const bool hit = i < cnt && tmin <= tmax;
            d[i] = hit ? tmin : __builtin_inff();
            hits += hit ? 1 : 0;
        }
        if (STATS) s.nv += 1; // one node record (128 B at width 8, 64 B at width 4)
        if constexpr (W == 8) {
            // nearest first: 19-comparator sorting network for 8 keys
            sort_pair(d[0], ref[0], d[2], ref[2]);
            sort_pair(d[1], ref[1], d[3], ref[3]);
            sort_pair(d[4], ref[4], d[6], ref[6]);
            sort_pair(d[5], ref[5], d[7], ref[7]);
            sort_pair(d[0], ref[0], d[4], ref[4]);
            sort_pair(d[1], ref[1], d[5], ref[5]);
            sort_pair(d[2], ref[2], d[6], ref[6]);
            sort_pair(d[3], ref[3], d[7], ref[7]);
            sort_pair(d[0], ref[0], d[1], ref[1]);
            sort_pair(d[2], ref[2], d[3], ref[3]);
            sort_pair(d[4], ref[4], d[5], ref[5]);
            sort_pair(d[6], ref[6], d[7], ref[7]);
            sort_pair(d[2], ref[2], d[4], ref[4]);
            sort_pair(d[3], ref[3], d[5], ref[5]);
            sort_pair(d[1], ref[1], d[4], ref[4]);
            sort_pair(d[3], ref[3], d[6], ref[6]);
            sort_pair(d[1], ref[1], d[2], ref[2]);
            sort_pair(d[3], ref[3], d[4], ref[4]);
            sort_pair(d[5], ref[5], d[6], ref[6]);
        } else {
            // 5-comparator sorting network for 4 keys
            sort_pair(d[0], ref[0], d[1], ref[1]);
            sort_pair(d[2], ref[2], d[3], ref[3]);
            sort_pair(d[0], ref[0], d[2], ref[2]);
            sort_pair(d[1], ref[1], d[3], ref[3]);
            sort_pair(d[1], ref[1], d[2], ref[2]);
        }
#pragma unroll
        for (int j = W - 1; j >= 1; --j) {
            if (j < hits) {
                stack_push<W>(lds, spill, s.sp, ref[j], d[j]);
                ++s.sp;
            }
        }
        if (hits > 0) {
            s.cur = ref[0];
            s.cur_d = d[0];
            next = true;
        }
    }
    return leaf_or_pop<STATS, W, EXT>(b, r, s, lds, spill, next);
}

// One step of the walk of node width W (2: the reference's binary walk).
template <bool STATS, bool EXT, int W>
__device__ __forceinline__ bool trace_any_width_step(const BvhView& b, const TraceRay& r, TraceState& s, uint2* lds,
                                                     StackEntry* spill)
{
    if constexpr (W == 2) return trace_step<STATS, EXT>(b, r, s, lds, spill);
    else return trace_wide_step<STATS, EXT, W>(b, r, s, lds, spill);
}

// Two lanes per ray (the binary walk of plain trees only: no deep-tree links, no spheres).
// The two lanes of a pair (sub 0 and 1) hold the same ray, state and stack and make the
// same decisions; what they split is the work of a step: at an inner node lane sub tests
// child sub's box (sub 0 the left child, sub 1 the right), at a leaf the pair tests two
// triangles at once. Each lane reads only its half of the node record, and the pair swaps
// the results with one shuffle per value before the reference's decision (near child by
// centroid distance, cull, strictly-nearer hits in leaf order), so the walk and every hit
// are the one-lane walk's. Returns true when the ray is finished.
template <bool STATS>
__device__ __forceinline__ bool trace_step_pair(const BvhView& b, const TraceRay& r, TraceState& s, uint2* lds,
                                                StackEntry* spill, int sub)
{
    bool next = false;
    if (s.leaf_k < s.leaf_end) {
        // in the middle of a leaf
    } else if (s.best >= 0 && !(s.cur & PT_REF_INSIDE) && s.best_t < s.cur_d) {
        // culled: a nearer hit exists and the origin is outside this box
    } else if (pt_ref_cnt(s.cur) != 0) {
        enter_leaf(b, s);
    } else {
        // node record: (llo, lhi.x), (lhi.yz, rlo.xy), (rlo.z, rhi), (ref_l, ref_r)
        const float4* nd = b.nodes + 4 * (size_t)pt_ref_index(s.cur);
        const float4 q0 = nd[sub], q1 = nd[sub + 1];
        const uint32_t m = reinterpret_cast<const uint32_t*>(nd + 3)[sub];
        const V3 lo = sub ? v3(q0.z, q0.w, q1.x) : v3(q0.x, q0.y, q0.z);
        const V3 hi = sub ? v3(q1.y, q1.z, q1.w) : v3(q0.w, q1.x, q1.y);
        const float dc = sq_length(0.5f * (lo + hi) - r.o);
        float t;
        const bool h = slab(r.o, r.inv, r.sx, r.sy, r.sz, lo, hi, &t);
        const uint32_t e = m | (box_contains(lo, hi, r.o) ? PT_REF_INSIDE : 0u);
        const float dc_o = __shfl_xor(dc, 1, 64), t_o = __shfl_xor(t, 1, 64);
        const int h_o = __shfl_xor(h ? 1 : 0, 1, 64);
        const uint32_t e_o = (uint32_t)__shfl_xor((int)e, 1, 64);
        if (STATS) s.nv += 2;
        const float dl = sub ? dc_o : dc, dr = sub ? dc : dc_o;
        const bool hl = sub ? h_o != 0 : h, hr = sub ? h : h_o != 0;
        const float tl = sub ? t_o : t, tr_ = sub ? t : t_o;
        const uint32_t el = sub ? e_o : e, er = sub ? e : e_o;
        // the child the reference pops first: left if dl < dr, else right
        const bool left_first = dl < dr;
        const bool h1 = left_first ? hl : hr, h2 = left_first ? hr : hl;
        const uint32_t e1 = left_first ? el : er, e2 = left_first ? er : el;
        const float d1 = left_first ? tl : tr_, d2 = left_first ? tr_ : tl;
        if (h1) {
            if (h2) {
                stack_push<2>(lds, spill, s.sp, e2, d2);
                ++s.sp;
            }
            s.cur = e1;
            s.cur_d = d1;
            next = true;
        } else if (h2) {
            s.cur = e2;
            s.cur_d = d2;
            next = true;
        }
    }
    if (s.leaf_k < s.leaf_end) {
        // two triangles of the leaf: lane sub tests slot leaf_k + sub, then both lanes take
        // the two results in leaf order with the reference's strictly-nearer rule
        const uint32_t k = s.leaf_k + (uint32_t)sub;
        const bool have = k < s.leaf_end;
        bool hit = false;
        float t = 0.0f;
        if (have) {
            const float4 t0 = b.tris[3 * (size_t)k];
            const float4 t1 = b.tris[3 * (size_t)k + 1];
            const float4 t2 = b.tris[3 * (size_t)k + 2];
            TriDev tr;
            tr.p0 = v3(t0.x, t0.y, t0.z);
            tr.e1 = v3(t0.w, t1.x, t1.y);
            tr.e2 = v3(t1.z, t1.w, t2.x);
            tr.n = v3(t2.y, t2.z, t2.w);
            hit = tri_intersect(tr, r.o, r.d, &t) && !(t < r.dmin || t > r.dmax);
        }
        const float t_o = __shfl_xor(t, 1, 64);
        const int hit_o = __shfl_xor(hit ? 1 : 0, 1, 64);
        const bool have1 = s.leaf_k + 1 < s.leaf_end;
        if (STATS) s.nt += have1 ? 2 : 1;
        const bool hit0 = sub ? hit_o != 0 : hit, hit1 = sub ? hit : hit_o != 0;
        const float ta = sub ? t_o : t, tb = sub ? t : t_o;
        if (hit0 && (s.best < 0 || ta < s.best_t)) {
            s.best = (int)s.leaf_k;
            s.best_t = ta;
            if (r.any) return true;
        }
        if (have1 && hit1 && (s.best < 0 || tb < s.best_t)) {
            s.best = (int)s.leaf_k + 1;
            s.best_t = tb;
            if (r.any) return true;
        }
        s.leaf_k = have1 ? s.leaf_k + 2 : s.leaf_end;
        if (s.leaf_k < s.leaf_end) return false;
        next = false;
    }
    if (!next) {
        if (s.sp == 0) return true;
        --s.sp;
        stack_pop<2>(lds, spill, s.sp, &s.cur, &s.cur_d);
    }
    return false;
}

__device__ __forceinline__ void trace_finish(const BvhView& b, const TraceSegment& seg, int i, const TraceState& s)
{
    const int p = seg.path ? seg.path[i] : i;
    if (seg.hit_flag) {
        seg.out_k[p] = s.best >= 0 ? 1 : 0;
        return;
    }
    int k = s.best;
    if (seg.map_prim && k >= 0) k = b.prim_ids[k];
    if (seg.out_t) seg.out_t[p] = s.best >= 0 ? s.best_t : (seg.map_prim ? __builtin_inff() : s.best_t);
    seg.out_k[p] = k;
}

} // namespace pt
