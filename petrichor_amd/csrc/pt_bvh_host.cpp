// pt_bvh_host.cpp — binned-SAH BVH construction on the host, in the reference's
// topology (BinnedSAHBVH::Build, Core/Accel/BinnedSAHBVH.cpp:15-225).
//
// Same node numbering (depth-first with an explicit stack; the right child is
// processed first), same bins, same SAH costs and same split. The reference sorts
// every node's primitive range by centroid before binning; bins are monotone in the
// centroid, so the child *sets* follow from the bins alone and only the order inside
// a leaf depends on the sort. By default a node's range is partitioned by bin and each
// leaf is sorted once by the axis of the last sort the reference applied to it (the
// parent's axis, or its own for a leaf chosen by the SAH), ties broken by triangle
// index. The leaf order then equals the reference's wherever the centroid keys are
// distinct. With reference_order, every node of more than leaf_max primitives is
// sorted as the reference sorts it (std::sort by the centroid on the widest axis, no
// tie-break, BinnedSAHBVH.cpp:104-107) and split at the SAH count: the leaf order is then
// the reference's also among equal keys, where it is whatever libstdc++'s introsort
// leaves (the same algorithm in every libstdc++ release the reference builds against).
#include "pt_bvh_host.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <numeric>
#include <stack>

namespace pt {

namespace {

struct Box {
    V3 lo = v3(FLT_MAX, FLT_MAX, FLT_MAX);
    V3 hi = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    void merge(V3 p) // AABB::Merge(point) (AABB.cpp:24-33)
    {
        lo.x = smin(lo.x, p.x);
        lo.y = smin(lo.y, p.y);
        lo.z = smin(lo.z, p.z);
        hi.x = smax(hi.x, p.x);
        hi.y = smax(hi.y, p.y);
        hi.z = smax(hi.z, p.z);
    }
    void merge(const Box& b) // AABB::Merge(AABB) (AABB.cpp:11-22)
    {
        lo.x = smin(lo.x, b.lo.x);
        lo.y = smin(lo.y, b.lo.y);
        lo.z = smin(lo.z, b.lo.z);
        hi.x = smax(hi.x, b.hi.x);
        hi.y = smax(hi.y, b.hi.y);
        hi.z = smax(hi.z, b.hi.z);
    }
    float area() const // AABB::GetSurfaceArea (AABB.h:76-81)
    {
        const V3 d = hi - lo;
        return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    int widest_axis() const // AABB::GetWidestAxis (AABB.cpp:35-52)
    {
        float w = 0.0f;
        int axis = 0;
        for (int a = 0; a < 3; ++a) {
            const float x = comp(hi, a) - comp(lo, a);
            if (x > w) {
                w = x;
                axis = a;
            }
        }
        return axis;
    }
};

} // namespace

bool build_bvh_host(const float* tri_pos, int num_tris, int num_bins, int leaf_max, HostBvh* out,
                    const float* spheres, int num_spheres, bool reference_order)
{
    const int n = num_tris + num_spheres;
    out->nodes.clear();
    out->order.resize(n);
    out->max_depth = 0;
    if (num_bins < 2 || num_bins > 256 || leaf_max < 1) return false;
    std::vector<Box> pbox(n);
    std::vector<V3> cen(n);
    for (int i = 0; i < num_tris; ++i) {
        const float* p = tri_pos + 9 * (size_t)i;
        const V3 p0 = v3(p[0], p[1], p[2]), p1 = v3(p[3], p[4], p[5]), p2 = v3(p[6], p[7], p[8]);
        Box b; // Triangle::CalcBoundary (Triangle.cpp:41-49)
        b.merge(p0);
        b.merge(p1);
        b.merge(p2);
        pbox[i] = b;
        cen[i] = (p0 + p1 + p2) / 3.0f; // Triangle::GetCentroid (Triangle.h:45-50)
    }
    for (int k = 0; k < num_spheres; ++k) {
        // Sphere::CalcBoundary (Sphere.cpp:24-30): origin -+ One * radius; GetCentroid: origin
        const float* q = spheres + 4 * (size_t)k;
        const V3 c = v3(q[0], q[1], q[2]), rr = v3(1.0f, 1.0f, 1.0f) * q[3];
        pbox[num_tris + k].lo = c - rr;
        pbox[num_tris + k].hi = c + rr;
        cen[num_tris + k] = c;
    }
    std::vector<int>& ids = out->order;
    std::iota(ids.begin(), ids.end(), 0);
    std::vector<int> bin(n, 0);
    std::vector<int> tmp;

    Box root;
    for (int i = 0; i < n; ++i) root.merge(pbox[i]);
    out->nodes.reserve(2 * (size_t)n + 1);
    HostNode r;
    r.lo = root.lo;
    r.hi = root.hi;
    r.begin = 0;
    r.end = n;
    out->nodes.push_back(r);
    std::vector<int> parent_axis(1, -1), depth(1, 1);

    auto sort_leaf = [&](int begin, int end, int axis) {
        if (axis < 0 || end - begin < 2) return;
        std::sort(ids.begin() + begin, ids.begin() + end, [&](int a, int b) {
            const float ca = comp(cen[a], axis), cb = comp(cen[b], axis);
            return ca < cb || (!(cb < ca) && a < b);
        });
    };

    std::vector<Box> bbox(num_bins), rbox(num_bins + 1);
    std::vector<int> bcnt(num_bins), rcnt(num_bins + 1);
    std::stack<int> st;
    st.push(0);
    while (!st.empty()) {
        const int ni = st.top();
        st.pop();
        const int begin = out->nodes[ni].begin, end = out->nodes[ni].end;
        const int cnt = end - begin;
        out->max_depth = std::max(out->max_depth, depth[ni]);
        if (cnt <= leaf_max) {
            out->nodes[ni].leaf = true;
            if (!reference_order) sort_leaf(begin, end, parent_axis[ni]);
            continue;
        }
        Box cb;
        for (int k = begin; k < end; ++k) cb.merge(cen[ids[k]]);
        const int axis = cb.widest_axis();
        const float len = comp(cb.hi, axis) - comp(cb.lo, axis);
        if (reference_order) // BinnedSAHBVH.cpp:104-107: bins are then contiguous runs of the range
            std::sort(ids.begin() + begin, ids.begin() + end,
                      [&](int a, int b) { return comp(cen[a], axis) < comp(cen[b], axis); });
        for (int b = 0; b < num_bins; ++b) {
            bbox[b] = Box();
            bcnt[b] = 0;
        }
        bool nan_bins = false;
        for (int k = begin; k < end; ++k) {
            const int id = ids[k];
            const float l = comp(cen[id], axis) - comp(cb.lo, axis);
            const float x = std::nextafter((float)num_bins * l / len, 0.0f); // BinnedSAHBVH.cpp:118-122
            if (x != x) { // len == 0: every bin position is NaN, the node becomes a leaf
                nan_bins = true;
                break;
            }
            // bins * l / len can round up to bins when bins is not a power of two (the
            // reference then indexes one past its bin array): clamped to the last bin
            const int b = std::min((int)x, num_bins - 1);
            bin[id] = b;
            bbox[b].merge(pbox[id]);
            bcnt[b]++;
        }
        int best_k = 0, best_left = 0;
        if (!nan_bins) {
            rbox[num_bins] = Box();
            rcnt[num_bins] = 0;
            for (int b = num_bins - 1; b >= 0; --b) {
                rbox[b] = rbox[b + 1];
                rbox[b].merge(bbox[b]);
                rcnt[b] = rcnt[b + 1] + bcnt[b];
            }
            float min_cost = FLT_MAX;
            Box lbox;
            int lcnt = 0;
            for (int k = 0; k < num_bins; ++k) { // GetSAHCost (BinnedSAHBVH.cpp:227-270)
                if (k > 0) {
                    lbox.merge(bbox[k - 1]);
                    lcnt += bcnt[k - 1];
                }
                const int rc = rcnt[k];
                const float sal = lcnt ? lbox.area() : 0;
                const float sar = rc ? rbox[k].area() : 0;
                const float cost = lcnt * sal + rc * sar;
                if (cost < min_cost) {
                    min_cost = cost;
                    best_k = k;
                    best_left = lcnt;
                }
            }
        }
        if (best_k == 0) {
            out->nodes[ni].leaf = true;
            if (!reference_order) sort_leaf(begin, end, axis);
            continue;
        }
        Box lb, rb;
        if (reference_order) { // sorted: the first best_left primitives are the left bins
            for (int k = begin; k < end; ++k) {
                const int id = ids[k];
                if (bin[id] < best_k) lb.merge(pbox[id]);
                else rb.merge(pbox[id]);
            }
        } else {
            tmp.clear();
            int w = begin;
            for (int k = begin; k < end; ++k) {
                const int id = ids[k];
                if (bin[id] < best_k) {
                    lb.merge(pbox[id]);
                    ids[w++] = id;
                } else {
                    rb.merge(pbox[id]);
                    tmp.push_back(id);
                }
            }
            std::copy(tmp.begin(), tmp.end(), ids.begin() + w);
        }
        HostNode left, right;
        left.lo = lb.lo;
        left.hi = lb.hi;
        left.begin = begin;
        left.end = begin + best_left;
        right.lo = rb.lo;
        right.hi = rb.hi;
        right.begin = begin + best_left;
        right.end = end;
        const int li = (int)out->nodes.size();
        out->nodes.push_back(left);
        const int ri = (int)out->nodes.size();
        out->nodes.push_back(right);
        out->nodes[ni].child[0] = li;
        out->nodes[ni].child[1] = ri;
        parent_axis.push_back(axis);
        parent_axis.push_back(axis);
        depth.push_back(depth[ni] + 1);
        depth.push_back(depth[ni] + 1);
        st.push(li);
        st.push(ri);
    }
    return true;
}

// Packs the tree for the traversal kernel. Inner nodes get indices in depth-first
// preorder (a node, then its left subtree, then its right subtree), so the first child
// usually follows its parent in memory. Each inner node is 4 float4: (lo_l, hi_l.x),
// (hi_l.yz, lo_r.xy), (lo_r.z, hi_r), (ref_l, ref_r, 0, 0) with child references
// encoded as in PT_REF_* (pt_device.h); leaves of more than PT_REF_MAX_SMALL_LEAF
// triangles go through the big-leaf table (first slot, count).
void pack_bvh2(const HostBvh& h, std::vector<float4>& nodes, std::vector<int2>& big_leaves, V3* root_lo,
               V3* root_hi, uint32_t* root_ref, std::vector<int2>* inner_range)
{
    nodes.clear();
    big_leaves.clear();
    const int nn = (int)h.nodes.size();
    std::vector<int> gidx(nn, -1);
    int inner = 0;
    std::vector<int> st;
    st.push_back(0);
    while (!st.empty()) { // preorder, left first
        const int i = st.back();
        st.pop_back();
        if (h.nodes[i].leaf) continue;
        gidx[i] = inner++;
        st.push_back(h.nodes[i].child[1]);
        st.push_back(h.nodes[i].child[0]);
    }
    nodes.resize(4 * (size_t)inner);
    if (inner_range) {
        inner_range->assign((size_t)inner, make_int2(0, 0));
        for (int i = 0; i < nn; ++i)
            if (gidx[i] >= 0) (*inner_range)[gidx[i]] = make_int2(h.nodes[i].begin, h.nodes[i].end);
    }
    // big leaves are numbered in leaf-slot order
    for (const HostNode& x : h.nodes)
        if (x.leaf && x.end - x.begin > PT_REF_MAX_SMALL_LEAF) big_leaves.push_back(make_int2(x.begin, x.end - x.begin));
    std::sort(big_leaves.begin(), big_leaves.end(), [](int2 a, int2 b) { return a.x < b.x; });
    auto ref_of = [&](int c) -> uint32_t {
        const HostNode& x = h.nodes[c];
        if (!x.leaf) return pt_ref_inner((uint32_t)gidx[c]);
        const int cnt = x.end - x.begin;
        if (cnt <= PT_REF_MAX_SMALL_LEAF) return pt_ref_leaf((uint32_t)x.begin, (uint32_t)cnt);
        const auto it = std::lower_bound(big_leaves.begin(), big_leaves.end(), x.begin,
                                         [](int2 a, int b) { return a.x < b; });
        return pt_ref_big_leaf((uint32_t)(it - big_leaves.begin()));
    };
    for (int i = 0; i < nn; ++i) {
        const HostNode& x = h.nodes[i];
        if (x.leaf) continue;
        const HostNode& l = h.nodes[x.child[0]];
        const HostNode& r = h.nodes[x.child[1]];
        float4* o = &nodes[4 * (size_t)gidx[i]];
        o[0] = make_float4(l.lo.x, l.lo.y, l.lo.z, l.hi.x);
        o[1] = make_float4(l.hi.y, l.hi.z, r.lo.x, r.lo.y);
        o[2] = make_float4(r.lo.z, r.hi.x, r.hi.y, r.hi.z);
        uint4 m = make_uint4(ref_of(x.child[0]), ref_of(x.child[1]), 0u, 0u);
        o[3] = *reinterpret_cast<float4*>(&m);
    }
    *root_lo = h.nodes[0].lo;
    *root_hi = h.nodes[0].hi;
    *root_ref = (h.nodes[0].end - h.nodes[0].begin == 0) ? PT_REF_EMPTY : ref_of(0);
}

} // namespace pt
