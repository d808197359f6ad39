// pt_wide.hip — the 4- and 8-wide BVHs: collapse of the binary tree into 4- or 8-ary
// nodes with quantised child boxes, built on the device.
//
// 8-wide node record (128 B, one cache line; 8 x uint4):
//   w0  origin p (3 x f32) | exponents ex, ey, ez (biased by 127) and child count (4 x u8)
//   w1  near-plane bytes qlo.x[0..7], qlo.y[0..7]
//   w2  qlo.z[0..7], qhi.x[0..7]
//   w3  qhi.y[0..7], qhi.z[0..7]
//   w4  child references 0..3, w5 child references 4..7 (PT_REF_* encoding: an inner
//       child is a wide-node index, a leaf child is the binary tree's leaf reference)
//   w6, w7 unused
// 4-wide node record (64 B; 4 x uint4):
//   w0  as above
//   w1  qlo.x[0..3], qlo.y[0..3], qlo.z[0..3], qhi.x[0..3]
//   w2  qhi.y[0..3], qhi.z[0..3], child references 0, 1
//   w3  child references 2, 3, unused
// Child box i is [p + qlo_i * s, p + qhi_i * s] with s = 2^(e - 127) per axis, rounded
// outwards so that it contains the exact box.
//
// Collapse: every wide node starts from one inner node of the binary tree and keeps
// opening the inner candidate with the largest surface area until it holds 4 or 8
// children or only leaves (Ylitie, Karras and Laine, "Efficient Incoherent Ray Traversal on
// GPUs Through Compressed Wide BVHs", HPG 2017, collapse by greedy expansion). Wide
// nodes are numbered level by level; the leaves and their triangle order are the
// binary tree's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <vector>

#include "pt_kernels.h"
#include "pt_wide.h"

namespace pt {

namespace {

__device__ __forceinline__ float area6(const float* lo, const float* hi)
{
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Candidate of a wide node: child box and reference (inner references index the binary
// node records).
struct Cand {
    float lo[3], hi[3];
    uint32_t ref;
};

__device__ __forceinline__ void read_children(const float4* bnodes, uint32_t idx, Cand& l, Cand& r)
{
    const float4* nd = bnodes + 4 * (size_t)idx;
    const float4 a = nd[0], b = nd[1], c = nd[2], m = nd[3];
    l.lo[0] = a.x; l.lo[1] = a.y; l.lo[2] = a.z;
    l.hi[0] = a.w; l.hi[1] = b.x; l.hi[2] = b.y;
    r.lo[0] = b.z; r.lo[1] = b.w; r.lo[2] = c.x;
    r.hi[0] = c.y; r.hi[1] = c.z; r.hi[2] = c.w;
    l.ref = __float_as_uint(m.x);
    r.ref = __float_as_uint(m.y);
}

// Pass A: greedy expansion of every wide node of the level into at most W candidates.
template <int W>
__global__ void k_collapse_expand(const float4* bnodes, const uint32_t* in_ref, int n, float4* cand, int* n_cand,
                                  int* n_inner)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    Cand c[W];
    read_children(bnodes, pt_ref_index(in_ref[w]), c[0], c[1]);
    int cnt = 2;
    while (cnt < W) {
        int best = -1;
        float best_a = -1.0f;
        for (int i = 0; i < cnt; ++i) {
            if (pt_ref_cnt(c[i].ref) != 0) continue; // a leaf stays
            const float a = area6(c[i].lo, c[i].hi);
            if (a > best_a) {
                best_a = a;
                best = i;
            }
        }
        if (best < 0) break;
        Cand l, r;
        read_children(bnodes, pt_ref_index(c[best].ref), l, r);
        c[best] = l;
        c[cnt++] = r;
    }
    int inner = 0;
    for (int i = 0; i < cnt; ++i) {
        cand[16 * (size_t)w + 2 * i] = make_float4(c[i].lo[0], c[i].lo[1], c[i].lo[2], __uint_as_float(c[i].ref));
        cand[16 * (size_t)w + 2 * i + 1] = make_float4(c[i].hi[0], c[i].hi[1], c[i].hi[2], 0.0f);
        inner += pt_ref_cnt(c[i].ref) == 0;
    }
    n_cand[w] = cnt;
    n_inner[w] = inner;
}

__device__ __forceinline__ float pow2f(int e) { return __uint_as_float((uint32_t)(e + 127) << 23); }

// Exponent of the quantisation step along one axis: 255 steps from lo reach hi.
__device__ __forceinline__ int axis_exponent(float lo, float hi)
{
    const float ext = hi - lo;
    int e = -126;
    if (ext > 0.0f) {
        frexpf(ext * (1.0f / 255.0f), &e);
        if (e < -126) e = -126;
    }
    while (e < 127 && lo + 255.0f * pow2f(e) < hi) ++e;
    return e;
}

__device__ __forceinline__ uint32_t quant_lo(float p, float s, float v)
{
    float q = floorf((v - p) / s);
    q = fminf(fmaxf(q, 0.0f), 255.0f);
    while (q > 0.0f && p + q * s > v) q -= 1.0f;
    return (uint32_t)q;
}

__device__ __forceinline__ uint32_t quant_hi(float p, float s, float v)
{
    float q = ceilf((v - p) / s);
    q = fminf(fmaxf(q, 0.0f), 255.0f);
    while (q < 255.0f && p + q * s < v) q += 1.0f;
    return (uint32_t)q;
}

// Pass B: node records, indices of the inner children (the next level, in order) and
// the next level's input.
template <int W>
__global__ void k_collapse_write(const float4* cand, const int* n_cand, const int* inner_base, int n, int node_base,
                                 int next_base, uint4* out, uint32_t* next_ref)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const int cnt = n_cand[w];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = 0; i < cnt; ++i) {
        const float4 a = cand[16 * (size_t)w + 2 * i], b = cand[16 * (size_t)w + 2 * i + 1];
        lo[0] = fminf(lo[0], a.x); lo[1] = fminf(lo[1], a.y); lo[2] = fminf(lo[2], a.z);
        hi[0] = fmaxf(hi[0], b.x); hi[1] = fmaxf(hi[1], b.y); hi[2] = fmaxf(hi[2], b.z);
    }
    int e[3];
    float s[3];
    for (int k = 0; k < 3; ++k) {
        e[k] = axis_exponent(lo[k], hi[k]);
        s[k] = pow2f(e[k]);
    }
    uint32_t qlo[3][2] = {{0, 0}, {0, 0}, {0, 0}}, qhi[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    uint32_t refs[8];
    int next = inner_base[w];
    for (int i = 0; i < W; ++i) {
        uint32_t bl[3] = {255, 255, 255}, bh[3] = {0, 0, 0};
        uint32_t ref = PT_REF_EMPTY;
        if (i < cnt) {
            const float4 a = cand[16 * (size_t)w + 2 * i], b = cand[16 * (size_t)w + 2 * i + 1];
            const float cl[3] = {a.x, a.y, a.z}, ch[3] = {b.x, b.y, b.z};
            for (int k = 0; k < 3; ++k) {
                bl[k] = quant_lo(lo[k], s[k], cl[k]);
                bh[k] = quant_hi(lo[k], s[k], ch[k]);
            }
            ref = __float_as_uint(a.w);
            if (pt_ref_cnt(ref) == 0) { // inner: becomes wide node next_base + next
                next_ref[next] = ref;
                ref = pt_ref_inner((uint32_t)(next_base + next));
                ++next;
            }
        }
        for (int k = 0; k < 3; ++k) {
            qlo[k][i >> 2] |= bl[k] << (8 * (i & 3));
            qhi[k][i >> 2] |= bh[k] << (8 * (i & 3));
        }
        refs[i] = ref;
    }
    const uint4 w0 = make_uint4(__float_as_uint(lo[0]), __float_as_uint(lo[1]), __float_as_uint(lo[2]),
                                (uint32_t)(e[0] + 127) | ((uint32_t)(e[1] + 127) << 8) |
                                    ((uint32_t)(e[2] + 127) << 16) | ((uint32_t)cnt << 24));
    if (W == 8) {
        uint4* o = out + 8 * (size_t)(node_base + w);
        o[0] = w0;
        o[1] = make_uint4(qlo[0][0], qlo[0][1], qlo[1][0], qlo[1][1]);
        o[2] = make_uint4(qlo[2][0], qlo[2][1], qhi[0][0], qhi[0][1]);
        o[3] = make_uint4(qhi[1][0], qhi[1][1], qhi[2][0], qhi[2][1]);
        o[4] = make_uint4(refs[0], refs[1], refs[2], refs[3]);
        o[5] = make_uint4(refs[4], refs[5], refs[6], refs[7]);
        o[6] = make_uint4(0, 0, 0, 0);
        o[7] = make_uint4(0, 0, 0, 0);
    } else {
        uint4* o = out + 4 * (size_t)(node_base + w);
        o[0] = w0;
        o[1] = make_uint4(qlo[0][0], qlo[1][0], qlo[2][0], qhi[0][0]);
        o[2] = make_uint4(qhi[1][0], qhi[2][0], refs[0], refs[1]);
        o[3] = make_uint4(refs[2], refs[3], 0, 0);
    }
}

inline unsigned blocks(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

#define W_CHECK(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            for (void* p_ : tmp) (void)hipFree(p_);                                                    \
            return e_;                                                                                 \
        }                                                                                              \
    } while (0)

} // namespace

hipError_t gpu_collapse_wide(int width, const float4* bnodes_d, int num_inner, uint32_t root_ref, hipStream_t st,
                             Wide8* out)
{
    *out = Wide8();
    if (width != 4 && width != 8) return hipErrorInvalidValue;
    const int words = width == 8 ? 8 : 4; // uint4 per node record
    if (pt_ref_cnt(root_ref) != 0 || root_ref == PT_REF_EMPTY || num_inner <= 0) {
        // a leaf (or empty) root: the traversal starts at the leaf; one unused node keeps
        // the layout non-empty
        hipError_t e = hipMalloc((void**)&out->nodes_d, words * sizeof(uint4));
        if (e == hipSuccess) e = hipMemsetAsync(out->nodes_d, 0, words * sizeof(uint4), st);
        if (e != hipSuccess) return e;
        out->root = root_ref;
        return hipStreamSynchronize(st);
    }
    std::vector<void*> tmp;
    // every wide node consumes at least one binary inner node
    const size_t cap = (size_t)num_inner;
    float4* cand;
    int *n_cand, *n_inner, *base;
    uint32_t* refs[2];
    W_CHECK(hipMalloc((void**)&cand, 16 * cap * sizeof(float4)));
    tmp.push_back(cand);
    W_CHECK(hipMalloc((void**)&n_cand, (cap + 1) * sizeof(int)));
    tmp.push_back(n_cand);
    W_CHECK(hipMalloc((void**)&n_inner, (cap + 1) * sizeof(int)));
    tmp.push_back(n_inner);
    W_CHECK(hipMalloc((void**)&base, (cap + 1) * sizeof(int)));
    tmp.push_back(base);
    W_CHECK(hipMalloc((void**)&refs[0], cap * sizeof(uint32_t)));
    tmp.push_back(refs[0]);
    W_CHECK(hipMalloc((void**)&refs[1], cap * sizeof(uint32_t)));
    tmp.push_back(refs[1]);
    size_t scan_bytes = 0;
    W_CHECK(exclusive_sum(nullptr, scan_bytes, n_inner, base, (size_t)((int)cap + 1), st));
    void* scan_tmp;
    W_CHECK(hipMalloc(&scan_tmp, scan_bytes));
    tmp.push_back(scan_tmp);
    uint4* nodes;
    W_CHECK(hipMalloc((void**)&nodes, words * cap * sizeof(uint4)));
    tmp.push_back(nodes); // freed with the rest on failure
    W_CHECK(hipMemcpyAsync(refs[0], &root_ref, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    int level_n = 1, total = 0, depth = 0, cur = 0;
    while (level_n > 0) {
        if (width == 8)
            hipLaunchKernelGGL(k_collapse_expand<8>, dim3(blocks(level_n)), dim3(256), 0, st, bnodes_d, refs[cur],
                               level_n, cand, n_cand, n_inner);
        else
            hipLaunchKernelGGL(k_collapse_expand<4>, dim3(blocks(level_n)), dim3(256), 0, st, bnodes_d, refs[cur],
                               level_n, cand, n_cand, n_inner);
        const int zero = 0;
        W_CHECK(hipMemcpyAsync(n_inner + level_n, &zero, sizeof(int), hipMemcpyHostToDevice, st));
        W_CHECK(exclusive_sum(scan_tmp, scan_bytes, n_inner, base, (size_t)(level_n + 1), st));
        int next_n = 0;
        W_CHECK(hipMemcpyAsync(&next_n, base + level_n, sizeof(int), hipMemcpyDeviceToHost, st));
        if (width == 8)
            hipLaunchKernelGGL(k_collapse_write<8>, dim3(blocks(level_n)), dim3(256), 0, st, cand, n_cand, base,
                               level_n, total, total + level_n, nodes, refs[1 - cur]);
        else
            hipLaunchKernelGGL(k_collapse_write<4>, dim3(blocks(level_n)), dim3(256), 0, st, cand, n_cand, base,
                               level_n, total, total + level_n, nodes, refs[1 - cur]);
        W_CHECK(hipGetLastError());
        W_CHECK(hipStreamSynchronize(st));
        total += level_n;
        depth += 1;
        level_n = next_n;
        cur = 1 - cur;
    }
    tmp.pop_back(); // the nodes are the result
    for (void* p : tmp) (void)hipFree(p);
    out->nodes_d = nodes;
    out->num_nodes = total;
    out->depth = depth;
    out->root = pt_ref_inner(0);
    return hipSuccess;
}

} // namespace pt
