// pt_capi.cpp — the C-ABI of include/petrichor_hip.h: contexts, scene upload, BVH
// build, trace entry points and the plain per-frame wavefront render loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/petrichor_hip.h"
#include "pt_build.h"
#include "pt_internal.h"
#include "pt_wide.h"

using namespace pt;

int64_t pt_scene_map_bytes(const pt_scene_desc* d)
{
    if (!(d->material_maps || d->material_normal_maps) || d->num_textures <= 0 || !d->texture_sizes) return 0;
    int64_t b = (int64_t)d->num_materials * (int64_t)sizeof(int2) + (int64_t)d->num_textures * (int64_t)sizeof(TexDev);
    if (d->material_normal_maps) b += (int64_t)d->num_materials * 4;
    if (d->material_map_strengths) b += (int64_t)d->num_materials * 8;
    for (int t = 0; t < d->num_textures; ++t)
        b += 12 * (int64_t)std::max(d->texture_sizes[2 * t], 0) * (int64_t)std::max(d->texture_sizes[2 * t + 1], 0);
    return b;
}

int pt_fail(pt_context* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

namespace {

int fail(pt_context* ctx, int code, const std::string& msg) { return pt_fail(ctx, code, msg); }

template <typename T>
hipError_t upload(T** dst, const T* src, size_t count, hipStream_t st)
{
    *dst = nullptr;
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc((void**)dst, count * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, st);
}

void free_all(std::initializer_list<void*> ptrs)
{
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

MaterialDev material_from(int type, const float* p)
{
    MaterialDev m;
    m.type = type;
    m.color = v3(p[0], p[1], p[2]);
    m.alpha = type == kMatGlossy ? p[3] * p[3] : 0.0f; // GGX::GGX (GGX.cpp:16-20)
    m.ior = type == kMatGlass ? p[3] : 1.333f;
    m.mix = type == kMatMix ? p[3] : 0.0f;
    m.mat0 = type == kMatMix ? (int)p[4] : -1;
    m.mat1 = type == kMatMix ? (int)p[5] : -1;
    m.is_enabled = p[6] != 0.0f ? 1 : 0;
    m.albedo = v3(1.0f, 1.0f, 1.0f);
    return m;
}

} // namespace

// ---- device-built trees ---------------------------------------------------------------

namespace {

// Trees deeper than the traversal stack (PT_MAX_STACK - 2 levels) are walked without one,
// by parent links (pt_traverse.h). The links are made here on the host from the packed
// nodes and checked to form one tree over every inner node, rooted at the root, so the
// walk up always ends; a tree that fails the check is reported as an internal error.
int add_parents(pt_bvh* b)
{
    if (b->max_depth <= PT_MAX_STACK - 2 || b->num_inner == 0) return PT_OK;
    pt_context* ctx = b->scene->ctx;
    const size_t ni = (size_t)b->num_inner;
    const uint32_t root = b->view.root_ref;
    std::vector<float4> nodes(4 * ni);
    PT_HIP(ctx, hipMemcpyAsync(nodes.data(), b->nodes_d, nodes.size() * sizeof(float4), hipMemcpyDeviceToHost,
                               ctx->stream));
    PT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (pt_ref_cnt(root) != 0 || pt_ref_index(root) >= ni) return fail(ctx, PT_ERR_STATE, "deep tree: bad root");
    std::vector<int32_t> parent(ni, -2);
    parent[pt_ref_index(root)] = -1;
    // breadth first from the root: every inner node is reached exactly once
    std::vector<uint32_t> level(1, pt_ref_index(root)), nextl;
    size_t reached = 1;
    while (!level.empty()) {
        nextl.clear();
        for (const uint32_t i : level) {
            const float4 m = nodes[4 * (size_t)i + 3];
            for (const float f : {m.x, m.y}) {
                uint32_t ref;
                std::memcpy(&ref, &f, 4);
                if (pt_ref_cnt(ref) != 0) continue; // a leaf
                const uint32_t c = pt_ref_index(ref);
                if (c >= ni || parent[c] != -2) return fail(ctx, PT_ERR_STATE, "deep tree: bad child link");
                parent[c] = (int32_t)i;
                nextl.push_back(c);
                ++reached;
            }
        }
        level.swap(nextl);
    }
    if (reached != ni) return fail(ctx, PT_ERR_STATE, "deep tree: unreachable inner nodes");
    PT_HIP(ctx, upload(&b->parents_d, parent.data(), ni, ctx->stream));
    PT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->view.parents = b->parents_d;
    b->device_bytes += (int64_t)ni * 4;
    return PT_OK;
}

// Sets BvhView::boxes_ordered from the packed binary nodes.
static int check_boxes(pt_bvh* b)
{
    pt_context* ctx = b->scene->ctx;
    bool ordered = false;
    PT_HIP(ctx, gpu_boxes_ordered(b->nodes_d, b->num_inner, ctx->stream, &ordered));
    b->view.boxes_ordered = ordered ? 1 : 0;
    return PT_OK;
}

int build_device(pt_scene* s, int bins, int leaf_max, pt_bvh** out)
{
    pt_context* ctx = s->ctx;
    PT_HIP(ctx, hipSetDevice(ctx->device));
    const auto t0 = std::chrono::steady_clock::now();
    GpuBvhResult r;
    hipError_t e = gpu_build_bvh(s->tri_pos_d, s->num_tris, s->spheres_d, s->num_spheres, bins, leaf_max, ctx->stream, &r);
    if (e != hipSuccess) {
        gpu_bvh_free(&r);
        return fail(ctx, e == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP,
                    std::string("device BVH build: ") + hipGetErrorString(e));
    }
    GpuPacked pk;
    e = gpu_pack_bvh(r, s->num_tris, s->tri_pos_d, s->spheres_d, s->num_spheres, s->tri_mat_d, s->mat_shading_of_d,
                     s->tri_smooth_d, s->tri_nrm_d, ctx->stream, &pk);
    const int num_nodes = r.num_nodes, max_depth = r.max_depth;
    gpu_bvh_free(&r);
    pt_bvh* b = new pt_bvh();
    b->scene = s;
    if (e == hipSuccess) e = hipMalloc((void**)&b->cursor_d, (PT_CURSOR_INTS + 1) * sizeof(int));
    if (e != hipSuccess) {
        gpu_packed_free(&pk);
        if (b->cursor_d) (void)hipFree(b->cursor_d);
        delete b;
        return fail(ctx, e == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP,
                    std::string("device BVH pack: ") + hipGetErrorString(e));
    }
    b->nodes_d = pk.nodes_d;
    b->big_leaves_d = pk.big_leaves_d;
    b->inner_range_d = pk.inner_range_d;
    b->tris_d = pk.tris_d;
    b->prim_ids_d = pk.prim_ids_d;
    b->tri_mat_d = pk.mat_d;
    b->tri_mat_shading_d = pk.mat_sh_d;
    b->tri_smooth_d = pk.smooth_d;
    b->tri_nrm_d = pk.nrm_d;
    b->kinds_d = pk.kinds_d;
    b->view.kinds = b->kinds_d;
    b->view.nodes = b->nodes_d;
    b->view.tris = b->tris_d;
    b->view.prim_ids = b->prim_ids_d;
    b->view.big_leaves = b->big_leaves_d;
    b->view.inner_range = b->inner_range_d;
    b->view.root_lo = pk.root_lo;
    b->view.root_hi = pk.root_hi;
    b->view.root_ref = pk.root_ref;
    b->num_nodes = num_nodes;
    b->max_depth = max_depth;
    b->num_inner = pk.num_inner;
    b->num_big = pk.num_big;
    b->host_ready = false;
    b->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const int64_t nt = s->num_prims();
    b->device_bytes = (int64_t)pk.num_inner * 72 + (int64_t)pk.num_big * 8 + nt * 60 +
                      (pk.smooth_d ? nt * 4 : 0) + (pk.nrm_d ? nt * 36 : 0) + (pk.kinds_d ? nt : 0);
    int rc = check_boxes(b);
    if (rc == PT_OK) rc = add_parents(b);
    if (rc != PT_OK) {
        pt_bvh_destroy(b);
        return rc;
    }
    *out = b;
    return PT_OK;
}

// Adds the wide nodes (width 4 or 8) to a built binary tree. On failure the tree is
// destroyed and *out cleared.
int make_wide(pt_bvh* b, int width, pt_bvh** out)
{
    pt_context* ctx = b->scene->ctx;
    const auto t0 = std::chrono::steady_clock::now();
    Wide8 w;
    hipError_t e = gpu_collapse_wide(width, b->nodes_d, b->num_inner, b->view.root_ref, ctx->stream, &w);
    // a wide level leaves at most width - 1 children pending on the stack
    const int stack = width == 8 ? PT_STACK8 : PT_STACK4;
    if (e == hipSuccess && (width - 1) * w.depth + 1 > stack) {
        if (w.nodes_d) (void)hipFree(w.nodes_d);
        pt_bvh_destroy(b);
        *out = nullptr;
        return fail(ctx, PT_ERR_UNSUPPORTED, "wide tree too deep for the traversal stack");
    }
    if (e != hipSuccess) {
        pt_bvh_destroy(b);
        *out = nullptr;
        return fail(ctx, e == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP,
                    std::string("wide collapse: ") + hipGetErrorString(e));
    }
    b->nodes8_d = w.nodes_d;
    b->num_nodes8 = w.num_nodes;
    b->depth8 = w.depth;
    b->width = width;
    b->view.nodes8 = w.nodes_d;
    b->view.root8 = w.root;
    b->view.width = width;
    b->device_bytes += (int64_t)w.num_nodes * wide_node_bytes(width);
    b->build_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = b;
    return PT_OK;
}

// Rebuilds the node table in the reference's numbering from the packed tree: boxes come
// from the parent's record, leaf ranges from the child references, inner ranges from the
// leftmost and rightmost leaves. Node ids follow BinnedSAHBVH::Build: the two children
// of a split get the next two ids when the parent leaves the stack, the right child is
// taken from the stack first.
int ensure_host(pt_bvh* b)
{
    if (b->host_ready) return PT_OK;
    pt_context* ctx = b->scene->ctx;
    const int nt = b->scene->num_prims();
    std::vector<float4> nodes(4 * (size_t)b->num_inner);
    std::vector<int2> big((size_t)b->num_big);
    HostBvh& h = b->host;
    h.order.resize(nt);
    PT_HIP(ctx, hipMemcpy(nodes.data(), b->nodes_d, nodes.size() * sizeof(float4), hipMemcpyDeviceToHost));
    if (!big.empty()) PT_HIP(ctx, hipMemcpy(big.data(), b->big_leaves_d, big.size() * sizeof(int2), hipMemcpyDeviceToHost));
    PT_HIP(ctx, hipMemcpy(h.order.data(), b->prim_ids_d, (size_t)nt * sizeof(int32_t), hipMemcpyDeviceToHost));
    auto ref_of = [&](int i, int side) {
        uint32_t r;
        std::memcpy(&r, (const char*)&nodes[4 * (size_t)i + 3] + 4 * side, 4);
        return r;
    };
    std::vector<int> beg(b->num_inner), end(b->num_inner);
    auto leaf_range = [&](uint32_t r, int* lo, int* hi) {
        const uint32_t c = pt_ref_cnt(r), idx = pt_ref_index(r);
        if (c == PT_REF_BIG) {
            *lo = big[idx].x;
            *hi = big[idx].x + big[idx].y;
        } else {
            *lo = (int)idx;
            *hi = (int)(idx + c);
        }
    };
    for (int i = b->num_inner - 1; i >= 0; --i) { // children follow their parent in preorder
        int lo, hi;
        const uint32_t rl = ref_of(i, 0), rr = ref_of(i, 1);
        if (pt_ref_cnt(rl) == 0) beg[i] = beg[pt_ref_index(rl)];
        else leaf_range(rl, &beg[i], &hi);
        if (pt_ref_cnt(rr) == 0) end[i] = end[pt_ref_index(rr)];
        else leaf_range(rr, &lo, &end[i]);
    }
    struct Item {
        int id;
        uint32_t ref;
        V3 lo, hi;
    };
    h.nodes.assign(b->num_nodes, HostNode());
    int next = 1;
    std::vector<Item> st;
    st.push_back({0, b->view.root_ref, b->view.root_lo, b->view.root_hi});
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        HostNode& x = h.nodes[it.id];
        x.lo = it.lo;
        x.hi = it.hi;
        if (pt_ref_cnt(it.ref) != 0) {
            x.leaf = true;
            leaf_range(it.ref, &x.begin, &x.end);
            continue;
        }
        const int i = (int)pt_ref_index(it.ref);
        const float4* o = &nodes[4 * (size_t)i];
        x.begin = beg[i];
        x.end = end[i];
        x.child[0] = next;
        x.child[1] = next + 1;
        next += 2;
        st.push_back({x.child[0], ref_of(i, 0), v3(o[0].x, o[0].y, o[0].z), v3(o[0].w, o[1].x, o[1].y)});
        st.push_back({x.child[1], ref_of(i, 1), v3(o[1].z, o[1].w, o[2].x), v3(o[2].y, o[2].z, o[2].w)});
    }
    h.max_depth = b->max_depth;
    b->host_ready = true;
    return PT_OK;
}

} // namespace

extern "C" {

int pt_ctx_create(int device, pt_context** out)
{
    if (!out) return PT_ERR_ARG;
    *out = nullptr;
    pt_context* c = new pt_context();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    // the device address of the allocation-free trace scratch, once per device
    if (e == hipSuccess) e = noalloc_scratch_init(device);
    if (e != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return PT_ERR_HIP;
    }
    *out = c;
    return PT_OK;
}

int pt_ctx_destroy(pt_context* ctx)
{
    if (!ctx) return PT_ERR_ARG;
    if (ctx->engine) pt_engine_destroy(ctx->engine);
    for (pt_engine* e : ctx->engine2)
        if (e) pt_engine_destroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PT_OK;
}

const char* pt_last_error(pt_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pt_scene_create(pt_context* ctx, const pt_scene_desc* d, pt_scene** out)
{
    if (!ctx || !d || !out) return fail(ctx, PT_ERR_ARG, "null argument");
    *out = nullptr;
    if (d->num_triangles < 0 || (d->num_triangles > 0 && (!d->positions || !d->material_ids)))
        return fail(ctx, PT_ERR_ARG, "bad triangle arrays");
    if (d->num_materials <= 0 || !d->material_types || !d->material_params)
        return fail(ctx, PT_ERR_ARG, "scene needs at least one material");
    if (d->num_lights < 0 || (d->num_lights > 0 && !d->light_triangles))
        return fail(ctx, PT_ERR_ARG, "bad light list");
    for (int i = 0; i < d->num_triangles; ++i)
        if (d->material_ids[i] < 0 || d->material_ids[i] >= d->num_materials)
            return fail(ctx, PT_ERR_ARG, "material id out of range at triangle " + std::to_string(i));
    if (d->num_spheres < 0 || (d->num_spheres > 0 && (!d->spheres || !d->sphere_material_ids)))
        return fail(ctx, PT_ERR_ARG, "bad sphere arrays");
    if ((int64_t)d->num_triangles + d->num_spheres > ((int64_t)1 << 27))
        return fail(ctx, PT_ERR_ARG, "too many primitives");
    for (int i = 0; i < d->num_spheres; ++i) {
        if (d->sphere_material_ids[i] < 0 || d->sphere_material_ids[i] >= d->num_materials)
            return fail(ctx, PT_ERR_ARG, "material id out of range at sphere " + std::to_string(i));
        for (int q = 0; q < 4; ++q)
            if (!std::isfinite(d->spheres[4 * (size_t)i + q])) return fail(ctx, PT_ERR_ARG, "sphere not finite");
    }
    for (int i = 0; i < d->num_lights; ++i)
        if (d->light_triangles[i] < 0 || d->light_triangles[i] >= d->num_triangles + d->num_spheres)
            return fail(ctx, PT_ERR_ARG, "light primitive out of range");
    for (int m = 0; m < d->num_materials; ++m) {
        const int t = d->material_types[m];
        if (t < 0 || t > PT_MAT_MIX) return fail(ctx, PT_ERR_UNSUPPORTED, "unknown material type");
        if (t == PT_MAT_MIX) {
            const int a = (int)d->material_params[8 * m + 4], b = (int)d->material_params[8 * m + 5];
            if (a < 0 || a >= d->num_materials || b < 0 || b >= d->num_materials ||
                d->material_types[a] == PT_MAT_MIX || d->material_types[b] == PT_MAT_MIX)
                return fail(ctx, PT_ERR_UNSUPPORTED, "mix material must reference two non-mix materials");
        }
    }
    if (d->env_rgb && (d->env_width <= 0 || d->env_height <= 0)) return fail(ctx, PT_ERR_ARG, "bad env size");
    if (d->material_maps || d->material_normal_maps) {
        if (d->num_textures < 0 || (d->num_textures > 0 && (!d->texture_sizes || !d->textures)))
            return fail(ctx, PT_ERR_ARG, "bad material maps");
        for (int t = 0; t < d->num_textures; ++t)
            if (d->texture_sizes[2 * t] <= 0 || d->texture_sizes[2 * t + 1] <= 0 || !d->textures[t])
                return fail(ctx, PT_ERR_ARG, "bad material map texture");
        for (int i = 0; d->material_maps && i < 2 * d->num_materials; ++i)
            if (d->material_maps[i] < -1 || d->material_maps[i] >= d->num_textures)
                return fail(ctx, PT_ERR_ARG, "material map index out of range");
        for (int i = 0; d->material_normal_maps && i < d->num_materials; ++i)
            if (d->material_normal_maps[i] < -1 || d->material_normal_maps[i] >= d->num_textures)
                return fail(ctx, PT_ERR_ARG, "normal map index out of range");
    }
    for (int i = 0; d->material_map_strengths && i < d->num_materials; ++i) {
        const float ns = d->material_map_strengths[2 * i], rs = d->material_map_strengths[2 * i + 1];
        if (!(ns >= 0.0f && ns <= 1.0f) || !std::isfinite(rs))
            return fail(ctx, PT_ERR_ARG, "normal map strength outside [0, 1] or roughness map strength not finite");
    }
    PT_HIP(ctx, hipSetDevice(ctx->device));

    pt_scene* s = new pt_scene();
    s->ctx = ctx;
    s->num_tris = d->num_triangles;
    s->num_spheres = d->num_spheres;
    const size_t nt = (size_t)d->num_triangles, ns = (size_t)d->num_spheres;
    s->tri_pos.assign(d->positions, d->positions + 9 * nt);
    // per-primitive arrays: the triangles' entries, then zeros (or the material) per sphere
    if (d->normals) {
        s->tri_nrm.assign(d->normals, d->normals + 9 * nt);
        s->tri_nrm.resize(9 * (nt + ns), 0.0f);
    }
    s->tri_mat.assign(d->material_ids, d->material_ids + nt);
    if (ns) s->tri_mat.insert(s->tri_mat.end(), d->sphere_material_ids, d->sphere_material_ids + ns);
    if (d->smooth) {
        s->tri_smooth.assign(d->smooth, d->smooth + nt);
        s->tri_smooth.resize(nt + ns, 0);
    }
    if (d->uvs) {
        s->tri_uv.assign(d->uvs, d->uvs + 6 * nt);
        s->tri_uv.resize(6 * (nt + ns), 0.0f);
    }
    for (size_t i = 0; i < ns; ++i) {
        const float* q = d->spheres + 4 * i;
        s->spheres.push_back(make_float4(q[0], q[1], q[2], q[3]));
    }
    for (int m = 0; m < d->num_materials; ++m)
        s->mats.push_back(material_from(d->material_types[m], d->material_params + 8 * m));
    // lights: a triangle record, or a sphere's (centre, radius) and zeros
    std::vector<float4> lights;
    std::vector<int8_t> light_kinds;
    bool sphere_lights = false;
    for (int i = 0; i < d->num_lights; ++i) {
        const int id = d->light_triangles[i];
        if (id >= d->num_triangles) {
            lights.push_back(s->spheres[(size_t)(id - d->num_triangles)]);
            lights.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            lights.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            light_kinds.push_back(1);
            sphere_lights = true;
            continue;
        }
        const float* p = d->positions + 9 * (size_t)id;
        const TriDev t = tri_from_points(v3(p[0], p[1], p[2]), v3(p[3], p[4], p[5]), v3(p[6], p[7], p[8]));
        lights.push_back(make_float4(t.p0.x, t.p0.y, t.p0.z, t.e1.x));
        lights.push_back(make_float4(t.e1.y, t.e1.z, t.e2.x, t.e2.y));
        lights.push_back(make_float4(t.e2.z, t.n.x, t.n.y, t.n.z));
        light_kinds.push_back(0);
    }
    s->num_lights = d->num_lights;
    const float* c = d->camera;
    s->cam.pos = v3(c[0], c[1], c[2]);
    s->cam.right = v3(c[3], c[4], c[5]);
    s->cam.up = v3(c[6], c[7], c[8]);
    s->cam.forward = v3(c[9], c[10], c[11]);
    s->cam.focus = c[12];
    s->cam.h_perf = c[13];
    s->cam.aperture = c[14];

    EnvDev& e = s->env;
    e.rgb = nullptr;
    e.w = e.h = 0;
    e.base = v3(d->env_base_color[0], d->env_base_color[1], d->env_base_color[2]);
    e.rot = d->env_z_rotation;
    e.use_is = d->env_importance_sampling;
    hipError_t err = upload(&s->mats_d, s->mats.data(), s->mats.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&s->lights_d, lights.data(), lights.size(), ctx->stream);
    if (err == hipSuccess && sphere_lights)
        err = upload(&s->light_kinds_d, light_kinds.data(), light_kinds.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&s->spheres_d, s->spheres.data(), s->spheres.size(), ctx->stream);
    if (err == hipSuccess && d->env_rgb) {
        e.w = d->env_width;
        e.h = d->env_height;
        err = upload(&s->env_rgb_d, d->env_rgb, 3 * (size_t)e.w * e.h, ctx->stream);
        e.rgb = s->env_rgb_d;
    }
    std::vector<int32_t> mat_sh(s->mats.size());
    for (size_t m = 0; m < s->mats.size(); ++m) { // MixMaterial::GetMaterial(0.0f)
        const MaterialDev& md = s->mats[m];
        mat_sh[m] = md.type == kMatMix ? (md.mix <= 0.0f ? md.mat0 : md.mat1) : (int32_t)m;
    }
    if (err == hipSuccess) err = upload(&s->mat_shading_of_d, mat_sh.data(), mat_sh.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&s->tri_pos_d, s->tri_pos.data(), s->tri_pos.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&s->tri_mat_d, s->tri_mat.data(), s->tri_mat.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&s->tri_smooth_d, s->tri_smooth.data(), s->tri_smooth.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&s->tri_nrm_d, s->tri_nrm.data(), s->tri_nrm.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&s->tri_uv_d, s->tri_uv.data(), s->tri_uv.size(), ctx->stream);
    if (err == hipSuccess && (d->material_maps || d->material_normal_maps)) {
        // material maps: the textures, their table and two map indices per material, the
        // normal map of each material and the map strengths
        std::vector<TexDev> tex;
        for (int t = 0; t < d->num_textures && err == hipSuccess; ++t) {
            const int tw = d->texture_sizes[2 * t], th = d->texture_sizes[2 * t + 1];
            float* p = nullptr;
            err = upload(&p, d->textures[t], 3 * (size_t)tw * th, ctx->stream);
            s->tex_d.push_back(p);
            tex.push_back(TexDev{p, tw, th});
        }
        std::vector<int2> maps(s->mats.size(), make_int2(-1, -1));
        for (size_t m = 0; d->material_maps && m < maps.size(); ++m)
            maps[m] = make_int2(d->material_maps[2 * m], d->material_maps[2 * m + 1]);
        if (err == hipSuccess) err = upload(&s->textures_d, tex.data(), tex.size(), ctx->stream);
        if (err == hipSuccess) err = upload(&s->mat_maps_d, maps.data(), maps.size(), ctx->stream);
        if (err == hipSuccess && d->material_normal_maps)
            err = upload(&s->mat_nmaps_d, d->material_normal_maps, s->mats.size(), ctx->stream);
    }
    if (err == hipSuccess && d->material_map_strengths)
        err = upload(&s->map_strength_d, (const float2*)d->material_map_strengths, s->mats.size(), ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    s->device_bytes = (int64_t)(s->mats.size() * (sizeof(MaterialDev) + 4) + lights.size() * sizeof(float4)) +
                      (d->env_rgb ? 3 * (int64_t)e.w * e.h * 4 : 0) +
                      (int64_t)nt * 40 + (int64_t)ns * 20 + (int64_t)(s->tri_smooth.size() * 4 + s->tri_nrm.size() * 4) +
                      (int64_t)s->tri_uv.size() * 4 + pt_scene_map_bytes(d);
    if (err != hipSuccess) {
        for (float* p : s->tex_d) (void)hipFree(p);
        free_all({s->mats_d, s->lights_d, s->env_rgb_d, s->mat_shading_of_d, s->tri_pos_d,
                  s->tri_mat_d, s->tri_smooth_d, s->tri_nrm_d, s->tri_uv_d, s->textures_d, s->mat_maps_d,
                  s->mat_nmaps_d, s->map_strength_d, s->spheres_d, s->light_kinds_d});
        delete s;
        return fail(ctx, err == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP,
                    std::string("scene upload: ") + hipGetErrorString(err));
    }
    *out = s;
    return PT_OK;
}

int pt_scene_destroy(pt_scene* s)
{
    if (!s) return PT_ERR_ARG;
    for (float* p : s->tex_d) (void)hipFree(p);
    free_all({s->mats_d, s->lights_d, s->env_rgb_d, s->mat_shading_of_d, s->tri_pos_d,
              s->tri_mat_d, s->tri_smooth_d, s->tri_nrm_d, s->tri_uv_d, s->textures_d, s->mat_maps_d,
              s->mat_nmaps_d, s->map_strength_d, s->spheres_d, s->light_kinds_d});
    delete s;
    return PT_OK;
}

static int bvh_build_impl(pt_scene* s, const pt_bvh_desc* desc, pt_bvh** out);

int pt_bvh_build(pt_scene* s, const pt_bvh_desc* desc, pt_bvh** out)
{
    const int rc = bvh_build_impl(s, desc, out);
    if (rc != PT_OK || !s->tri_uv_d || s->num_prims() == 0) return rc;
    // texture coordinates per leaf slot, gathered through the leaf order
    pt_bvh* b = *out;
    pt_context* ctx = s->ctx;
    hipError_t e = hipMalloc((void**)&b->tri_uv_d, 6 * sizeof(float) * (size_t)s->num_prims());
    if (e == hipSuccess) e = launch_gather_rows(s->tri_uv_d, b->prim_ids_d, s->num_prims(), 6, b->tri_uv_d, ctx->stream);
    if (e != hipSuccess) {
        pt_bvh_destroy(b);
        *out = nullptr;
        return fail(ctx, e == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP, hipGetErrorString(e));
    }
    b->device_bytes += 24 * (int64_t)s->num_prims();
    return PT_OK;
}

static int bvh_build_impl(pt_scene* s, const pt_bvh_desc* desc, pt_bvh** out)
{
    if (!s || !out) return PT_ERR_ARG;
    pt_context* ctx = s->ctx;
    *out = nullptr;
    const int bins = desc ? desc->num_bins : 16;
    const int leaf_max = desc ? desc->leaf_max : 4;
    const int width = desc ? desc->width : 2;
    const int builder = desc ? desc->builder : PT_BUILD_DEVICE;
    if (width != 2 && width != 4 && width != 8) return fail(ctx, PT_ERR_UNSUPPORTED, "BVH width must be 2, 4 or 8");
    if (bins < 2 || bins > 256 || leaf_max < 1) return fail(ctx, PT_ERR_ARG, "bad BVH parameters");
    if (builder != PT_BUILD_DEVICE && builder != PT_BUILD_HOST && builder != PT_BUILD_HOST_REFERENCE_ORDER)
        return fail(ctx, PT_ERR_ARG, "unknown builder");
    if (builder == PT_BUILD_DEVICE && bins > 64)
        return fail(ctx, PT_ERR_UNSUPPORTED, "the device builder takes at most 64 bins");
    if (builder == PT_BUILD_DEVICE && s->num_prims() > 0) {
        const int rc = build_device(s, bins, leaf_max, out);
        if (rc != PT_OK || width == 2) return rc;
        return make_wide(*out, width, out);
    }
    pt_bvh* b = new pt_bvh();
    b->scene = s;
    const auto t0 = std::chrono::steady_clock::now();
    if (!build_bvh_host(s->tri_pos.data(), s->num_tris, bins, leaf_max, &b->host, (const float*)s->spheres.data(),
                        s->num_spheres, builder == PT_BUILD_HOST_REFERENCE_ORDER)) {
        delete b;
        return fail(ctx, PT_ERR_ARG, "bad BVH parameters");
    }
    b->host_ready = true;
    b->num_nodes = (int)b->host.nodes.size();
    b->max_depth = b->host.max_depth;
    std::vector<float4> nodes;
    std::vector<int2> big, ranges;
    V3 rlo, rhi;
    uint32_t rref;
    pack_bvh2(b->host, nodes, big, &rlo, &rhi, &rref, &ranges);
    const size_t nt = (size_t)s->num_prims();
    std::vector<float4> tris(3 * nt);
    std::vector<int32_t> mat(nt), mat_sh(nt), smooth;
    std::vector<float> nrm;
    std::vector<int8_t> kinds;
    if (!s->tri_smooth.empty()) smooth.resize(nt);
    if (!s->tri_nrm.empty()) nrm.resize(9 * nt);
    if (s->num_spheres > 0) kinds.resize(nt);
    for (size_t k = 0; k < nt; ++k) {
        const int id = b->host.order[k];
        if (id >= s->num_tris) { // a sphere: (centre, radius), then zeros
            tris[3 * k] = s->spheres[(size_t)(id - s->num_tris)];
            tris[3 * k + 1] = tris[3 * k + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            kinds[k] = 1;
        } else {
            const float* p = &s->tri_pos[9 * (size_t)id];
            const TriDev t = tri_from_points(v3(p[0], p[1], p[2]), v3(p[3], p[4], p[5]), v3(p[6], p[7], p[8]));
            tris[3 * k] = make_float4(t.p0.x, t.p0.y, t.p0.z, t.e1.x);
            tris[3 * k + 1] = make_float4(t.e1.y, t.e1.z, t.e2.x, t.e2.y);
            tris[3 * k + 2] = make_float4(t.e2.z, t.n.x, t.n.y, t.n.z);
        }
        const int m = s->tri_mat[id];
        mat[k] = m;
        const MaterialDev& md = s->mats[m];
        mat_sh[k] = md.type == kMatMix ? (md.mix <= 0.0f ? md.mat0 : md.mat1) : m; // GetMaterial(0.0f)
        if (!smooth.empty()) smooth[k] = s->tri_smooth[id];
        if (!nrm.empty()) std::memcpy(&nrm[9 * k], &s->tri_nrm[9 * (size_t)id], 9 * sizeof(float));
    }
    hipError_t err = hipSetDevice(ctx->device);
    if (err == hipSuccess) err = upload(&b->nodes_d, nodes.data(), nodes.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&b->big_leaves_d, big.data(), big.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&b->inner_range_d, ranges.data(), ranges.size(), ctx->stream);
    if (err == hipSuccess) err = hipMalloc((void**)&b->cursor_d, (PT_CURSOR_INTS + 1) * sizeof(int));
    if (err == hipSuccess) err = upload(&b->tris_d, tris.data(), tris.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&b->prim_ids_d, b->host.order.data(), nt, ctx->stream);
    if (err == hipSuccess) err = upload(&b->tri_mat_d, mat.data(), nt, ctx->stream);
    if (err == hipSuccess) err = upload(&b->tri_mat_shading_d, mat_sh.data(), nt, ctx->stream);
    if (err == hipSuccess) err = upload(&b->tri_smooth_d, smooth.data(), smooth.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&b->tri_nrm_d, nrm.data(), nrm.size(), ctx->stream);
    if (err == hipSuccess) err = upload(&b->kinds_d, kinds.data(), kinds.size(), ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    if (err != hipSuccess) {
        free_all({b->nodes_d, b->big_leaves_d, b->inner_range_d, b->cursor_d, b->tris_d, b->prim_ids_d, b->tri_mat_d,
                  b->tri_mat_shading_d, b->tri_smooth_d, b->tri_nrm_d, b->kinds_d});
        delete b;
        return fail(ctx, err == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP,
                    std::string("bvh upload: ") + hipGetErrorString(err));
    }
    b->view.nodes = b->nodes_d;
    b->view.tris = b->tris_d;
    b->view.prim_ids = b->prim_ids_d;
    b->view.kinds = b->kinds_d;
    b->view.root_lo = rlo;
    b->view.root_hi = rhi;
    b->view.root_ref = rref;
    b->view.big_leaves = b->big_leaves_d;
    b->view.inner_range = b->inner_range_d;
    b->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    b->device_bytes = (int64_t)(nodes.size() * sizeof(float4) + big.size() * sizeof(int2) + tris.size() * sizeof(float4)) +
                      (int64_t)ranges.size() * 8 +
                      (int64_t)nt * 12 + (int64_t)(smooth.size() * 4 + nrm.size() * 4 + kinds.size());
    b->num_inner = (int)(nodes.size() / 4);
    int rc = check_boxes(b);
    if (rc == PT_OK) rc = add_parents(b);
    if (rc != PT_OK) {
        pt_bvh_destroy(b);
        return rc;
    }
    *out = b;
    if (width != 2) return make_wide(b, width, out);
    return PT_OK;
}

int pt_bvh_destroy(pt_bvh* b)
{
    if (!b) return PT_ERR_ARG;
    free_all({b->nodes8_d, b->nodes_d, b->big_leaves_d, b->inner_range_d, b->cursor_d, b->tris_d, b->prim_ids_d, b->tri_mat_d,
              b->tri_mat_shading_d, b->tri_smooth_d, b->tri_nrm_d, b->order_keys_d, b->order_idx_d, b->order_temp_d,
              b->tri_uv_d, b->parents_d, b->kinds_d, b->host_io_d, b->soa_rays_d});
    delete b;
    return PT_OK;
}

int pt_bvh_export_order(pt_bvh* b, int32_t* prim_order)
{
    if (!b || !prim_order) return PT_ERR_ARG;
    if (int rc = ensure_host(b)) return rc;
    std::memcpy(prim_order, b->host.order.data(), b->host.order.size() * sizeof(int32_t));
    return PT_OK;
}

int pt_bvh_export_nodes(pt_bvh* b, float* boxes, int32_t* child, int32_t* range, int32_t* leaf)
{
    if (!b) return PT_ERR_ARG;
    if (int rc = ensure_host(b)) return rc;
    for (size_t i = 0; i < b->host.nodes.size(); ++i) {
        const HostNode& n = b->host.nodes[i];
        if (boxes) {
            float* o = boxes + 6 * i;
            o[0] = n.lo.x; o[1] = n.lo.y; o[2] = n.lo.z;
            o[3] = n.hi.x; o[4] = n.hi.y; o[5] = n.hi.z;
        }
        if (child) { child[2 * i] = n.child[0]; child[2 * i + 1] = n.child[1]; }
        if (range) { range[2 * i] = n.begin; range[2 * i + 1] = n.end; }
        if (leaf) leaf[i] = n.leaf ? 1 : 0;
    }
    return PT_OK;
}

int pt_bvh_build_host(const float* tri_pos, int32_t num_triangles, const pt_bvh_desc* desc, int32_t* prim_order,
                      float* boxes, int32_t* child, int32_t* range, int32_t* leaf, int32_t* max_depth)
{
    if (num_triangles < 0 || (num_triangles > 0 && !tri_pos)) return PT_ERR_ARG;
    HostBvh h;
    const int builder = desc ? desc->builder : PT_BUILD_HOST;
    if (builder != PT_BUILD_DEVICE && builder != PT_BUILD_HOST && builder != PT_BUILD_HOST_REFERENCE_ORDER)
        return PT_ERR_ARG;
    if (!build_bvh_host(tri_pos, num_triangles, desc ? desc->num_bins : 16, desc ? desc->leaf_max : 4, &h, nullptr, 0,
                        builder == PT_BUILD_HOST_REFERENCE_ORDER))
        return PT_ERR_ARG;
    if (prim_order) std::memcpy(prim_order, h.order.data(), h.order.size() * sizeof(int32_t));
    for (size_t i = 0; i < h.nodes.size(); ++i) {
        const HostNode& n = h.nodes[i];
        if (boxes) {
            float* o = boxes + 6 * i;
            o[0] = n.lo.x; o[1] = n.lo.y; o[2] = n.lo.z;
            o[3] = n.hi.x; o[4] = n.hi.y; o[5] = n.hi.z;
        }
        if (child) { child[2 * i] = n.child[0]; child[2 * i + 1] = n.child[1]; }
        if (range) { range[2 * i] = n.begin; range[2 * i + 1] = n.end; }
        if (leaf) leaf[i] = n.leaf ? 1 : 0;
    }
    if (max_depth) *max_depth = h.max_depth;
    return (int)h.nodes.size();
}

int pt_bvh_info(pt_bvh* b, int32_t* num_nodes, int32_t* max_depth, double* build_ms)
{
    if (!b) return PT_ERR_ARG;
    if (num_nodes) *num_nodes = b->num_nodes;
    if (max_depth) *max_depth = b->max_depth;
    if (build_ms) *build_ms = b->build_ms;
    return PT_OK;
}

} // extern "C"

namespace {

// API launches of at least this many rays are traced in coherence order (pt_order.hip).
constexpr int64_t kOrderMinRays = (int64_t)1 << 20;

TraceSegment api_segment(const int* count_d, const float* rays, bool any, float* t_out, int32_t* k_out)
{
    TraceSegment g;
    g.rays = (const float4*)rays;
    g.path = nullptr;
    g.count = count_d;
    g.any = any ? 1 : 0;
    g.map_prim = any ? 0 : 1;
    g.hit_flag = any ? 1 : 0;
    g.out_t = t_out;
    g.out_k = k_out;
    return g;
}

// cursor_d holds the work cursors (PT_CURSOR_STRIDE apart), cursor_d[PT_CURSOR_INTS] the ray
// count of the launch.
// slots: closest hits give the leaf slot (and t on a miss as the walk leaves it) instead
// of the primitive index.
int api_trace(pt_bvh* b, int64_t n, const float* rays, bool any, float* t_out, int32_t* k_out,
              unsigned long long* stats, hipStream_t st, bool slots = false)
{
    pt_context* ctx = b->scene->ctx;
    TraceJob job;
    job.seg[0] = api_segment(b->cursor_d + PT_CURSOR_INTS, rays, any, t_out, k_out);
    if (slots) job.seg[0].map_prim = 0;
    job.num_seg = 1;
    job.views = nullptr;
    job.job_of_path = nullptr;
    job.queries = nullptr;
    const int n32 = (int)n;
    if (n >= kOrderMinRays && b->view.root_ref != PT_REF_EMPTY) {
        // large launches trace their rays in coherence order (pt_order.hip)
        if (b->order_cap < n) {
            const int64_t cap = n;
            const size_t tmp = ray_order_temp_bytes((int)cap);
            PT_HIP(ctx, hipStreamSynchronize(st)); // the old buffers may still be read
            (void)hipFree(b->order_keys_d);
            (void)hipFree(b->order_idx_d);
            (void)hipFree(b->order_temp_d);
            b->order_keys_d = nullptr;
            b->order_idx_d = nullptr;
            b->order_temp_d = nullptr;
            b->order_cap = 0;
            PT_HIP(ctx, hipMalloc((void**)&b->order_keys_d, 2 * sizeof(uint32_t) * (size_t)cap));
            PT_HIP(ctx, hipMalloc((void**)&b->order_idx_d, 2 * sizeof(int32_t) * (size_t)cap));
            PT_HIP(ctx, hipMalloc(&b->order_temp_d, tmp > 0 ? tmp : 1));
            b->order_temp_bytes = tmp;
            b->order_cap = cap;
        }
        int32_t* sorted = nullptr;
        PT_HIP(ctx, launch_ray_order((const float4*)rays, n32, b->view.root_lo, b->view.root_hi, b->order_keys_d,
                                     b->order_idx_d, b->order_temp_d, b->order_temp_bytes, st, &sorted));
        job.seg[0].path = sorted; // sorted indices: ray and hit of item i at path[i]
        job.seg[0].rays_by_path = 1;
    }
    PT_HIP(ctx, launch_zero(b->cursor_d, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
    PT_HIP(ctx, hipMemcpyAsync(b->cursor_d + PT_CURSOR_INTS, &n32, sizeof(int), hipMemcpyHostToDevice, st));
    PT_HIP(ctx, launch_trace_job(b->view, job, b->cursor_d, stats, st, b->trace_tune));
    return PT_OK;
}

// pt_trace_closest on the GPU's fixed scratch (pt_order.hip) instead of the BVH's buffers:
// no allocation, and the ray count is written by a kernel rather than copied from the host.
int api_trace_noalloc(pt_bvh* b, int64_t n, const float* rays, float* t_out, int32_t* k_out, hipStream_t st)
{
    pt_context* ctx = b->scene->ctx;
    const NoallocScratch sc = noalloc_scratch(ctx->device);
    if (!sc.cursor) return fail(ctx, PT_ERR_STATE, "the fixed trace scratch of this device is not set up");
    TraceJob job;
    job.seg[0] = api_segment(sc.cursor + PT_CURSOR_INTS, rays, false, t_out, k_out);
    job.num_seg = 1;
    job.views = nullptr;
    job.job_of_path = nullptr;
    job.queries = nullptr;
    const int n32 = (int)n;
    if (n >= kOrderMinRays && b->view.root_ref != PT_REF_EMPTY) {
        if (ray_order_temp_bytes(n32) > sc.temp_bytes)
            return fail(ctx, PT_ERR_UNSUPPORTED, "the sort storage of this ray count exceeds the fixed scratch");
        int32_t* sorted = nullptr;
        PT_HIP(ctx, launch_ray_order((const float4*)rays, n32, b->view.root_lo, b->view.root_hi, sc.keys, sc.idx,
                                     sc.temp, sc.temp_bytes, st, &sorted));
        job.seg[0].path = sorted;
        job.seg[0].rays_by_path = 1;
    }
    PT_HIP(ctx, launch_zero(sc.cursor, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
    PT_HIP(ctx, launch_set(sc.cursor + PT_CURSOR_INTS, n32, st));
    PT_HIP(ctx, launch_trace_job(b->view, job, sc.cursor, nullptr, st, b->trace_tune));
    return PT_OK;
}

} // namespace

extern "C" {

int pt_trace_closest_noalloc(pt_bvh* b, int64_t n, const float* rays, float* t_out, int32_t* prim_out, void* stream)
{
    if (!b) return PT_ERR_ARG;
    pt_context* ctx = b->scene->ctx;
    if (n < 0 || (n > 0 && (!rays || !t_out || !prim_out))) return fail(ctx, PT_ERR_ARG, "bad trace arguments");
    if (n > PT_NOALLOC_MAX_RAYS)
        return fail(ctx, PT_ERR_ARG, "pt_trace_closest_noalloc takes at most PT_NOALLOC_MAX_RAYS (2^21) rays, got " +
                                         std::to_string(n));
    if (n == 0) return PT_OK;
    PT_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return api_trace_noalloc(b, n, rays, t_out, prim_out, st);
}

int pt_bvh_set_trace_tuning(pt_bvh* b, const pt_tuning* t)
{
    if (!b || !t) return PT_ERR_ARG;
    pt_context* ctx = b->scene->ctx;
    const int tpr = t->threads_per_ray == 0 ? 1 : t->threads_per_ray;
    if (t->trace_blocks_per_cu < 0 || t->trace_blocks_per_cu > 8 || t->trace_chunk < 64 || t->trace_chunk > 4096 ||
        (tpr != 1 && tpr != 2) || t->leaf_batch < -1 || t->leaf_batch > 4)
        return fail(ctx, PT_ERR_ARG, "trace tuning out of range");
    b->trace_tune.trace_blocks_per_cu = t->trace_blocks_per_cu;
    b->trace_tune.trace_chunk = t->trace_chunk;
    b->trace_tune.threads_per_ray = tpr;
    b->trace_tune.leaf_batch = t->leaf_batch == 0 ? pt::LaunchTuning{}.leaf_batch : t->leaf_batch;
    return PT_OK;
}

int pt_trace_closest(pt_bvh* b, int64_t n, const float* rays, float* t_out, int32_t* prim_out, void* stream)
{
    if (!b) return PT_ERR_ARG;
    pt_context* ctx = b->scene->ctx;
    if (n < 0 || n > 0x7fffff00 || (n > 0 && (!rays || !t_out || !prim_out)))
        return fail(ctx, PT_ERR_ARG, "bad trace arguments");
    if (n == 0) return PT_OK;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return api_trace(b, n, rays, false, t_out, prim_out, nullptr, st);
}

int pt_trace_closest_host(pt_bvh* b, int64_t n, const float* rays, float* t_out, int32_t* prim_out)
{
    if (!b) return PT_ERR_ARG;
    pt_context* ctx = b->scene->ctx;
    if (n < 0 || n > 0x7fffff00 || (n > 0 && (!rays || !t_out || !prim_out)))
        return fail(ctx, PT_ERR_ARG, "bad trace arguments");
    if (n == 0) return PT_OK;
    PT_HIP(ctx, hipSetDevice(ctx->device));
    if (b->host_io_cap < n) {
        if (b->host_io_d) PT_HIP(ctx, hipFree(b->host_io_d));
        b->host_io_d = nullptr;
        b->host_io_cap = 0;
        PT_HIP(ctx, hipMalloc(&b->host_io_d, 40 * (size_t)n));
        b->host_io_cap = n;
    }
    float* rays_d = (float*)b->host_io_d;
    float* t_d = rays_d + 8 * (size_t)n;
    int32_t* prim_d = (int32_t*)(t_d + n);
    PT_HIP(ctx, hipMemcpyAsync(rays_d, rays, 32 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    const int rc = api_trace(b, n, rays_d, false, t_d, prim_d, nullptr, ctx->stream);
    if (rc != PT_OK) return rc;
    PT_HIP(ctx, hipMemcpyAsync(t_out, t_d, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(ctx, hipMemcpyAsync(prim_out, prim_d, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PT_OK;
}

// The structure-of-arrays entry points: rays packed into the BVH's staging records, traced
// as by pt_trace_closest / pt_trace_any, hits written per array.
static int trace_soa(pt_bvh* b, int64_t n, const pt_rays_soa* r, bool any, float* t, int32_t* k, float* u, float* v,
                     void* stream)
{
    pt_context* ctx = b->scene->ctx;
    if (n == 0) return PT_OK;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    if (b->soa_rays_cap < n) {
        PT_HIP(ctx, hipStreamSynchronize(st)); // the old records may still be read
        (void)hipFree(b->soa_rays_d);
        b->soa_rays_d = nullptr;
        b->soa_rays_cap = 0;
        PT_HIP(ctx, hipMalloc((void**)&b->soa_rays_d, 2 * sizeof(float4) * (size_t)n));
        b->soa_rays_cap = n;
    }
    PT_HIP(ctx, launch_pack_rays_soa(n, r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->dist_min, r->dist_max,
                                     b->soa_rays_d, st));
    const int rc = api_trace(b, n, (const float*)b->soa_rays_d, any, t, k, nullptr, st, !any);
    if (rc != PT_OK || any) return rc;
    PT_HIP(ctx, launch_hits_soa(n, b->view, b->soa_rays_d, t, k, u, v, st));
    return PT_OK;
}

static bool soa_rays_ok(const pt_rays_soa* r)
{
    return r && r->ox && r->oy && r->oz && r->dx && r->dy && r->dz;
}

int pt_trace_closest_soa(pt_bvh* b, int64_t n, const pt_rays_soa* rays, const pt_hits_soa* hits, void* stream)
{
    if (!b) return PT_ERR_ARG;
    pt_context* ctx = b->scene->ctx;
    if (n < 0 || n > 0x7fffff00 || (n > 0 && (!soa_rays_ok(rays) || !hits || !hits->t || !hits->prim)))
        return fail(ctx, PT_ERR_ARG, "bad trace arguments");
    return trace_soa(b, n, rays, false, hits->t, hits->prim, hits->u, hits->v, stream);
}

int pt_trace_any_soa(pt_bvh* b, int64_t n, const pt_rays_soa* rays, int32_t* hit_out, void* stream)
{
    if (!b) return PT_ERR_ARG;
    pt_context* ctx = b->scene->ctx;
    if (n < 0 || n > 0x7fffff00 || (n > 0 && (!soa_rays_ok(rays) || !hit_out)))
        return fail(ctx, PT_ERR_ARG, "bad trace arguments");
    return trace_soa(b, n, rays, true, nullptr, hit_out, nullptr, nullptr, stream);
}

int pt_trace_any(pt_bvh* b, int64_t n, const float* rays, int32_t* hit_out, void* stream)
{
    if (!b) return PT_ERR_ARG;
    pt_context* ctx = b->scene->ctx;
    if (n < 0 || n > 0x7fffff00 || (n > 0 && (!rays || !hit_out))) return fail(ctx, PT_ERR_ARG, "bad trace arguments");
    if (n == 0) return PT_OK;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return api_trace(b, n, rays, true, nullptr, hit_out, nullptr, st);
}

int pt_trace_stats(pt_bvh* b, int64_t n, const float* rays, float* t_out, int32_t* prim_out, int64_t* host_stats)
{
    if (!b || !host_stats) return PT_ERR_ARG;
    pt_context* ctx = b->scene->ctx;
    if (n < 0 || n > 0x7fffff00 || (n > 0 && (!rays || !t_out || !prim_out)))
        return fail(ctx, PT_ERR_ARG, "bad trace arguments");
    unsigned long long* st_d = nullptr;
    PT_HIP(ctx, hipMalloc((void**)&st_d, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(st_d, h, sizeof h, hipMemcpyHostToDevice, ctx->stream);
    int rc = PT_OK;
    if (e == hipSuccess && n > 0) rc = api_trace(b, n, rays, false, t_out, prim_out, st_d, ctx->stream);
    if (e == hipSuccess && rc == PT_OK) e = hipMemcpyAsync(h, st_d, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && rc == PT_OK) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(st_d);
    if (rc != PT_OK) return rc;
    if (e != hipSuccess) return fail(ctx, PT_ERR_HIP, hipGetErrorString(e));
    host_stats[0] = (int64_t)h[0];
    host_stats[1] = (int64_t)h[1];
    return PT_OK;
}

int pt_tonemap_rgb8(pt_context* ctx, int64_t num_pixels, const float* rgb, uint8_t* rgb8_out, int32_t on_device,
                    void* stream)
{
    if (!ctx) return PT_ERR_ARG;
    if (num_pixels < 0 || num_pixels > ((int64_t)1 << 40) || (num_pixels > 0 && (!rgb || !rgb8_out)))
        return fail(ctx, PT_ERR_ARG, "bad tonemap arguments");
    if (num_pixels == 0) return PT_OK;
    PT_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    if (on_device) {
        PT_HIP(ctx, pt::launch_tonemap_rgb8(rgb, rgb8_out, num_pixels, st));
        return PT_OK;
    }
    // host buffers: staged through device memory, synchronous
    const size_t in_b = (size_t)num_pixels * 3 * sizeof(float), out_b = (size_t)num_pixels * 3;
    float* in_d = nullptr;
    uint8_t* out_d = nullptr;
    PT_HIP(ctx, hipMalloc((void**)&in_d, in_b));
    hipError_t e = hipMalloc((void**)&out_d, out_b);
    if (e == hipSuccess) e = hipMemcpyAsync(in_d, rgb, in_b, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = pt::launch_tonemap_rgb8(in_d, out_d, num_pixels, st);
    if (e == hipSuccess) e = hipMemcpyAsync(rgb8_out, out_d, out_b, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(in_d);
    if (out_d) (void)hipFree(out_d);
    if (e != hipSuccess) return fail(ctx, e == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP, hipGetErrorString(e));
    return PT_OK;
}

} // extern "C"

