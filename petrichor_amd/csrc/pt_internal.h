// pt_internal.h — library-internal object definitions shared by the C-ABI files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/petrichor_hip.h"
#include "pt_bvh_host.h"
#include "pt_kernels.h"

struct pt_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    pt_engine* engine = nullptr; // single-job engine of pt_render, grown on demand
    pt_engine* engine2[2] = {nullptr, nullptr}; // the other parts' engines of a large image rendered in parts
};

struct pt_scene {
    pt_context* ctx = nullptr;
    int num_tris = 0;    // triangles: primitives 0 .. num_tris - 1
    int num_spheres = 0; // spheres: primitives num_tris .. num_prims() - 1
    int num_prims() const { return num_tris + num_spheres; }
    std::vector<float4> spheres;        // centre, radius
    float4* spheres_d = nullptr;
    int8_t* light_kinds_d = nullptr;    // per light: 1 for a sphere; null without sphere lights
    std::vector<float> tri_pos, tri_nrm;
    std::vector<int32_t> tri_mat, tri_smooth;
    std::vector<pt::MaterialDev> mats;
    pt::MaterialDev* mats_d = nullptr;
    float4* lights_d = nullptr;
    int num_lights = 0;
    pt::CameraDev cam;
    pt::EnvDev env;
    float* env_rgb_d = nullptr;
    // primitive data resident for the device BVH builder; the per-primitive arrays
    // (materials, smooth flags, normals, texture coordinates) cover triangles and then
    // spheres (zeros for a sphere)
    float* tri_pos_d = nullptr;        // 9 floats per triangle
    int32_t* tri_mat_d = nullptr;
    int32_t* tri_smooth_d = nullptr;   // null without per-vertex normals
    float* tri_nrm_d = nullptr;
    std::vector<float> tri_uv;         // 6 floats per primitive, or empty
    float* tri_uv_d = nullptr;
    std::vector<float*> tex_d;         // material map textures
    pt::TexDev* textures_d = nullptr;
    int2* mat_maps_d = nullptr;        // per material, or null without maps
    int* mat_nmaps_d = nullptr;        // per material: GGX normal map, or null without normal maps
    float2* map_strength_d = nullptr;  // per material: normal / roughness map strength, or null (1, 1)
    int32_t* mat_shading_of_d = nullptr; // material -> material that shades it (mix resolved)
    int64_t device_bytes = 0; // materials, lights, environment, triangles
};

struct pt_bvh {
    pt_scene* scene = nullptr;
    pt::HostBvh host;
    float4* nodes_d = nullptr;
    int2* big_leaves_d = nullptr;
    int2* inner_range_d = nullptr; // leaf-slot range of every inner node (preorder)
    int* cursor_d = nullptr; // trace work cursor of the API entry points
    // coherence order of large API launches: keys and ray indices (2 x order_cap each) and
    // the radix sort's temporary storage, grown on demand
    uint32_t* order_keys_d = nullptr;
    int32_t* order_idx_d = nullptr;
    void* order_temp_d = nullptr;
    size_t order_temp_bytes = 0;
    int64_t order_cap = 0;
    float4* tris_d = nullptr;
    int32_t* prim_ids_d = nullptr;
    int32_t* tri_mat_d = nullptr;
    int32_t* tri_mat_shading_d = nullptr;
    int32_t* tri_smooth_d = nullptr;
    float* tri_nrm_d = nullptr;
    uint4* nodes8_d = nullptr; // wide nodes (width 4 or 8 only)
    float* tri_uv_d = nullptr; // 6 floats per leaf slot, when the scene has texture coordinates
    int32_t* parents_d = nullptr; // inner nodes' parents, for trees deeper than the traversal stack
    int8_t* kinds_d = nullptr;    // per leaf slot: 1 for a sphere; only when the scene has spheres
    void* host_io_d = nullptr;    // pt_trace_closest_host's device staging (40 bytes per ray)
    int64_t host_io_cap = 0;
    pt::LaunchTuning trace_tune;  // launch shapes of this BVH's traces (pt_bvh_set_trace_tuning)
    float4* soa_rays_d = nullptr; // the structure-of-arrays entry points' ray records (32 bytes per ray)
    int64_t soa_rays_cap = 0;
    int width = 2;
    int num_nodes8 = 0;
    int depth8 = 0;
    pt::BvhView view;
    int num_nodes = 0;
    int max_depth = 0;
    int num_inner = 0;
    int num_big = 0;
    bool host_ready = false; // `host` holds the node table (built lazily after a device build)
    double build_ms = 0.0;
    int64_t device_bytes = 0; // nodes, leaf-ordered triangles and per-triangle data
};


// Sets the context's error text and returns the code.
int pt_fail(pt_context* ctx, int code, const std::string& msg);
// Device bytes of a scene's material-map textures and tables (pt_scene_create).
int64_t pt_scene_map_bytes(const pt_scene_desc* d);

#define PT_HIP(ctx, call)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return pt_fail((ctx), e_ == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP,                 \
                           std::string(#call) + ": " + hipGetErrorString(e_));                         \
    } while (0)
