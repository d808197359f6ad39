// pt_build.h — binned-SAH BVH construction and packing on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pt_device.h"

namespace pt {

// One node of the binary tree as the device builder creates it. Nodes are numbered
// level by level (root 0; the two children of a split are adjacent, left first).
struct BuildNode {
    float lo[3], hi[3];
    int32_t begin, end; // leaf-slot range
    int32_t child;      // left child id (right = child + 1), -1 for a leaf
    int32_t leaf;
    int32_t sort_axis;  // centroid axis of the leaf's final order, -1: input order
    int32_t depth;      // root = 1
};

struct GpuBvhResult {
    BuildNode* nodes_d = nullptr; // num_nodes entries
    int32_t* ids_d = nullptr;     // leaf slot -> triangle
    int num_nodes = 0;
    int max_depth = 0;
    std::vector<int> level_begin; // first node id of every level, plus the end
};

struct GpuPacked {
    float4* nodes_d = nullptr;    // 4 float4 per inner node, depth-first preorder
    int2* big_leaves_d = nullptr; // (first slot, count) of leaves over PT_REF_MAX_SMALL_LEAF
    int2* inner_range_d = nullptr; // leaf-slot range [begin, end) of every inner node (preorder)
    float4* tris_d = nullptr;     // 3 float4 per leaf slot (TriDev)
    int32_t* prim_ids_d = nullptr;
    int32_t* mat_d = nullptr;
    int32_t* mat_sh_d = nullptr;
    int32_t* smooth_d = nullptr;  // only when the scene has per-vertex normals
    float* nrm_d = nullptr;
    int8_t* kinds_d = nullptr;    // per leaf slot: 1 for a sphere; only when the scene has spheres
    int num_inner = 0;
    int num_big = 0;
    V3 root_lo, root_hi;
    uint32_t root_ref = 0;
};

// Builds the tree over num_tris triangles (9 floats each, device memory) followed by
// num_spheres spheres (centre, radius), primitives num_tris... The result's device arrays
// belong to the caller (gpu_bvh_free). Synchronises the stream once per level of the tree.
hipError_t gpu_build_bvh(const float* tri_pos_d, int num_tris, const float4* spheres_d, int num_spheres, int num_bins,
                         int leaf_max, hipStream_t st, GpuBvhResult* out);
void gpu_bvh_free(GpuBvhResult* r);

// Packs a built tree into the traversal layout and gathers the per-slot triangle data.
// smooth_d / nrm_d may be null (flat shading only). The packed arrays belong to the caller.
// The per-primitive arrays (materials, smooth flags, normals) cover triangles and spheres.
hipError_t gpu_pack_bvh(const GpuBvhResult& r, int num_tris, const float* tri_pos_d, const float4* spheres_d,
                        int num_spheres, const int32_t* tri_mat_d, const int32_t* mat_shading_of_d,
                        const int32_t* smooth_d, const float* nrm_d, hipStream_t st, GpuPacked* out);
void gpu_packed_free(GpuPacked* p);
// Whether every child box of the num_inner binary nodes has lo <= hi on every axis (false
// when a coordinate is NaN): the trace kernels' branch-free slab test relies on it.
hipError_t gpu_boxes_ordered(const float4* nodes_d, int num_inner, hipStream_t st, bool* ordered);

} // namespace pt
