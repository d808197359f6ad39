// pt_wide.h — the 8-wide BVH with quantised child boxes (layout in pt_wide.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pt_device.h"

namespace pt {

struct Wide8 {
    uint4* nodes_d = nullptr; // 8 x uint4 per node, level order, root first
    int num_nodes = 0;
    int depth = 0;            // levels of wide nodes
    uint32_t root = PT_REF_EMPTY; // pt_ref_inner(0), or the binary root when it is a leaf
};

// Collapses a packed binary tree (4 float4 per inner node, preorder) into 8-wide nodes.
hipError_t gpu_collapse_bvh8(const float4* bnodes_d, int num_inner, uint32_t root_ref, hipStream_t st, Wide8* out);

} // namespace pt
