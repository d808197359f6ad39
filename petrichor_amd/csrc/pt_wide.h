// pt_wide.h — the 4- and 8-wide BVHs with quantised child boxes (layouts in pt_wide.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pt_device.h"

namespace pt {

struct Wide8 {
    uint4* nodes_d = nullptr; // 8 (width 8) or 4 (width 4) x uint4 per node, level order, root first
    int num_nodes = 0;
    int depth = 0;            // levels of wide nodes
    uint32_t root = PT_REF_EMPTY; // pt_ref_inner(0), or the binary root when it is a leaf
};

// Bytes of one wide node record: 128 at width 8, 64 at width 4.
inline int wide_node_bytes(int width) { return width == 8 ? 128 : 64; }

// Collapses a packed binary tree (4 float4 per inner node, preorder) into nodes of
// `width` (4 or 8) children.
hipError_t gpu_collapse_wide(int width, const float4* bnodes_d, int num_inner, uint32_t root_ref, hipStream_t st,
                             Wide8* out);

} // namespace pt
