// pt_bvh_host.h — host construction of the binary BVH (reference topology).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pt_device.h"

namespace pt {

struct HostNode {
    V3 lo, hi;
    int child[2] = {-1, -1};
    int begin = 0, end = 0;
    bool leaf = false;
};

struct HostBvh {
    std::vector<HostNode> nodes; // the reference's node order (BinnedSAHBVH::m_nodes)
    std::vector<int> order;      // leaf slot -> triangle (BinnedSAHBVH::m_primitiveIDs)
    int max_depth = 0;
};

// Primitives are the triangles (9 floats each) and then the spheres (centre, radius).
// reference_order: sort every split node's range as the reference does, so the leaf order
// is the reference's also among equal centroid keys (slower: one sort per node).
bool build_bvh_host(const float* tri_pos, int num_tris, int num_bins, int leaf_max, HostBvh* out,
                    const float* spheres = nullptr, int num_spheres = 0, bool reference_order = false);
void pack_bvh2(const HostBvh& h, std::vector<float4>& nodes, std::vector<int2>& big_leaves, V3* root_lo,
               V3* root_hi, uint32_t* root_ref, std::vector<int2>* inner_range = nullptr);

} // namespace pt
