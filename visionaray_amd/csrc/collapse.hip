// collapse.hip -- all 4-wide traversal nodes of a tree in one fully
// parallel kernel (quad p for every pair p; the quad itself is made by
// make_quad, collapse.h).  LBVH builds make their quads in the bottom-up
// pass instead; this kernel serves rotated LBVHs, trees adopted from the
// host and refits.
#include "collapse.h"

#include <algorithm>

namespace vr {

namespace {

constexpr int BLOCK = 256;

template <bool RANGES>
__global__ __launch_bounds__(BLOCK) void k_collapse(const pair_node* __restrict__ pairs, uint32_t num_pairs,
                                                    uint32_t leaf_max, quad_node* __restrict__ quads)
{
    const global_pairs src{pairs};
    for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < num_pairs; p += gridDim.x * BLOCK) {
        if (leaf_max > 0 && !pair_gets_quad(p, pairs[p].e[3], leaf_max)) continue;
        make_quad<RANGES>(src, p, num_pairs, leaf_max, quads);
    }
}

}  // namespace

int collapse_quads(bvh_state& s, uint32_t leaf_max, hipStream_t st)
{
    if (s.num_prims == 0 || s.num_pairs == 0) return VR_OK;  // empty, or the root is a leaf
    VR_CHECK(leaf_max <= MAX_LEAF_COUNT, VR_ERR_INVALID_ARGUMENT, "collapse: leaf_max above 32");
    const size_t np = s.num_pairs;
    if (!s.quads) VR_HIP_CHECK(hipMalloc(&s.quads, np * sizeof(quad_node)));
    s.collapse_leaf_max = leaf_max;
    const unsigned grid = std::max(1u, std::min(div_up(np, BLOCK), 16384u));
    // an LBVH's own Karras ranges give every count, unless rotations moved them
    if (leaf_max > 0 && s.is_lbvh && !s.rotated)
        hipLaunchKernelGGL(k_collapse<true>, dim3(grid), dim3(BLOCK), 0, st, s.pairs, (uint32_t)np, leaf_max, s.quads);
    else
        hipLaunchKernelGGL(k_collapse<false>, dim3(grid), dim3(BLOCK), 0, st, s.pairs, (uint32_t)np, leaf_max,
                           s.quads);
    VR_HIP_CHECK(hipGetLastError());
    s.root = 0u;
    return VR_OK;
}

}  // namespace vr
