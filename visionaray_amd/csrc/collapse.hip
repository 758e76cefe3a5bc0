// collapse.hip -- the 4-wide traversal nodes built from the binary tree.
//
// Every binary pair p gets the quad it would root: its children are found
// by opening binary children, largest surface area first, until there are
// four (or none can be opened).  Every child box is a binary node's box,
// unchanged.  For an LBVH, a subtree of at most leaf_max triangles becomes
// one leaf over its sorted range (the triangles of a Karras subtree are
// contiguous in traversal order).  An inner child refers to the quad of its
// own pair index, so all quads are built by one fully parallel kernel with
// no queue or level order; traversal reaches only the quads hanging from
// quad 0 (the root pair's), the others are never read.  The reference
// traverses its binary tree directly (detail/bvh/intersect.inl:166-356);
// fewer, wider nodes halve the chain of dependent node fetches per ray.
#include "vr_internal.h"

#include <algorithm>

namespace vr {

namespace {

constexpr int BLOCK = 256;

struct child_ref {
    uint32_t e;  // pair entry
    uint32_t n;  // LBVH: triangles below (leaf_max > 0)
    f3 lo, hi;
};

__device__ inline float half_area(f3 lo, f3 hi)
{
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// triangles under an LBVH pair: its range [min(p, j), max(p, j)], j in e[3]
__device__ inline uint32_t lbvh_count(const pair_node* __restrict__ pairs, uint32_t p)
{
    const uint32_t j = pairs[p].e[3];
    return max(p, j) - min(p, j) + 1u;
}

// the two children of pair p (their entries, boxes, box locations and, for an
// LBVH, triangle counts)
__device__ inline void load_pair(const pair_node* __restrict__ pairs, uint32_t num_pairs, uint32_t p,
                                 bool lbvh, child_ref& a, child_ref& b)
{
    const float4* q = reinterpret_cast<const float4*>(pairs + p);
    const float4 q0 = q[0], q1 = q[1], q2 = q[2];
    const uint4 e = *reinterpret_cast<const uint4*>(q + 3);
    a.e = e.x;
    a.lo = mk3(q0.x, q0.y, q0.z);
    a.hi = mk3(q0.w, q1.x, q1.y);
    b.e = e.y;
    b.lo = mk3(q1.z, q1.w, q2.x);
    b.hi = mk3(q2.y, q2.z, q2.w);
    a.n = 1u;
    b.n = 1u;
    if (lbvh) {
        if (!(a.e & LEAF_BIT) && a.e < num_pairs) a.n = lbvh_count(pairs, a.e);
        if (!(b.e & LEAF_BIT) && b.e < num_pairs) b.n = lbvh_count(pairs, b.e);
    }
}

__global__ __launch_bounds__(BLOCK) void k_collapse(const pair_node* __restrict__ pairs, uint32_t num_pairs,
                                                    uint32_t leaf_max, quad_node* __restrict__ quads)
{
    const bool lbvh = leaf_max > 0;
    // an inner child that can be opened: an inner pair above the leaf size
    auto openable = [&](const child_ref& c) {
        return !(c.e & LEAF_BIT) && c.e < num_pairs && (!lbvh || c.n > leaf_max);
    };
    for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < num_pairs; p += gridDim.x * BLOCK) {
        // a subtree of at most leaf_max triangles is always a leaf of its
        // parent's quad, never a quad itself (the root pair is kept: a tree
        // that small still starts at quad 0)
        if (lbvh && p != 0 && lbvh_count(pairs, p) <= leaf_max) continue;
        child_ref c[4];
        load_pair(pairs, num_pairs, p, lbvh, c[0], c[1]);
        int nc = 2;
        while (nc < 4) {
            int best = -1;
            float best_a = -1.0f;
            for (int k = 0; k < nc; ++k) {
                if (!openable(c[k])) continue;
                const float a = half_area(c[k].lo, c[k].hi);
                if (a > best_a) {
                    best_a = a;
                    best = k;
                }
            }
            if (best < 0) break;
            child_ref a, b;
            load_pair(pairs, num_pairs, c[best].e, lbvh, a, b);
            c[best] = a;
            c[nc] = b;
            ++nc;
        }
        quad_node qn;
        const float inf = __builtin_inff();
        // unused slots: lo = hi = +inf, a box the slab test never enters
        float lx[4] = {inf, inf, inf, inf}, ly[4] = {inf, inf, inf, inf}, lz[4] = {inf, inf, inf, inf};
        float hx[4] = {inf, inf, inf, inf}, hy[4] = {inf, inf, inf, inf}, hz[4] = {inf, inf, inf, inf};
        uint32_t ce[4] = {LEAF_BIT, LEAF_BIT, LEAF_BIT, LEAF_BIT};
        for (int k = 0; k < nc; ++k) {
            lx[k] = c[k].lo.x; ly[k] = c[k].lo.y; lz[k] = c[k].lo.z;
            hx[k] = c[k].hi.x; hy[k] = c[k].hi.y; hz[k] = c[k].hi.z;
            // leaves as they are, small LBVH subtrees as leaf ranges, other
            // inner children as the quad of their own pair
            uint32_t e = c[k].e;
            if (!(e & LEAF_BIT) && lbvh && e < num_pairs && c[k].n <= leaf_max) {
                const uint32_t j = pairs[e].e[3];
                e = leaf_entry(min(e, j), c[k].n);
            }
            ce[k] = e;
        }
        qn.lo_x = make_float4(lx[0], lx[1], lx[2], lx[3]);
        qn.lo_y = make_float4(ly[0], ly[1], ly[2], ly[3]);
        qn.lo_z = make_float4(lz[0], lz[1], lz[2], lz[3]);
        qn.hi_x = make_float4(hx[0], hx[1], hx[2], hx[3]);
        qn.hi_y = make_float4(hy[0], hy[1], hy[2], hy[3]);
        qn.hi_z = make_float4(hz[0], hz[1], hz[2], hz[3]);
        qn.child = make_uint4(ce[0], ce[1], ce[2], ce[3]);
        qn.meta = make_uint4((uint32_t)nc, 0u, 0u, 0u);
        quads[p] = qn;
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
    hipLaunchKernelGGL(k_collapse, dim3(grid), dim3(BLOCK), 0, st, s.pairs, (uint32_t)np, leaf_max, s.quads);
    VR_HIP_CHECK(hipGetLastError());
    s.root = 0u;
    return VR_OK;
}

int refit_quads(bvh_state& s, hipStream_t st)
{
    // the quads follow the refitted pair boxes (the greedy choices may move)
    return collapse_quads(s, s.collapse_leaf_max, st);
}

}  // namespace vr
