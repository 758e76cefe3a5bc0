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
    uint32_t e;      // pair entry
    uint32_t n;      // LBVH: triangles below (leaf_max > 0)
    uint32_t first;  // LBVH: first triangle of an inner child's range
    f3 lo, hi;
};

__device__ inline float half_area(f3 lo, f3 hi)
{
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// the two children of pair p: entries and boxes, and for an LBVH their
// triangle counts and range starts.  A Karras pair p covers [first, last] =
// [min(p, j), max(p, j)], j in e[3]; its left child (inner: pair `split`)
// covers [first, split] and its right child (inner: pair split + 1) covers
// [split + 1, last], so an unchanged Karras tree (RANGES) gives both counts
// from p's own record.  After tree rotations the ranges only hold for the
// untouched subtrees of at most leaf_max triangles, and the counts are read
// from each child's record: a changed pair keeps a count above leaf_max
// either way, which is all the collapse asks of it.
template <bool RANGES>
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
    a.first = 0u;
    b.first = 0u;
    if (!lbvh) return;
    const bool ai = !(a.e & LEAF_BIT) && a.e < num_pairs;
    const bool bi = !(b.e & LEAF_BIT) && b.e < num_pairs;
    if constexpr (RANGES) {
        const uint32_t first = min(p, e.w), last = max(p, e.w);
        if (ai) {
            a.n = a.e - first + 1u;
            a.first = first;
        }
        if (bi) {
            b.n = last - b.e + 1u;
            b.first = b.e;
        }
    } else {
        if (ai) {
            const uint32_t j = pairs[a.e].e[3];
            a.n = max(a.e, j) - min(a.e, j) + 1u;
            a.first = min(a.e, j);
        }
        if (bi) {
            const uint32_t j = pairs[b.e].e[3];
            b.n = max(b.e, j) - min(b.e, j) + 1u;
            b.first = min(b.e, j);
        }
    }
}

template <bool RANGES>
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
        if (lbvh && p != 0) {
            const uint32_t j = pairs[p].e[3];
            if (max(p, j) - min(p, j) + 1u <= leaf_max) continue;
        }
        // every index into c[] is a compile-time constant once the loops
        // are unrolled, so the four children stay in registers
        child_ref c[4];
        load_pair<RANGES>(pairs, num_pairs, p, lbvh, c[0], c[1]);
        int nc = 2;
#pragma unroll
        for (int step = 0; step < 2; ++step) {
            if (nc != 2 + step) break;
            int best = -1;
            float best_a = -1.0f;
#pragma unroll
            for (int k = 0; k < 2 + step; ++k) {
                if (!openable(c[k])) continue;
                const float a = half_area(c[k].lo, c[k].hi);
                if (a > best_a) {
                    best_a = a;
                    best = k;
                }
            }
            if (best < 0) break;
            uint32_t eb = c[0].e;
#pragma unroll
            for (int k = 1; k < 2 + step; ++k)
                if (k == best) eb = c[k].e;
            child_ref a, b;
            load_pair<RANGES>(pairs, num_pairs, eb, lbvh, a, b);
#pragma unroll
            for (int k = 0; k < 2 + step; ++k)
                if (k == best) c[k] = a;
            c[2 + step] = b;
            ++nc;
        }
        quad_node qn;
        const float inf = __builtin_inff();
        // unused slots: lo = hi = +inf, a box the slab test never enters
        float lx[4], ly[4], lz[4], hx[4], hy[4], hz[4];
        uint32_t ce[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool used = k < nc;
            lx[k] = used ? c[k].lo.x : inf;
            ly[k] = used ? c[k].lo.y : inf;
            lz[k] = used ? c[k].lo.z : inf;
            hx[k] = used ? c[k].hi.x : inf;
            hy[k] = used ? c[k].hi.y : inf;
            hz[k] = used ? c[k].hi.z : inf;
            // leaves as they are, small LBVH subtrees as leaf ranges, other
            // inner children as the quad of their own pair
            uint32_t e = c[k].e;
            if (!(e & LEAF_BIT) && lbvh && e < num_pairs && c[k].n <= leaf_max) e = leaf_entry(c[k].first, c[k].n);
            ce[k] = used ? e : LEAF_BIT;
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

int refit_quads(bvh_state& s, hipStream_t st)
{
    // the quads follow the refitted pair boxes (the greedy choices may move)
    return collapse_quads(s, s.collapse_leaf_max, st);
}

}  // namespace vr
