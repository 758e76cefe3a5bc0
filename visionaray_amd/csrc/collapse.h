// collapse.h -- one 4-wide traversal node from the binary tree.
//
// Quad p is the collapse rooted at binary pair p: its children are found by
// opening binary children, largest surface area first, until there are four
// (or none can be opened).  Every child box is a binary node's box,
// unchanged.  For an LBVH, a subtree of at most leaf_max triangles becomes
// one leaf over its sorted range (the triangles of a Karras subtree are
// contiguous in traversal order).  An inner child refers to the quad of its
// own pair index, so quads can be made in any order: by a kernel over all
// pairs (collapse.hip), or by the bottom-up pass as soon as a pair's subtree
// is complete (lbvh.hip).  The reference traverses its binary tree directly
// (detail/bvh/intersect.inl:166-356); fewer, wider nodes halve the chain of
// dependent node fetches per ray.
//
// A source of pair records provides ent(q), the entries of pair q (child 0,
// child 1, parent link, range end), and box(loc, lo, hi), the box of child
// slot loc = 2 q + side.  The collapse keeps only entries, counts and box
// locations in registers and reads a box when it compares or writes it.
#pragma once

#include "vr_internal.h"

namespace vr {

struct child_ref {
    uint32_t e;      // pair entry
    uint32_t n;      // LBVH: triangles below (leaf_max > 0)
    uint32_t first;  // LBVH: first triangle of an inner child's range
    uint32_t loc;    // the child's box: slot loc & 1 of pair loc >> 1
};

__device__ inline float half_area(f3 lo, f3 hi)
{
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// pair records in global memory
struct global_pairs {
    const pair_node* pairs;
    __device__ uint4 ent(uint32_t q) const { return *reinterpret_cast<const uint4*>(pairs[q].e); }
    __device__ void box(uint32_t loc, f3& lo, f3& hi) const
    {
        const float2* f = reinterpret_cast<const float2*>(pairs[loc >> 1].f + 6 * (loc & 1u));
        const float2 a = f[0], b = f[1], c = f[2];
        lo = mk3(a.x, a.y, b.x);
        hi = mk3(b.y, c.x, c.y);
    }
};

// the two children of pair p: entries and box locations, and for an LBVH
// their triangle counts and range starts.  A Karras pair p covers
// [first, last] = [min(p, j), max(p, j)], j = e[3]; its left child (inner:
// pair `split`) covers [first, split] and its right child (inner: pair
// split + 1) covers [split + 1, last], so an unchanged Karras tree (RANGES)
// gives both counts from p's own entries.  After tree rotations the ranges
// only hold for the untouched subtrees of at most leaf_max triangles, and the
// counts are read from each child's entries: a changed pair keeps a count
// above leaf_max either way, which is all the collapse asks of it.
template <bool RANGES, typename Src>
__device__ inline void pair_children(const Src& src, uint32_t num_pairs, uint32_t p, bool lbvh, child_ref& a,
                                     child_ref& b)
{
    const uint4 e = src.ent(p);
    a.e = e.x;
    b.e = e.y;
    a.loc = 2u * p;
    b.loc = 2u * p + 1u;
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
            const uint32_t j = src.ent(a.e).w;
            a.n = max(a.e, j) - min(a.e, j) + 1u;
            a.first = min(a.e, j);
        }
        if (bi) {
            const uint32_t j = src.ent(b.e).w;
            b.n = max(b.e, j) - min(b.e, j) + 1u;
            b.first = min(b.e, j);
        }
    }
}

// quad p (leaf_max > 0: an LBVH, small subtrees become leaf ranges)
template <bool RANGES, typename Src>
__device__ inline void make_quad(const Src& src, uint32_t p, uint32_t num_pairs, uint32_t leaf_max,
                                 quad_node* __restrict__ quads)
{
    const bool lbvh = leaf_max > 0;
    // an inner child that can be opened: an inner pair above the leaf size
    auto openable = [&](const child_ref& c) {
        return !(c.e & LEAF_BIT) && c.e < num_pairs && (!lbvh || c.n > leaf_max);
    };
    // every index into c[] is a compile-time constant once the loops are
    // unrolled, so the four children stay in registers
    child_ref c[4];
    pair_children<RANGES>(src, num_pairs, p, lbvh, c[0], c[1]);
    int nc = 2;
#pragma unroll
    for (int step = 0; step < 2; ++step) {
        if (nc != 2 + step) break;
        int best = -1;
        float best_a = -1.0f;
#pragma unroll
        for (int k = 0; k < 2 + step; ++k) {
            if (!openable(c[k])) continue;
            f3 lo, hi;
            src.box(c[k].loc, lo, hi);
            const float a = half_area(lo, hi);
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
        pair_children<RANGES>(src, num_pairs, eb, lbvh, a, b);
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
        f3 lo = mk3(inf, inf, inf), hi = mk3(inf, inf, inf);
        if (used) src.box(c[k].loc, lo, hi);
        lx[k] = lo.x;
        ly[k] = lo.y;
        lz[k] = lo.z;
        hx[k] = hi.x;
        hy[k] = hi.y;
        hz[k] = hi.z;
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

// The up to eight children of the 8-wide node rooted at pair p, by the
// rules of the reference's bvh_collapser (detail/bvh/collapse.h:37-110):
// open the inner child of largest surface area (strictly above zero, the
// first of equal ones) while the node has fewer than eight children; the
// root never opens a child whose two children are both leaves.  The opened
// child's first child takes its slot and its second child is appended.
// With leaf_max > 0 (an LBVH) a subtree of at most leaf_max triangles counts
// as a leaf, as in the 4-wide collapse.  Returns the number of children.
template <bool RANGES, typename Src>
__device__ inline int make_wide8(const Src& src, uint32_t p, uint32_t num_pairs, uint32_t leaf_max, child_ref* c)
{
    const bool lbvh = leaf_max > 0;
    auto openable = [&](const child_ref& x) {
        return !(x.e & LEAF_BIT) && x.e < num_pairs && (!lbvh || x.n > leaf_max);
    };
    pair_children<RANGES>(src, num_pairs, p, lbvh, c[0], c[1]);
    int nc = 2;
#pragma unroll
    for (int step = 0; step < 6; ++step) {
        if (nc != 2 + step) break;
        int best = -1;
        float best_a = 0.0f;
#pragma unroll
        for (int k = 0; k < 2 + step; ++k) {
            if (!openable(c[k])) continue;
            if (p == 0) {
                child_ref a, b;
                pair_children<RANGES>(src, num_pairs, c[k].e, lbvh, a, b);
                if (!openable(a) && !openable(b)) continue;
            }
            f3 lo, hi;
            src.box(c[k].loc, lo, hi);
            const float a = half_area(lo, hi);
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
        pair_children<RANGES>(src, num_pairs, eb, lbvh, a, b);
#pragma unroll
        for (int k = 0; k < 2 + step; ++k)
            if (k == best) c[k] = a;
        c[2 + step] = b;
        ++nc;
    }
    return nc;
}

// a pair of at most leaf_max triangles is a leaf of its parent's quad, never
// a quad itself (the root pair is kept: a tree that small still starts at
// quad 0)
__device__ inline bool pair_gets_quad(uint32_t p, uint32_t e3, uint32_t leaf_max)
{
    return p == 0 || leaf_max == 0 || max(p, e3) - min(p, e3) + 1u > leaf_max;
}

}  // namespace vr
