// collapse.hip -- the 4-wide traversal nodes built from the binary tree.
//
// A binary tree node becomes a quad node whose children are found by opening
// binary children, largest surface area first, until there are four (or
// none can be opened).  Every child box is a binary node's box, unchanged.
// For an LBVH, a subtree of at most leaf_max triangles becomes one leaf over
// its sorted range (the triangles of a Karras subtree are contiguous in
// traversal order).  The reference traverses its binary tree directly
// (detail/bvh/intersect.inl:166-356); fewer, wider nodes halve the chain of
// dependent node fetches per ray on gfx950, where the traversal waits on
// memory latency.
//
// Built top-down one level per launch: level L reads the (binary node, quad
// index) pairs of its queue, writes those quads and appends its inner
// children to the next queue.  Quad indices come from one counter, a node's
// children taking consecutive indices.  Three queue counters rotate so that
// each launch reads one, appends to the next and clears the third.
#include "vr_internal.h"

#include <algorithm>

namespace vr {

namespace {

constexpr int BLOCK = 256;
constexpr unsigned GRID = 1024;
constexpr uint32_t NO_SOURCE = 0xFFFFFFFFu;  // unused child slot

struct child_ref {
    uint32_t e;    // pair entry
    uint32_t n;    // LBVH: triangles below (leaf_max > 0)
    uint32_t src;  // where its box lives: pair << 1 | side
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
    a.src = p << 1;
    a.lo = mk3(q0.x, q0.y, q0.z);
    a.hi = mk3(q0.w, q1.x, q1.y);
    b.e = e.y;
    b.src = (p << 1) | 1u;
    b.lo = mk3(q1.z, q1.w, q2.x);
    b.hi = mk3(q2.y, q2.z, q2.w);
    a.n = 1u;
    b.n = 1u;
    if (lbvh) {
        if (!(a.e & LEAF_BIT) && a.e < num_pairs) a.n = lbvh_count(pairs, a.e);
        if (!(b.e & LEAF_BIT) && b.e < num_pairs) b.n = lbvh_count(pairs, b.e);
    }
}

__global__ void k_collapse_init(uint32_t* ctr, uint2* q0)
{
    // ctr[0] quads allocated, ctr[1..3] queue counters
    ctr[0] = 1u;
    ctr[1] = 1u;
    ctr[2] = 0u;
    ctr[3] = 0u;
    q0[0] = make_uint2(0u, 0u);
}

__global__ __launch_bounds__(BLOCK) void k_collapse_level(const pair_node* __restrict__ pairs, uint32_t num_pairs,
                                                          uint32_t leaf_max, quad_node* __restrict__ quads,
                                                          uint4* __restrict__ srcs, uint32_t* __restrict__ ctr,
                                                          int level, const uint2* __restrict__ in,
                                                          uint2* __restrict__ out)
{
    const uint32_t* in_count = ctr + 1 + level % 3;
    uint32_t* out_count = ctr + 1 + (level + 1) % 3;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[1 + (level + 2) % 3] = 0u;
    const uint32_t cnt = *in_count;
    const bool lbvh = leaf_max > 0;
    // an inner child that can be opened: an inner pair above the leaf size
    auto openable = [&](const child_ref& c) {
        return !(c.e & LEAF_BIT) && c.e < num_pairs && (!lbvh || c.n > leaf_max);
    };
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < cnt; i += GRID * BLOCK) {
        const uint2 item = in[i];
        const uint32_t bin = item.x, qi = item.y;
        if (bin >= num_pairs || qi >= num_pairs) continue;  // corrupt queue: never write out of bounds
        child_ref c[4];
        load_pair(pairs, num_pairs, bin, lbvh, c[0], c[1]);
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
        // final entries: leaves as they are, small LBVH subtrees as leaf ranges,
        // the other inner children as new quads (consecutive indices)
        uint32_t ent[4];
        uint32_t n_inner = 0;
        for (int k = 0; k < nc; ++k) {
            const uint32_t e = c[k].e;
            ent[k] = e;
            if (e & LEAF_BIT) continue;
            if (lbvh && e < num_pairs && c[k].n <= leaf_max) {
                const uint32_t j = pairs[e].e[3];
                ent[k] = leaf_entry(min(e, j), c[k].n);
                continue;
            }
            ++n_inner;
        }
        if (n_inner > 0) {
            const uint32_t base = atomicAdd(&ctr[0], n_inner);
            const uint32_t qbase = atomicAdd(out_count, n_inner);
            uint32_t j = 0;
            for (int k = 0; k < nc; ++k) {
                if (ent[k] & LEAF_BIT) continue;
                out[qbase + j] = make_uint2(ent[k], base + j);
                ent[k] = base + j;
                ++j;
            }
        }
        quad_node qn;
        const float inf = __builtin_inff();
        // unused slots: lo = hi = +inf, a box the slab test never enters
        float lx[4] = {inf, inf, inf, inf}, ly[4] = {inf, inf, inf, inf}, lz[4] = {inf, inf, inf, inf};
        float hx[4] = {inf, inf, inf, inf}, hy[4] = {inf, inf, inf, inf}, hz[4] = {inf, inf, inf, inf};
        uint32_t ce[4] = {LEAF_BIT, LEAF_BIT, LEAF_BIT, LEAF_BIT};
        uint32_t cs[4] = {NO_SOURCE, NO_SOURCE, NO_SOURCE, NO_SOURCE};
        for (int k = 0; k < nc; ++k) {
            lx[k] = c[k].lo.x; ly[k] = c[k].lo.y; lz[k] = c[k].lo.z;
            hx[k] = c[k].hi.x; hy[k] = c[k].hi.y; hz[k] = c[k].hi.z;
            ce[k] = ent[k];
            cs[k] = c[k].src;
        }
        qn.lo_x = make_float4(lx[0], lx[1], lx[2], lx[3]);
        qn.lo_y = make_float4(ly[0], ly[1], ly[2], ly[3]);
        qn.lo_z = make_float4(lz[0], lz[1], lz[2], lz[3]);
        qn.hi_x = make_float4(hx[0], hx[1], hx[2], hx[3]);
        qn.hi_y = make_float4(hy[0], hy[1], hy[2], hy[3]);
        qn.hi_z = make_float4(hz[0], hz[1], hz[2], hz[3]);
        qn.child = make_uint4(ce[0], ce[1], ce[2], ce[3]);
        qn.meta = make_uint4((uint32_t)nc, 0u, 0u, 0u);
        quads[qi] = qn;
        srcs[qi] = make_uint4(cs[0], cs[1], cs[2], cs[3]);
    }
}

// Refit: the collapse is kept, every child box is copied again from the pair
// record that holds it (the binary boxes were just refitted).
__global__ __launch_bounds__(BLOCK) void k_refit_quads(const pair_node* __restrict__ pairs, uint32_t num_pairs,
                                                       quad_node* __restrict__ quads, const uint4* __restrict__ srcs,
                                                       const uint32_t* __restrict__ ctr)
{
    const uint32_t nq = min(ctr[0], num_pairs);
    for (uint32_t q = blockIdx.x * BLOCK + threadIdx.x; q < nq; q += gridDim.x * BLOCK) {
        const uint4 s4 = srcs[q];
        const uint32_t sv[4] = {s4.x, s4.y, s4.z, s4.w};
        const float inf = __builtin_inff();
        float v[6][4];
        for (int k = 0; k < 4; ++k) {
            const uint32_t src = sv[k];
            if (src == NO_SOURCE || (src >> 1) >= num_pairs) {
                for (int c = 0; c < 6; ++c) v[c][k] = inf;
                continue;
            }
            const float* f = pairs[src >> 1].f + (src & 1u) * 6;
            for (int c = 0; c < 6; ++c) v[c][k] = f[c];
        }
        quad_node& qn = quads[q];
        qn.lo_x = make_float4(v[0][0], v[0][1], v[0][2], v[0][3]);
        qn.lo_y = make_float4(v[1][0], v[1][1], v[1][2], v[1][3]);
        qn.lo_z = make_float4(v[2][0], v[2][1], v[2][2], v[2][3]);
        qn.hi_x = make_float4(v[3][0], v[3][1], v[3][2], v[3][3]);
        qn.hi_y = make_float4(v[4][0], v[4][1], v[4][2], v[4][3]);
        qn.hi_z = make_float4(v[5][0], v[5][1], v[5][2], v[5][3]);
    }
}

}  // namespace

int collapse_quads(bvh_state& s, uint32_t leaf_max, hipStream_t st)
{
    if (s.num_prims == 0 || s.num_pairs == 0) return VR_OK;  // empty, or the root is a leaf
    VR_CHECK(leaf_max <= MAX_LEAF_COUNT, VR_ERR_INVALID_ARGUMENT, "collapse: leaf_max above 32");
    const size_t np = s.num_pairs;
    if (!s.quads) VR_HIP_CHECK(hipMalloc(&s.quads, np * sizeof(quad_node)));
    // counters (4 words, padded to 256 bytes), the child box sources of every
    // quad, and two queues of np (bin, quad) pairs
    if (!s.collapse) VR_HIP_CHECK(hipMalloc(&s.collapse, 256 + np * sizeof(uint4) + 2 * np * sizeof(uint2)));
    uint32_t* ctr = s.collapse;
    uint4* srcs = reinterpret_cast<uint4*>(reinterpret_cast<char*>(s.collapse) + 256);
    uint2* q[2] = {reinterpret_cast<uint2*>(srcs + np), reinterpret_cast<uint2*>(srcs + np) + np};
    hipLaunchKernelGGL(k_collapse_init, dim3(1), dim3(1), 0, st, ctr, q[0]);
    // one quad level consumes at least one binary level
    const int levels = (int)std::max<uint32_t>(1u, s.depth);
    for (int L = 0; L < levels; ++L)
        hipLaunchKernelGGL(k_collapse_level, dim3(GRID), dim3(BLOCK), 0, st, s.pairs, (uint32_t)np, leaf_max, s.quads,
                           srcs, ctr, L, q[L & 1], q[(L + 1) & 1]);
    VR_HIP_CHECK(hipGetLastError());
    s.root = 0u;
    return VR_OK;
}

int refit_quads(bvh_state& s, hipStream_t st)
{
    if (s.num_prims == 0 || s.num_pairs == 0) return VR_OK;
    VR_CHECK(s.quads && s.collapse, VR_ERR_INVALID_ARGUMENT, "refit: the tree has no traversal nodes");
    const size_t np = s.num_pairs;
    const uint4* srcs = reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(s.collapse) + 256);
    const unsigned grid = std::max(1u, std::min(div_up(np, BLOCK), 8192u));
    hipLaunchKernelGGL(k_refit_quads, dim3(grid), dim3(BLOCK), 0, st, s.pairs, (uint32_t)np, s.quads, srcs,
                       s.collapse);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

}  // namespace vr
