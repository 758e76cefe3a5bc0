// collapse.hip -- all 4-wide traversal nodes of a tree in one fully
// parallel kernel (quad p for every pair p; the quad itself is made by
// make_quad, collapse.h).  LBVH builds make their quads in the bottom-up
// pass instead; this kernel serves rotated LBVHs, trees adopted from the
// host and refits.  Also the optional 8-wide nodes (float or 8-bit planes),
// made top down from the root, one launch per level.
#include "collapse.h"

#include <algorithm>
#include <vector>

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

// ---------------------------------------------------------------------------
// 8-wide nodes
// ---------------------------------------------------------------------------

__device__ inline float pow2f(int e) { return __int_as_float((e + 127) << 23); }

// Smallest exponent e in [-126, 119] with fmaf(255, 2^e, o) >= m: the step
// that lets eight bits reach from the node's lower corner o to its upper m.
__device__ inline int quant_exp(float o, float m)
{
    if (!(m > o)) return -126;
    int e = max(ilogbf(m - o) - 8, -126);
    while (e < 119 && fmaf(255.0f, pow2f(e), o) < m) ++e;
    return e;
}

// the largest q with fmaf(q, s, o) <= v (a lower plane, rounded down), and
// the smallest q with fmaf(q, s, o) >= v (an upper plane, rounded up): the
// traversal decodes a plane as fmaf(q, s, o), so the decoded box contains
// the exact one
__device__ inline uint32_t quant_lo(float v, float o, float s)
{
    int q = (int)fminf(fmaxf(floorf((v - o) / s), 0.0f), 255.0f);
    while (q > 0 && fmaf((float)q, s, o) > v) --q;
    while (q < 255 && fmaf((float)(q + 1), s, o) <= v) ++q;
    return (uint32_t)q;
}
__device__ inline uint32_t quant_hi(float v, float o, float s)
{
    int q = (int)fminf(fmaxf(ceilf((v - o) / s), 0.0f), 255.0f);
    while (q < 255 && fmaf((float)q, s, o) < v) ++q;
    while (q > 0 && fmaf((float)(q - 1), s, o) >= v) --q;
    return (uint32_t)q;
}

// eight bytes (slots 0-7) as two words
__device__ inline uint2 pack8(const uint32_t* b)
{
    return make_uint2(b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24), b[4] | (b[5] << 8) | (b[6] << 16) | (b[7] << 24));
}

__global__ void k_wide_init(uint32_t* counts, uint32_t n, uint32_t* frontier)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) counts[i] = i == 0 ? 1u : 0u;
    if (threadIdx.x == 0) frontier[0] = 0u;  // the root pair
}

// One level of the top-down 8-wide collapse: the nodes of this level's
// frontier (pair indices) are made, and their inner children are the next
// level's frontier.  Only nodes reachable from the root are written.
template <bool RANGES, int KIND>
__global__ __launch_bounds__(BLOCK) void k_wide_level(const pair_node* __restrict__ pairs, uint32_t num_pairs,
                                                      uint32_t leaf_max, const uint32_t* __restrict__ fin,
                                                      const uint32_t* __restrict__ nin, uint32_t* __restrict__ fout,
                                                      uint32_t* __restrict__ nout, void* __restrict__ nodes)
{
    const global_pairs src{pairs};
    const uint32_t n = *nin;
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const bool lbvh = leaf_max > 0;
    const float inf = __builtin_inff();
    // the trip count is uniform per block, so every lane of a wave takes
    // part in the wave's frontier append
    for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t inner = 0;  // slots holding an inner child with a node of its own
        uint32_t ce[8];
        if (i < n) {
            const uint32_t p = fin[i];
            child_ref c[8];
            const int nc = make_wide8<RANGES>(src, p, num_pairs, leaf_max, c);
            float bl[3][8], bh[3][8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const bool used = k < nc;
                f3 lo = mk3(inf, inf, inf), hi = mk3(inf, inf, inf);
                if (used) src.box(c[k].loc, lo, hi);
                bl[0][k] = lo.x;
                bl[1][k] = lo.y;
                bl[2][k] = lo.z;
                bh[0][k] = hi.x;
                bh[1][k] = hi.y;
                bh[2][k] = hi.z;
                uint32_t e = c[k].e;
                const bool in = !(e & LEAF_BIT) && e < num_pairs;
                if (in && lbvh && c[k].n <= leaf_max) e = leaf_entry(c[k].first, c[k].n);
                else if (used && in) inner |= 1u << k;
                ce[k] = used ? e : LEAF_BIT;
            }
            if constexpr (KIND == NODE_OCT) {
                oct_node o;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    quad_node& q = o.h[h];
                    const int b = 4 * h;
                    q.lo_x = make_float4(bl[0][b], bl[0][b + 1], bl[0][b + 2], bl[0][b + 3]);
                    q.lo_y = make_float4(bl[1][b], bl[1][b + 1], bl[1][b + 2], bl[1][b + 3]);
                    q.lo_z = make_float4(bl[2][b], bl[2][b + 1], bl[2][b + 2], bl[2][b + 3]);
                    q.hi_x = make_float4(bh[0][b], bh[0][b + 1], bh[0][b + 2], bh[0][b + 3]);
                    q.hi_y = make_float4(bh[1][b], bh[1][b + 1], bh[1][b + 2], bh[1][b + 3]);
                    q.hi_z = make_float4(bh[2][b], bh[2][b + 1], bh[2][b + 2], bh[2][b + 3]);
                    q.child = make_uint4(ce[b], ce[b + 1], ce[b + 2], ce[b + 3]);
                    q.meta = make_uint4((uint32_t)nc, 0u, 0u, 0u);
                }
                static_cast<oct_node*>(nodes)[p] = o;
            } else {
                float org[3], sc[3];
                uint32_t ql[3][8], qh[3][8];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float lo = bl[a][0], hi = bh[a][0];
#pragma unroll
                    for (int k = 1; k < 8; ++k)
                        if (k < nc) {
                            lo = fminf(lo, bl[a][k]);
                            hi = fmaxf(hi, bh[a][k]);
                        }
                    org[a] = lo;
                    sc[a] = pow2f(quant_exp(lo, hi));
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        ql[a][k] = k < nc ? quant_lo(bl[a][k], org[a], sc[a]) : 0u;
                        qh[a][k] = k < nc ? quant_hi(bh[a][k], org[a], sc[a]) : 0u;
                    }
                }
                octq_node o;
                o.org = make_float4(org[0], org[1], org[2], __uint_as_float((uint32_t)nc));
                o.scale = make_float4(sc[0], sc[1], sc[2], 0.0f);
                const uint2 lx = pack8(ql[0]), ly = pack8(ql[1]), lz = pack8(ql[2]);
                const uint2 hx = pack8(qh[0]), hy = pack8(qh[1]), hz = pack8(qh[2]);
                o.q[0] = make_uint4(lx.x, lx.y, ly.x, ly.y);
                o.q[1] = make_uint4(lz.x, lz.y, hx.x, hx.y);
                o.q[2] = make_uint4(hy.x, hy.y, hz.x, hz.y);
                o.child[0] = make_uint4(ce[0], ce[1], ce[2], ce[3]);
                o.child[1] = make_uint4(ce[4], ce[5], ce[6], ce[7]);
                o.pad = make_uint4(0u, 0u, 0u, 0u);
                static_cast<octq_node*>(nodes)[p] = o;
            }
        }
        // append the inner children to the next frontier: one atomic per
        // wave, each lane's run at its exclusive prefix
        const uint32_t m = (uint32_t)__popc(inner);
        uint32_t incl = m;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d);
            if (lane >= (uint32_t)d) incl += t;
        }
        const uint32_t total = __shfl(incl, 63);
        uint32_t base = 0;
        if (lane == 63 && total) base = atomicAdd(nout, total);
        base = __shfl(base, 63);
        uint32_t at = base + incl - m;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (inner & (1u << k)) fout[at++] = ce[k];
    }
}

template <bool RANGES, int KIND>
void launch_wide_levels(const bvh_state& s, uint32_t leaf_max, uint32_t levels, uint32_t* fa, uint32_t* fb,
                        uint32_t* counts, hipStream_t st)
{
    const unsigned grid = std::max(1u, std::min(div_up(s.num_pairs, BLOCK), 2048u));
    for (uint32_t l = 0; l < levels; ++l) {
        const uint32_t* fin = (l & 1u) ? fb : fa;
        uint32_t* fout = (l & 1u) ? fa : fb;
        hipLaunchKernelGGL((k_wide_level<RANGES, KIND>), dim3(grid), dim3(BLOCK), 0, st, s.pairs,
                           (uint32_t)s.num_pairs, leaf_max, fin, counts + l, fout, counts + l + 1, s.wide);
    }
}

}  // namespace

// The 8-wide nodes of s.wide_kind, top down from the root pair: one launch
// per level of the binary tree (which bounds the 8-wide levels).  The
// frontier lists and level counters are stream-ordered temporaries; the
// level counts come back to the host (the 8-wide depth sizes the traversal
// stack), so this waits for `st`.
int collapse_wide(bvh_state& s, hipStream_t st)
{
    s.wide_depth = 0;
    if (s.wide_kind == NODE_QUAD || s.num_prims == 0 || s.num_pairs == 0) return VR_OK;
    // an LBVH keeps an upper bound as its depth; the exact one is on the device
    uint32_t depth = s.depth;
    if (s.is_lbvh && !s.rotated && s.num_prims > 1) {
        VR_HIP_CHECK(hipMemcpyAsync(&depth, s.root_box + 6, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        VR_HIP_CHECK(hipStreamSynchronize(st));
        VR_CHECK(depth >= 2 && depth <= s.depth, VR_ERR_HIP, "collapse: unexpected LBVH depth");
    }
    const size_t node_bytes = s.wide_kind == NODE_OCT ? sizeof(oct_node) : sizeof(octq_node);
    const size_t bytes = s.num_pairs * node_bytes;
    if (s.wide_bytes < bytes) {
        if (s.wide) VR_HIP_CHECK(hipFreeAsync(s.wide, st));
        s.wide = nullptr;
        s.wide_bytes = 0;
        VR_HIP_CHECK(hipMallocAsync(&s.wide, bytes, st));
        s.wide_bytes = bytes;
    }
    const uint32_t leaf_max = s.wide_binary_leaves ? 0u : s.collapse_leaf_max;
    const bool ranges = leaf_max > 0 && s.is_lbvh && !s.rotated;
    const uint32_t levels = std::max(1u, depth);
    uint32_t* t = nullptr;
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&t), (2 * s.num_pairs + levels + 2) * sizeof(uint32_t), st));
    uint32_t* fa = t;
    uint32_t* fb = t + s.num_pairs;
    uint32_t* counts = t + 2 * s.num_pairs;
    hipLaunchKernelGGL(k_wide_init, dim3(1), dim3(BLOCK), 0, st, counts, levels + 2, fa);
    if (s.wide_kind == NODE_OCT) {
        if (ranges) launch_wide_levels<true, NODE_OCT>(s, leaf_max, levels, fa, fb, counts, st);
        else launch_wide_levels<false, NODE_OCT>(s, leaf_max, levels, fa, fb, counts, st);
    } else {
        if (ranges) launch_wide_levels<true, NODE_OCTQ>(s, leaf_max, levels, fa, fb, counts, st);
        else launch_wide_levels<false, NODE_OCTQ>(s, leaf_max, levels, fa, fb, counts, st);
    }
    VR_HIP_CHECK(hipGetLastError());
    std::vector<uint32_t> h(levels + 2);
    VR_HIP_CHECK(hipMemcpyAsync(h.data(), counts, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipFreeAsync(t, st));
    VR_HIP_CHECK(hipStreamSynchronize(st));
    uint32_t wd = 0;
    while (wd < levels && h[wd] > 0) ++wd;
    VR_CHECK(h[levels] == 0, VR_ERR_HIP, "collapse: 8-wide levels beyond the tree's depth");
    s.wide_depth = wd;
    return VR_OK;
}

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
