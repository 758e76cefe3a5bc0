// quality.hip -- the SAH cost of a BVH (sah_cost, detail/bvh/statistics.h:38-68)
// evaluated on the device over its binary nodes.
//
// Every pair record holds the boxes of its two children, so one pass over the
// pair records visits every node but the root once: inner children add their
// surface area to A_i, leaves to A_l and, times their primitive count, to
// A_l·N.  Each surface area is computed in float as the reference does
// (2 (s.x s.y + s.y s.z + s.z s.x)); the sums are FP64 (the reference sums in
// float, node by node).  A grid-stride loop of 64-byte records: HBM bound.
#include "vr_internal.h"

#include <algorithm>

namespace vr {

namespace {

constexpr int BLOCK = 256;

__device__ inline float surface_area6(float lx, float ly, float lz, float hx, float hy, float hz)
{
    const float sx = hx - lx, sy = hy - ly, sz = hz - lz;
    return 2.0f * (sx * sy + sy * sz + sz * sx);
}

__global__ void k_zero_f64(double* p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

__global__ __launch_bounds__(BLOCK) void k_sah_sums(const pair_node* __restrict__ pairs, uint32_t num_pairs,
                                                    double* __restrict__ sums)
{
    double ai = 0.0, al = 0.0, aln = 0.0;
    for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < num_pairs; p += gridDim.x * BLOCK) {
        const float4* q = reinterpret_cast<const float4*>(pairs + p);
        const float4 q0 = q[0], q1 = q[1], q2 = q[2];
        const uint4 e = *reinterpret_cast<const uint4*>(q + 3);
        const float sa0 = surface_area6(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y);
        const float sa1 = surface_area6(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w);
        const uint32_t ent[2] = {e.x, e.y};
        const float sa[2] = {sa0, sa1};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (ent[k] & LEAF_BIT) {
                const uint32_t cnt = ((ent[k] >> 26) & 31u) + 1u;
                al += (double)sa[k];
                aln += (double)(sa[k] * (float)cnt);
            } else {
                ai += (double)sa[k];
            }
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        ai += __shfl_xor(ai, off);
        al += __shfl_xor(al, off);
        aln += __shfl_xor(aln, off);
    }
    __shared__ double s[3][BLOCK / 64];
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        s[0][w] = ai;
        s[1][w] = al;
        s[2][w] = aln;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int k = 0; k < BLOCK / 64; ++k) t += s[threadIdx.x][k];
        atomicAdd(&sums[threadIdx.x], t);
    }
}

// ---------------------------------------------------------------------------
// tree rotations on the device (Kensler) for an LBVH
// ---------------------------------------------------------------------------

constexpr float ROT_CT = 1.2f;            // sah_cost's inner-node weight
constexpr float ROT_CI = 1.0f;            // and primitive weight
constexpr uint32_t ROT_MAX_DEPTH = 65;    // levels the traversal stack covers
constexpr uint32_t ROT_HBINS = 128;

__device__ inline float sa6(const float* f) { return surface_area6(f[0], f[1], f[2], f[3], f[4], f[5]); }

// triangles below a child entry: a leaf's count, or the pair's Karras range
// (exact for every pair the rotations leave alone; for a changed pair it
// stays above the leaf size, which is all the collapse asks of it)
__device__ inline uint32_t entry_count(const pair_node* __restrict__ pairs, uint32_t e)
{
    if (e & LEAF_BIT) return ((e >> 26) & 31u) + 1u;
    const uint32_t j = pairs[e].e[3];
    return max(e, j) - min(e, j) + 1u;
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// subtree height of every pair, and how many pairs have each height
__global__ __launch_bounds__(BLOCK) void k_pair_heights(const uint32_t* __restrict__ heights, uint32_t np,
                                                        uint32_t* __restrict__ hpair, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t s_hist[ROT_HBINS];
    for (uint32_t i = threadIdx.x; i < ROT_HBINS; i += BLOCK) s_hist[i] = 0u;
    __syncthreads();
    for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < np; p += gridDim.x * BLOCK) {
        const uint32_t h = min(1u + max(heights[2 * p], heights[2 * p + 1]), ROT_HBINS - 1u);
        hpair[p] = h;
        atomicAdd(&s_hist[h], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ROT_HBINS; i += BLOCK)
        if (s_hist[i]) atomicAdd(&hist[i], s_hist[i]);
}

// pairs grouped by height: a block counts its tile in LDS, reserves each
// height's run with one global atomic, then places its pairs (the global
// counters see one atomic per block and height, not one per pair)
constexpr uint32_t ORDER_PER_THREAD = 16;

__global__ __launch_bounds__(BLOCK) void k_order_by_height(const uint32_t* __restrict__ hpair, uint32_t np,
                                                           uint32_t* __restrict__ cursor, uint32_t* __restrict__ order)
{
    __shared__ uint32_t s_cnt[ROT_HBINS], s_base[ROT_HBINS];
    for (uint32_t i = threadIdx.x; i < ROT_HBINS; i += BLOCK) s_cnt[i] = 0u;
    __syncthreads();
    const uint32_t t0 = blockIdx.x * BLOCK * ORDER_PER_THREAD;
    uint32_t rank[ORDER_PER_THREAD];
#pragma unroll
    for (uint32_t k = 0; k < ORDER_PER_THREAD; ++k) {
        const uint32_t p = t0 + k * BLOCK + threadIdx.x;
        rank[k] = p < np ? atomicAdd(&s_cnt[hpair[p]], 1u) : 0u;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ROT_HBINS; i += BLOCK)
        s_base[i] = s_cnt[i] ? atomicAdd(&cursor[i], s_cnt[i]) : 0u;
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < ORDER_PER_THREAD; ++k) {
        const uint32_t p = t0 + k * BLOCK + threadIdx.x;
        if (p < np) order[s_base[hpair[p]] + rank[k]] = p;
    }
}

// SAH cost of the subtree below a child slot
__device__ inline float entry_cost(const float* __restrict__ cost, uint32_t e, const float* box)
{
    if (e & LEAF_BIT) return (ROT_CI * sa6(box)) * (float)(((e >> 26) & 31u) + 1u);
    return cost[e];
}

// cost(p) = C_t SA(p) + cost(child 0) + cost(child 1), for the pairs of one height
__global__ __launch_bounds__(BLOCK) void k_cost_level(const pair_node* __restrict__ pairs,
                                                      const uint32_t* __restrict__ order, uint32_t count,
                                                      float* __restrict__ cost)
{
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < count; i += gridDim.x * BLOCK) {
        const uint32_t p = order[i];
        const pair_node P = pairs[p];
        float u[6];
        for (int k = 0; k < 3; ++k) {
            u[k] = tmin_f(P.f[k], P.f[6 + k]);
            u[3 + k] = tmax_f(P.f[3 + k], P.f[9 + k]);
        }
        cost[p] = ROT_CT * sa6(u) + entry_cost(cost, P.e[0], P.f) + entry_cost(cost, P.e[1], P.f + 6);
    }
}

__global__ void k_frontier_root(uint32_t* front, uint32_t* counters)
{
    front[0] = 0u;  // the root pair
    counters[0] = 1u;
}

// One level of the top-down pass.  At each pair p of the level: exchange one
// child's subtree O with a grandchild X under its other child C (C's other
// child Y stays), when that lowers cost(C) + cost(O) -> cost(C') + cost(X),
// C' = {O, Y}.  Only pairs above the leaf size are opened and C' stays above
// it, so every subtree of at most leaf_max triangles is left whole (its
// triangles stay one contiguous range for the collapse), and O may go one
// level down only while its leaves stay within ROT_MAX_DEPTH levels.  The
// pairs of a level root disjoint subtrees, so their exchanges do not
// interact; the changed pairs' children go to the next level's list.
__global__ __launch_bounds__(BLOCK) void k_rotate_level(pair_node* __restrict__ pairs, uint32_t* __restrict__ heights,
                                                        float* __restrict__ cost, const uint32_t* __restrict__ front,
                                                        uint32_t* __restrict__ next, uint32_t* __restrict__ counters,
                                                        uint32_t level, uint32_t leaf_max,
                                                        uint32_t* __restrict__ result)
{
    const uint32_t cnt = counters[level];
    uint32_t depth = 0, rotated = 0;
    // whole waves run every iteration (the list appends are wave-wide)
    const uint32_t span = (cnt + 63u) & ~63u;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < span; i += gridDim.x * BLOCK) {
        const bool live = i < cnt;
        const uint32_t p = live ? front[i] : 0u;
        // the big child pairs go to the next level's list: one atomic per wave
        bool push[2] = {false, false};
        uint32_t pushed[2] = {0u, 0u};
        if (live) {
            pair_node P = pairs[p];
            uint32_t h[2] = {heights[2 * p], heights[2 * p + 1]};
            float c[2] = {entry_cost(cost, P.e[0], P.f), entry_cost(cost, P.e[1], P.f + 6)};
            uint32_t n[2] = {entry_count(pairs, P.e[0]), entry_count(pairs, P.e[1])};
            float best = c[0] + c[1];
            const float base = best;
            int choice = -1;
            float best_box[6];
            float best_cc = 0.0f;
            for (int s = 0; s < 2; ++s) {
                const uint32_t eC = P.e[s];
                if ((eC & LEAF_BIT) || n[s] <= leaf_max) continue;
                const int o = 1 - s;
                if (level + 2u + h[o] > ROT_MAX_DEPTH) continue;
                const pair_node& Cr = pairs[eC];
                for (int g = 0; g < 2; ++g) {
                    const uint32_t eY = Cr.e[1 - g];
                    const uint32_t nY = entry_count(pairs, eY);
                    if (n[o] + nY <= leaf_max) continue;
                    const float* bY = Cr.f + 6 * (1 - g);
                    float u[6];
                    for (int k = 0; k < 3; ++k) {
                        u[k] = tmin_f(P.f[6 * o + k], bY[k]);
                        u[3 + k] = tmax_f(P.f[6 * o + 3 + k], bY[3 + k]);
                    }
                    const float cc = ROT_CT * sa6(u) + c[o] + entry_cost(cost, eY, bY);
                    const float tot = cc + entry_cost(cost, Cr.e[g], Cr.f + 6 * g);
                    if (tot < best) {
                        best = tot;
                        choice = 2 * s + g;
                        best_cc = cc;
                        for (int k = 0; k < 6; ++k) best_box[k] = u[k];
                    }
                }
            }
            if (choice >= 0 && best < base * (1.0f - 1e-6f)) {
                const int s = choice >> 1, g = choice & 1, o = 1 - s;
                const uint32_t eC = P.e[s];
                pair_node Cr = pairs[eC];
                const uint32_t eO = P.e[o], eX = Cr.e[g];
                const uint32_t hO = h[o], hX = heights[2 * eC + g], hY = heights[2 * eC + 1 - g];
                float bX[6];
                for (int k = 0; k < 6; ++k) bX[k] = Cr.f[6 * g + k];
                // C' = {O, Y}: O takes X's slot
                Cr.e[g] = eO;
                for (int k = 0; k < 6; ++k) Cr.f[6 * g + k] = P.f[6 * o + k];
                pairs[eC] = Cr;
                heights[2 * eC + g] = hO;
                cost[eC] = best_cc;
                // p: C's slot gets C's new box, O's slot gets X
                P.e[o] = eX;
                for (int k = 0; k < 6; ++k) {
                    P.f[6 * s + k] = best_box[k];
                    P.f[6 * o + k] = bX[k];
                }
                pairs[p] = P;
                h[s] = 1u + max(hO, hY);
                h[o] = hX;
                heights[2 * p + s] = h[s];
                heights[2 * p + o] = h[o];
                n[o] = entry_count(pairs, eX);
                if (!(eO & LEAF_BIT)) pairs[eO].e[2] = (eC << 1) | (uint32_t)g;
                if (!(eX & LEAF_BIT)) pairs[eX].e[2] = (p << 1) | (uint32_t)o;
                float u[6];
                for (int k = 0; k < 3; ++k) {
                    u[k] = tmin_f(P.f[k], P.f[6 + k]);
                    u[3 + k] = tmax_f(P.f[3 + k], P.f[9 + k]);
                }
                cost[p] = ROT_CT * sa6(u) + best;
                ++rotated;
            }
            depth = max(depth, level + 1u + max(h[0], h[1]));
            for (int k = 0; k < 2; ++k) {
                push[k] = !(P.e[k] & LEAF_BIT) && n[k] > leaf_max;
                pushed[k] = P.e[k];
            }
        }
        const uint32_t mine = (uint32_t)push[0] + (uint32_t)push[1];
        const uint64_t m1 = __ballot(mine >= 1u), m2 = __ballot(mine == 2u);
        const uint32_t total = (uint32_t)(__popcll(m1) + __popcll(m2));
        const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        const uint32_t before = (uint32_t)(__popcll(m1 & below) + __popcll(m2 & below));
        const int leader = __ffsll((unsigned long long)m1) - 1;
        uint32_t base = 0;
        if (total && (int)lane == leader) base = atomicAdd(&counters[level + 1], total);
        if (total) base = __shfl(base, leader);
        uint32_t at = base + before;
        for (int k = 0; k < 2; ++k)
            if (push[k]) next[at++] = pushed[k];
    }
    // the level's depth and rotation count: one atomic per wave
    for (int off = 32; off >= 1; off >>= 1) {
        depth = max(depth, (uint32_t)__shfl_xor((int)depth, off));
        rotated += (uint32_t)__shfl_xor((int)rotated, off);
    }
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    if (lane == 0) {
        if (depth) atomicMax(&result[0], depth);
        if (rotated) atomicAdd(&result[1], rotated);
    }
}

}  // namespace

// One top-down pass of SAH-driven tree rotations over a freshly built LBVH
// (boxes and heights from the bottom-up pass): subtree costs bottom up by
// height, then the levels from the root.  Sets bs.depth to the resulting
// number of levels.  Synchronizes the stream (two small read-backs).
int lbvh_rotate(bvh_state& bs, uint32_t leaf_max, uint32_t* rotations, hipStream_t st)
{
    *rotations = 0;
    const size_t np = bs.num_pairs;
    if (bs.num_prims < 16 || np == 0) return VR_OK;
    VR_CHECK(np < (1ull << 31), VR_ERR_UNSUPPORTED, "lbvh_rotate: too many nodes");
    const size_t words = 4 * np + 2 * ROT_HBINS + (ROT_MAX_DEPTH + 2) + 2;
    uint32_t* tmp = nullptr;
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&tmp), words * 4 + np * sizeof(float), st));
    uint32_t* hpair = tmp;
    uint32_t* order = hpair + np;
    uint32_t* front0 = order + np;
    uint32_t* front1 = front0 + np;
    uint32_t* hist = front1 + np;
    uint32_t* cursor = hist + ROT_HBINS;
    uint32_t* counters = cursor + ROT_HBINS;
    uint32_t* result = counters + ROT_MAX_DEPTH + 2;  // [0] depth, [1] rotations
    float* cost = reinterpret_cast<float*>(tmp + words);
    const uint32_t small = 2 * ROT_HBINS + ROT_MAX_DEPTH + 4;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(BLOCK), 0, st, hist, small);
    const unsigned grid = std::max(1u, std::min(div_up(np, BLOCK), 2048u));
    hipLaunchKernelGGL(k_pair_heights, dim3(grid), dim3(BLOCK), 0, st, bs.heights, (uint32_t)np, hpair, hist);
    uint32_t h_hist[ROT_HBINS], h_off[ROT_HBINS];
    VR_HIP_CHECK(hipMemcpyAsync(h_hist, hist, sizeof(h_hist), hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipStreamSynchronize(st));
    uint32_t acc = 0, hmax = 0;
    for (uint32_t k = 0; k < ROT_HBINS; ++k) {
        h_off[k] = acc;
        acc += h_hist[k];
        if (h_hist[k]) hmax = k;
    }
    VR_CHECK(acc == np, VR_ERR_HIP, "lbvh_rotate: height histogram does not cover the nodes");
    VR_HIP_CHECK(hipMemcpyAsync(cursor, h_off, sizeof(h_off), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_order_by_height, dim3(div_up(np, (size_t)BLOCK * ORDER_PER_THREAD)), dim3(BLOCK), 0, st,
                       hpair, (uint32_t)np, cursor, order);
    for (uint32_t hh = 2; hh <= hmax; ++hh) {
        if (!h_hist[hh]) continue;
        const unsigned g = std::max(1u, std::min(div_up(h_hist[hh], BLOCK), 2048u));
        hipLaunchKernelGGL(k_cost_level, dim3(g), dim3(BLOCK), 0, st, bs.pairs, order + h_off[hh], h_hist[hh], cost);
    }
    hipLaunchKernelGGL(k_frontier_root, dim3(1), dim3(1), 0, st, front0, counters);
    for (uint32_t level = 0; level < ROT_MAX_DEPTH; ++level) {
        uint32_t* cur = (level & 1) ? front1 : front0;
        uint32_t* nxt = (level & 1) ? front0 : front1;
        hipLaunchKernelGGL(k_rotate_level, dim3(grid), dim3(BLOCK), 0, st, bs.pairs, bs.heights, cost, cur, nxt,
                           counters, level, leaf_max, result);
    }
    VR_HIP_CHECK(hipGetLastError());
    uint32_t h_res[2];
    VR_HIP_CHECK(hipMemcpyAsync(h_res, result, sizeof(h_res), hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipFreeAsync(tmp, st));
    VR_HIP_CHECK(hipStreamSynchronize(st));
    VR_CHECK(h_res[0] <= ROT_MAX_DEPTH, VR_ERR_HIP, "lbvh_rotate: tree deeper than the traversal stack");
    bs.depth = h_res[0];
    *rotations = h_res[1];
    return VR_OK;
}

int sah_cost_device(const bvh_state& s, float ci, float cl, float cp, double* out, hipStream_t st)
{
    VR_CHECK(s.num_prims > 0, VR_ERR_INVALID_ARGUMENT, "sah_cost: empty BVH");
    if (s.num_pairs == 0) {
        // a single leaf: A_i = 0, A_l = A_r
        *out = (double)cl + (double)cp * (double)s.num_prims;
        return VR_OK;
    }
    double* d = nullptr;
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&d), 3 * sizeof(double), st));
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, d, 3);
    const unsigned grid = std::max(1u, std::min(div_up(s.num_pairs, BLOCK), 4096u));
    hipLaunchKernelGGL(k_sah_sums, dim3(grid), dim3(BLOCK), 0, st, s.pairs, (uint32_t)s.num_pairs, d);
    double h[3];
    pair_node root;
    VR_HIP_CHECK(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipMemcpyAsync(&root, s.pairs + s.root, sizeof(root), hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipFreeAsync(d, st));
    VR_HIP_CHECK(hipStreamSynchronize(st));
    // the root: the union of its pair's child boxes, an inner node
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = tmin_f(root.f[k], root.f[6 + k]);
        hi[k] = tmax_f(root.f[3 + k], root.f[9 + k]);
    }
    const float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
    const float a_r = 2.0f * (sx * sy + sy * sz + sz * sx);
    const double a_i = h[0] + (double)a_r;
    *out = (double)ci * (a_i / a_r) + (double)cl * (h[1] / a_r) + (double)cp * (h[2] / a_r);
    return VR_OK;
}

}  // namespace vr
