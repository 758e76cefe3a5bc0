// lbvh.hip -- LBVH construction on gfx950.
//
// Same tree as the reference GPU builder (detail/bvh/lbvh.h:681-844), built
// MI355X-first:
//   1. centroid bounds: wave64 shuffle reduction, one ordered-int atomic per
//      component per block (the reference does 6 float CAS loops per
//      primitive on 6 global words, lbvh.h:63-91, 233-238)
//   2. 30-bit Morton codes (lbvh.h:242-269), bit-exact
//   3. stable LSD radix sort of (code, id), 4 passes of 8 bits with wave64
//      ballot-match ranking and LDS staging so the scatter writes runs of one
//      digit contiguously (the reference calls CUB's merge sort, lbvh.h:740-755;
//      a stable sort of ids 0..n-1 has exactly one result)
//   4. the Karras 2012 hierarchy (lbvh.h:120-185, 281-330) found bottom-up:
//      a climb from each leaf takes the parent on the side whose boundary
//      keys share the longer prefix, and the second thread to reach a node
//      computes its box (one atomic per node instead of lbvh.h:332-372's six
//      float CAS atomics per level per leaf) -- the same nodes, numbers and
//      exact min/max boxes -- fused with the collapse to the reference
//      bvh_node layout (lbvh.h:374-451), the copy of triangles into
//      traversal order and the 4-wide traversal nodes.
//      Nodes inside a 1024-leaf chunk meet in LDS, nodes inside a 65536-leaf
//      block in one workgroup, the rest in a single workgroup: no
//      device-scope fence is needed.
#include "collapse.h"
#include "vr_internal.h"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace vr {

namespace {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int RADIX = 256;
constexpr int ITEMS = 8;
constexpr int TILE = BLOCK * ITEMS;  // keys per sort tile
// leaves per workgroup of the first bottom-up pass
#ifndef VR_LBVH_CHUNK
#define VR_LBVH_CHUNK 1024
#endif
constexpr int CHUNK = VR_LBVH_CHUNK;
// split slots: no child has arrived / both have (LDS only); otherwise the
// first arrival's outer bound (< n)
constexpr uint32_t SLOT_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t SLOT_DONE = 0xFFFFFFFEu;

// float <-> order-preserving unsigned (for integer atomic min/max)
__device__ inline uint32_t f2o(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float o2f(uint32_t o)
{
    uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(u);
}

__device__ inline uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// number of set bits of m below this lane
__device__ inline uint32_t mbcnt64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// load a reference triangle (44-byte stride, 4-byte aligned)
struct tri_in {
    uint32_t geom_id, prim_id;
    f3 v1, e1, e2;
};
__device__ inline tri_in tri_from_words(const float* q)
{
    tri_in t;
    t.geom_id = __float_as_uint(q[0]);
    t.prim_id = __float_as_uint(q[1]);
    t.v1 = mk3(q[2], q[3], q[4]);
    t.e1 = mk3(q[5], q[6], q[7]);
    t.e2 = mk3(q[8], q[9], q[10]);
    return t;
}
__device__ inline tri_in load_tri(const vr_triangle* __restrict__ p, size_t i)
{
    return tri_from_words(reinterpret_cast<const float*>(p + i));
}

// get_bounds(triangle) (math/detail/triangle.inl:41-53)
__device__ inline void tri_box(const tri_in& t, f3& lo, f3& hi)
{
    lo = mk3(FLT_MAX, FLT_MAX, FLT_MAX);
    hi = mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    const f3 p0 = t.v1, p1 = add3(t.v1, t.e1), p2 = add3(t.v1, t.e2);
    lo = min3v(lo, p0); hi = max3v(hi, p0);
    lo = min3v(lo, p1); hi = max3v(hi, p1);
    lo = min3v(lo, p2); hi = max3v(hi, p2);
}

// aabb::center (math/detail/aabb.inl:56-61)
__device__ inline f3 tri_centroid(const tri_in& t)
{
    f3 lo, hi;
    tri_box(t, lo, hi);
    return muls3(add3(hi, lo), 0.5f);
}

// The centroid box in one pass over the triangles (the build reads the
// caller's array; index_bvh_t's copy, bvh.h:891-896, is not made): a block
// stages CB_U * CB_BLOCK triangles (11 words each) in LDS, all of a thread's
// 16-byte loads issued before the first store when the source is 16-byte
// aligned, and reduces their centroids; every block leaves its six extremes
// in `part` (no contended atomics), and k_bounds_final folds them.  min / max
// are exact, so any reduction order gives the same box.
constexpr int CB_BLOCK = 256;
constexpr int CB_U = 4;
constexpr int CB_GRID = 2048;
constexpr int TRI_WORDS = 11;

__device__ inline void block_minmax6(float v[6], float* __restrict__ out)
{
    for (int off = 32; off >= 1; off >>= 1) {
        for (int k = 0; k < 3; ++k) v[k] = fminf(v[k], __shfl_xor(v[k], off));
        for (int k = 3; k < 6; ++k) v[k] = fmaxf(v[k], __shfl_xor(v[k], off));
    }
    __shared__ float s[WAVES][6];
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; ++k) s[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float r = s[0][k];
        for (int q = 1; q < WAVES; ++q) r = k < 3 ? fminf(r, s[q][k]) : fmaxf(r, s[q][k]);
        out[k] = r;
    }
}

__global__ __launch_bounds__(CB_BLOCK) void k_centroid_bounds(const vr_triangle* __restrict__ src, size_t n,
                                                              int vec, float* __restrict__ part,
                                                              float* __restrict__ cen)
{
    constexpr uint32_t GROUP = CB_U * CB_BLOCK;            // triangles per step
    constexpr uint32_t VEC = GROUP * TRI_WORDS / 4;        // their 16-byte words
    constexpr uint32_t PER_THREAD = VEC / CB_BLOCK;        // 11: VEC is a multiple of CB_BLOCK
    __shared__ float s_tri[GROUP * TRI_WORDS];
    float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    const uint32_t tid = threadIdx.x;
    const size_t groups = (n + GROUP - 1) / GROUP;
    for (size_t g = blockIdx.x; g < groups; g += gridDim.x) {
        const size_t t0 = g * GROUP;
        const uint32_t cnt = (uint32_t)min((size_t)GROUP, n - t0);
        const float* sp = reinterpret_cast<const float*>(src + t0);
        if (vec && cnt == GROUP) {
            // a whole group starts on 16 bytes (GROUP * 44 is a multiple of 16)
            float4 x[PER_THREAD];
#pragma unroll
            for (uint32_t q = 0; q < PER_THREAD; ++q) x[q] = reinterpret_cast<const float4*>(sp)[tid + q * CB_BLOCK];
#pragma unroll
            for (uint32_t q = 0; q < PER_THREAD; ++q) reinterpret_cast<float4*>(s_tri)[tid + q * CB_BLOCK] = x[q];
        } else {
            for (uint32_t q = tid; q < cnt * TRI_WORDS; q += CB_BLOCK) s_tri[q] = sp[q];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < CB_U; ++u) {
            const uint32_t t = u * CB_BLOCK + tid;
            if (t < cnt) {
                const f3 m = tri_centroid(tri_from_words(s_tri + t * TRI_WORDS));
                v[0] = fminf(v[0], m.x); v[1] = fminf(v[1], m.y); v[2] = fminf(v[2], m.z);
                v[3] = fmaxf(v[3], m.x); v[4] = fmaxf(v[4], m.y); v[5] = fmaxf(v[5], m.z);
                if (cen) {
                    float* o = cen + 3 * (t0 + t);
                    o[0] = m.x;
                    o[1] = m.y;
                    o[2] = m.z;
                }
            }
        }
        __syncthreads();
    }
    block_minmax6(v, part + (size_t)blockIdx.x * 6);
}

// the centroid box from the blocks' extremes, as order-preserving words
__global__ __launch_bounds__(CB_BLOCK) void k_bounds_final(const float* __restrict__ part, uint32_t parts,
                                                           uint32_t* __restrict__ cb)
{
    float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t b = threadIdx.x; b < parts; b += CB_BLOCK) {
        for (int k = 0; k < 3; ++k) v[k] = fminf(v[k], part[(size_t)b * 6 + k]);
        for (int k = 3; k < 6; ++k) v[k] = fmaxf(v[k], part[(size_t)b * 6 + k]);
    }
    __shared__ float r[6];
    block_minmax6(v, r);
    __syncthreads();
    if (threadIdx.x < 6) cb[threadIdx.x] = f2o(r[threadIdx.x]);
}

// morton_encode3D (morton.h:33-47)
__device__ inline uint32_t separate_bits3(uint32_t n)
{
    n &= 0x000003FFu;
    n = (n ^ (n << 16)) & 0xFF0000FFu;
    n = (n ^ (n << 8)) & 0x0300F00Fu;
    n = (n ^ (n << 4)) & 0x030C30C3u;
    n = (n ^ (n << 2)) & 0x09249249u;
    return n;
}

// ---------------------------------------------------------------------------
// stable LSD radix sort, 8-bit digits
// ---------------------------------------------------------------------------

// The sort tile of this workgroup.  Workgroups go to the 8 XCDs round
// robin (workgroup b on XCD b mod 8), and the per-tile digit counts are
// digit-major words (hist[d * tiles + t]): with tile = b, the 16 tiles whose
// counts share a 64-byte line ran on 8 XCDs, and each XCD's L2 read and
// wrote that line on its own.  Here each XCD takes one contiguous range of
// tiles (a bijection of [0, tiles)), so a line's tiles share an L2.
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t tiles)
{
    constexpr uint32_t X = 8;
    const uint32_t xcd = b % X, local = b / X;
    const uint32_t base = tiles / X, rem = tiles % X;  // the first rem XCDs take base + 1 tiles
    return xcd < rem ? xcd * (base + 1) + local : rem * (base + 1) + (xcd - rem) * base + local;
}

// lanes of this wave holding the same 8-bit digit (among `valid` lanes)
__device__ inline uint64_t match_digit(uint32_t d, bool valid)
{
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// per-tile digit histogram, digit-major: hist[d * num_tiles + tile]
__global__ __launch_bounds__(BLOCK) void k_radix_hist(const uint32_t* __restrict__ keys, size_t n, int shift,
                                                      uint32_t* __restrict__ hist, uint32_t num_tiles)
{
    __shared__ uint32_t h[RADIX];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t tile = xcd_tile(blockIdx.x, num_tiles);
    const size_t base = (size_t)tile * TILE;
    // one LDS increment per key: a count does not depend on the order of its
    // increments (ranking the wave's equal digits first, as the scatter must,
    // cost eight ballots per key and left the pass VALU-bound)
    static_assert(ITEMS == 8, "two 16-byte loads per thread");
    if (base + TILE <= n && (reinterpret_cast<uintptr_t>(keys) & 15u) == 0) {
        const uint4* p = reinterpret_cast<const uint4*>(keys + base) + 2u * threadIdx.x;
        const uint4 a = p[0], b = p[1];
        const uint32_t k[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) atomicAdd(&h[(k[j] >> shift) & 0xFFu], 1u);
    } else {
        for (int k = 0; k < ITEMS; ++k) {
            const size_t idx = base + (size_t)k * BLOCK + threadIdx.x;
            if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 0xFFu], 1u);
        }
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * num_tiles + tile] = h[threadIdx.x];
}

// block-wide exclusive scan of one value per thread (BLOCK threads)
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp /* WAVES + 1 */, uint32_t& total)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x / 64;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) s_tmp[w] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (int q = 0; q < WAVES; ++q) {
        const uint32_t s = s_tmp[q];
        if (q < (int)w) wbase += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wbase + x - v;
}

__global__ __launch_bounds__(BLOCK) void k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                         const uint32_t* __restrict__ vals_in, int first_pass,
                                                         size_t n, int shift, const uint32_t* __restrict__ hist,
                                                         const uint32_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ totals, uint32_t num_tiles,
                                                         uint32_t* __restrict__ keys_out,
                                                         uint32_t* __restrict__ vals_out)
{
    __shared__ uint32_t s_key[TILE];
    __shared__ uint32_t s_val[TILE];
    __shared__ uint32_t s_dstart[RADIX];
    __shared__ uint32_t s_goff[RADIX];
    // per (item, wave) row and digit: the keys of that digit the wave holds
    // in that item, then (in place) their first slot in the tile
    __shared__ uint16_t s_cnt[ITEMS * WAVES][RADIX];
    __shared__ uint32_t s_tmp[WAVES + 1];

    const uint32_t tid = threadIdx.x, w = tid / 64;
    const uint32_t tile = xcd_tile(blockIdx.x, num_tiles);
    const size_t base = (size_t)tile * TILE;
    const uint32_t cnt = hist[(size_t)tid * num_tiles + tile];
    uint32_t total;
    // digit tid's first output slot: all smaller digits, then this digit in
    // the tiles before this one
    const uint32_t dbase = block_exclusive_scan(totals[tid], s_tmp, total);
    s_goff[tid] = dbase + offsets[(size_t)tid * num_tiles + tile];
    const uint32_t dstart = block_exclusive_scan(cnt, s_tmp, total);
    s_dstart[tid] = dstart;
#pragma unroll
    for (int r = 0; r < ITEMS * WAVES; ++r) s_cnt[r][tid] = 0;

    // every key and value of the tile in flight at once, before the barriers
    uint32_t keys[ITEMS], vals[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const size_t idx = base + (size_t)k * BLOCK + tid;
        const bool valid = idx < n;
        keys[k] = valid ? keys_in[idx] : 0u;
        vals[k] = valid ? (first_pass ? (uint32_t)idx : vals_in[idx]) : 0u;
    }
    __syncthreads();
    // tile order is (item, wave, lane): a key's slot is its digit's prefix
    // over the earlier (item, wave) rows plus its rank among the wave's
    // equal digits of the item (stable)
    uint32_t rank[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const bool valid = base + (size_t)k * BLOCK + tid < n;
        const uint32_t d = (keys[k] >> shift) & 0xFFu;
        const uint64_t m = match_digit(d, valid);
        rank[k] = mbcnt64(m);
        if (valid && rank[k] == 0) s_cnt[k * WAVES + w][d] = (uint16_t)__popcll(m);
    }
    __syncthreads();
    {
        uint32_t run = dstart;  // at most TILE
#pragma unroll
        for (int r = 0; r < ITEMS * WAVES; ++r) {
            const uint32_t c = s_cnt[r][tid];
            s_cnt[r][tid] = (uint16_t)run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (base + (size_t)k * BLOCK + tid < n) {
            const uint32_t d = (keys[k] >> shift) & 0xFFu;
            const uint32_t pos = min((uint32_t)s_cnt[k * WAVES + w][d] + rank[k], (uint32_t)TILE - 1u);
            s_key[pos] = keys[k];
            s_val[pos] = vals[k];
        }
    }
    __syncthreads();
    const size_t tile_n = std::min<size_t>(TILE, n - base);
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t p = (uint32_t)k * BLOCK + tid;
        if (p < tile_n) {
            const uint32_t key = s_key[p];
            const uint32_t d = (key >> shift) & 0xFFu;
            const size_t dst = (size_t)s_goff[d] + (p - s_dstart[d]);
            keys_out[dst] = key;
            vals_out[dst] = s_val[p];
        }
    }
}

// The spine of one radix pass: block d scans digit d's row of per-tile
// counts (hist[d * tiles + t]) into offs and leaves the row's total; the
// scatter adds the exclusive scan of the 256 totals (one per thread).
__global__ __launch_bounds__(BLOCK) void k_radix_spine(const uint32_t* __restrict__ hist, uint32_t tiles,
                                                       uint32_t* __restrict__ offs, uint32_t* __restrict__ totals)
{
    __shared__ uint32_t s_tmp[WAVES + 1];
    const uint32_t* row = hist + (size_t)blockIdx.x * tiles;
    uint32_t* orow = offs + (size_t)blockIdx.x * tiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < tiles; base += TILE) {
        const uint32_t b = base + threadIdx.x * ITEMS;
        uint32_t v[ITEMS];
        uint32_t sum = 0;
        for (int k = 0; k < ITEMS; ++k) {
            v[k] = b + k < tiles ? row[b + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = block_exclusive_scan(sum, s_tmp, total) + carry;
        for (int k = 0; k < ITEMS; ++k) {
            if (b + k < tiles) orow[b + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// assign_morton_codes (detail/bvh/lbvh.h:242-269): centroid relative to the
// centroid box, scaled by 1024, clamped to [0, 1023] with the ternary min/max
// (a degenerate axis gives 0/0 = NaN, which the clamp maps to 1023), then
// truncated and interleaved.
//
// One block per sort tile (TILE keys), staged through LDS in BLOCK-triangle
// chunks (16-byte loads when the triangles are 16-byte aligned).  In a build
// the same kernel leaves the first radix pass's per-tile digit counts and
// clears the bottom-up pass's ready-list counters and the split slots on the
// chunks' edges.
struct morton_extras {
    uint32_t* hist;         // digit-0 counts per tile (first radix pass), or null
    uint32_t tiles;
    uint32_t* edges;        // split slots: those of splits CHUNK * k - 1 set empty, or null
    uint32_t* counts;       // ncounts words cleared
    uint32_t ncounts;
};

__global__ __launch_bounds__(BLOCK) void k_morton(const vr_triangle* __restrict__ prims, const float* __restrict__ cen,
                                                  size_t n, int vec, const uint32_t* __restrict__ cb,
                                                  uint32_t* __restrict__ codes, morton_extras x)
{
    __shared__ float s_tri[BLOCK * TRI_WORDS];
    __shared__ uint32_t h[RADIX];
    const uint32_t tid = threadIdx.x;
    h[tid] = 0u;
    const f3 lo = mk3(o2f(cb[0]), o2f(cb[1]), o2f(cb[2]));
    const f3 hi = mk3(o2f(cb[3]), o2f(cb[4]), o2f(cb[5]));
    const f3 center = muls3(add3(hi, lo), 0.5f);
    const f3 size = sub3(hi, lo);
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);  // (gridDim.x = x.tiles when x.hist is set)
    const size_t base = (size_t)tile * TILE;
    for (int k = 0; k < ITEMS; ++k) {
        const size_t t0 = base + (size_t)k * BLOCK;
        if (t0 >= n) break;  // uniform over the block
        const uint32_t cnt = (uint32_t)min((size_t)BLOCK, n - t0);
        if (!cen) {
            const float* sp = reinterpret_cast<const float*>(prims + t0);
            if (vec && cnt == BLOCK) {
                for (uint32_t q = tid; q < BLOCK * TRI_WORDS / 4; q += BLOCK)
                    reinterpret_cast<float4*>(s_tri)[q] = reinterpret_cast<const float4*>(sp)[q];
            } else {
                for (uint32_t q = tid; q < cnt * TRI_WORDS; q += BLOCK) s_tri[q] = sp[q];
            }
        }
        __syncthreads();
        const bool valid = tid < cnt;
        const size_t i = t0 + tid;
        uint32_t code = 0u;
        if (valid) {
            // the centroid the bounds pass kept (the same float operations),
            // or from the triangle
            f3 c = cen ? mk3(cen[3 * i], cen[3 * i + 1], cen[3 * i + 2])
                       : tri_centroid(tri_from_words(s_tri + tid * TRI_WORDS));
            c = sub3(c, center);
            c = add3(c, muls3(size, 0.5f));
            c = f3{c.x / size.x, c.y / size.y, c.z / size.z};
            c = min3v(max3v(muls3(c, 1024.0f), mk3(0.0f, 0.0f, 0.0f)), mk3(1023.0f, 1023.0f, 1023.0f));
            code = separate_bits3((uint32_t)c.x) | (separate_bits3((uint32_t)c.y) << 1) |
                   (separate_bits3((uint32_t)c.z) << 2);
            codes[i] = code;
            if (x.edges && i % CHUNK == CHUNK - 1 && i + 1 < n) x.edges[i] = SLOT_EMPTY;
        }
        if (x.hist && valid) atomicAdd(&h[code & 0xFFu], 1u);
        __syncthreads();
    }
    if (x.hist) x.hist[(size_t)tid * x.tiles + tile] = h[tid];
    for (size_t c = (size_t)blockIdx.x * BLOCK + tid; c < x.ncounts; c += (size_t)gridDim.x * BLOCK) x.counts[c] = 0u;
}

// ---------------------------------------------------------------------------
// Hierarchy and boxes, bottom-up (detail/bvh/lbvh.h:120-185, 281-372)
// ---------------------------------------------------------------------------
//
// The binary radix tree over the sorted keys, with equal codes ordered by
// index (the reference's delta: common-prefix length of the codes, or 32 +
// that of the indices when the codes are equal), is unique.  The reference
// finds each node's range and split top-down (determine_range and a binary
// search, one thread per node), then climbs from the leaves with float
// atomics for the boxes.  Here one climb from the leaves does both: a node
// covering [l, r] has its parent on the side where its boundary keys share
// the longer prefix, delta(l - 1, l) against delta(r, r + 1) (the two are
// never equal for a node of the tree); the parent's split is then r (this
// node is its left child) or l - 1 (its right child).  The first child to
// reach a split leaves its box, height and outer bound there; the second
// completes the parent.  The reference's node numbering follows from the
// side alone: a left child is node `split`, which is its own r, a right child
// node split + 1, its own l, and the root node 0, so every record, entry and
// box equals the top-down construction's.
//
// Every box is a min / max union computed once, so the boxes are the
// reference's exactly.  Three passes, each finishing the nodes inside blocks
// of a growing size with workgroup-scope ordering only (on gfx950 a
// device-scope fence writes back the XCD's L2, far too slow to pay per node):
//   1. k_bottom_up_local: one workgroup per CHUNK leaves; the splits whose
//      two children both lie in the chunk meet in LDS;
//   2. k_bottom_up_mid: one workgroup per SUPER leaves, in global memory;
//   3. k_bottom_up_top: one workgroup for the rest.
// A climb stops at a split outside its pass's block: it leaves its data in
// the split's global stage and exchanges its bound into the split's global
// slot; the second arrival there appends the split to the next pass's ready
// list (bucketed by the block the split is inside of).  A node's parent link
// (entry word 2) is known only when its parent's own side is: pass 1 sets it
// in LDS before writing the chunk's records, a later pass writes it into
// the child's record.  Kernel boundaries make every earlier write visible to
// the next pass.
constexpr int TOP_BLOCK = 1024;
// subtrees of at most this many triangles become one traversal leaf
#ifndef VR_LBVH_LEAF_MAX
#define VR_LBVH_LEAF_MAX 2
#endif
constexpr uint32_t LBVH_LEAF_MAX = VR_LBVH_LEAF_MAX;
constexpr uint32_t SUPER = 64u * CHUNK;
// pairs completed by one workgroup of the mid or top pass whose quads are
// made after its climbs, one per thread
constexpr uint32_t DONE_CAP = 4096;

// One split's hand-over in global memory: each child's box, height and
// outer bound (the left child's l, the right child's r) on its side, and the
// prefix length at that outer edge (delta(l - 1) for the left child,
// delta(r) for the right one: the two the parent's own side is decided by),
// so a later pass climbs without reading keys.
struct alignas(64) split_stage {
    float box[12];
    uint32_t hd[2];  // height | (outer delta + 1) << 8
    uint32_t bound[2];
};

__device__ inline uint32_t pack_hd(uint32_t h, int d) { return h | ((uint32_t)(d + 1) << 8); }
__device__ inline uint32_t hd_height(uint32_t hd) { return hd & 0xFFu; }
__device__ inline int hd_delta(uint32_t hd) { return (int)(hd >> 8) - 1; }
static_assert(sizeof(split_stage) == 64, "split_stage is one 64-byte record");

// common-prefix length of the sorted keys k and k + 1 (codes a and b)
__device__ inline int key_delta(uint32_t a, uint32_t b, uint32_t k)
{
    const uint32_t x = a ^ b;
    return x ? (int)__clz((int)x) : 32 + (int)__clz((int)(k ^ (k + 1u)));
}

__device__ inline uint32_t entry_index(uint32_t e)
{
    return e & LEAF_BIT ? e & 0x03FFFFFFu : e;
}

__device__ inline void store_box(float* f, f3 lo, f3 hi)
{
    f[0] = lo.x; f[1] = lo.y; f[2] = lo.z;
    f[3] = hi.x; f[4] = hi.y; f[5] = hi.z;
}

__device__ inline void store_ref_node(vr_bvh_node* o, f3 lo, f3 hi, uint32_t first, uint32_t num)
{
    // bvh_node::set_inner / set_leaf (bvh.h:122-138); axis = sign = 0
    uint4* q = reinterpret_cast<uint4*>(o);
    q[0] = make_uint4(__float_as_uint(lo.x), __float_as_uint(lo.y), __float_as_uint(lo.z), __float_as_uint(hi.x));
    q[1] = make_uint4(__float_as_uint(hi.y), __float_as_uint(hi.z), first, num);
}

// the reference-layout nodes of pair p's two children (slots 1 + 2p + side):
// an inner child c points at its own children (1 + 2c), a leaf at its
// primitive's original index
__device__ inline void store_ref_children(vr_bvh_node* ref_nodes, const int32_t* __restrict__ sorted_ids, uint32_t p,
                                          const float* b, uint32_t e0, uint32_t e1)
{
    for (int side = 0; side < 2; ++side) {
        const uint32_t e = side ? e1 : e0;
        const float* c = b + 6 * side;
        const bool leaf = e & LEAF_BIT;
        store_ref_node(&ref_nodes[1 + 2 * p + side], mk3(c[0], c[1], c[2]), mk3(c[3], c[4], c[5]),
                       leaf ? (uint32_t)sorted_ids[entry_index(e)] : 1u + 2u * e, leaf ? 1u : 0u);
    }
}

// one pair record: child boxes b[12] and entries
__device__ inline void store_pair(pair_node* __restrict__ pairs, uint32_t p, const float* b, uint4 e)
{
    float4* dst = reinterpret_cast<float4*>(pairs[p].f);
    dst[0] = make_float4(b[0], b[1], b[2], b[3]);
    dst[1] = make_float4(b[4], b[5], b[6], b[7]);
    dst[2] = make_float4(b[8], b[9], b[10], b[11]);
    *reinterpret_cast<uint4*>(pairs[p].e) = e;
}

__device__ inline void finish_root(float* root_box, vr_bvh_node* ref_nodes, uint32_t* depth_out, f3 lo, f3 hi,
                                   uint32_t h)
{
    store_box(root_box, lo, hi);
    if (ref_nodes) store_ref_node(&ref_nodes[0], lo, hi, 1u, 0u);
    *depth_out = h;
}

// the split's node is complete (both children arrived) and goes to the pass
// that owns the block the split is inside of: its SUPER block's list, or the
// top list for a split on the edge of two SUPER blocks
__device__ inline void push_ready(uint32_t s, uint32_t* __restrict__ ready, uint32_t* __restrict__ bucket_count,
                                  uint32_t* __restrict__ top, uint32_t* __restrict__ top_count)
{
    const uint32_t b = s / SUPER;
    if ((s + 1u) / SUPER == b)
        ready[(size_t)b * SUPER + atomicAdd(&bucket_count[b], 1u)] = s;
    else
        top[atomicAdd(top_count, 1u)] = s;
}

// the chunk's pair records in LDS: entries by pair index, child boxes by the
// pair's split (the index in its first entry)
struct lds_split_pairs {
    const float* stage_;
    const uint4* ent_;
    uint32_t c0;
    __device__ uint4 ent(uint32_t q) const { return ent_[q - c0]; }
    __device__ void box(uint32_t loc, f3& lo, f3& hi) const
    {
        const uint32_t s = entry_index(ent_[(loc >> 1) - c0].x);
        const float* b = stage_ + (size_t)(s - c0) * 12 + 6 * (loc & 1u);
        lo = mk3(b[0], b[1], b[2]);
        hi = mk3(b[3], b[4], b[5]);
    }
};

__global__ __launch_bounds__(CHUNK) void k_bottom_up_local(const vr_triangle* __restrict__ prims, size_t n,
                                                           const uint32_t* __restrict__ codes,
                                                           const int32_t* __restrict__ sorted_ids,
                                                           pair_node* __restrict__ pairs,
                                                           split_stage* __restrict__ gstage,
                                                           uint32_t* __restrict__ gslot,
                                                           uint32_t* __restrict__ heights, tri_node* __restrict__ tris,
                                                           vr_bvh_node* __restrict__ ref_nodes,
                                                           float* __restrict__ root_box, uint32_t* __restrict__ depth_out,
                                                           uint32_t* __restrict__ ready, uint32_t* __restrict__ bucket_count,
                                                           uint32_t* __restrict__ top, uint32_t* __restrict__ top_count,
                                                           int slot_heights, quad_node* __restrict__ quads,
                                                           uint32_t leaf_max)
{
    // by split - c0: child boxes, heights, slot
    __shared__ float s_stage[CHUNK * 12];
    __shared__ uint8_t s_h[CHUNK * 2];
    __shared__ uint32_t s_slot[CHUNK];
    // by pair index - c0: entries of the pairs completed here (w = SLOT_EMPTY:
    // not completed here)
    __shared__ uint4 s_ent[CHUNK];
    // keys c0 - 1 .. c1
    __shared__ uint32_t s_code[CHUNK + 2];
    const uint32_t tid = threadIdx.x;
    const uint32_t N = (uint32_t)n;
    const uint32_t c0 = blockIdx.x * (uint32_t)CHUNK;
    const uint32_t c1 = min(N, c0 + (uint32_t)CHUNK);
    const uint32_t j = c0 + tid;
    // the leaf's triangle, requested before the chunk's keys and the barrier
    // so that the two dependent reads overlap them
    const int32_t id = j < N ? sorted_ids[j] : 0;
    const tri_in t = load_tri(prims, (size_t)id);
    s_slot[tid] = SLOT_EMPTY;
    s_ent[tid] = make_uint4(0u, 0u, SLOT_EMPTY, SLOT_EMPTY);
    for (uint32_t t = tid; t < CHUNK + 2; t += CHUNK) {
        const int64_t g = (int64_t)c0 - 1 + t;
        if (g >= 0 && g < (int64_t)N) s_code[t] = codes[g];
    }
    __syncthreads();
    // keys k and k + 1 for c0 - 1 <= k < c1
    auto delta = [&](uint32_t k) { return key_delta(s_code[k + 1 - c0], s_code[k + 2 - c0], k); };
    // the leaf and its climb; every thread then meets at the barrier below
    auto leaf_and_climb = [&]() {
        tri_node tn;
        tn.a = make_float4(t.v1.x, t.v1.y, t.v1.z, __uint_as_float(t.prim_id));
        tn.b = make_float4(t.e1.x, t.e1.y, t.e1.z, __uint_as_float(t.geom_id));
        tn.c = make_float4(t.e2.x, t.e2.y, t.e2.z, __uint_as_float((uint32_t)id));
        tris[j] = tn;
        f3 lo, hi;
        tri_box(t, lo, hi);
        if (N == 1) {
            store_box(root_box, lo, hi);
            if (ref_nodes) store_ref_node(&ref_nodes[0], lo, hi, (uint32_t)id, 1u);
            *depth_out = 1;
            return;
        }
        uint32_t l = j, r = j, h = 1, split = 0, e0 = 0, e1 = 0;
        bool inner = false;
        // a radix tree over n < 2^26 keys of 30 bits has fewer than 64 levels;
        // the bound only guards against a corrupt key array
        for (int level = 0; level < 128; ++level) {
            const bool root = l == 0 && r == N - 1;
            bool right = false;
            int dl = -1, dr = -1;
            if (!root) {
                dl = l > 0 ? delta(l - 1) : -1;
                dr = r + 1 < N ? delta(r) : -1;
                right = dr > dl;  // parent's split r: this node is its left child
            }
            if (inner) {
                // this node's number, its entries, and its children's links
                const uint32_t p = root ? 0u : (right ? r : l);
                s_ent[p - c0] = make_uint4(e0, e1, SLOT_EMPTY, p == l ? r : l);
                if (!(e0 & LEAF_BIT)) s_ent[split - c0].z = (p << 1) | 0u;
                if (!(e1 & LEAF_BIT)) s_ent[split + 1 - c0].z = (p << 1) | 1u;
            }
            if (root) {
                finish_root(root_box, ref_nodes, depth_out, lo, hi, h);
                return;
            }
            const uint32_t s = right ? r : l - 1;
            const uint32_t side = right ? 0u : 1u;
            const uint32_t bound = right ? l : r;
            if (s < c0 || s + 2 > c1) {
                // the parent spans the chunk's edge: hand over in global memory
                split_stage* g = gstage + s;
                store_box(g->box + 6 * side, lo, hi);
                g->hd[side] = pack_hd(h, right ? dl : dr);
                g->bound[side] = bound;
                if (__hip_atomic_exchange(&gslot[s], bound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
                    SLOT_EMPTY)
                    push_ready(s, ready, bucket_count, top, top_count);
                return;
            }
            const uint32_t k = s - c0;
            store_box(&s_stage[k * 12 + side * 6], lo, hi);
            s_h[2 * k + side] = (uint8_t)h;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            const uint32_t other = atomicExch(&s_slot[k], bound);
            if (other == SLOT_EMPTY) return;  // first: the sibling's thread completes the parent
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            s_slot[k] = SLOT_DONE;
            if (right) r = other;
            else l = other;
            float b[12];
            for (int q = 0; q < 12; ++q) b[q] = s_stage[k * 12 + q];
            lo = min3v(mk3(b[0], b[1], b[2]), mk3(b[6], b[7], b[8]));
            hi = max3v(mk3(b[3], b[4], b[5]), mk3(b[9], b[10], b[11]));
            h = max((uint32_t)s_h[2 * k], (uint32_t)s_h[2 * k + 1]) + 1u;
            split = s;
            e0 = l == s ? leaf_entry(s, 1u) : s;
            e1 = s + 1 == r ? leaf_entry(s + 1, 1u) : s + 1;
            inner = true;
        }
    };
    if (j < N) leaf_and_climb();
    __syncthreads();
    if (N < 2) return;
    // split c0 + tid inside the chunk: a lone arrival goes to global memory
    // for the pass that completes its sibling; the slot is set either way
    if (tid + 2 <= c1 - c0) {
        const uint32_t s = c0 + tid, v = s_slot[tid];
        uint32_t g = SLOT_EMPTY;
        if (v != SLOT_EMPTY && v != SLOT_DONE) {
            const uint32_t side = v > s ? 1u : 0u;  // a right child's bound is its r > s
            split_stage* st = gstage + s;
            for (int q = 0; q < 6; ++q) st->box[6 * side + q] = s_stage[tid * 12 + 6 * side + q];
            // the outer edge's prefix length: delta(l - 1) of a left child
            // (l = v), delta(r) of a right one (r = v); both keys are in s_code
            const int od = side ? (v + 1 < N ? delta(v) : -1) : (v > 0 ? delta(v - 1) : -1);
            st->hd[side] = pack_hd(s_h[2 * tid + side], od);
            st->bound[side] = v;
            g = v;
        }
        gslot[s] = g;
    }
    // the records of the pairs completed here: 16-byte word f of the
    // chunk's records from thread f mod CHUNK, so each store instruction
    // writes a contiguous kilobyte per wave
    for (uint32_t f = tid; f < 4u * CHUNK; f += CHUNK) {
        const uint32_t qq = f >> 2, part = f & 3u;
        if (c0 + qq + 1 >= N) break;
        const uint4 e = s_ent[qq];
        if (e.w == SLOT_EMPTY) continue;
        uint4 v = e;
        if (part < 3) {
            const float* b = s_stage + (entry_index(e.x) - c0) * 12 + 4 * part;
            v = make_uint4(__float_as_uint(b[0]), __float_as_uint(b[1]), __float_as_uint(b[2]),
                           __float_as_uint(b[3]));
        }
        reinterpret_cast<uint4*>(pairs)[(size_t)(c0 + qq) * 4 + part] = v;
    }
    // then per pair: slot heights, reference nodes, and the quads (every
    // pair below a completed pair is in the chunk)
    const uint32_t q = j;
    if (q + 1 >= N) return;
    const uint4 e = s_ent[tid];
    if (e.w == SLOT_EMPTY) return;
    const uint32_t k = entry_index(e.x) - c0;
    if (slot_heights) {
        heights[2 * q] = s_h[2 * k];
        heights[2 * q + 1] = s_h[2 * k + 1];
    }
    if (ref_nodes) store_ref_children(ref_nodes, sorted_ids, q, s_stage + k * 12, e.x, e.y);
    if (quads && pair_gets_quad(q, e.w, leaf_max))
        make_quad<true>(lds_split_pairs{s_stage, s_ent, c0}, q, N - 1, leaf_max, quads);
}

// quads of the pairs a workgroup completed (the first DONE_CAP listed in
// LDS, the rest in its global overflow list), in parallel after its climbs:
// the barrier orders every record the workgroup wrote before them
__device__ inline void quads_after_climbs(const pair_node* __restrict__ pairs, uint32_t npairs, uint32_t leaf_max,
                                          quad_node* __restrict__ quads, const uint32_t* s_done,
                                          const uint32_t* s_ndone, const uint32_t* ovf)
{
    __syncthreads();
    if (!quads) return;
    const uint32_t count = *s_ndone;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x)
        make_quad<true>(global_pairs{pairs}, i < DONE_CAP ? s_done[i] : ovf[i - DONE_CAP], npairs, leaf_max, quads);
}

// Complete the node at split s, whose two children have arrived, and climb
// until an arrival is first at its split, the root is done, or (when the
// block [c0, c1) is not the whole tree) the next split is not inside the
// block: that arrival is exchanged for the next pass and, as the second,
// appended to the top list.
__device__ inline void climb_split(uint32_t s, uint32_t N, uint32_t c0, uint32_t c1, bool whole,
                                   const int32_t* __restrict__ sorted_ids,
                                   pair_node* __restrict__ pairs, split_stage* __restrict__ gstage,
                                   uint32_t* __restrict__ gslot, uint32_t* __restrict__ heights, int slot_heights,
                                   vr_bvh_node* __restrict__ ref_nodes, float* __restrict__ root_box,
                                   uint32_t* __restrict__ depth_out, uint32_t* __restrict__ top,
                                   uint32_t* __restrict__ top_count, quad_node* __restrict__ quads,
                                   uint32_t leaf_max, uint32_t* s_done, uint32_t* s_ndone, uint32_t* ovf)
{
    // after a second arrival: this thread's child (box, height, outer prefix
    // length, side) and the parent's range from the exchanged bound, so only
    // the sibling's half of the stage is read
    bool carried = false;
    f3 clo = {}, chi = {};
    uint32_t chd = 0, cside = 0, cl = 0, cr = 0;
    for (int level = 0; level < 128 && s + 1 < N; ++level) {
        const split_stage* g = gstage + s;
        float b[12];
        uint32_t hd0, hd1, l, r;
        if (!carried) {
            const float4* gb = reinterpret_cast<const float4*>(g->box);
            const float4 q0 = gb[0], q1 = gb[1], q2 = gb[2];
            const uint4 meta = *reinterpret_cast<const uint4*>(g->hd);
            b[0] = q0.x; b[1] = q0.y; b[2] = q0.z; b[3] = q0.w;
            b[4] = q1.x; b[5] = q1.y; b[6] = q1.z; b[7] = q1.w;
            b[8] = q2.x; b[9] = q2.y; b[10] = q2.z; b[11] = q2.w;
            hd0 = meta.x;
            hd1 = meta.y;
            l = meta.z;
            r = meta.w;
        } else {
            const uint32_t os = 1u - cside;
            const float2* ob = reinterpret_cast<const float2*>(g->box + 6 * os);
            const float2 a = ob[0], c = ob[1], d = ob[2];
            const uint32_t ohd = g->hd[os];
            const float mine[6] = {clo.x, clo.y, clo.z, chi.x, chi.y, chi.z};
            const float theirs[6] = {a.x, a.y, c.x, c.y, d.x, d.y};
            for (int q = 0; q < 6; ++q) {
                b[q] = cside ? theirs[q] : mine[q];
                b[6 + q] = cside ? mine[q] : theirs[q];
            }
            hd0 = cside ? ohd : chd;
            hd1 = cside ? chd : ohd;
            l = cl;
            r = cr;
        }
        const uint32_t h0 = hd_height(hd0), h1 = hd_height(hd1);
        const f3 lo = min3v(mk3(b[0], b[1], b[2]), mk3(b[6], b[7], b[8]));
        const f3 hi = max3v(mk3(b[3], b[4], b[5]), mk3(b[9], b[10], b[11]));
        const uint32_t h = max(h0, h1) + 1u;
        const uint32_t e0 = l == s ? leaf_entry(s, 1u) : s;
        const uint32_t e1 = s + 1 == r ? leaf_entry(s + 1, 1u) : s + 1;
        const bool root = l == 0 && r == N - 1;
        // the prefix lengths at the node's edges: delta(l - 1) is the left
        // child's outer one, delta(r) the right child's
        const int dl = hd_delta(hd0), dr = hd_delta(hd1);
        const bool right = !root && dr > dl;
        const uint32_t p = root ? 0u : (right ? r : l);
        const uint32_t e3 = p == l ? r : l;
        // this thread's arrival at the parent's split first (its release
        // waits only for the hand-over's stores), then the node's record and
        // the rest
        uint32_t s2 = 0, side = 0, bound = 0, other = SLOT_EMPTY;
        bool climb = false;
        if (!root) {
            s2 = right ? r : l - 1;
            side = right ? 0u : 1u;
            bound = right ? l : r;
            split_stage* g2 = gstage + s2;
            store_box(g2->box + 6 * side, lo, hi);
            g2->hd[side] = pack_hd(h, right ? dl : dr);
            g2->bound[side] = bound;
            if (!whole && (s2 < c0 || s2 + 2 > c1)) {
                if (__hip_atomic_exchange(&gslot[s2], bound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
                    SLOT_EMPTY)
                    top[atomicAdd(top_count, 1u)] = s2;
            } else {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                other = __hip_atomic_exchange(&gslot[s2], bound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                climb = other != SLOT_EMPTY;
            }
        }
        // the record but for word 14, its parent link, which the pass that
        // completes the parent writes (the root's stays empty); the
        // children's links now.  No word is written twice, so these stores
        // need no order among the passes' threads.
        {
            float4* dst = reinterpret_cast<float4*>(pairs[p].f);
            dst[0] = make_float4(b[0], b[1], b[2], b[3]);
            dst[1] = make_float4(b[4], b[5], b[6], b[7]);
            dst[2] = make_float4(b[8], b[9], b[10], b[11]);
            *reinterpret_cast<uint2*>(pairs[p].e) = make_uint2(e0, e1);
            pairs[p].e[3] = e3;
            if (root) pairs[p].e[2] = SLOT_EMPTY;
        }
        if (!(e0 & LEAF_BIT)) pairs[s].e[2] = (p << 1) | 0u;
        if (!(e1 & LEAF_BIT)) pairs[s + 1].e[2] = (p << 1) | 1u;
        if (slot_heights) {
            heights[2 * p] = h0;
            heights[2 * p + 1] = h1;
        }
        if (ref_nodes) store_ref_children(ref_nodes, sorted_ids, p, b, e0, e1);
        // p spans chunks (far above the leaf size) and its subtree is
        // complete: its quad after the workgroup's climbs
        if (quads) {
            const uint32_t slot = atomicAdd(s_ndone, 1u);
            if (slot < DONE_CAP) s_done[slot] = p;
            else ovf[slot - DONE_CAP] = p;
        }
        if (root) {
            finish_root(root_box, ref_nodes, depth_out, lo, hi, h);
            return;
        }
        if (!climb) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        carried = true;
        clo = lo;
        chi = hi;
        chd = pack_hd(h, right ? dl : dr);
        cside = side;
        cl = right ? l : other;
        cr = right ? other : r;
        s = s2;
    }
}

// pass 2: workgroup b finishes the nodes inside leaves [b * SUPER, (b + 1) * SUPER)
__global__ __launch_bounds__(TOP_BLOCK) void k_bottom_up_mid(size_t n, const int32_t* __restrict__ sorted_ids,
                                                             pair_node* __restrict__ pairs,
                                                             split_stage* __restrict__ gstage,
                                                             uint32_t* __restrict__ gslot,
                                                             uint32_t* __restrict__ heights, int slot_heights,
                                                             vr_bvh_node* __restrict__ ref_nodes,
                                                             float* __restrict__ root_box,
                                                             uint32_t* __restrict__ depth_out,
                                                             const uint32_t* __restrict__ ready,
                                                             const uint32_t* __restrict__ bucket_count,
                                                             uint32_t* __restrict__ top, uint32_t* __restrict__ top_count,
                                                             uint32_t* __restrict__ ovf, quad_node* __restrict__ quads,
                                                             uint32_t leaf_max)
{
    const uint32_t N = (uint32_t)n;
    const uint32_t c0 = blockIdx.x * SUPER;
    const uint32_t c1 = min(N, c0 + SUPER);
    __shared__ uint32_t s_done[DONE_CAP];
    __shared__ uint32_t s_ndone;
    if (threadIdx.x == 0) s_ndone = 0u;
    __syncthreads();
    const uint32_t count = min(bucket_count[blockIdx.x], SUPER);
    for (uint32_t i = threadIdx.x; i < count; i += TOP_BLOCK)
        climb_split(ready[(size_t)c0 + i], N, c0, c1, false, sorted_ids, pairs, gstage, gslot, heights,
                    slot_heights, ref_nodes, root_box, depth_out, top, top_count, quads, leaf_max, s_done, &s_ndone,
                    ovf + c0);
    quads_after_climbs(pairs, N - 1, leaf_max, quads, s_done, &s_ndone, ovf + c0);
}

// pass 3: one workgroup finishes the nodes that span SUPER blocks
__global__ __launch_bounds__(TOP_BLOCK) void k_bottom_up_top(size_t n, const int32_t* __restrict__ sorted_ids,
                                                             pair_node* __restrict__ pairs,
                                                             split_stage* __restrict__ gstage,
                                                             uint32_t* __restrict__ gslot,
                                                             uint32_t* __restrict__ heights, int slot_heights,
                                                             vr_bvh_node* __restrict__ ref_nodes,
                                                             float* __restrict__ root_box,
                                                             uint32_t* __restrict__ depth_out,
                                                             const uint32_t* __restrict__ top,
                                                             const uint32_t* __restrict__ top_count,
                                                             uint32_t* __restrict__ done, uint32_t* __restrict__ done_count,
                                                             quad_node* __restrict__ quads, uint32_t leaf_max)
{
    const uint32_t N = (uint32_t)n;
    __shared__ uint32_t s_ndone;
    if (threadIdx.x == 0) s_ndone = 0u;
    __syncthreads();
    const uint32_t count = min(*top_count, N - 1);
    // the completed pairs go straight to the global list (`done`, no LDS
    // part): their quads are made by k_list_quads, many workgroups wide
    for (uint32_t i = threadIdx.x; i < count; i += TOP_BLOCK)
        climb_split(top[i], N, 0, N, true, sorted_ids, pairs, gstage, gslot, heights, slot_heights, ref_nodes,
                    root_box, depth_out, nullptr, nullptr, quads, leaf_max, done, &s_ndone, done + DONE_CAP);
    __syncthreads();
    if (threadIdx.x == 0) *done_count = s_ndone;
}

// the quads of the pairs the top pass completed (its one workgroup would
// make them in rounds of dependent record reads, one chain per thread)
__global__ __launch_bounds__(BLOCK) void k_list_quads(const pair_node* __restrict__ pairs, uint32_t npairs,
                                                      uint32_t leaf_max, quad_node* __restrict__ quads,
                                                      const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count)
{
    const uint32_t c = min(*count, npairs);
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < c; i += gridDim.x * BLOCK)
        make_quad<true>(global_pairs{pairs}, list[i], npairs, leaf_max, quads);
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v)
{
    const size_t i = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

// the split slots on the chunks' edges (CHUNK * k - 1 for 0 < k, below n - 1) set empty
__global__ void k_clear_edges(uint32_t* slots, size_t n)
{
    const size_t k = blockIdx.x * (size_t)BLOCK + threadIdx.x + 1;
    const size_t s = k * CHUNK - 1;
    if (s + 1 < n) slots[s] = SLOT_EMPTY;
}


__global__ __launch_bounds__(BLOCK) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) dst[i] = src[i];
}

// Counter calibration for gathers: thread i reads the 128-byte line
// lines[i] of src with eight 16-byte loads, as a traversal lane reads one
// 4-wide node, and the xor of everything read goes to *out (one atomic per
// block), so a permutation of distinct lines moves a known number of bytes.
__global__ __launch_bounds__(BLOCK) void k_gather_lines(const uint4* __restrict__ src,
                                                        const uint32_t* __restrict__ lines, size_t n,
                                                        uint32_t* __restrict__ out)
{
    uint32_t acc = 0;
    for (size_t i = blockIdx.x * (size_t)BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) {
        const uint4* p = src + (size_t)lines[i] * 8u;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint4 v = p[k];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) acc ^= __shfl_xor(acc, off);
    __shared__ uint32_t part[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int q = 0; q < BLOCK / 64; ++q) t ^= part[q];
        atomicXor(out, t);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// scratch layout for one build of n primitives
struct build_scratch {
    uint32_t* keys0;
    uint32_t* vals0;
    uint32_t* keys1;
    uint32_t* vals1;
    uint32_t* hist;   // per-tile digit counts, digit-major
    uint32_t* offs;   // their exclusive scan along each digit's row
    uint32_t* totals; // per-digit totals
    uint32_t* cb;     // 6 words
    float* part;      // CB_GRID x 6 per-block centroid extremes
    split_stage* stage;  // bottom-up hand-overs in global memory, by split (n - 1)
    uint32_t* counts;  // bottom-up ready-list counters: one per SUPER block, then the top list's,
                       // then the number of pairs the top pass completed (the ready lists reuse
                       // keys0 and vals0 after the sort; the mid pass's quad overflow lists and
                       // the top pass's completed pairs keys1)
    size_t bytes;
};

build_scratch plan_scratch(size_t n, char* base)
{
    const size_t tiles = std::max<size_t>(1, (n + TILE - 1) / TILE);
    const size_t hist_n = tiles * RADIX;
    size_t off = 0;
    auto take = [&](size_t elems) {
        char* p = base ? base + off : nullptr;
        off += align_up(std::max<size_t>(elems, 1) * 4, 256);
        return reinterpret_cast<uint32_t*>(p);
    };
    build_scratch s;
    s.keys0 = take(n);
    s.vals0 = take(n);
    s.keys1 = take(n);
    s.vals1 = take(n);
    s.hist = take(hist_n);
    s.offs = take(hist_n);
    s.totals = take(RADIX);
    s.cb = take(8);
    s.part = reinterpret_cast<float*>(take((size_t)CB_GRID * 6));
    s.stage = reinterpret_cast<split_stage*>(take(n * (sizeof(split_stage) / 4)));
    s.counts = take((n + SUPER - 1) / SUPER + 2);  // + the top pass's done count
    s.bytes = off;
    return s;
}

// centroid box, then the codes with the kernel's extras
int launch_morton(const vr_triangle* prims, size_t n, float* part, uint32_t* cb, uint32_t* codes,
                  const morton_extras& x, float* cen, hipStream_t st)
{
    const unsigned grid = std::max(1u, std::min(div_up(n, CB_U * CB_BLOCK), (unsigned)CB_GRID));
    const int vec = reinterpret_cast<uintptr_t>(prims) % 16 == 0;
    hipLaunchKernelGGL(k_centroid_bounds, dim3(grid), dim3(CB_BLOCK), 0, st, prims, n, vec, part, cen);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(CB_BLOCK), 0, st, part, grid, cb);
    hipLaunchKernelGGL(k_morton, dim3(div_up(n, TILE)), dim3(BLOCK), 0, st, prims, cen, n, vec, cb, codes, x);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

// the bottom-up passes over the sorted keys (hierarchy, boxes, records,
// quads); a build has cleared the ready-list counters and the chunk-edge
// split slots already (k_morton), a refit clears them here
void launch_bottom_up(const bvh_state& bs, const build_scratch& s, const vr_triangle* prims, vr_bvh_node* ref_nodes,
                      bool build, hipStream_t st)
{
    const size_t n = bs.num_prims;
    const size_t nb = (n + SUPER - 1) / SUPER;
    uint32_t* depth = reinterpret_cast<uint32_t*>(bs.root_box + 6);
    uint32_t* slots = bs.flags;
    const int slot_heights = (int)(build && bs.want_rotate);
    if (!build) {
        hipLaunchKernelGGL(k_fill_u32, dim3(div_up(nb + 1, BLOCK)), dim3(BLOCK), 0, st, s.counts, nb + 1, 0u);
        if (n > (size_t)CHUNK)
            hipLaunchKernelGGL(k_clear_edges, dim3(div_up(n / CHUNK, BLOCK)), dim3(BLOCK), 0, st, slots, n);
    }
    hipLaunchKernelGGL(k_bottom_up_local, dim3(div_up(n, CHUNK)), dim3(CHUNK), 0, st, prims, n, bs.sorted_codes,
                       bs.sorted_ids, bs.pairs, s.stage, slots, bs.heights, bs.tris, ref_nodes, bs.root_box, depth,
                       s.keys0, s.counts, s.vals0, s.counts + nb, slot_heights, bs.quads, LBVH_LEAF_MAX);
    if (n > (size_t)CHUNK)
        hipLaunchKernelGGL(k_bottom_up_mid, dim3((unsigned)nb), dim3(TOP_BLOCK), 0, st, n, bs.sorted_ids, bs.pairs, s.stage, slots, bs.heights, slot_heights, ref_nodes, bs.root_box,
                           depth, s.keys0, s.counts, s.vals0, s.counts + nb, s.keys1, bs.quads, LBVH_LEAF_MAX);
    if (n > (size_t)SUPER) {
        hipLaunchKernelGGL(k_bottom_up_top, dim3(1), dim3(TOP_BLOCK), 0, st, n, bs.sorted_ids,
                           bs.pairs, s.stage, slots, bs.heights, slot_heights, ref_nodes, bs.root_box, depth, s.vals0,
                           s.counts + nb, s.keys1, s.counts + nb + 1, bs.quads, LBVH_LEAF_MAX);
        if (bs.quads)
            hipLaunchKernelGGL(k_list_quads, dim3(64), dim3(BLOCK), 0, st, bs.pairs, (uint32_t)(n - 1),
                               LBVH_LEAF_MAX, bs.quads, s.keys1, s.counts + nb + 1);
    }
}

}  // namespace

// Build (or rebuild in place) the LBVH of `bs` from n primitives.
int lbvh_build_into(bvh_state& bs, const vr_triangle* d_prims, size_t n, bool keep_nodes, bool rotate,
                    hipStream_t st)
{
    VR_CHECK(n <= (size_t)MAX_LEAF_FIRST, VR_ERR_UNSUPPORTED, "lbvh: more than 2^26 - 1 primitives");
    VR_CHECK(!(keep_nodes && rotate), VR_ERR_UNSUPPORTED,
             "lbvh: the reference node layout is not kept for a rotated tree");
    const size_t npairs = n > 1 ? n - 1 : 0;
    // (re)allocate owned arrays when the size changes
    if (bs.num_prims != n || !bs.is_lbvh || (keep_nodes && !bs.ref_nodes)) {
        // the handle's arrays: one block, stream ordered from the device's
        // pool (a build after a destroyed one of similar size reuses it)
        if (bs.arena) {
            VR_HIP_CHECK(hipFreeAsync(bs.arena, st));
        } else {
            void* owned[] = {bs.pairs, bs.quads, bs.tris, bs.prims, bs.sorted_codes, bs.sorted_ids,
                             bs.flags, bs.heights, bs.ref_nodes, bs.indices, bs.root_box, bs.scratch};
            for (void* q : owned)
                if (q) VR_HIP_CHECK(bs.pooled ? hipFreeAsync(q, st) : hipFree(q));
        }
        const int device = bs.device;
        bs = bvh_state{};
        bs.device = device;
        bs.num_prims = n;
        bs.num_pairs = npairs;
        bs.is_lbvh = true;
        bs.want_rotate = rotate;
        bs.pooled = true;
        if (n == 0) return VR_OK;
        const size_t np1 = std::max<size_t>(npairs, 1);
        if (keep_nodes) bs.num_ref_nodes = 2 * n - 1;
        build_scratch plan = plan_scratch(n, nullptr);
        bs.scratch_bytes = plan.bytes;
        // no copy of the primitives and no index list: the build reads the
        // caller's triangles (valid until its work on the stream is done), a
        // refit the new ones, and an LBVH's index list is the identity
        const size_t sizes[] = {np1 * sizeof(pair_node), np1 * sizeof(quad_node), n * sizeof(tri_node),
                                0, n * 4, n * 4, np1 * 4, 2 * np1 * 4, 8 * sizeof(float),
                                0, bs.num_ref_nodes * sizeof(vr_bvh_node), plan.bytes};
        size_t total = 0;
        for (size_t b : sizes) total += align_up(b, 256);
        VR_HIP_CHECK(hipMallocAsync(&bs.arena, total, st));
        char* at = static_cast<char*>(bs.arena);
        auto carve = [&](auto*& ptr, size_t bytes) {
            ptr = bytes ? reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(at) : nullptr;
            at += align_up(bytes, 256);
        };
        carve(bs.pairs, sizes[0]);
        carve(bs.quads, sizes[1]);
        carve(bs.tris, sizes[2]);
        carve(bs.prims, sizes[3]);
        carve(bs.sorted_codes, sizes[4]);
        carve(bs.sorted_ids, sizes[5]);
        carve(bs.flags, sizes[6]);
        carve(bs.heights, sizes[7]);
        carve(bs.root_box, sizes[8]);
        carve(bs.indices, sizes[9]);
        carve(bs.ref_nodes, sizes[10]);
        carve(bs.scratch, sizes[11]);
    }
    if (n == 0) return VR_OK;
    build_scratch s = plan_scratch(n, reinterpret_cast<char*>(bs.scratch));

    // 1-2: centroid bounds, Morton codes (with the first sort pass's digit
    // counts, the cleared ready-list counters and chunk-edge split slots)
    const uint32_t tiles = (uint32_t)((n + TILE - 1) / TILE);
    const size_t nb = (n + SUPER - 1) / SUPER;
    const morton_extras extras{s.hist, tiles, bs.flags, s.counts, (uint32_t)(nb + 1)};
    // the centroids go to the bottom-up passes' stage (64 bytes per split,
    // unused until then): 12 bytes per primitive written and read once
    // instead of the triangles (44 bytes) read again
    int rc = launch_morton(d_prims, n, s.part, s.cb, s.keys0, extras, reinterpret_cast<float*>(s.stage), st);
    if (rc) return rc;

    // 3: radix sort (code, id) -- 30-bit keys, four 8-bit passes
    uint32_t* kin = s.keys0;
    uint32_t* vin = s.vals0;
    uint32_t* kout = s.keys1;
    uint32_t* vout = s.vals1;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = pass * 8;
        if (pass == 3) {  // last pass writes straight into the BVH's arrays
            kout = bs.sorted_codes;
            vout = reinterpret_cast<uint32_t*>(bs.sorted_ids);
        }
        if (pass > 0)
            hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(BLOCK), 0, st, kin, n, shift, s.hist, tiles);
        hipLaunchKernelGGL(k_radix_spine, dim3(RADIX), dim3(BLOCK), 0, st, s.hist, tiles, s.offs, s.totals);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(BLOCK), 0, st, kin, vin, pass == 0 ? 1 : 0, n, shift,
                           s.hist, s.offs, s.totals, tiles, kout, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    VR_HIP_CHECK(hipGetLastError());

    // 4: hierarchy and boxes from the leaves up, traversal-order triangles,
    // reference layout, and the 4-wide traversal nodes (each pair's quad as
    // soon as its subtree is complete)
    launch_bottom_up(bs, s, d_prims, keep_nodes ? bs.ref_nodes : nullptr, true, st);
    VR_HIP_CHECK(hipGetLastError());
    bs.root = n == 1 ? leaf_entry(0, 1) : 0u;
    bs.max_leaf = 1;
    bs.collapse_leaf_max = LBVH_LEAF_MAX;

    // The exact depth stays on the device (vr_bvh_get_info reads it); 65
    // bounds every Karras tree, which the traversal stack covers, so the
    // build never waits on the device.
    bs.depth = n <= 1 ? (uint32_t)n : 65u;
    bs.rotated = false;
    bs.want_rotate = rotate;
    if (rotate && n > 1) {
        // optional: SAH-driven tree rotations, then the quads of the new tree
        uint32_t count = 0;
        rc = lbvh_rotate(bs, LBVH_LEAF_MAX, &count, st);
        if (rc) return rc;
        bs.rotated = count > 0;
        if (bs.rotated) return collapse_quads(bs, LBVH_LEAF_MAX, st);
    }
    return VR_OK;
}

// Refit after the vertices moved: same ids, new boxes (detail/bvh/refit.h).
int lbvh_refit(bvh_state& bs, const vr_triangle* d_prims, hipStream_t st)
{
    const size_t n = bs.num_prims;
    if (n == 0) return VR_OK;
    VR_CHECK(!bs.rotated, VR_ERR_UNSUPPORTED, "refit: a rotated LBVH is rebuilt (vr_lbvh_rebuild), not refitted");
    build_scratch s = plan_scratch(n, reinterpret_cast<char*>(bs.scratch));
    // the sorted keys are kept, so the bottom-up pass rebuilds the same
    // hierarchy with the new boxes, and the quads from them (the greedy
    // choices may move)
    launch_bottom_up(bs, s, d_prims, bs.ref_nodes, false, st);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

namespace {
__global__ void k_iota(uint32_t* __restrict__ d, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) d[i] = (uint32_t)i;
}
}  // namespace

// an LBVH's index list (tree.indices()): leaves name their primitives by
// original index, so the list is the identity
int lbvh_indices(uint32_t* d, size_t n, hipStream_t st)
{
    if (n == 0) return VR_OK;
    hipLaunchKernelGGL(k_iota, dim3(std::max(1u, std::min(div_up(n, BLOCK), 4096u))), dim3(BLOCK), 0, st, d, n);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

int device_copy(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    VR_CHECK(bytes % 16 == 0, VR_ERR_INVALID_ARGUMENT, "vr_device_copy: bytes must be a multiple of 16");
    const size_t n = bytes / 16;
    if (n == 0) return VR_OK;
    const unsigned grid = std::max(1u, std::min(div_up(n, BLOCK), 65536u));
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(BLOCK), 0, st, static_cast<uint4*>(dst),
                       static_cast<const uint4*>(src), n);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

int device_gather_lines(const void* src, size_t src_bytes, const uint32_t* lines, size_t n, uint32_t* out,
                        hipStream_t st)
{
    VR_CHECK(src_bytes % 128 == 0, VR_ERR_INVALID_ARGUMENT, "vr_device_gather_lines: src_bytes must be whole lines");
    if (n == 0) return VR_OK;
    const unsigned grid = std::max(1u, std::min(div_up(n, BLOCK), 65536u));
    hipLaunchKernelGGL(k_gather_lines, dim3(grid), dim3(BLOCK), 0, st, static_cast<const uint4*>(src), lines, n, out);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

int lbvh_morton_only(const vr_triangle* d_prims, size_t n, uint32_t* d_codes, hipStream_t st)
{
    if (n == 0) return VR_OK;
    void* tmp = nullptr;
    VR_HIP_CHECK(hipMallocAsync(&tmp, 256 + (size_t)CB_GRID * 6 * sizeof(float), st));
    uint32_t* cb = static_cast<uint32_t*>(tmp);
    float* part = reinterpret_cast<float*>(static_cast<char*>(tmp) + 256);
    int rc = launch_morton(d_prims, n, part, cb, d_codes, morton_extras{}, nullptr, st);
    (void)hipFreeAsync(tmp, st);
    return rc;
}

}  // namespace vr
