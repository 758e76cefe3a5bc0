// sah.hip -- the reference's binned SAH builder on the device
// (binned_sah_builder with object splits, detail/bvh/sah.h:176-807, driven by
// build_top_down, detail/bvh/build_top_down.h:27-174).
//
// The reference builds depth first on one core: a node's references are the
// range [first, end) of one array, split at the best of 16 centroid bins
// along the largest centroid extent, reordered by std::partition, the right
// subtree built (and numbered) before the left.  Here the tree is built one
// level at a time, every node of a level at once, and then renumbered:
//
//  * a node's references stay one contiguous range of the reference array in
//    the order the reference's partitions leave them (libstdc++'s
//    std::partition on bidirectional iterators swaps the k-th reference that
//    belongs right, counted from the left, with the k-th that belongs left,
//    counted from the right; those swaps are reproduced exactly), so every
//    leaf lists its primitives in the reference's order;
//  * bins hold exact minima, maxima and counts, so the bins, the sweep over
//    them (the same float operations in the same order as the reference),
//    the split and the leaf decisions equal the reference's; boxes may differ
//    from the reference's only in the sign of a zero (its sequential ternary
//    min of +0 and -0 depends on the order, a device minimum does not);
//  * nodes of more than LARGE references are binned by many workgroups into
//    global bins and partitioned through device-wide scans; smaller ones by
//    one workgroup each, in LDS;
//  * the final pass numbers the nodes as the reference's recursion does
//    (children appended in pairs, right subtree first: a node's children
//    start at 1 + 2 r, r = the inner nodes visited before it right first in
//    pre-order) and writes the reference layout, the index list, and the
//    traversal arrays (pair records, triangles in index order).
#include "vr_internal.h"

#include <algorithm>
#include <vector>

namespace vr {

namespace {

constexpr int NUM_BINS = 16;
constexpr int BIN_WORDS = 13;  // prim lo xyz (min), prim hi xyz (max), cent lo xyz (min), cent hi xyz (max), count
constexpr int NODE_BINS = NUM_BINS * BIN_WORDS;
constexpr int BLOCK = 256;
// nodes of at most this many references are split by one workgroup
#ifndef VR_SAH_LARGE
#define VR_SAH_LARGE 2048
#endif
constexpr uint32_t LARGE = VR_SAH_LARGE;

// a reference: the primitive's box and its index (sah.h:181-193)
struct alignas(32) sref {
    float lo[3];
    float hi[3];
    int32_t index;
    int32_t pad;
};

// one node, numbered in the order the levels make them (children in pairs)
struct alignas(16) snode {
    float pb[6];      // primitive bounds lo xyz, hi xyz
    float cb[6];      // centroid bounds
    uint32_t first;   // range in the reference array
    uint32_t count;
    uint32_t child;   // inner: left child (right = child + 1)
    uint32_t info;    // bit 0 inner, bits 8-9 split axis, bit 16 sign; bits 17+ parent + 1 (0: root)
};

// per large node of the level: projection, and the split once decided
struct slot_t {
    float k0, k1;
    uint32_t axis, node;
    uint32_t split;          // 0: leaf; else the split bin (1..15)
    uint32_t first, count, pivot;
    int32_t lslot, rslot;    // the children's slots at the next level (-1: not large)
    uint32_t lnode, rnode;
    uint32_t sf_a, st_a, st_b;  // scan bases of the node's range
    uint32_t pad;
};

struct counters {
    uint32_t nodes;       // nodes made so far
    uint32_t large_next;  // large nodes of the next level
    uint32_t small_next;  // small nodes of the next level
    uint32_t max_leaf;
    uint32_t bad;         // a leaf the traversal cannot hold (more than 32 primitives)
    uint32_t pad[3];
};

// order-preserving words of floats, for exact atomic minima and maxima
__device__ inline uint32_t ford(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float fdec(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u); }

__device__ inline float axis3(f3 v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; }

// safe_half_surface_area (math/detail/aabb.inl:211-217)
__device__ inline float safe_hsa(f3 lo, f3 hi)
{
    f3 s = sub3(hi, lo);
    s = f3{tmax_f(0.0f, s.x), tmax_f(0.0f, s.y), tmax_f(0.0f, s.z)};
    return s.x * s.y + s.y * s.z + s.z * s.x;
}

// max_index (math/detail/vector.inl:874-884): ties go to the later axis
__device__ inline int max_index3(f3 u)
{
    const int i = u.y < u.x ? 0 : 1;
    return axis3(u, 2) < axis3(u, i) ? i : 2;
}

// aabb center: (max + min) * 0.5
__device__ inline f3 center_of(const sref& r)
{
    return muls3(add3(mk3(r.hi[0], r.hi[1], r.hi[2]), mk3(r.lo[0], r.lo[1], r.lo[2])), 0.5f);
}

// projection::project_unsafe / project (sah.h:281-312)
__device__ inline int project_unsafe(float k0, float k1, int axis, f3 c) { return (int)(k1 * (axis3(c, axis) - k0)); }
__device__ inline int project(float k0, float k1, int axis, f3 c)
{
    return min(max(project_unsafe(k0, k1, axis, c), 0), NUM_BINS - 1);
}

__device__ inline void bins_identity(uint32_t* w, int i)
{
    const int k = i % BIN_WORDS;
    w[i] = k < 3 || (k >= 6 && k < 9) ? ford(FLT_MAX) : k < 12 ? ford(-FLT_MAX) : 0u;
}

// add a reference to a bin's words (shared or global memory)
__device__ inline void bin_add(uint32_t* b, const sref& r, f3 c)
{
    for (int k = 0; k < 3; ++k) {
        atomicMin(&b[k], ford(r.lo[k]));
        atomicMax(&b[3 + k], ford(r.hi[k]));
    }
    atomicMin(&b[6], ford(c.x));
    atomicMin(&b[7], ford(c.y));
    atomicMin(&b[8], ford(c.z));
    atomicMax(&b[9], ford(c.x));
    atomicMax(&b[10], ford(c.y));
    atomicMax(&b[11], ford(c.z));
    atomicAdd(&b[12], 1u);
}

struct sbox {
    f3 lo, hi;
};
__device__ inline void insert(sbox& a, const sbox& b)
{
    a.lo = min3v(a.lo, b.lo);
    a.hi = max3v(a.hi, b.hi);
}

struct sbin {
    sbox p, c;
    uint32_t n;
};
__device__ inline sbin read_bin(const uint32_t* w)
{
    sbin b;
    b.p.lo = mk3(fdec(w[0]), fdec(w[1]), fdec(w[2]));
    b.p.hi = mk3(fdec(w[3]), fdec(w[4]), fdec(w[5]));
    b.c.lo = mk3(fdec(w[6]), fdec(w[7]), fdec(w[8]));
    b.c.hi = mk3(fdec(w[9]), fdec(w[10]), fdec(w[11]));
    b.n = w[12];
    return b;
}
__device__ inline void merge(sbin& a, const sbin& b)
{
    insert(a.p, b.p);
    insert(a.c, b.c);
    a.n += b.n;
}

struct split_out {
    int best;  // -1: a leaf
    sbin L, R;
};

// find_split (sah.h:343-400) and the leaf test of split (sah.h:771-776):
// sweep left to right, then right to left, the rightmost of equal costs
__device__ inline split_out sah_sweep(const uint32_t* w, f3 plo, f3 phi, uint32_t count)
{
    split_out o;
    sbin acc_l[NUM_BINS];
    acc_l[0] = read_bin(w);
    for (int i = 1; i < NUM_BINS; ++i) {
        acc_l[i] = acc_l[i - 1];
        merge(acc_l[i], read_bin(w + i * BIN_WORDS));
    }
    const float hsa_parent = safe_hsa(plo, phi);
    sbin acc_r = read_bin(w + (NUM_BINS - 1) * BIN_WORDS), best_r = acc_r;
    float best_cost = FLT_MAX;
    int best = -1;
    for (int i = NUM_BINS - 1; i > 0; --i) {
        if (i < NUM_BINS - 1) merge(acc_r, read_bin(w + i * BIN_WORDS));
        const sbin& l = acc_l[i - 1];
        const float cost = 1.0f + (safe_hsa(l.p.lo, l.p.hi) / hsa_parent) * (3.0f * (float)l.n) +
                           (safe_hsa(acc_r.p.lo, acc_r.p.hi) / hsa_parent) * (3.0f * (float)acc_r.n);
        if (cost < best_cost) {
            best_cost = cost;
            best = i;
            best_r = acc_r;
        }
    }
    if (best < 0 || best_cost > 3.0f * (float)count) {
        o.best = -1;
        return o;
    }
    o.best = best;
    o.L = acc_l[best - 1];
    o.R = best_r;
    return o;
}

__device__ inline void set_box(float* d, const sbox& b)
{
    d[0] = b.lo.x;
    d[1] = b.lo.y;
    d[2] = b.lo.z;
    d[3] = b.hi.x;
    d[4] = b.hi.y;
    d[5] = b.hi.z;
}

// a child node: its record, and whether the next level splits it (more than
// max_leaf references and a centroid extent above zero, sah.h:741-757)
// large or small
__device__ inline void make_child(snode* nodes, uint32_t id, uint32_t parent, const sbin& b, uint32_t first,
                                  uint32_t max_leaf, counters* C, uint32_t* large_next, slot_t* slots_next,
                                  uint32_t* small_next, int32_t& slot)
{
    snode c;
    set_box(c.pb, b.p);
    set_box(c.cb, b.c);
    c.first = first;
    c.count = b.n;
    c.child = 0;
    c.info = (parent + 1u) << 17;
    nodes[id] = c;
    slot = -1;
    const f3 cs = sub3(b.c.hi, b.c.lo);
    const int axis = max_index3(cs);
    if (b.n <= max_leaf || !(axis3(cs, axis) > 0.0f)) return;
    if (b.n > LARGE) {
        const uint32_t j = atomicAdd(&C->large_next, 1u);
        slot_t s{};
        s.k0 = axis3(b.c.lo, axis);
        s.k1 = (float)NUM_BINS / (axis3(b.c.hi, axis) - s.k0);
        s.axis = (uint32_t)axis;
        s.node = id;
        slots_next[j] = s;
        large_next[j] = id;
        slot = (int32_t)j;
    } else {
        small_next[atomicAdd(&C->small_next, 1u)] = id;
    }
}

// ---------------------------------------------------------------------------
// references and the root
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_sah_refs(const vr_triangle* __restrict__ prims, size_t n,
                                                    sref* __restrict__ refs, uint32_t* __restrict__ root_words)
{
    // this thread's extremes of the primitive and centroid bounds
    float v[12];
    for (int k = 0; k < 3; ++k) {
        v[k] = FLT_MAX;
        v[3 + k] = -FLT_MAX;
        v[6 + k] = FLT_MAX;
        v[9 + k] = -FLT_MAX;
    }
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) {
        const vr_triangle t = prims[i];
        // get_bounds of a triangle: v1, v1 + e1, v1 + e2 (math/detail/triangle.inl:41-53)
        const f3 v1 = mk3(t.v1[0], t.v1[1], t.v1[2]);
        f3 lo = mk3(FLT_MAX, FLT_MAX, FLT_MAX), hi = mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        lo = min3v(lo, v1);
        hi = max3v(hi, v1);
        const f3 a = add3(v1, mk3(t.e1[0], t.e1[1], t.e1[2])), b = add3(v1, mk3(t.e2[0], t.e2[1], t.e2[2]));
        lo = min3v(lo, a);
        hi = max3v(hi, a);
        lo = min3v(lo, b);
        hi = max3v(hi, b);
        sref r;
        r.lo[0] = lo.x;
        r.lo[1] = lo.y;
        r.lo[2] = lo.z;
        r.hi[0] = hi.x;
        r.hi[1] = hi.y;
        r.hi[2] = hi.z;
        r.index = (int32_t)i;
        r.pad = 0;
        refs[i] = r;
        const f3 c = center_of(r);
        v[0] = fminf(v[0], lo.x);
        v[1] = fminf(v[1], lo.y);
        v[2] = fminf(v[2], lo.z);
        v[3] = fmaxf(v[3], hi.x);
        v[4] = fmaxf(v[4], hi.y);
        v[5] = fmaxf(v[5], hi.z);
        v[6] = fminf(v[6], c.x);
        v[7] = fminf(v[7], c.y);
        v[8] = fminf(v[8], c.z);
        v[9] = fmaxf(v[9], c.x);
        v[10] = fmaxf(v[10], c.y);
        v[11] = fmaxf(v[11], c.z);
    }
    __shared__ uint32_t s[12];
    if (threadIdx.x < 12) s[threadIdx.x] = threadIdx.x % 6 < 3 ? ford(FLT_MAX) : ford(-FLT_MAX);
    __syncthreads();
    for (int k = 0; k < 12; ++k) {
        if (k % 6 < 3) atomicMin(&s[k], ford(v[k]));
        else atomicMax(&s[k], ford(v[k]));
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        if (threadIdx.x % 6 < 3) atomicMin(&root_words[threadIdx.x], s[threadIdx.x]);
        else atomicMax(&root_words[threadIdx.x], s[threadIdx.x]);
    }
}

__global__ void k_sah_init_words(uint32_t* w)
{
    if (threadIdx.x < 12) w[threadIdx.x] = threadIdx.x % 6 < 3 ? ford(FLT_MAX) : ford(-FLT_MAX);
}

// the root node (level 0)
__global__ void k_sah_root(const uint32_t* __restrict__ root_words, uint32_t n, uint32_t max_leaf, snode* nodes,
                           counters* C, uint32_t* large_next, slot_t* slots_next, uint32_t* small_next)
{
    sbin b;
    b.p.lo = mk3(fdec(root_words[0]), fdec(root_words[1]), fdec(root_words[2]));
    b.p.hi = mk3(fdec(root_words[3]), fdec(root_words[4]), fdec(root_words[5]));
    b.c.lo = mk3(fdec(root_words[6]), fdec(root_words[7]), fdec(root_words[8]));
    b.c.hi = mk3(fdec(root_words[9]), fdec(root_words[10]), fdec(root_words[11]));
    b.n = n;
    C->nodes = 1;
    C->large_next = 0;
    C->small_next = 0;
    C->max_leaf = 0;
    C->bad = 0;
    int32_t slot;
    make_child(nodes, 0u, 0u, b, 0u, max_leaf, C, large_next, slots_next, small_next, slot);
    nodes[0].info = 0;  // no parent
}

__global__ __launch_bounds__(BLOCK) void k_sah_fill_i32(int32_t* __restrict__ p, size_t n, int32_t v)
{
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) p[i] = v;
}

__global__ __launch_bounds__(BLOCK) void k_sah_bins_clear(uint32_t* __restrict__ w, size_t words)
{
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < words; i += (size_t)gridDim.x * BLOCK) {
        const int k = (int)(i % BIN_WORDS);
        w[i] = k < 3 || (k >= 6 && k < 9) ? ford(FLT_MAX) : k < 12 ? ford(-FLT_MAX) : 0u;
    }
}

// ---------------------------------------------------------------------------
// large nodes: global bins, device-wide partition
// ---------------------------------------------------------------------------

// Each workgroup bins a contiguous stretch of the reference array; a node's
// range is contiguous, so a stretch mostly feeds one node, whose bins it
// gathers in LDS (flushed when the stretch moves on to another node);
// references of other nodes go to the global bins directly.
__global__ __launch_bounds__(BLOCK) void k_sah_bin_large(const sref* __restrict__ refs, size_t n,
                                                         const int32_t* __restrict__ slot_of,
                                                         const slot_t* __restrict__ slots, uint32_t* __restrict__ bins,
                                                         size_t stretch)
{
    __shared__ uint32_t lb[NODE_BINS];
    __shared__ int32_t cur;
    const size_t b0 = (size_t)blockIdx.x * stretch, b1 = min(n, b0 + stretch);
    if (threadIdx.x == 0) cur = -1;
    for (int i = threadIdx.x; i < NODE_BINS; i += BLOCK) bins_identity(lb, i);
    __syncthreads();
    auto flush = [&]() {
        if (cur >= 0)
            for (int i = threadIdx.x; i < NODE_BINS; i += BLOCK) {
                const int k = i % BIN_WORDS;
                uint32_t* g = bins + (size_t)cur * NODE_BINS + i;
                if (k == 12) {
                    if (lb[i]) atomicAdd(g, lb[i]);
                } else if (k < 3 || (k >= 6 && k < 9)) {
                    if (lb[i] != ford(FLT_MAX)) atomicMin(g, lb[i]);
                } else {
                    if (lb[i] != ford(-FLT_MAX)) atomicMax(g, lb[i]);
                }
            }
        __syncthreads();
        for (int i = threadIdx.x; i < NODE_BINS; i += BLOCK) bins_identity(lb, i);
    };
    for (size_t base = b0; base < b1; base += BLOCK) {
        const int32_t s0 = slot_of[base];  // the same for every thread
        if (s0 != cur) {
            flush();
            __syncthreads();
            if (threadIdx.x == 0) cur = s0;
            __syncthreads();
        }
        const size_t p = base + threadIdx.x;
        if (p < b1) {
            const int32_t s = slot_of[p];
            if (s >= 0) {
                const slot_t& sl = slots[s];
                const sref r = refs[p];
                const f3 c = center_of(r);
                const int bin = project(sl.k0, sl.k1, (int)sl.axis, c);
                bin_add(s == cur ? lb + bin * BIN_WORDS : bins + (size_t)s * NODE_BINS + bin * BIN_WORDS, r, c);
            }
        }
        __syncthreads();
    }
    flush();
}

// one thread per large node: the split (or a leaf) and the children
__global__ __launch_bounds__(64) void k_sah_split_large(uint32_t m, slot_t* __restrict__ slots,
                                                        const uint32_t* __restrict__ bins, snode* __restrict__ nodes,
                                                        counters* C, uint32_t max_leaf, uint32_t* large_next,
                                                        slot_t* slots_next, uint32_t* small_next)
{
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= m) return;
    slot_t& sl = slots[i];
    snode nd = nodes[sl.node];
    const split_out so = sah_sweep(bins + (size_t)i * NODE_BINS, mk3(nd.pb[0], nd.pb[1], nd.pb[2]),
                                   mk3(nd.pb[3], nd.pb[4], nd.pb[5]), nd.count);
    sl.first = nd.first;
    sl.count = nd.count;
    if (so.best < 0) {
        sl.split = 0;
        sl.lslot = sl.rslot = -1;
        return;
    }
    const uint32_t c = atomicAdd(&C->nodes, 2u);
    const int axis = (int)sl.axis;
    const uint32_t sign = axis3(so.L.p.lo, axis) < axis3(so.R.p.lo, axis) ? 0u : 1u;
    nodes[sl.node].child = c;
    nodes[sl.node].info = (nd.info & ~0x1FFFFu) | 1u | ((uint32_t)axis << 8) | (sign << 16);
    sl.split = (uint32_t)so.best;
    sl.pivot = nd.first + so.L.n;
    sl.lnode = c;
    sl.rnode = c + 1u;
    make_child(nodes, c, sl.node, so.L, nd.first, max_leaf, C, large_next, slots_next, small_next, sl.lslot);
    make_child(nodes, c + 1u, sl.node, so.R, sl.pivot, max_leaf, C, large_next, slots_next, small_next, sl.rslot);
}

// Partition of the large nodes (std::partition, bidirectional form): f marks
// a reference left of the pivot that belongs right, t one right of the
// pivot that belongs left; the k-th f from the left swaps with the k-th t
// from the right.  This pass marks them, counts them per workgroup, and
// moves the next level's slots and each reference's node along.
__global__ __launch_bounds__(BLOCK) void k_sah_flags_large(const sref* __restrict__ refs, size_t n,
                                                           const int32_t* __restrict__ slot_of,
                                                           const slot_t* __restrict__ slots,
                                                           int32_t* __restrict__ slot_next,
                                                           uint32_t* __restrict__ node_of, uint8_t* __restrict__ flags,
                                                           uint32_t* __restrict__ bsum)
{
    const size_t p = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    uint32_t f = 0, t = 0;
    if (p < n) {
        const int32_t s = slot_of[p];
        int32_t nxt = -1;
        if (s >= 0) {
            const slot_t& sl = slots[s];
            if (sl.split) {
                const bool left = p < sl.pivot;
                const bool pred = project_unsafe(sl.k0, sl.k1, (int)sl.axis, center_of(refs[p])) < (int)sl.split;
                f = left && !pred;
                t = !left && pred;
                nxt = left ? sl.lslot : sl.rslot;
                node_of[p] = left ? sl.lnode : sl.rnode;
            }
        }
        slot_next[p] = nxt;
        flags[p] = (uint8_t)(f | (t << 1));
    }
    __shared__ uint32_t sf[BLOCK / 64], stt[BLOCK / 64];
    const uint32_t cf = (uint32_t)__popcll(__ballot(f)), ct = (uint32_t)__popcll(__ballot(t));
    if ((threadIdx.x & 63) == 0) {
        sf[threadIdx.x / 64] = cf;
        stt[threadIdx.x / 64] = ct;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, b = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            a += sf[w];
            b += stt[w];
        }
        bsum[2 * blockIdx.x] = a;
        bsum[2 * blockIdx.x + 1] = b;
    }
}

// exclusive scan of the workgroup counts (one workgroup); totals at the end
__global__ __launch_bounds__(1024) void k_sah_scan(uint32_t* __restrict__ bsum, uint32_t nb, uint32_t* __restrict__ totals)
{
    __shared__ uint32_t carry[2];
    __shared__ uint32_t wsum[2][16];
    if (threadIdx.x < 2) carry[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x / 64;
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        uint32_t v[2] = {i < nb ? bsum[2 * i] : 0u, i < nb ? bsum[2 * i + 1] : 0u};
        uint32_t incl[2] = {v[0], v[1]};
        for (int d = 1; d < 64; d <<= 1)
            for (int k = 0; k < 2; ++k) {
                const uint32_t o = __shfl_up(incl[k], d);
                if (lane >= (uint32_t)d) incl[k] += o;
            }
        if (lane == 63) {
            wsum[0][w] = incl[0];
            wsum[1][w] = incl[1];
        }
        __syncthreads();
        uint32_t off[2] = {carry[0], carry[1]};
        for (uint32_t q = 0; q < w; ++q) {
            off[0] += wsum[0][q];
            off[1] += wsum[1][q];
        }
        if (i < nb) {
            bsum[2 * i] = off[0] + incl[0] - v[0];
            bsum[2 * i + 1] = off[1] + incl[1] - v[1];
        }
        __syncthreads();
        if (threadIdx.x == 1023) {
            carry[0] = off[0] + incl[0];
            carry[1] = off[1] + incl[1];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2) totals[threadIdx.x] = carry[threadIdx.x];
}

// each f and t reference's place among all of them (order kept), and each
// large node's scan bases
__global__ __launch_bounds__(BLOCK) void k_sah_apply_large(const uint8_t* __restrict__ flags, size_t n,
                                                           const uint32_t* __restrict__ bsum,
                                                           const int32_t* __restrict__ slot_of,
                                                           slot_t* __restrict__ slots, uint32_t* __restrict__ pf,
                                                           uint32_t* __restrict__ pt)
{
    const size_t p = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t fl = p < n ? flags[p] : 0u;
    const uint32_t f = fl & 1u, t = (fl >> 1) & 1u;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x / 64;
    const uint64_t bf = __ballot(f), bt = __ballot(t);
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    __shared__ uint32_t sf[BLOCK / 64], stt[BLOCK / 64];
    if (lane == 0) {
        sf[w] = (uint32_t)__popcll(bf);
        stt[w] = (uint32_t)__popcll(bt);
    }
    __syncthreads();
    uint32_t of = bsum[2 * blockIdx.x], ot = bsum[2 * blockIdx.x + 1];
    for (uint32_t q = 0; q < w; ++q) {
        of += sf[q];
        ot += stt[q];
    }
    const uint32_t sfp = of + (uint32_t)__popcll(bf & below), stp = ot + (uint32_t)__popcll(bt & below);
    if (p >= n) return;
    if (f) pf[sfp] = (uint32_t)p;
    if (t) pt[stp] = (uint32_t)p;
    const int32_t s = slot_of[p];
    if (s >= 0) {
        slot_t& sl = slots[s];
        if (sl.split) {
            if (p == sl.first) {
                sl.sf_a = sfp;
                sl.st_a = stp;
            }
            if (p + 1 == (size_t)sl.first + sl.count) sl.st_b = stp + t;
        }
    }
}

// the swaps: the k-th f of a node with its k-th t from the right
__global__ __launch_bounds__(BLOCK) void k_sah_swap_large(const uint32_t* __restrict__ pf,
                                                          const uint32_t* __restrict__ pt,
                                                          const uint32_t* __restrict__ totals,
                                                          const int32_t* __restrict__ slot_of,
                                                          const slot_t* __restrict__ slots, sref* __restrict__ refs)
{
    const uint32_t nf = totals[0];
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < nf; i += gridDim.x * BLOCK) {
        const uint32_t p = pf[i];
        const slot_t& sl = slots[slot_of[p]];
        const uint32_t k = i - sl.sf_a;
        if (k >= sl.st_b - sl.st_a) continue;  // (a node has as many f as t references)
        const uint32_t q = pt[sl.st_b - 1u - k];
        const sref a = refs[p], b = refs[q];
        refs[p] = b;
        refs[q] = a;
    }
}

// ---------------------------------------------------------------------------
// small nodes: one workgroup per node
// ---------------------------------------------------------------------------

// the wave's inclusive scan of one bin per lane (lanes 0..15) towards higher
// lanes (UP) or lower ones: exact minima, maxima and sums
template <bool UP>
__device__ inline sbin wave_bin_scan(sbin v, uint32_t lane)
{
    for (int d = 1; d < NUM_BINS; d <<= 1) {
        sbin o;
        float* of = reinterpret_cast<float*>(&o);
        const float* vf = reinterpret_cast<const float*>(&v);
        for (int k = 0; k < 12; ++k) of[k] = UP ? __shfl_up(vf[k], d) : __shfl_down(vf[k], d);
        o.n = UP ? __shfl_up(v.n, d) : __shfl_down(v.n, d);
        const bool take = UP ? lane >= (uint32_t)d : lane + d < (uint32_t)NUM_BINS;
        if (take) merge(v, o);
    }
    return v;
}

__device__ inline sbin shfl_bin(const sbin& v, int src)
{
    sbin o;
    float* of = reinterpret_cast<float*>(&o);
    const float* vf = reinterpret_cast<const float*>(&v);
    for (int k = 0; k < 12; ++k) of[k] = __shfl(vf[k], src);
    o.n = __shfl(v.n, src);
    return o;
}

// Small nodes: one wave per node.  Bins in the wave's LDS; the sweep of
// find_split with one bin per lane (prefix and suffix unions by shuffles,
// every split cost by its own lane with the reference's expression, the
// lowest cost winning and the rightmost of equal ones, as the reference's
// strict < from the right); the partition's f and t references ranked by
// ballots, in range order.
__global__ __launch_bounds__(BLOCK) void k_sah_small(const uint32_t* __restrict__ list, uint32_t m,
                                                     sref* __restrict__ refs, snode* __restrict__ nodes,
                                                     uint32_t* __restrict__ node_of, counters* C, uint32_t max_leaf,
                                                     uint32_t* large_next, slot_t* slots_next, uint32_t* small_next)
{
    constexpr int WAVES = BLOCK / 64;
    __shared__ uint32_t lbs[WAVES][NODE_BINS];
    __shared__ uint16_t s_pfs[WAVES][LARGE], s_pts[WAVES][LARGE];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x / 64;
    uint32_t* lb = lbs[w];
    uint16_t* s_pf = s_pfs[w];
    uint16_t* s_pt = s_pts[w];
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint32_t nid = blockIdx.x * WAVES + w; nid < m; nid += gridDim.x * WAVES) {
        const uint32_t node = list[nid];
        const snode nd = nodes[node];
        const f3 clo = mk3(nd.cb[0], nd.cb[1], nd.cb[2]), chi = mk3(nd.cb[3], nd.cb[4], nd.cb[5]);
        const int axis = max_index3(sub3(chi, clo));
        const float k0 = axis3(clo, axis), k1 = (float)NUM_BINS / (axis3(chi, axis) - k0);
        for (int i = lane; i < NODE_BINS; i += 64) bins_identity(lb, i);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        for (uint32_t e = lane; e < nd.count; e += 64) {
            const sref r = refs[nd.first + e];
            const f3 c = center_of(r);
            bin_add(lb + project(k0, k1, axis, c) * BIN_WORDS, r, c);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // the sweep, one bin per lane
        sbin b;
        if (lane < (uint32_t)NUM_BINS) b = read_bin(lb + lane * BIN_WORDS);
        else b = read_bin(lb);  // any bin: lanes >= 16 take no part
        const sbin accl = wave_bin_scan<true>(b, lane);
        const sbin accr = wave_bin_scan<false>(b, lane);
        const sbin left = shfl_bin(accl, (int)(lane > 0 ? lane - 1 : 0));  // acc_l[i - 1] on lane i
        const float hsa_parent = safe_hsa(mk3(nd.pb[0], nd.pb[1], nd.pb[2]), mk3(nd.pb[3], nd.pb[4], nd.pb[5]));
        float cost = 1.0f + (safe_hsa(left.p.lo, left.p.hi) / hsa_parent) * (3.0f * (float)left.n) +
                     (safe_hsa(accr.p.lo, accr.p.hi) / hsa_parent) * (3.0f * (float)accr.n);
        // a split must cost less than FLT_MAX (the sweep's start) to count
        const bool valid = lane >= 1 && lane < (uint32_t)NUM_BINS && cost < FLT_MAX;
        uint32_t key_lane = lane;
        float key = valid ? cost : FLT_MAX;
        for (int d = 32; d >= 1; d >>= 1) {
            const float ok = __shfl_xor(key, d);
            const uint32_t ol = __shfl_xor(key_lane, d);
            if (ok < key || (ok == key && ol > key_lane)) {
                key = ok;
                key_lane = ol;
            }
        }
        const int best = key < FLT_MAX && !(key > 3.0f * (float)nd.count) ? (int)key_lane : -1;
        if (best < 0) continue;  // a leaf: the node keeps its record
        const sbin L = shfl_bin(accl, best - 1), R = shfl_bin(accr, best);
        const uint32_t pivot = nd.first + L.n;
        // mark and rank the f and t references in range order
        uint32_t cf = 0, ct = 0;
        for (uint32_t e0 = 0; e0 < nd.count; e0 += 64) {
            const uint32_t e = e0 + lane;
            uint32_t f = 0, t = 0;
            if (e < nd.count) {
                const bool l = nd.first + e < pivot;
                const bool pred = project_unsafe(k0, k1, axis, center_of(refs[nd.first + e])) < best;
                f = l && !pred;
                t = !l && pred;
            }
            const uint64_t bf = __ballot(f), bt = __ballot(t);
            if (f) s_pf[cf + (uint32_t)__popcll(bf & below)] = (uint16_t)e;
            if (t) s_pt[ct + (uint32_t)__popcll(bt & below)] = (uint16_t)e;
            cf += (uint32_t)__popcll(bf);
            ct += (uint32_t)__popcll(bt);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (uint32_t k = lane; k < min(cf, ct); k += 64) {
            const uint32_t p = nd.first + s_pf[k], q = nd.first + s_pt[ct - 1u - k];
            const sref a = refs[p], bb = refs[q];
            refs[p] = bb;
            refs[q] = a;
        }
        uint32_t c = 0;
        if (lane == 0) {
            c = atomicAdd(&C->nodes, 2u);
            const uint32_t sign = axis3(L.p.lo, axis) < axis3(R.p.lo, axis) ? 0u : 1u;
            nodes[node].child = c;
            nodes[node].info = (nd.info & ~0x1FFFFu) | 1u | ((uint32_t)axis << 8) | (sign << 16);
            int32_t slot;
            make_child(nodes, c, node, L, nd.first, max_leaf, C, large_next, slots_next, small_next, slot);
            make_child(nodes, c + 1u, node, R, pivot, max_leaf, C, large_next, slots_next, small_next, slot);
        }
        c = __shfl(c, 0);
        for (uint32_t e = lane; e < nd.count; e += 64) node_of[nd.first + e] = nd.first + e < pivot ? c : c + 1u;
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------
// numbering and output
// ---------------------------------------------------------------------------

// inner nodes below (and including) each node of a level, deepest level first
__global__ __launch_bounds__(BLOCK) void k_sah_inner_count(const snode* __restrict__ nodes, uint32_t b, uint32_t e,
                                                           uint32_t* __restrict__ ic)
{
    const uint32_t i = b + blockIdx.x * BLOCK + threadIdx.x;
    if (i >= e) return;
    const snode& nd = nodes[i];
    ic[i] = (nd.info & 1u) ? 1u + ic[nd.child] + ic[nd.child + 1u] : 0u;
}

// the reference's numbering, top down: r = inner nodes before this one in
// right-first pre-order, the node's own index (its parent's first child
// 1 + 2 r(parent), + 1 on the right), and the primitives listed before it
__global__ __launch_bounds__(BLOCK) void k_sah_number(const snode* __restrict__ nodes, uint32_t b, uint32_t e,
                                                      const uint32_t* __restrict__ ic, uint32_t* __restrict__ rk,
                                                      uint32_t* __restrict__ refidx, uint32_t* __restrict__ before)
{
    const uint32_t i = b + blockIdx.x * BLOCK + threadIdx.x;
    if (i >= e) return;
    const snode& nd = nodes[i];
    if (!(nd.info & 1u)) return;
    const uint32_t L = nd.child, R = nd.child + 1u;
    const uint32_t fc = 1u + 2u * rk[i];
    refidx[L] = fc;
    refidx[R] = fc + 1u;
    rk[R] = rk[i] + 1u;
    rk[L] = rk[i] + 1u + ic[R];
    before[R] = before[i];
    before[L] = before[i] + nodes[R].count;
}

__global__ void k_sah_number_root(uint32_t* rk, uint32_t* refidx, uint32_t* before)
{
    rk[0] = 0;
    refidx[0] = 0;
    before[0] = 0;
}

__device__ inline uint32_t entry_of(const snode& c, uint32_t rk, uint32_t before)
{
    return (c.info & 1u) ? rk : leaf_entry(before, c.count);
}

// reference-layout nodes and pair records
__global__ __launch_bounds__(BLOCK) void k_sah_write_nodes(const snode* __restrict__ nodes, uint32_t num,
                                                           const uint32_t* __restrict__ rk,
                                                           const uint32_t* __restrict__ refidx,
                                                           const uint32_t* __restrict__ before,
                                                           vr_bvh_node* __restrict__ ref, pair_node* __restrict__ pairs,
                                                           counters* C)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= num) return;
    const snode nd = nodes[i];
    const bool inner = nd.info & 1u;
    vr_bvh_node o;
    for (int k = 0; k < 3; ++k) {
        o.lo[k] = nd.pb[k];
        o.hi[k] = nd.pb[3 + k];
    }
    o.first = inner ? 1u + 2u * rk[i] : before[i];
    o.num_prims = inner ? 0 : (uint16_t)min(nd.count, 65535u);
    o.axis = inner ? (uint8_t)((nd.info >> 8) & 3u) : 0;
    o.sign = inner ? (uint8_t)((nd.info >> 16) & 1u) : 0;
    ref[refidx[i]] = o;
    if (!inner) {
        atomicMax(&C->max_leaf, nd.count);
        if (nd.count > MAX_LEAF_COUNT) atomicAdd(&C->bad, 1u);
        return;
    }
    const snode L = nodes[nd.child], R = nodes[nd.child + 1u];
    pair_node pn;
    for (int k = 0; k < 6; ++k) {
        pn.f[k] = L.pb[k];
        pn.f[6 + k] = R.pb[k];
    }
    pn.e[0] = entry_of(L, rk[nd.child], before[nd.child]);
    pn.e[1] = entry_of(R, rk[nd.child + 1u], before[nd.child + 1u]);
    const uint32_t parent = nd.info >> 17;
    if (parent == 0u) {
        pn.e[2] = 0xFFFFFFFFu;
    } else {
        const uint32_t pa = parent - 1u;
        pn.e[2] = (rk[pa] << 1) | (i == nodes[pa].child + 1u ? 1u : 0u);
    }
    pn.e[3] = 0u;
    pairs[rk[i]] = pn;
}

// the index list and the triangles in its order
__global__ __launch_bounds__(BLOCK) void k_sah_write_prims(const sref* __restrict__ refs, size_t n,
                                                           const uint32_t* __restrict__ node_of,
                                                           const snode* __restrict__ nodes,
                                                           const uint32_t* __restrict__ before,
                                                           const vr_triangle* __restrict__ prims,
                                                           uint32_t* __restrict__ indices, tri_node* __restrict__ tris)
{
    const size_t p = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n) return;
    const uint32_t x = node_of[p];
    const uint32_t pos = before[x] + (uint32_t)(p - nodes[x].first);
    const uint32_t id = (uint32_t)refs[p].index;
    indices[pos] = id;
    const vr_triangle t = prims[id];
    tri_node tn;
    tn.a = make_float4(t.v1[0], t.v1[1], t.v1[2], __uint_as_float(t.prim_id));
    tn.b = make_float4(t.e1[0], t.e1[1], t.e1[2], __uint_as_float(t.geom_id));
    tn.c = make_float4(t.e2[0], t.e2[1], t.e2[2], __uint_as_float(id));
    tris[pos] = tn;
}

unsigned grid_for(size_t n) { return std::max(1u, std::min(div_up(n, BLOCK), 65535u)); }
size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

int sah_build_device(const vr_triangle* d_prims, size_t n, int max_leaf_size, bvh_state& s, hipStream_t st)
{
    VR_CHECK(n <= (size_t)MAX_LEAF_FIRST, VR_ERR_UNSUPPORTED, "sah: more than 2^26 - 1 primitives");
    if (max_leaf_size <= 0) max_leaf_size = 4;  // build_top_down.h:153-156
    VR_CHECK(max_leaf_size <= (int)MAX_LEAF_COUNT, VR_ERR_UNSUPPORTED, "sah: leaf size above 32");
    s.num_prims = n;
    s.is_lbvh = false;
    s.pooled = true;
    if (n == 0) return VR_OK;
    const uint32_t max_leaf = (uint32_t)max_leaf_size;
    const size_t max_nodes = 2 * n - 1;
    const size_t max_large = n / LARGE + 2;
    const size_t nblk = div_up(n, BLOCK);
    // temporaries, stream ordered
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += round_up(bytes, 256);
        return at;
    };
    const size_t o_refs = take(n * sizeof(sref)), o_nodes = take(max_nodes * sizeof(snode)),
                 o_slot0 = take(n * 4), o_slot1 = take(n * 4), o_nodeof = take(n * 4), o_flags = take(n),
                 o_pf = take(n * 4), o_pt = take(n * 4), o_bsum = take(2 * nblk * 4), o_tot = take(16),
                 o_bins = take(max_large * NODE_BINS * 4), o_sl0 = take(max_large * sizeof(slot_t)),
                 o_sl1 = take(max_large * sizeof(slot_t)), o_l0 = take(max_large * 4), o_l1 = take(max_large * 4),
                 o_sm0 = take(n * 4), o_sm1 = take(n * 4), o_cnt = take(sizeof(counters)), o_root = take(64),
                 o_ic = take(max_nodes * 4), o_rk = take(max_nodes * 4), o_ri = take(max_nodes * 4),
                 o_bf = take(max_nodes * 4);
    char* tmp = nullptr;
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&tmp), off, st));
    auto at = [&](size_t o) { return tmp + o; };
    sref* refs = reinterpret_cast<sref*>(at(o_refs));
    snode* nodes = reinterpret_cast<snode*>(at(o_nodes));
    int32_t* slot_of = reinterpret_cast<int32_t*>(at(o_slot0));
    int32_t* slot_next = reinterpret_cast<int32_t*>(at(o_slot1));
    uint32_t* node_of = reinterpret_cast<uint32_t*>(at(o_nodeof));
    uint8_t* flags = reinterpret_cast<uint8_t*>(at(o_flags));
    uint32_t* pf = reinterpret_cast<uint32_t*>(at(o_pf));
    uint32_t* pt = reinterpret_cast<uint32_t*>(at(o_pt));
    uint32_t* bsum = reinterpret_cast<uint32_t*>(at(o_bsum));
    uint32_t* totals = reinterpret_cast<uint32_t*>(at(o_tot));
    uint32_t* bins = reinterpret_cast<uint32_t*>(at(o_bins));
    slot_t* slots = reinterpret_cast<slot_t*>(at(o_sl0));
    slot_t* slots_next = reinterpret_cast<slot_t*>(at(o_sl1));
    uint32_t* large = reinterpret_cast<uint32_t*>(at(o_l0));
    uint32_t* large_next = reinterpret_cast<uint32_t*>(at(o_l1));
    uint32_t* small = reinterpret_cast<uint32_t*>(at(o_sm0));
    uint32_t* small_next = reinterpret_cast<uint32_t*>(at(o_sm1));
    counters* C = reinterpret_cast<counters*>(at(o_cnt));
    uint32_t* root_words = reinterpret_cast<uint32_t*>(at(o_root));
    uint32_t* ic = reinterpret_cast<uint32_t*>(at(o_ic));
    uint32_t* rk = reinterpret_cast<uint32_t*>(at(o_rk));
    uint32_t* refidx = reinterpret_cast<uint32_t*>(at(o_ri));
    uint32_t* before = reinterpret_cast<uint32_t*>(at(o_bf));

    int rc = VR_OK;
    auto fail_free = [&](int code) {
        (void)hipFreeAsync(tmp, st);
        return code;
    };
    // references, root bounds, the root
    hipLaunchKernelGGL(k_sah_init_words, dim3(1), dim3(64), 0, st, root_words);
    hipLaunchKernelGGL(k_sah_refs, dim3(std::min(grid_for(n), 4096u)), dim3(BLOCK), 0, st, d_prims, n, refs,
                       root_words);
    hipLaunchKernelGGL(k_sah_root, dim3(1), dim3(1), 0, st, root_words, (uint32_t)n, max_leaf, nodes, C, large_next,
                       slots_next, small_next);
    hipLaunchKernelGGL(k_sah_fill_i32, dim3(grid_for(n)), dim3(BLOCK), 0, st, reinterpret_cast<int32_t*>(node_of), n,
                       0);
    VR_HIP_CHECK(hipGetLastError());
    counters h{};
    std::vector<uint32_t> level_start{0u, 1u};  // node ranges of the levels
    for (int level = 0;; ++level) {
        VR_HIP_CHECK(hipMemcpyAsync(&h, C, sizeof(h), hipMemcpyDeviceToHost, st));
        VR_HIP_CHECK(hipStreamSynchronize(st));
        const uint32_t n_large = h.large_next, n_small = h.small_next;
        if (n_large == 0 && n_small == 0) break;
        if (level > 4096 || n_large > max_large || n_small > n)
            return fail_free(fail(VR_ERR_HIP, "sah: level bookkeeping overflow"));
        std::swap(large, large_next);
        std::swap(slots, slots_next);
        std::swap(small, small_next);
        // the large nodes' references carry their slot: the first level sets
        // them (every reference belongs to the root)
        if (level == 0)
            hipLaunchKernelGGL(k_sah_fill_i32, dim3(grid_for(n)), dim3(BLOCK), 0, st, slot_of, n,
                               n_large ? 0 : -1);
        hipLaunchKernelGGL(k_sah_fill_i32, dim3(1), dim3(64), 0, st, reinterpret_cast<int32_t*>(&C->large_next), 2,
                           0);
        if (n_large) {
            const size_t words = (size_t)n_large * NODE_BINS;
            hipLaunchKernelGGL(k_sah_bins_clear, dim3(grid_for(words)), dim3(BLOCK), 0, st, bins, words);
            const unsigned bgrid = std::min(grid_for(n), 2048u);
            const size_t stretch = round_up(div_up(n, bgrid), BLOCK);
            hipLaunchKernelGGL(k_sah_bin_large, dim3(div_up(n, stretch)), dim3(BLOCK), 0, st, refs, n, slot_of,
                               slots, bins, stretch);
            hipLaunchKernelGGL(k_sah_split_large, dim3(div_up(n_large, 64)), dim3(64), 0, st, n_large, slots, bins,
                               nodes, C, max_leaf, large_next, slots_next, small_next);
            hipLaunchKernelGGL(k_sah_flags_large, dim3((unsigned)nblk), dim3(BLOCK), 0, st, refs, n, slot_of, slots,
                               slot_next, node_of, flags, bsum);
            hipLaunchKernelGGL(k_sah_scan, dim3(1), dim3(1024), 0, st, bsum, (uint32_t)nblk, totals);
            hipLaunchKernelGGL(k_sah_apply_large, dim3((unsigned)nblk), dim3(BLOCK), 0, st, flags, n, bsum, slot_of,
                               slots, pf, pt);
            hipLaunchKernelGGL(k_sah_swap_large, dim3(grid_for(n)), dim3(BLOCK), 0, st, pf, pt, totals, slot_of, slots,
                               refs);
            std::swap(slot_of, slot_next);
        }
        if (n_small)
            hipLaunchKernelGGL(k_sah_small, dim3(std::min(div_up(n_small, BLOCK / 64), 65535u)), dim3(BLOCK), 0, st, small, n_small, refs,
                               nodes, node_of, C, max_leaf, large_next, slots_next, small_next);
        VR_HIP_CHECK(hipGetLastError());
        VR_HIP_CHECK(hipMemcpyAsync(&h, C, sizeof(h), hipMemcpyDeviceToHost, st));
        VR_HIP_CHECK(hipStreamSynchronize(st));
        level_start.push_back(h.nodes);
    }
    const uint32_t num_nodes = h.nodes;
    // levels: [level_start[l], level_start[l + 1]); the last entry may repeat
    while (level_start.size() > 2 && level_start[level_start.size() - 1] == level_start[level_start.size() - 2])
        level_start.pop_back();
    const int levels = (int)level_start.size() - 1;
    for (int l = levels - 1; l >= 0; --l) {
        const uint32_t b = level_start[l], e = level_start[l + 1];
        hipLaunchKernelGGL(k_sah_inner_count, dim3(div_up(e - b, BLOCK)), dim3(BLOCK), 0, st, nodes, b, e, ic);
    }
    hipLaunchKernelGGL(k_sah_number_root, dim3(1), dim3(1), 0, st, rk, refidx, before);
    for (int l = 0; l < levels; ++l) {
        const uint32_t b = level_start[l], e = level_start[l + 1];
        hipLaunchKernelGGL(k_sah_number, dim3(div_up(e - b, BLOCK)), dim3(BLOCK), 0, st, nodes, b, e, ic, rk, refidx,
                           before);
    }
    VR_HIP_CHECK(hipGetLastError());
    // the handle's arrays
    uint32_t num_inner = (num_nodes - 1u) / 2u;
    s.num_pairs = num_inner;
    s.num_ref_nodes = num_nodes;
    s.depth = (uint32_t)levels;
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&s.ref_nodes), num_nodes * sizeof(vr_bvh_node), st));
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&s.indices), n * 4, st));
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&s.tris), n * sizeof(tri_node), st));
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&s.pairs), std::max<size_t>(num_inner, 1) * sizeof(pair_node),
                                st));
    VR_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&s.quads), std::max<size_t>(num_inner, 1) * sizeof(quad_node),
                                st));
    hipLaunchKernelGGL(k_sah_write_nodes, dim3(div_up(num_nodes, BLOCK)), dim3(BLOCK), 0, st, nodes, num_nodes, rk,
                       refidx, before, s.ref_nodes, s.pairs, C);
    hipLaunchKernelGGL(k_sah_write_prims, dim3((unsigned)nblk), dim3(BLOCK), 0, st, refs, n, node_of, nodes, before,
                       d_prims, s.indices, s.tris);
    VR_HIP_CHECK(hipGetLastError());
    VR_HIP_CHECK(hipMemcpyAsync(&h, C, sizeof(h), hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipStreamSynchronize(st));
    VR_HIP_CHECK(hipFreeAsync(tmp, st));
    s.max_leaf = std::max(1u, h.max_leaf);
    VR_CHECK(h.bad == 0, VR_ERR_UNSUPPORTED, "sah: a leaf of more than 32 primitives (the traversal's limit)");
    if (num_nodes == 1) {
        s.root = leaf_entry(0u, (uint32_t)n);
        s.num_pairs = 0;
        return VR_OK;
    }
    s.root = 0u;
    VR_CHECK(stack_fits(s), VR_ERR_UNSUPPORTED, "sah: tree deeper than the traversal stack");
    rc = collapse_quads(s, 0u, st);
    (void)rc;
    return rc;
}

}  // namespace vr
