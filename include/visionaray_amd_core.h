// visionaray_amd_core.h -- device data layouts, exact-arithmetic helpers,
// BVH traversal, random numbers, camera rays and pixel order of the HIP
// library (gfx950).  The library's own kernels and user kernels written
// against the device API (visionaray_amd_device.hpp) compile this same code,
// so a user kernel computes exactly what the library's kernels compute.
//
// Compile with -ffp-contract=off (as the library is): each float operation
// below then rounds exactly as written, in the order the reference writes it,
// so results are bit-identical to the reference's CPU arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "visionaray_amd.h"

namespace vr {

// ---------------------------------------------------------------------------
// device layouts
// ---------------------------------------------------------------------------

// Traversal node: the two children of one inner node in one 64-byte record,
// so an inner visit is a single cache-line fetch of four 16-byte loads.
//   f[0..5]   child 0 box  lo.x lo.y lo.z hi.x hi.y hi.z
//   f[6..11]  child 1 box
//   e[0], e[1] child entries (see below); e[2] parent pair (-1 at the root
//   pair); e[3] for an LBVH the far end j of the pair's leaf range
//   [min(index, j), max(index, j)], unused otherwise.
// Child entry: bit 31 clear -> inner node, the index of its pair record;
// bit 31 set -> leaf covering triangles [first, first + count) with
// first = entry & 0x03FFFFFF and count = ((entry >> 26) & 31) + 1.
struct alignas(64) pair_node {
    float f[12];
    uint32_t e[4];
};
static_assert(sizeof(pair_node) == 64, "pair_node is one 64-byte line");

constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t LEAF_FIRST_MASK = 0x03FFFFFFu;
constexpr uint32_t MAX_LEAF_FIRST = LEAF_FIRST_MASK;
constexpr uint32_t MAX_LEAF_COUNT = 32;

__host__ __device__ inline uint32_t leaf_entry(uint32_t first, uint32_t count)
{
    return LEAF_BIT | ((count - 1u) << 26) | first;
}

// Traversal node: up to four children in one 128-byte line (eight 16-byte
// loads), boxes as structure of arrays so each load carries one coordinate
// of all four children.  It is a collapse of the binary tree (pair records):
// every child box is the box of a binary node, the same exact union.
//   lo_x .. hi_z  child c's box component in lane c; an unused slot has
//                 lo = hi = +inf, a box no ray enters
//   child[c]      entry as in pair_node (inner: quad index)
//   meta.x        number of children (2..4); meta.y..w unused
struct alignas(128) quad_node {
    float4 lo_x, lo_y, lo_z, hi_x, hi_y, hi_z;
    uint4 child;
    uint4 meta;
};
static_assert(sizeof(quad_node) == 128, "quad_node is one 128-byte line");

// 8-wide traversal node (bvh_multi_node<8>, bvh.h:167-312: the collapse of
// the binary tree into nodes of up to eight children): two quad_node halves,
// children 0-3 in h[0] and 4-7 in h[1], 256 bytes; h[0].meta.x = number of
// children.  An inner child refers to the oct_node of its own pair index.
struct alignas(128) oct_node {
    quad_node h[2];
};
static_assert(sizeof(oct_node) == 256, "oct_node is two 128-byte lines");

// The same node with 8-bit child planes (the idea of bvh_compressed_node<8>,
// bvh.h:316-426), one 128-byte line.  Plane = fmaf(q, scale, org) in float;
// the quantization rounds outward through exactly this decode, so every
// decoded child box contains the child's exact box.
//   org.xyz    origin (lower corner of the node's box); org.w = number of
//              children (uint bits); slots >= that number are unused
//   scale.xyz  power-of-two step per axis
//   q          bytes 32..79: lo_x[8] lo_y[8] lo_z[8] hi_x[8] hi_y[8] hi_z[8]
//   child      entries of slots 0-7 (as quad_node)
struct alignas(128) octq_node {
    float4 org;
    float4 scale;
    uint4 q[3];
    uint4 child[2];
    uint4 pad;
};
static_assert(sizeof(octq_node) == 128, "octq_node is one 128-byte line");

// traversal node kinds of a BVH (bvh_view::wide holds the 8-wide ones)
enum node_kind { NODE_QUAD = 0, NODE_OCT = 1, NODE_OCTQ = 2 };

// Triangle in traversal order, 48 bytes = three 16-byte loads:
//   a = v1.xyz, prim_id    b = e1.xyz, geom_id    c = e2.xyz, original index
struct alignas(16) tri_node {
    float4 a;
    float4 b;
    float4 c;
};
static_assert(sizeof(tri_node) == 48, "tri_node layout");

// Device view of a BVH, passed by value into kernels (the analogue of
// index_bvh_ref_t, bvh.h:495-569).
struct bvh_view {
    const quad_node* quads;
    const tri_node* tris;
    const pair_node* pairs;  // the binary tree (traverse_binary)
    uint32_t num_prims;
    uint32_t root;  // entry of the root: 0 (quad 0 / pair 0) or a leaf entry; unused when num_prims == 0
    uint32_t num_quads;  // also the number of pair records
    const void* wide;    // oct_node / octq_node per pair index, or null
};


// ---------------------------------------------------------------------------
// exact scalar helpers (reference semantics)
// ---------------------------------------------------------------------------

// math/detail/math.h:116-128 ternary min/max
__host__ __device__ inline float tmin_f(float x, float y) { return x < y ? x : y; }
__host__ __device__ inline float tmax_f(float x, float y) { return x < y ? y : x; }

struct f3 {
    float x, y, z;
};
__host__ __device__ inline f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 mul3(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline f3 muls3(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 smul3(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline f3 neg3(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline f3 min3v(f3 a, f3 b) { return f3{tmin_f(a.x, b.x), tmin_f(a.y, b.y), tmin_f(a.z, b.z)}; }
__host__ __device__ inline f3 max3v(f3 a, f3 b) { return f3{tmax_f(a.x, b.x), tmax_f(a.y, b.y), tmax_f(a.z, b.z)}; }
// math/detail/vector3.inl:292-307
__host__ __device__ inline f3 cross3(f3 u, f3 v)
{
    return f3{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
__host__ __device__ inline float dot3(f3 u, f3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
// normalize = v * (1 / sqrt(dot(v, v))), both correctly rounded
// (math/detail/vector3.inl:332-335, math/detail/math.h:596-599)
__host__ __device__ inline f3 normalize3(f3 v) { return muls3(v, 1.0f / sqrtf(dot3(v, v))); }
// math/detail/vector3.inl:432-447
__host__ __device__ inline float min_el3(f3 u)
{
    float r = u.x;
    if (u.y < r) r = u.y;
    if (u.z < r) r = u.z;
    return r;
}
__host__ __device__ inline float max_el3(f3 u)
{
    float r = u.x;
    if (u.y > r) r = u.y;
    if (u.z > r) r = u.z;
    return r;
}
// math/detail/vector.inl:679-686
__host__ __device__ inline f3 faceforward3(f3 n, f3 i, f3 nref) { return dot3(nref, i) < 0.0f ? neg3(n) : n; }

// ---------------------------------------------------------------------------
// ray / triangle: Moller-Trumbore, two-sided, inclusive edges
// (math/intersect.h:176-229); returns hit and writes t, b1, b2
// ---------------------------------------------------------------------------
__device__ inline bool intersect_tri(f3 ori, f3 dir, f3 v1, f3 e1, f3 e2, float& t, float& b1, float& b2)
{
    f3 s1 = cross3(dir, e2);
    float div = dot3(s1, e1);
    if (!(div != 0.0f)) return false;
    float inv_div = 1.0f / div;
    f3 d = sub3(ori, v1);
    b1 = dot3(d, s1) * inv_div;
    if (!(b1 >= 0.0f && b1 <= 1.0f)) return false;
    f3 s2 = cross3(d, e1);
    b2 = dot3(dir, s2) * inv_div;
    if (!(b2 >= 0.0f && b1 + b2 <= 1.0f)) return false;
    t = dot3(e2, s2) * inv_div;
    return true;
}

struct hit_t {
    float t;
    float u;
    float v;
    int32_t tri;  // traversal position of the hit triangle, -1 = miss
};

// ---------------------------------------------------------------------------
// BVH traversal (detail/bvh/intersect.inl:166-356) over the 4-wide collapse.
//
// The root box is not tested.  At a node every child box is slab tested
// exactly as the reference tests a binary node's children: tnear =
// max(tmin, max_el(min(tlo, thi))) and tfar = min(tmax, min_el(max(tlo,
// thi))) with tlo = (lo - ori) * inv_dir, and the child is entered when
// tfar >= tnear and tnear < t_closest.  Entered children are visited nearest
// first (a sorting network on tnear); the others are deferred on the stack,
// farthest deepest.  Primitives are accepted with is_closer (update_if.h:
// 68-107): hit && t < t_closest && tmin <= t <= tmax.  The closest hit is the
// reference's; only which of several triangles hit at exactly the same t is
// reported can differ (the reference keeps the first found in its binary
// near-first order).
//
// The stack keeps its top LDS_DEPTH entries in LDS, one column per lane
// (stack[k * STRIDE], so the 64 lanes of a wave touch 64 consecutive words);
// deeper entries go to a per-lane private array that only rays on unusually
// deep paths touch.  A node defers at most three children, so MAX_STACK
// entries cover trees of depth (MAX_STACK - 1) / 3 (checked on the host).
// ---------------------------------------------------------------------------

struct trace_counters {
    uint32_t nodes;
    uint32_t tris;
};

// a Karras tree has at most 65 levels (the common prefix of a node's range,
// 30 code bits + 32 index bits, grows strictly from parent to child), so its
// collapse at most 64 quad levels: 3 * 64 + 1 entries always suffice
constexpr int MAX_STACK = 196;
// 8-wide nodes defer up to 7 children per level: trees of at most 48 binary
// levels
constexpr int MAX_STACK_WIDE = 7 * 48 + 1;
template <int NODE>
constexpr int max_stack_of() { return NODE == NODE_QUAD ? MAX_STACK : MAX_STACK_WIDE; }

// Slab test of one child box; returns tnear, or +inf when the child is not
// entered.  The slab distances round exactly as the reference's; the min/max
// reductions use the hardware min/max (v_min3/v_max3), which agree with the
// reference's ternary forms on every non-NaN input up to the sign of a zero,
// and a signed zero changes no comparison below.
__device__ __forceinline__ float slab_enter(float lx, float ly, float lz, float hx, float hy, float hz, f3 ori,
                                            f3 inv, float tmin, float tmax, float tclosest)
{
    const float x0 = (lx - ori.x) * inv.x, x1 = (hx - ori.x) * inv.x;
    const float y0 = (ly - ori.y) * inv.y, y1 = (hy - ori.y) * inv.y;
    const float z0 = (lz - ori.z) * inv.z, z1 = (hz - ori.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    return (tf >= tn && tn < tclosest) ? tn : __builtin_inff();
}

__device__ __forceinline__ void cswap(float& ka, uint32_t& ea, float& kb, uint32_t& eb)
{
    const bool s = kb < ka;
    const float k = s ? kb : ka;
    kb = s ? ka : kb;
    ka = k;
    const uint32_t e = s ? eb : ea;
    eb = s ? ea : eb;
    ea = e;
}

// One query in flight.  A query advances one iteration at a time
// (trav_iter) so that a lane can hold its query while other lanes of the
// wave shade or start new ones; traverse() runs one query to the end.
struct trav_state {
    f3 ori, dir, inv;
    float tmin, tmax;
    uint32_t cur;     // entry being visited
    int sp;           // deferred entries on the stack
    uint32_t budget;  // visits left: a ray visits each node at most once
    uint32_t lf, ln;  // trav_iter: triangles [lf, lf + ln) of a leaf still to test
    // byte offsets in a quad_node of the near slab planes per axis, one byte
    // each (x, y, z from the low byte): lo (0, 16, 32) where inv >= 0, hi
    // (48, 64, 80) where inv < 0; the far planes are at near ^ 48, ^ 80, ^ 112
    uint32_t near;
    hit_t r;
};

// quad_node byte offsets: lo_x 0, lo_y 16, lo_z 32, hi_x 48, hi_y 64, hi_z 80
constexpr uint32_t QN_FAR_X = 48u, QN_FAR_Y = 80u, QN_FAR_Z = 112u;
// quads addressable with 32-bit byte offsets (node << 7)
constexpr uint32_t QN_NEAR_FAR_MAX = 1u << 25;
// oct_node: both halves as quad_node (node << 8, + 128 for the second half)
constexpr uint32_t OCT_NEAR_FAR_MAX = 1u << 24;
// octq_node byte offsets: lo_x 32, lo_y 40, lo_z 48, hi_x 56, hi_y 64,
// hi_z 72; near plane 32 + 8 axis (+ 24 where inv < 0), half the quad_node
// offset plus 32; the far plane at near ^ 24, ^ 104, ^ 120
constexpr uint32_t OQ_FAR_X = 24u, OQ_FAR_Y = 104u, OQ_FAR_Z = 120u;

__device__ __forceinline__ void trav_begin(trav_state& T, const bvh_view& b, f3 ori, f3 dir, float tmin, float tmax)
{
    T.ori = ori;
    T.dir = dir;
    T.inv = mk3(dir.x != 0.0f ? 1.0f / dir.x : FLT_MAX, dir.y != 0.0f ? 1.0f / dir.y : FLT_MAX,
                dir.z != 0.0f ? 1.0f / dir.z : FLT_MAX);
    T.tmin = tmin;
    T.tmax = tmax;
    T.near = (T.inv.x < 0.0f ? 48u : 0u) | ((T.inv.y < 0.0f ? 64u : 16u) << 8) | ((T.inv.z < 0.0f ? 80u : 32u) << 16);
    T.cur = b.root;
    T.sp = 0;
    T.budget = 2u * b.num_prims + 2u;
    T.lf = 0;
    T.ln = 0;
    T.r.t = FLT_MAX;
    T.r.u = 0.0f;
    T.r.v = 0.0f;
    T.r.tri = -1;
}

// entry of a finished query (never a valid leaf: first + count would exceed
// any primitive count)
constexpr uint32_t TRAV_DONE = 0xFFFFFFFFu;

// node iterations are chosen while at least this many times as many lanes
// have node work as have a triangle to test (a node visit costs about twice a
// triangle test)
// persistent ray batches: idle lanes take new rays once this many of their
// wave wait (fewer refills, each serving more lanes); measured best on the
// large scene's bounce rays: 32 for the one-level trace, 16 for the
// two-level one
#ifndef VR_TRACE_REFILL_LANES
#define VR_TRACE_REFILL_LANES 32
#endif
#ifndef VR_TLAS_REFILL_LANES
#define VR_TLAS_REFILL_LANES 16
#endif
#ifndef VR_TRAV_NODE_BIAS
#define VR_TRAV_NODE_BIAS 2u
#endif

// One iteration of the query, either one node visit or one triangle test,
// whichever more lanes of the wave can do (the wave votes).  A lane keeps
// one leaf's triangle range and, after taking it, goes on descending (the
// leaf's triangles wait for a triangle iteration), so lanes that meet leaves
// at different depths stay busy; only the order in which triangles are
// tested changes (Aila and Laine 2009, speculative traversal).  Returns true
// when this lane's query is finished (T.r is its result).
// any_rt: this query is an any-hit query (a shadow ray) although the kernel
// traces closest hits otherwise.
// The traversal's box and triangle tests: the library's own, the slab test
// of one child box (tnear, or +inf when the child is not entered) and
// Moller-Trumbore on the triangle's three 16-byte records.  Any type with
// these two members replaces them (the intersectors of the device API).
struct default_isect {
    static constexpr bool near_far = true;  // box_nf below is the same test as box
    __device__ __forceinline__ float box(float lx, float ly, float lz, float hx, float hy, float hz,
                                         const trav_state& T, float tclosest) const
    {
        return slab_enter(lx, ly, lz, hx, hy, hz, T.ori, T.inv, T.tmin, T.tmax, tclosest);
    }
    // The same slab test given each axis' near and far plane (the planes the
    // ray's direction sign selects): with inv >= 0, (lo - o) * inv <=
    // (hi - o) * inv (subtraction and multiplication round monotonically), so
    // min and max of the two distances are the near and far ones exactly;
    // with inv < 0 the other way round.
    __device__ __forceinline__ float box_nf(float nx, float ny, float nz, float fx, float fy, float fz,
                                            const trav_state& T, float tclosest) const
    {
        const float x0 = (nx - T.ori.x) * T.inv.x, x1 = (fx - T.ori.x) * T.inv.x;
        const float y0 = (ny - T.ori.y) * T.inv.y, y1 = (fy - T.ori.y) * T.inv.y;
        const float z0 = (nz - T.ori.z) * T.inv.z, z1 = (fz - T.ori.z) * T.inv.z;
        const float tn = fmaxf(fmaxf(x0, y0), fmaxf(z0, T.tmin));
        const float tf = fminf(fminf(x1, y1), fminf(z1, T.tmax));
        return (tf >= tn && tn < tclosest) ? tn : __builtin_inff();
    }
    __device__ __forceinline__ bool tri(const trav_state& T, float4 a, float4 b, float4 c, float& t, float& u,
                                        float& v) const
    {
        return intersect_tri(T.ori, T.dir, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), t, u, v);
    }
};

// The library's tests, taking two triangles of a leaf per triangle iteration
// (in leaf order, so with the same result): the persistent ray batches use
// it (measured on the large scene's bounce rays: 2.70 -> 2.86-2.93 Grays/s);
// the path tracer keeps one per iteration (33.0 -> 33.5 ms per frame with
// two)
struct paired_isect : default_isect {
};
template <typename I>
struct isect_pairs {
    static constexpr bool value = false;
};
template <>
struct isect_pairs<paired_isect> {
    static constexpr bool value = true;
};

// Order of one query's node visits and triangle tests.  TRAV_VOTED: each
// iteration the wave votes (below), so the order depends on which queries
// share the wave; the fast kernels use it.  TRAV_LANE: a lane tests the
// triangles of the leaf it holds before it visits the next node, an order
// that depends on the ray and the tree alone (the closest hit is the same
// either way, up to which of several triangles hit at exactly the same t is
// kept); counting passes, the AO kernel and the device API use it, so their
// per-ray counts and results do not depend on the launch.
enum trav_order { TRAV_VOTED = 0, TRAV_LANE = 1 };

// The 8 (tnear, entry) pairs of a wide node's children in ascending tnear
// (a 19-comparator network, depth 6; ties keep no particular order)
__device__ __forceinline__ void sort8(float* k, uint32_t* e)
{
    cswap(k[0], e[0], k[2], e[2]);
    cswap(k[1], e[1], k[3], e[3]);
    cswap(k[4], e[4], k[6], e[6]);
    cswap(k[5], e[5], k[7], e[7]);
    cswap(k[0], e[0], k[4], e[4]);
    cswap(k[1], e[1], k[5], e[5]);
    cswap(k[2], e[2], k[6], e[6]);
    cswap(k[3], e[3], k[7], e[7]);
    cswap(k[0], e[0], k[1], e[1]);
    cswap(k[2], e[2], k[3], e[3]);
    cswap(k[4], e[4], k[5], e[5]);
    cswap(k[6], e[6], k[7], e[7]);
    cswap(k[2], e[2], k[4], e[4]);
    cswap(k[3], e[3], k[5], e[5]);
    cswap(k[1], e[1], k[4], e[4]);
    cswap(k[3], e[3], k[6], e[6]);
    cswap(k[1], e[1], k[2], e[2]);
    cswap(k[3], e[3], k[4], e[4]);
    cswap(k[5], e[5], k[6], e[6]);
}

// Slab tests of the eight children of wide node `node` (NODE_OCT or
// NODE_OCTQ): tnear per slot (+inf: not entered) and the slots' entries.
template <int NODE, typename ISECT>
__device__ __forceinline__ void wide_children(const trav_state& T, const bvh_view& b, uint32_t node, float tc,
                                              const ISECT& is, float* k, uint32_t* e)
{
    const char* nb = reinterpret_cast<const char*>(b.wide);
    if constexpr (NODE == NODE_OCT) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            // each half is a quad_node: the same near/far plane offsets
            const uint32_t vo = (node << 8) + 128u * h;
            const uint32_t ax = vo | (T.near & 0xFFu), ay = vo | ((T.near >> 8) & 0xFFu), az = vo | (T.near >> 16);
            const float4 nx = *reinterpret_cast<const float4*>(nb + ax);
            const float4 ny = *reinterpret_cast<const float4*>(nb + ay);
            const float4 nz = *reinterpret_cast<const float4*>(nb + az);
            const float4 fx = *reinterpret_cast<const float4*>(nb + (ax ^ QN_FAR_X));
            const float4 fy = *reinterpret_cast<const float4*>(nb + (ay ^ QN_FAR_Y));
            const float4 fz = *reinterpret_cast<const float4*>(nb + (az ^ QN_FAR_Z));
            uint4 ch = *reinterpret_cast<const uint4*>(nb + vo + 96u);
            asm volatile("" : "+v"(ch.x), "+v"(ch.y), "+v"(ch.z), "+v"(ch.w));
            k[4 * h + 0] = is.box_nf(nx.x, ny.x, nz.x, fx.x, fy.x, fz.x, T, tc);
            k[4 * h + 1] = is.box_nf(nx.y, ny.y, nz.y, fx.y, fy.y, fz.y, T, tc);
            k[4 * h + 2] = is.box_nf(nx.z, ny.z, nz.z, fx.z, fy.z, fz.z, T, tc);
            k[4 * h + 3] = is.box_nf(nx.w, ny.w, nz.w, fx.w, fy.w, fz.w, T, tc);
            e[4 * h + 0] = ch.x;
            e[4 * h + 1] = ch.y;
            e[4 * h + 2] = ch.z;
            e[4 * h + 3] = ch.w;
        }
    } else {
        const uint32_t vo = node << 7;
        const float4 org = *reinterpret_cast<const float4*>(nb + vo);
        const float4 sc = *reinterpret_cast<const float4*>(nb + vo + 16u);
        const uint32_t nq = ((T.near >> 1) & 0x7F7F7Fu) + 0x202020u;
        const uint32_t ax = vo | (nq & 0xFFu), ay = vo | ((nq >> 8) & 0xFFu), az = vo | (nq >> 16);
        const uint2 qnx = *reinterpret_cast<const uint2*>(nb + ax);
        const uint2 qny = *reinterpret_cast<const uint2*>(nb + ay);
        const uint2 qnz = *reinterpret_cast<const uint2*>(nb + az);
        const uint2 qfx = *reinterpret_cast<const uint2*>(nb + (ax ^ OQ_FAR_X));
        const uint2 qfy = *reinterpret_cast<const uint2*>(nb + (ay ^ OQ_FAR_Y));
        const uint2 qfz = *reinterpret_cast<const uint2*>(nb + (az ^ OQ_FAR_Z));
        uint4 c0 = *reinterpret_cast<const uint4*>(nb + vo + 80u);
        uint4 c1 = *reinterpret_cast<const uint4*>(nb + vo + 96u);
        asm volatile("" : "+v"(c0.x), "+v"(c0.y), "+v"(c0.z), "+v"(c0.w));
        asm volatile("" : "+v"(c1.x), "+v"(c1.y), "+v"(c1.z), "+v"(c1.w));
        const uint32_t cnt = __float_as_uint(org.w);
        e[0] = c0.x;
        e[1] = c0.y;
        e[2] = c0.z;
        e[3] = c0.w;
        e[4] = c1.x;
        e[5] = c1.y;
        e[6] = c1.z;
        e[7] = c1.w;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int sh = 8 * (i & 3);
            auto plane = [&](uint2 w, float s, float o) {
                const float q = (float)(((i < 4 ? w.x : w.y) >> sh) & 0xFFu);
                return fmaf(q, s, o);  // exact q * s (a power of two), one rounding
            };
            const float t = is.box_nf(plane(qnx, sc.x, org.x), plane(qny, sc.y, org.y), plane(qnz, sc.z, org.z),
                                      plane(qfx, sc.x, org.x), plane(qfy, sc.y, org.y), plane(qfz, sc.z, org.z),
                                      T, tc);
            k[i] = (uint32_t)i < cnt ? t : __builtin_inff();
        }
    }
}

template <bool ANY_HIT, bool COUNT, int LDS_DEPTH, int STRIDE, int ORDER = TRAV_VOTED, typename ISECT = default_isect,
          int NODE = NODE_QUAD>
__device__ __forceinline__ bool trav_iter(trav_state& T, const bvh_view& b, uint32_t* __restrict__ stack,
                                          uint32_t* __restrict__ spill, trace_counters* cnt, bool any_rt = false,
                                          const ISECT& is = ISECT{})
{
    if (b.num_prims == 0) return true;
    auto push = [&](uint32_t e) {
        if (T.sp < LDS_DEPTH) stack[T.sp * STRIDE] = e;
        else spill[T.sp - LDS_DEPTH] = e;
        ++T.sp;
    };
    auto pop = [&]() -> uint32_t {
        if (T.sp == 0) return TRAV_DONE;
        --T.sp;
        // read the LDS column unconditionally (clamped) and the private
        // overflow only when needed, so the two reads never merge into one
        // generic-address load
        uint32_t e = stack[(T.sp < LDS_DEPTH ? T.sp : LDS_DEPTH - 1) * STRIDE];
        if (T.sp >= LDS_DEPTH) {
            e = spill[T.sp - LDS_DEPTH];
            // an empty asm on the value (no instruction): keeps the compiler
            // from sinking both reads into one generic-address load
            asm volatile("" : "+v"(e));
        }
        return e;
    };
    // a leaf at cur goes to the free triangle range
    if (T.ln == 0 && (T.cur & LEAF_BIT) && T.cur != TRAV_DONE) {
        T.lf = T.cur & LEAF_FIRST_MASK;
        T.ln = ((T.cur >> 26) & 31u) + 1u;
        if (T.lf + T.ln > b.num_prims) return true;  // corrupt entry
        T.cur = pop();
    }
    const bool node_work = !(T.cur & LEAF_BIT);
    if (!node_work && T.ln == 0) return true;  // cur is TRAV_DONE: finished
    bool do_node;
    if (ORDER == TRAV_VOTED) {
        const uint32_t n_node = (uint32_t)__popcll(__ballot(node_work));
        const uint32_t n_tri = (uint32_t)__popcll(__ballot(T.ln > 0));
        do_node = n_node >= VR_TRAV_NODE_BIAS * n_tri;
    } else {
        do_node = T.ln == 0;
    }
    if (do_node) {
        if (!node_work) return false;
        if (--T.budget == 0 || T.cur >= b.num_quads) return true;
        if (COUNT) cnt->nodes++;
        const float tc = T.r.t;
        if constexpr (NODE != NODE_QUAD) {
            static_assert(ISECT::near_far, "wide nodes take the library's own box test");
            float k[8];
            uint32_t e[8];
            wide_children<NODE>(T, b, T.cur, tc, is, k, e);
            sort8(k, e);
            if (!(k[0] < __builtin_inff())) {
                T.cur = pop();
            } else {
                // deferred children, farthest deepest: e[j] goes to sp + n - j
                int n = 0;
#pragma unroll
                for (int j = 1; j < 8; ++j) n += k[j] < __builtin_inff();
                if (T.sp + 7 <= LDS_DEPTH) {
#pragma unroll
                    for (int j = 1; j < 8; ++j)
                        if (n >= j) stack[(T.sp + n - j) * STRIDE] = e[j];
                    T.sp += n;
                } else {
#pragma unroll
                    for (int j = 7; j >= 1; --j)
                        if (n >= j) push(e[j]);
                }
                T.cur = e[0];
            }
            return false;
        }
        float k0, k1, k2, k3;
        uint4 ch;
        bool near_far = false;
        if constexpr (ISECT::near_far) near_far = b.num_quads <= QN_NEAR_FAR_MAX;
        if (near_far) {
            // per-lane byte offsets from the node's base: with a wave-uniform
            // BVH base the loads take the base in scalar registers and one
            // 32-bit add each
            const char* qb = reinterpret_cast<const char*>(b.quads);
            const uint32_t vo = T.cur << 7;
            const uint32_t ax = vo | (T.near & 0xFFu), ay = vo | ((T.near >> 8) & 0xFFu), az = vo | (T.near >> 16);
            const float4 nx = *reinterpret_cast<const float4*>(qb + ax);
            const float4 ny = *reinterpret_cast<const float4*>(qb + ay);
            const float4 nz = *reinterpret_cast<const float4*>(qb + az);
            const float4 fx = *reinterpret_cast<const float4*>(qb + (ax ^ QN_FAR_X));
            const float4 fy = *reinterpret_cast<const float4*>(qb + (ay ^ QN_FAR_Y));
            const float4 fz = *reinterpret_cast<const float4*>(qb + (az ^ QN_FAR_Z));
            ch = *reinterpret_cast<const uint4*>(qb + vo + 96u);
            // used here: the child entries are loaded with the planes (one
            // memory round trip per visit), not sunk below the box tests
            asm volatile("" : "+v"(ch.x), "+v"(ch.y), "+v"(ch.z), "+v"(ch.w));
            if constexpr (ISECT::near_far) {
                k0 = is.box_nf(nx.x, ny.x, nz.x, fx.x, fy.x, fz.x, T, tc);
                k1 = is.box_nf(nx.y, ny.y, nz.y, fx.y, fy.y, fz.y, T, tc);
                k2 = is.box_nf(nx.z, ny.z, nz.z, fx.z, fy.z, fz.z, T, tc);
                k3 = is.box_nf(nx.w, ny.w, nz.w, fx.w, fy.w, fz.w, T, tc);
            }
        } else {
            const float4* p = reinterpret_cast<const float4*>(b.quads + T.cur);
            const float4 lx = p[0], ly = p[1], lz = p[2], hx = p[3], hy = p[4], hz = p[5];
            ch = *reinterpret_cast<const uint4*>(p + 6);
            asm volatile("" : "+v"(ch.x), "+v"(ch.y), "+v"(ch.z), "+v"(ch.w));
            k0 = is.box(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, T, tc);
            k1 = is.box(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, T, tc);
            k2 = is.box(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, T, tc);
            k3 = is.box(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, T, tc);
        }
        uint32_t e0 = ch.x, e1 = ch.y, e2 = ch.z, e3 = ch.w;
        cswap(k0, e0, k1, e1);
        cswap(k2, e2, k3, e3);
        cswap(k0, e0, k2, e2);
        cswap(k1, e1, k3, e3);
        cswap(k1, e1, k2, e2);
        if (!(k0 < __builtin_inff())) {
            T.cur = pop();
        } else {
            // deferred children, farthest deepest: e_k goes to sp + n - k
            const int n = (k1 < __builtin_inff()) + (k2 < __builtin_inff()) + (k3 < __builtin_inff());
            if (T.sp + 3 <= LDS_DEPTH) {
                if (n >= 1) stack[(T.sp + n - 1) * STRIDE] = e1;
                if (n >= 2) stack[(T.sp + n - 2) * STRIDE] = e2;
                if (n >= 3) stack[T.sp * STRIDE] = e3;
                T.sp += n;
            } else {
                if (n >= 3) push(e3);
                if (n >= 2) push(e2);
                if (n >= 1) push(e1);
            }
            T.cur = e0;
        }
        return false;
    }
    if (T.ln == 0) return false;
    const uint32_t k = T.lf;
    // the library's own test takes two triangles of the leaf per iteration
    // (in leaf order; a lane with one left tests it twice, and the second
    // test of the same triangle never replaces the hit: not strictly closer)
    constexpr bool PAIRED = isect_pairs<ISECT>::value;
    const uint32_t take = (PAIRED && T.ln >= 2u) ? 2u : 1u;
    T.lf += take;
    T.ln -= take;
    const float4* tp = reinterpret_cast<const float4*>(b.tris + k);
    float4 ta = tp[0];
    const float4 tb = tp[1], tq = tp[2];
    // v1 is used only after the determinant test: pinned here so that the
    // three records arrive in one memory round trip
    asm volatile("" : "+v"(ta.x), "+v"(ta.y), "+v"(ta.z));
    float4 sa, sb, sq;
    if constexpr (PAIRED) {
        const float4* sp = reinterpret_cast<const float4*>(b.tris + k + take - 1u);
        sa = sp[0];
        sb = sp[1];
        sq = sp[2];
        asm volatile("" : "+v"(sa.x), "+v"(sa.y), "+v"(sa.z), "+v"(sb.x), "+v"(sb.y), "+v"(sb.z));
        asm volatile("" : "+v"(sq.x), "+v"(sq.y), "+v"(sq.z));
    }
    if (COUNT) cnt->tris += take;
    float t, u, v;
    if (is.tri(T, ta, tb, tq, t, u, v)) {
        if (t < T.r.t && t >= T.tmin && t <= T.tmax) {
            T.r.t = t;
            T.r.u = u;
            T.r.v = v;
            T.r.tri = (int32_t)k;
            if (ANY_HIT || any_rt) return true;
        }
    }
    if constexpr (PAIRED) {
        if (is.tri(T, sa, sb, sq, t, u, v)) {
            if (t < T.r.t && t >= T.tmin && t <= T.tmax) {
                T.r.t = t;
                T.r.u = u;
                T.r.v = v;
                T.r.tri = (int32_t)(k + take - 1u);
                if (ANY_HIT || any_rt) return true;
            }
        }
    }
    return false;
}

template <bool ANY_HIT, bool COUNT, int LDS_DEPTH, int STRIDE, int ORDER = TRAV_VOTED, typename ISECT = default_isect,
          int NODE = NODE_QUAD>
__device__ __forceinline__ hit_t traverse(const bvh_view& b, f3 ori, f3 dir, float tmin, float tmax,
                                          uint32_t* __restrict__ stack, trace_counters* cnt,
                                          const ISECT& is = ISECT{})
{
    trav_state T;
    uint32_t spill[max_stack_of<NODE>() - LDS_DEPTH];
    trav_begin(T, b, ori, dir, tmin, tmax);
    while (!trav_iter<ANY_HIT, COUNT, LDS_DEPTH, STRIDE, ORDER, ISECT, NODE>(T, b, stack, spill, cnt, false, is)) {
    }
    return T.r;
}

// The reference's traversal of the binary tree as written
// (detail/bvh/intersect.inl:166-356, device order): at a pair both child
// boxes are slab tested with the reference's exact arithmetic, the nearer
// entered child (tnear0 < tnear1 ? 0 : 1) is visited first and the other is
// deferred; triangles are accepted with is_closer.  Used for the
// reference-scheme comparison launch, not by the fast paths.
template <bool ANY_HIT, int LDS_DEPTH, int STRIDE>
__device__ __forceinline__ hit_t traverse_binary(const bvh_view& b, f3 ori, f3 dir, float tmin, float tmax,
                                                 uint32_t* __restrict__ stack)
{
    hit_t r;
    r.t = FLT_MAX;
    r.u = 0.0f;
    r.v = 0.0f;
    r.tri = -1;
    if (b.num_prims == 0) return r;
    const f3 inv = mk3(dir.x != 0.0f ? 1.0f / dir.x : FLT_MAX, dir.y != 0.0f ? 1.0f / dir.y : FLT_MAX,
                       dir.z != 0.0f ? 1.0f / dir.z : FLT_MAX);
    int sp = 0;
    uint32_t cur = b.root;
    uint32_t spill[MAX_STACK - LDS_DEPTH];
    auto push = [&](uint32_t e) {
        if (sp < LDS_DEPTH) stack[sp * STRIDE] = e;
        else spill[sp - LDS_DEPTH] = e;
        ++sp;
    };
    auto pop = [&]() -> uint32_t {
        --sp;
        return sp < LDS_DEPTH ? stack[sp * STRIDE] : spill[sp - LDS_DEPTH];
    };
    // a ray visits each node at most once; a longer walk means a corrupt tree
    uint32_t budget = 2u * b.num_prims + 2u;
    for (;;) {
        while (!(cur & LEAF_BIT)) {
            if (--budget == 0 || cur >= b.num_quads) return r;
            const float4* p = reinterpret_cast<const float4*>(b.pairs + cur);
            const float4 a0 = p[0], q1 = p[1], q2 = p[2];
            const uint4 q3 = *reinterpret_cast<const uint4*>(p + 3);
            float tn0, tf0, tn1, tf1;
            {
                const f3 tlo = mul3(sub3(mk3(a0.x, a0.y, a0.z), ori), inv);
                const f3 thi = mul3(sub3(mk3(a0.w, q1.x, q1.y), ori), inv);
                tn0 = tmax_f(tmin, max_el3(min3v(tlo, thi)));
                tf0 = tmin_f(tmax, min_el3(max3v(tlo, thi)));
            }
            {
                const f3 tlo = mul3(sub3(mk3(q1.z, q1.w, q2.x), ori), inv);
                const f3 thi = mul3(sub3(mk3(q2.y, q2.z, q2.w), ori), inv);
                tn1 = tmax_f(tmin, max_el3(min3v(tlo, thi)));
                tf1 = tmin_f(tmax, min_el3(max3v(tlo, thi)));
            }
            const bool h0 = tf0 >= tn0 && tn0 < r.t;
            const bool h1 = tf1 >= tn1 && tn1 < r.t;
            if (h0 && h1) {
                const bool near1 = !(tn0 < tn1);
                if (sp >= MAX_STACK) return r;  // cannot happen for trees that pass stack_fits
                push(near1 ? q3.x : q3.y);
                cur = near1 ? q3.y : q3.x;
            } else if (h0) {
                cur = q3.x;
            } else if (h1) {
                cur = q3.y;
            } else {
                if (sp == 0) return r;
                cur = pop();
            }
        }
        const uint32_t first = cur & LEAF_FIRST_MASK;
        const uint32_t count = ((cur >> 26) & 31u) + 1u;
        if (first + count > b.num_prims) return r;  // corrupt entry
        for (uint32_t k = first; k < first + count; ++k) {
            const float4* tp = reinterpret_cast<const float4*>(b.tris + k);
            const float4 ta = tp[0], tb = tp[1], tq = tp[2];
            float t, u, v;
            if (intersect_tri(ori, dir, mk3(ta.x, ta.y, ta.z), mk3(tb.x, tb.y, tb.z), mk3(tq.x, tq.y, tq.z), t, u, v)) {
                if (t < r.t && t >= tmin && t <= tmax) {
                    r.t = t;
                    r.u = u;
                    r.v = v;
                    r.tri = (int32_t)k;
                    if (ANY_HIT) return r;
                }
            }
        }
        if (sp == 0) return r;
        cur = pop();
    }
}

// LDS part of the traversal stack (entries per lane)
constexpr int LDS_STACK = 16;

// ---------------------------------------------------------------------------
// seeds, random numbers, camera rays, sampling (the reference's host build)
// ---------------------------------------------------------------------------

// make_random_seed (make_random_seed.h:23-37)
__device__ inline uint32_t make_random_seed(int32_t val0, int32_t val1)
{
    uint32_t v0 = (uint32_t)val0, v1 = (uint32_t)val1, s0 = 0;
#pragma unroll
    for (uint32_t n = 0; n < 4; ++n) {
        s0 += 0x9E3779B9u;
        v0 += ((v1 << 4) + 0xA341316Cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xC8013EA4u);
        v1 += ((v0 << 4) + 0xAd90777Du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7E95761Eu);
    }
    return v0;
}

// random_generator<float> as built for the host (random_generator.h:29-62):
// minstd_rand0 (x <- 16807 x mod 2^31 - 1; seed mod 2^31 - 1, zero -> 1) with
// uniform_real_distribution<float>(0, 1) through generate_canonical<float, 24>,
// one engine draw per sample: float(x - 1) / 2^31, clamped below 1.
struct rng_t {
    uint32_t x;
};
__device__ inline void rng_seed(rng_t& g, uint32_t s)
{
    const uint32_t x = s % 2147483647u;
    g.x = x == 0 ? 1u : x;
}
__device__ inline float rng_next(rng_t& g)
{
    const uint64_t t = (uint64_t)g.x * 16807u;
    uint32_t r = (uint32_t)(t & 0x7FFFFFFFu) + (uint32_t)(t >> 31);  // 2^31 == 1 mod 2^31 - 1
    if (r >= 0x7FFFFFFFu) r -= 0x7FFFFFFFu;
    g.x = r;
    const float f = (float)(r - 1u) * 4.656612873077392578125e-10f;  // exact 2^-31 scaling
    return f >= 1.0f ? 0x1.fffffep-1f : f;
}
// The reference host build's compiler (GCC) evaluates the two draws of
// f(gen.next(), gen.next()) right to left: arg_order 1.
constexpr uint32_t REFERENCE_ARG_ORDER = 1;

// two draws for one argument list, in the order the host compiler evaluated
// the reference's `f(gen.next(), gen.next())`: arg_order 1 = right first
__device__ inline void rng_pair(rng_t& g, uint32_t arg_order, float& first_arg, float& second_arg)
{
    const float a = rng_next(g);
    const float b = rng_next(g);
    first_arg = arg_order ? b : a;
    second_arg = arg_order ? a : b;
}

// make_primary_ray with jitter (detail/sched_common.h:209-233) and
// pinhole_camera::primary_ray (detail/pinhole_camera.inl:103-121), image region
// (0,0)-(1,1)
__device__ inline void primary_ray(const vr_camera& cam, rng_t& g, uint32_t arg_order, int x, int y, int w, int h,
                                   f3& o, f3& d)
{
    float a, b;
    rng_pair(g, arg_order, a, b);
    const float jx = a - 0.5f, jy = b - 0.5f;
    const float fx = (float)x + jx, fy = (float)y + jy;
    float sx = (fx + 0.5f) / (float)w;
    float sy = (fy + 0.5f) / (float)h;
    sx = (1.0f - sx) * 0.0f + sx * 1.0f;
    sy = (1.0f - sy) * 0.0f + sy * 1.0f;
    const float u = 2.0f * sx - 1.0f;
    const float v = 2.0f * sy - 1.0f;
    const f3 U = mk3(cam.U[0], cam.U[1], cam.U[2]);
    const f3 V = mk3(cam.V[0], cam.V[1], cam.V[2]);
    const f3 W = mk3(cam.W[0], cam.W[1], cam.W[2]);
    d = normalize3(add3(add3(muls3(U, u), muls3(V, v)), W));
    o = mk3(cam.eye[0], cam.eye[1], cam.eye[2]);
}

// cosine_sample_hemisphere (sampling.h:117-127)
__device__ inline f3 cosine_hemisphere(float u1, float u2)
{
    const float r = sqrtf(u1);
    const float theta = 6.28318530717958623199592693709e+00f * u2;
    float sn, cs;
    sincosf(theta, &sn, &cs);
    return mk3(r * cs, r * sn, sqrtf(1.0f - u1));
}

// make_orthonormal_basis (math/detail/vector3.inl:356-366; brdf.h:62-66)
__device__ inline void onb(f3 w, f3& u, f3& v)
{
    v = fabsf(w.x) > fabsf(w.y) ? normalize3(mk3(-w.z, 0.0f, w.x)) : normalize3(mk3(0.0f, w.z, -w.y));
    u = cross3(v, w);
}

// Pixel order of the work queue: tile x tile squares row by row (tile <= 1:
// scanlines).  The queue runs over whole tiles; indices in the padding of a
// partial tile map to no pixel.
__host__ __device__ inline uint32_t padded_pixels(uint32_t w, uint32_t h, uint32_t tile)
{
    if (tile <= 1) return w * h;
    return ((w + tile - 1) / tile) * ((h + tile - 1) / tile) * tile * tile;
}
__device__ inline bool tile_pixel(uint32_t idx, uint32_t w, uint32_t h, uint32_t tile, int& x, int& y)
{
    if (tile <= 1) {
        x = (int)(idx % w);
        y = (int)(idx / w);
        return true;
    }
    const uint32_t tiles_x = (w + tile - 1) / tile, tt = tile * tile;
    const uint32_t t = idx / tt, o = idx % tt;
    x = (int)((t % tiles_x) * tile + o % tile);
    y = (int)((t / tiles_x) * tile + o / tile);
    return (uint32_t)x < w && (uint32_t)y < h;
}

}  // namespace vr
