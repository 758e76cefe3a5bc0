// instance.hip -- instanced two-level BVHs (index_bvh_inst_t, bvh.h:679-782;
// the instance intersect overload, detail/bvh/intersect.inl:362-394).
//
// The top level is small (one entry per instance) and is built on the host:
// each instance's box is its BVH's root box, corner by corner through the
// instance transform (get_bounds, detail/bvh/get_bounds.inl:30-56), and the
// tree over those boxes is the reference's CPU lbvh_builder tree
// (host_build.cpp), stored as pair records whose leaves are ranges of
// instance slots, then collapsed 4-wide like the bottom levels.  On the
// device one lane follows one ray through both levels in one voted loop: at
// an instance of a top-level leaf it tests the instance's box and moves the
// ray into the instance's space (ori' = A (ori - T), dir' = A dir with A the
// stored inverse) for that instance's BVH.  t is the same parameter in both
// spaces (the direction is not renormalized), so the closest hit over all
// instances is the smallest t, and it culls boxes in both levels.
#include "vr_internal.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <vector>

namespace vr {

namespace {

constexpr int BLOCK = 256;

// one instance on the device
struct alignas(16) inst_rec {
    float inv[9];  // the inverse 3x3 map, row by row
    float ti[3];   // the negated translation
    uint32_t blas;
    int32_t inst_id;
    float box[6];  // world box (lo, hi): a leaf's instances are tested one by one before entering
};
static_assert(sizeof(inst_rec) == 80, "instance record");

struct tlas_args {
    const quad_node* quads;  // the top level, 4-wide
    uint32_t num_quads;
    uint32_t root;
    const uint32_t* slots;  // instance per leaf slot
    uint32_t num_slots;
    const inst_rec* insts;
    const bvh_view* views;
    const vr_ray* rays;
    uint32_t n;
    vr_inst_hit* hits;
};

// m(r, c) of a 3x3 matrix, row by row
struct mat3h {
    float m[3][3];
};

// inverse(matrix<3, 3>) (math/detail/matrix3.inl:300-321): cofactors by
// det2(m00, m01, m10, m11) = m00 m11 - m10 m01 (math/detail/math.h:616-620),
// the determinant along row 0, every entry cofactor / det; -ffp-contract=off
// keeps each product and sum rounded as written
mat3h inverse3(const mat3h& a)
{
    auto det2 = [](float m00, float m01, float m10, float m11) { return m00 * m11 - m10 * m01; };
    const auto& m = a.m;
    const float a00 = det2(m[1][1], m[1][2], m[2][1], m[2][2]);
    const float a01 = det2(m[1][0], m[1][2], m[2][0], m[2][2]);
    const float a02 = det2(m[1][0], m[1][1], m[2][0], m[2][1]);
    const float a10 = det2(m[0][1], m[0][2], m[2][1], m[2][2]);
    const float a11 = det2(m[0][0], m[0][2], m[2][0], m[2][2]);
    const float a12 = det2(m[0][0], m[0][1], m[2][0], m[2][1]);
    const float a20 = det2(m[0][1], m[0][2], m[1][1], m[1][2]);
    const float a21 = det2(m[0][0], m[0][2], m[1][0], m[1][2]);
    const float a22 = det2(m[0][0], m[0][1], m[1][0], m[1][1]);
    const float det = m[0][0] * a00 - m[0][1] * a01 + m[0][2] * a02;
    // the constructor takes the entries column by column
    mat3h r;
    r.m[0][0] = a00 / det;
    r.m[1][0] = -a01 / det;
    r.m[2][0] = a02 / det;
    r.m[0][1] = -a10 / det;
    r.m[1][1] = a11 / det;
    r.m[2][1] = -a12 / det;
    r.m[0][2] = a20 / det;
    r.m[1][2] = -a21 / det;
    r.m[2][2] = a22 / det;
    return r;
}

// matrix<3, 3> * vector<3> (matrix3.inl:226-235): each row summed left to right
inline void mat_vec(const mat3h& a, const float v[3], float out[3])
{
    for (int r = 0; r < 3; ++r) out[r] = a.m[r][0] * v[0] + a.m[r][1] * v[1] + a.m[r][2] * v[2];
}

__device__ inline f3 row_mul(const float* inv, f3 v)
{
    return mk3(inv[0] * v.x + inv[1] * v.y + inv[2] * v.z, inv[3] * v.x + inv[4] * v.y + inv[5] * v.z,
               inv[6] * v.x + inv[7] * v.y + inv[8] * v.z);
}

// deferred entries per lane: both levels' stacks and the level markers
constexpr int TWO_LEVEL_STACK = 2 * MAX_STACK + 8;
// on the stack: back to the top level (restore the world-space ray)
constexpr uint32_t TO_TOP = 0xFFFFFFFEu;
// a lane's next entry when the marker is next on its stack: nothing more at
// this level (the marker is taken once the lane's triangle range is done)
constexpr uint32_t LEVEL_END = 0xFFFFFFFDu;

// One voted loop over both levels.  Each iteration the wave either visits
// nodes (a 4-wide top-level node or a 4-wide bottom-level node, whichever
// level each lane is in) or works on leaves (a triangle test in an instance,
// or a top-level instance: its box test and, on a hit, the entry into its
// space), whichever more lanes need -- the same vote as the one-level
// traversal, so lanes in different levels still advance together.  Entering
// an instance pushes the rest of its top-level leaf and a TO_TOP marker; the
// marker's pop restores the world-space ray.  Boxes and triangles are culled
// by the closest hit over all instances so far.
constexpr uint32_t RAY_CHUNK = 64;  // rays a wave reserves at once

// Persistent lanes: a resident grid; a lane whose query ends takes the next
// ray of its wave's reserved chunk (one atomic per 64 rays), so a wave keeps
// all lanes busy until the rays run out.
template <bool ANY_HIT>
__global__ __launch_bounds__(BLOCK) void k_tlas_trace(tlas_args A, uint32_t* __restrict__ work)
{
    __shared__ uint32_t stack[LDS_STACK * BLOCK];
    uint32_t* st = stack + threadIdx.x;
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    // the query
    f3 ow = mk3(0.0f, 0.0f, 0.0f), dw = mk3(0.0f, 0.0f, 1.0f), iw = dw;
    float tmin = 0.0f, tmax = 0.0f;
    // the current space's ray and level
    f3 o = ow, d = dw, iv = iw;
    bool bottom = false;
    const quad_node* q = A.quads;
    const tri_node* tris = nullptr;
    uint32_t nq = A.num_quads, np = A.num_slots;
    uint32_t cur = TRAV_DONE, lf = 0, ln = 0;
    int sp = 0;
    uint32_t spill[TWO_LEVEL_STACK - LDS_STACK];
    float best_t = FLT_MAX, best_u = 0.0f, best_v = 0.0f;
    int32_t best_tri = -1, best_inst = -1, cur_inst = -1;
    uint32_t budget = 0;
    uint32_t idx = 0;
    bool tracing = false, done = false;
    uint32_t chunk_base = 0, chunk_left = 0;  // the wave's reserved ray indices (uniform)
    auto push = [&](uint32_t e) {
        if (sp < LDS_STACK) st[sp * BLOCK] = e;
        else spill[sp - LDS_STACK] = e;
        ++sp;
    };
    auto top = [&]() -> uint32_t { return sp - 1 < LDS_STACK ? st[(sp - 1) * BLOCK] : spill[sp - 1 - LDS_STACK]; };
    // the next deferred entry of this level (LEVEL_END at the marker)
    auto pop = [&]() -> uint32_t {
        if (sp == 0) return TRAV_DONE;
        const uint32_t e = top();
        if (e == TO_TOP) return LEVEL_END;
        --sp;
        return e;
    };
    auto to_top = [&]() {
        o = ow;
        d = dw;
        iv = iw;
        bottom = false;
        q = A.quads;
        nq = A.num_quads;
        np = A.num_slots;
    };
    auto abandon = [&]() {  // a corrupt tree: end the query
        cur = TRAV_DONE;
        ln = 0;
        sp = 0;
        to_top();
    };
    for (;;) {
        // lanes without a query take the next rays of the wave's chunk, once
        // VR_TLAS_REFILL_LANES of them wait (or no lane has a query left)
        const bool need = !done && !tracing;
        const uint64_t needm = __ballot(need);
        if (needm && (__popcll(needm) >= VR_TLAS_REFILL_LANES || __ballot(tracing) == 0)) {
            const uint32_t k = (uint32_t)__popcll(needm);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(needm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)needm, 0u));
            uint32_t fresh = 0;
            if (chunk_left < k) {
                const int leader = __ffsll((unsigned long long)needm) - 1;
                if ((int)lane == leader) fresh = atomicAdd(work, RAY_CHUNK);
                fresh = __shfl(fresh, leader);
            }
            if (need) {
                const uint32_t my = rank < chunk_left ? chunk_base + rank : fresh + (rank - chunk_left);
                if (my < A.n) {
                    const float4* rp = reinterpret_cast<const float4*>(A.rays + my);
                    const float4 r0 = rp[0], r1 = rp[1];
                    ow = mk3(r0.x, r0.y, r0.z);
                    dw = mk3(r0.w, r1.x, r1.y);
                    tmin = r1.z;
                    tmax = r1.w;
                    iw = mk3(dw.x != 0.0f ? 1.0f / dw.x : FLT_MAX, dw.y != 0.0f ? 1.0f / dw.y : FLT_MAX,
                             dw.z != 0.0f ? 1.0f / dw.z : FLT_MAX);
                    to_top();
                    cur = A.num_slots > 0 ? A.root : TRAV_DONE;
                    lf = ln = 0;
                    sp = 0;
                    best_t = FLT_MAX;
                    best_u = best_v = 0.0f;
                    best_tri = best_inst = cur_inst = -1;
                    budget = 1u << 26;  // a guard against corrupt trees only
                    idx = my;
                    tracing = true;
                } else {
                    done = true;
                }
            }
            if (chunk_left < k) {
                chunk_base = fresh + (k - chunk_left);
                chunk_left = RAY_CHUNK - (k - chunk_left);
            } else {
                chunk_base += k;
                chunk_left -= k;
            }
        }
        if (__ballot(!done) == 0) break;
        if (tracing) {
            // the instance is done (its range too): back to the top level
            if (ln == 0 && cur == LEVEL_END) {
                --sp;  // the marker
                to_top();
                cur = pop();
            }
            // a leaf entry becomes the lane's range: triangles in an
            // instance, instance slots at the top
            if (ln == 0 && (cur & LEAF_BIT) && cur != TRAV_DONE && cur != LEVEL_END) {
                lf = cur & LEAF_FIRST_MASK;
                ln = ((cur >> 26) & 31u) + 1u;
                if (lf + ln > np) abandon();
                else cur = pop();
            }
        }
        const bool node_work = tracing && !(cur & LEAF_BIT);
        const bool leaf_work = tracing && ln > 0;
        if (tracing && !node_work && !leaf_work) {
            // the query is done: its hit record
            vr_inst_hit out;
            out.pad = 0;
            if (best_tri >= 0) {
                const inst_rec* rec = A.insts + best_inst;
                const tri_node* tn = A.views[rec->blas].tris + best_tri;
                out.prim_id = (int32_t)__float_as_uint(tn->a.w);
                out.geom_id = (int32_t)__float_as_uint(tn->b.w);
                out.t = best_t;
                out.u = best_u;
                out.v = best_v;
                out.inst_id = rec->inst_id;
                out.instance = best_inst;
            } else {
                out.prim_id = -1;
                out.geom_id = -1;
                out.t = FLT_MAX;
                out.u = 0.0f;
                out.v = 0.0f;
                out.inst_id = -1;
                out.instance = -1;
            }
            A.hits[idx] = out;
            tracing = false;
            continue;
        }
        if ((node_work || leaf_work) && --budget == 0) {
            abandon();
            continue;
        }
        const uint32_t n_node = (uint32_t)__popcll(__ballot(node_work));
        const uint32_t n_leaf = (uint32_t)__popcll(__ballot(leaf_work));
        if ((float)n_node >= VR_TRAV_NODE_BIAS * (float)n_leaf) {
            if (!node_work) continue;
            if (cur >= nq) {
                abandon();
                continue;
            }
            const float4* p = reinterpret_cast<const float4*>(q + cur);
            const float4 lx = p[0], ly = p[1], lz = p[2], hx = p[3], hy = p[4], hz = p[5];
            const uint4 ch = *reinterpret_cast<const uint4*>(p + 6);
            float k0 = slab_enter(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, o, iv, tmin, tmax, best_t);
            float k1 = slab_enter(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, o, iv, tmin, tmax, best_t);
            float k2 = slab_enter(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, o, iv, tmin, tmax, best_t);
            float k3 = slab_enter(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, o, iv, tmin, tmax, best_t);
            uint32_t e0 = ch.x, e1 = ch.y, e2 = ch.z, e3 = ch.w;
            cswap(k0, e0, k1, e1);
            cswap(k2, e2, k3, e3);
            cswap(k0, e0, k2, e2);
            cswap(k1, e1, k3, e3);
            cswap(k1, e1, k2, e2);
            if (!(k0 < __builtin_inff())) {
                cur = pop();
            } else {
                if (sp + 3 > TWO_LEVEL_STACK) {  // cannot happen for trees the host accepted
                    abandon();
                    continue;
                }
                if (k3 < __builtin_inff()) push(e3);
                if (k2 < __builtin_inff()) push(e2);
                if (k1 < __builtin_inff()) push(e1);
                cur = e0;
            }
            continue;
        }
        if (!leaf_work) continue;
        const uint32_t k = lf++;
        --ln;
        if (bottom) {
            // the leaf's next triangle in the same iteration (in leaf order;
            // a lane with one left tests it twice, and the second test of the
            // same triangle never replaces the hit: not strictly closer)
            const uint32_t k2 = ln > 0 ? k + 1u : k;
            if (ln > 0) {
                ++lf;
                --ln;
            }
            const float4* tp = reinterpret_cast<const float4*>(tris + k);
            const float4* sp2 = reinterpret_cast<const float4*>(tris + k2);
            const float4 ta = tp[0], tb = tp[1], tq = tp[2];
            const float4 sa = sp2[0], sb = sp2[1], sq = sp2[2];
            float t, u, v;
            bool ended = false;
            if (intersect_tri(o, d, mk3(ta.x, ta.y, ta.z), mk3(tb.x, tb.y, tb.z), mk3(tq.x, tq.y, tq.z), t, u, v) &&
                t < best_t && t >= tmin && t <= tmax) {
                best_t = t;
                best_u = u;
                best_v = v;
                best_tri = (int32_t)k;
                best_inst = cur_inst;
                ended = ANY_HIT;
            }
            if (!ended &&
                intersect_tri(o, d, mk3(sa.x, sa.y, sa.z), mk3(sb.x, sb.y, sb.z), mk3(sq.x, sq.y, sq.z), t, u, v) &&
                t < best_t && t >= tmin && t <= tmax) {
                best_t = t;
                best_u = u;
                best_v = v;
                best_tri = (int32_t)k2;
                best_inst = cur_inst;
                ended = ANY_HIT;
            }
            if (ended) {  // any hit ends the query
                cur = TRAV_DONE;
                ln = 0;
                sp = 0;
                to_top();
            }
            continue;
        }
        // an instance of the top-level leaf: its own box, then its space
        const uint32_t ii = A.slots[k];
        const inst_rec* rec = A.insts + ii;
        if (!(slab_enter(rec->box[0], rec->box[1], rec->box[2], rec->box[3], rec->box[4], rec->box[5], ow, iw, tmin,
                         tmax, best_t) < __builtin_inff()))
            continue;
        if (sp + 3 > TWO_LEVEL_STACK) {
            abandon();
            continue;
        }
        // the rest of this leaf waits under the marker; cur (deferred at the
        // top level) goes below both
        if (cur != TRAV_DONE) push(cur);
        if (ln > 0) push(leaf_entry(lf, ln));
        push(TO_TOP);
        ln = 0;
        float m[9], ti[3];
        for (int c = 0; c < 9; ++c) m[c] = rec->inv[c];
        for (int c = 0; c < 3; ++c) ti[c] = rec->ti[c];
        // index_bvh_inst_t::transform_ray (bvh.h:769-776)
        o = row_mul(m, mk3(ow.x + ti[0], ow.y + ti[1], ow.z + ti[2]));
        d = row_mul(m, dw);
        iv = mk3(d.x != 0.0f ? 1.0f / d.x : FLT_MAX, d.y != 0.0f ? 1.0f / d.y : FLT_MAX,
                 d.z != 0.0f ? 1.0f / d.z : FLT_MAX);
        const bvh_view& bv = A.views[rec->blas];
        bottom = true;
        q = bv.quads;
        tris = bv.tris;
        nq = bv.num_quads;
        np = bv.num_prims;
        cur = bv.root;
        cur_inst = (int32_t)ii;
    }
}

__global__ void k_reset_work(uint32_t* work) { *work = 0u; }

}  // namespace

}  // namespace vr

using namespace vr;

struct vr_tlas_t {
    int device = 0;
    uint32_t num_instances = 0;
    uint32_t num_pairs = 0;
    uint32_t root = 0;
    quad_node* quads = nullptr;
    uint32_t* slots = nullptr;
    inst_rec* insts = nullptr;
    bvh_view* views = nullptr;
    std::vector<vr_bvh_node> nodes;  // the reference-layout top level
    std::vector<uint32_t> indices;
};

int vr_tlas_destroy(vr_tlas t)
{
    if (!t) return VR_OK;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(t->device);
    void* ptrs[] = {t->quads, t->slots, t->insts, t->views};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    (void)hipSetDevice(cur);
    delete t;
    return VR_OK;
}

int vr_tlas_build(const vr_bvh* blas, uint32_t num_blas, const vr_instance* h_instances, uint32_t num_instances,
                  void* stream, vr_tlas* out)
{
    VR_TRY_BEGIN
    VR_CHECK(out, VR_ERR_INVALID_ARGUMENT, "vr_tlas_build: null out handle");
    VR_CHECK(num_instances == 0 || (h_instances && blas), VR_ERR_INVALID_ARGUMENT, "vr_tlas_build: null array");
    VR_CHECK(num_instances <= (uint32_t)MAX_LEAF_FIRST, VR_ERR_UNSUPPORTED, "vr_tlas_build: too many instances");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    // the bottom levels: their traversal views and root boxes
    std::vector<bvh_view> views(num_blas);
    std::vector<float> root_boxes(6 * (size_t)num_blas);
    for (uint32_t b = 0; b < num_blas; ++b) {
        VR_CHECK(blas[b], VR_ERR_INVALID_ARGUMENT, "vr_tlas_build: null BVH");
        const bvh_state& s = blas[b]->s;
        VR_CHECK(s.num_prims > 0, VR_ERR_INVALID_ARGUMENT, "vr_tlas_build: an empty BVH has no box to instance");
        // (a tree whose root is a leaf has no traversal nodes; its root entry
        // is the leaf)
        VR_CHECK(s.quads || (s.root & LEAF_BIT), VR_ERR_INVALID_ARGUMENT, "vr_tlas_build: BVH without traversal nodes");
        views[b] = s.view();
        // the BVH's root box (node(0).get_bounds()): kept by LBVH builds, the
        // first reference node of adopted trees
        float* rb = &root_boxes[6 * (size_t)b];
        if (s.root_box) {
            VR_HIP_CHECK(hipMemcpyAsync(rb, s.root_box, 6 * sizeof(float), hipMemcpyDeviceToHost, st));
        } else if (s.ref_nodes) {
            VR_HIP_CHECK(hipMemcpyAsync(rb, s.ref_nodes, 6 * sizeof(float), hipMemcpyDeviceToHost, st));
        } else {
            return fail(VR_ERR_INVALID_ARGUMENT, "vr_tlas_build: BVH without a root box");
        }
    }
    VR_HIP_CHECK(hipStreamSynchronize(st));
    // the instances: inverse maps and world boxes
    std::vector<inst_rec> recs(num_instances);
    std::vector<float> boxes(6 * (size_t)num_instances);
    for (uint32_t i = 0; i < num_instances; ++i) {
        const vr_instance& in = h_instances[i];
        VR_CHECK(in.blas < num_blas, VR_ERR_INVALID_ARGUMENT, "vr_tlas_build: instance refers to no BVH");
        // top_left(mat4x3)(r, c) = M(c, r) = column r's entry c (matrix4x3.inl:206-218)
        mat3h lin;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) lin.m[r][c] = in.transform[3 * r + c];
        const mat3h ainv = inverse3(lin);
        inst_rec& rec = recs[i];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) rec.inv[3 * r + c] = ainv.m[r][c];
        for (int k = 0; k < 3; ++k) rec.ti[k] = -in.transform[9 + k];
        rec.blas = in.blas;
        rec.inst_id = in.inst_id;
        // get_bounds(inst): affine = inverse(affine_inv), trans = -trans_inv,
        // the BVH box's corners in compute_vertices order (aabb.inl:305-324)
        const mat3h aff = inverse3(ainv);
        const float* bb = &root_boxes[6 * (size_t)in.blas];
        const float lo[3] = {bb[0], bb[1], bb[2]}, hi[3] = {bb[3], bb[4], bb[5]};
        const float corners[8][3] = {{hi[0], hi[1], hi[2]}, {lo[0], hi[1], hi[2]}, {lo[0], lo[1], hi[2]},
                                     {hi[0], lo[1], hi[2]}, {lo[0], hi[1], lo[2]}, {hi[0], hi[1], lo[2]},
                                     {hi[0], lo[1], lo[2]}, {lo[0], lo[1], lo[2]}};
        float blo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (const auto& v : corners) {
            float w[3];
            mat_vec(aff, v, w);
            for (int k = 0; k < 3; ++k) {
                const float x = w[k] + -rec.ti[k];
                blo[k] = x < blo[k] ? x : blo[k];  // aabb::insert: min(a, b) = a < b ? a : b
                bhi[k] = bhi[k] < x ? x : bhi[k];
            }
        }
        for (int k = 0; k < 3; ++k) {
            boxes[6 * (size_t)i + k] = rec.box[k] = blo[k];
            boxes[6 * (size_t)i + 3 + k] = rec.box[3 + k] = bhi[k];
        }
    }
    vr_tlas t = new vr_tlas_t;
    t->device = current_device();
    t->num_instances = num_instances;
    auto bail = [&](int code) {
        vr_tlas_destroy(t);
        return code;
    };
    // the top level: the CPU lbvh_builder tree over the boxes (leaves of up to 4)
    std::vector<pair_node> pairs;
    std::vector<uint32_t> slots;
    if (num_instances > 0) {
        t->nodes.resize(2 * (size_t)num_instances - 1);
        t->indices.resize(num_instances);
        size_t nn = 0;
        rc = lbvh_build_host_boxes(boxes.data(), num_instances, 4, t->nodes.data(), &nn, t->indices.data());
        if (rc) return bail(rc);
        t->nodes.resize(nn);
        // pair records in depth-first order; leaves become slot ranges
        struct item {
            uint32_t node, depth;
            int64_t parent;
            uint32_t side;
        };
        std::vector<item> work{{0u, 1u, -1, 0u}};
        uint32_t depth = 0;
        while (!work.empty()) {
            const item it = work.back();
            work.pop_back();
            if (it.node >= nn) return bail(fail(VR_ERR_HIP, "vr_tlas_build: child index out of range"));
            const vr_bvh_node& nd = t->nodes[it.node];
            depth = std::max(depth, it.depth);
            uint32_t entry;
            if (nd.num_prims) {
                entry = leaf_entry((uint32_t)slots.size(), nd.num_prims);
                for (uint32_t k = nd.first; k < nd.first + nd.num_prims; ++k) slots.push_back(t->indices[k]);
            } else {
                entry = (uint32_t)pairs.size();
                pair_node pn{};
                const vr_bvh_node& l = t->nodes[nd.first];
                const vr_bvh_node& r = t->nodes[nd.first + 1];
                for (int k = 0; k < 3; ++k) {
                    pn.f[k] = l.lo[k];
                    pn.f[3 + k] = l.hi[k];
                    pn.f[6 + k] = r.lo[k];
                    pn.f[9 + k] = r.hi[k];
                }
                pairs.push_back(pn);
                work.push_back({nd.first + 1, it.depth + 1, (int64_t)entry, 1u});
                work.push_back({nd.first, it.depth + 1, (int64_t)entry, 0u});
            }
            if (it.parent < 0) t->root = entry;
            else pairs[(size_t)it.parent].e[it.side] = entry;
        }
        if (3 * depth + 1 > (uint32_t)MAX_STACK)
            return bail(fail(VR_ERR_UNSUPPORTED, "vr_tlas_build: top level deeper than its traversal stack"));
    }
    t->num_pairs = (uint32_t)pairs.size();
    // the 4-wide top level: quad p is the collapse rooted at pair p (open the
    // largest-area inner child until four children), as for the bottom levels
    std::vector<quad_node> quads(pairs.size());
    for (size_t p = 0; p < pairs.size(); ++p) {
        struct child {
            uint32_t e;
            float lo[3], hi[3];
        };
        auto from = [](const pair_node& r, int side) {
            child c;
            c.e = r.e[side];
            for (int k = 0; k < 3; ++k) {
                c.lo[k] = r.f[6 * side + k];
                c.hi[k] = r.f[6 * side + 3 + k];
            }
            return c;
        };
        child c[4] = {from(pairs[p], 0), from(pairs[p], 1), {}, {}};
        int nc = 2;
        while (nc < 4) {
            int best = -1;
            float best_a = -1.0f;
            for (int k = 0; k < nc; ++k) {
                if ((c[k].e & LEAF_BIT) || c[k].e >= pairs.size()) continue;
                const float dx = c[k].hi[0] - c[k].lo[0], dy = c[k].hi[1] - c[k].lo[1], dz = c[k].hi[2] - c[k].lo[2];
                const float a = dx * dy + dy * dz + dz * dx;
                if (a > best_a) {
                    best_a = a;
                    best = k;
                }
            }
            if (best < 0) break;
            const pair_node& r = pairs[c[best].e];
            c[nc] = from(r, 1);
            c[best] = from(r, 0);
            ++nc;
        }
        const float inf = std::numeric_limits<float>::infinity();
        float v[6][4];
        uint32_t ce[4];
        for (int k = 0; k < 4; ++k) {
            for (int a = 0; a < 3; ++a) {
                v[a][k] = k < nc ? c[k].lo[a] : inf;
                v[3 + a][k] = k < nc ? c[k].hi[a] : inf;
            }
            ce[k] = k < nc ? c[k].e : LEAF_BIT;
        }
        quad_node& qn = quads[p];
        qn.lo_x = make_float4(v[0][0], v[0][1], v[0][2], v[0][3]);
        qn.lo_y = make_float4(v[1][0], v[1][1], v[1][2], v[1][3]);
        qn.lo_z = make_float4(v[2][0], v[2][1], v[2][2], v[2][3]);
        qn.hi_x = make_float4(v[3][0], v[3][1], v[3][2], v[3][3]);
        qn.hi_y = make_float4(v[4][0], v[4][1], v[4][2], v[4][3]);
        qn.hi_z = make_float4(v[5][0], v[5][1], v[5][2], v[5][3]);
        qn.child = make_uint4(ce[0], ce[1], ce[2], ce[3]);
        qn.meta = make_uint4((uint32_t)nc, 0u, 0u, 0u);
    }
    auto upload = [&](auto** dst, const auto* src, size_t count) -> bool {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(**dst);
        if (hipMalloc(reinterpret_cast<void**>(dst), bytes) != hipSuccess) return false;
        return count == 0 || hipMemcpy(*dst, src, count * sizeof(**dst), hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!upload(&t->quads, quads.data(), quads.size()) || !upload(&t->slots, slots.data(), slots.size()) ||
        !upload(&t->insts, recs.data(), recs.size()) || !upload(&t->views, views.data(), views.size()))
        return bail(fail(VR_ERR_OUT_OF_MEMORY, "vr_tlas_build: device allocation or upload failed"));
    *out = t;
    return VR_OK;
    VR_TRY_END
}

namespace {

int tlas_trace(vr_tlas t, bool any_hit, const vr_ray* d_rays, size_t n, vr_inst_hit* d_hits, void* stream)
{
    VR_CHECK(t, VR_ERR_INVALID_ARGUMENT, "vr_tlas trace: null handle");
    VR_CHECK(n == 0 || (d_rays && d_hits), VR_ERR_INVALID_ARGUMENT, "vr_tlas trace: null array");
    VR_CHECK(n <= 0xFFFFFFFFull, VR_ERR_UNSUPPORTED, "vr_tlas trace: too many rays");
    if (n == 0) return VR_OK;
    int rc = ensure_device();
    if (rc) return rc;
    tlas_args A{t->quads, t->num_pairs, t->root, t->slots, t->num_instances, t->insts, t->views,
                d_rays, (uint32_t)n, d_hits};
    // a resident grid (the blocks the device holds at once), no more waves
    // than there are rays
    static int grid[2] = {0, 0};
    int& g = grid[any_hit ? 1 : 0];
    if (g == 0) {
        int dev = 0, per_cu = 0, cus = 0;
        VR_HIP_CHECK(hipGetDevice(&dev));
        if (any_hit) VR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tlas_trace<true>, BLOCK, 0));
        else VR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tlas_trace<false>, BLOCK, 0));
        VR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        g = std::max(1, per_cu) * std::max(1, cus);
    }
    const unsigned blocks = std::max(1u, std::min((unsigned)g, div_up(n, BLOCK)));
    uint32_t* work = next_work_counter();
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(1), 0, st, work);
    if (any_hit) hipLaunchKernelGGL(k_tlas_trace<true>, dim3(blocks), dim3(BLOCK), 0, st, A, work);
    else hipLaunchKernelGGL(k_tlas_trace<false>, dim3(blocks), dim3(BLOCK), 0, st, A, work);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

}  // namespace

int vr_tlas_closest_hit(vr_tlas t, const vr_ray* d_rays, size_t n, vr_inst_hit* d_hits, void* stream)
{
    return tlas_trace(t, false, d_rays, n, d_hits, stream);
}

int vr_tlas_any_hit(vr_tlas t, const vr_ray* d_rays, size_t n, vr_inst_hit* d_hits, void* stream)
{
    return tlas_trace(t, true, d_rays, n, d_hits, stream);
}

int vr_tlas_nodes(vr_tlas t, vr_bvh_node* h_nodes, uint32_t* h_indices, size_t* num_nodes)
{
    VR_CHECK(t && num_nodes, VR_ERR_INVALID_ARGUMENT, "vr_tlas_nodes: null argument");
    *num_nodes = t->nodes.size();
    if (h_nodes && !t->nodes.empty()) std::memcpy(h_nodes, t->nodes.data(), t->nodes.size() * sizeof(vr_bvh_node));
    if (h_indices && !t->indices.empty()) std::memcpy(h_indices, t->indices.data(), t->indices.size() * 4);
    return VR_OK;
}
