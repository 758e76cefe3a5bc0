// instance.hip -- instanced two-level BVHs (index_bvh_inst_t, bvh.h:679-782;
// the instance intersect overload, detail/bvh/intersect.inl:362-394).
//
// The top level is small (one entry per instance) and is built on the host:
// each instance's box is its BVH's root box, corner by corner through the
// instance transform (get_bounds, detail/bvh/get_bounds.inl:30-56), and the
// tree over those boxes is the reference's CPU lbvh_builder tree
// (host_build.cpp), stored as pair records whose leaves are ranges of
// instance slots.  On the device one lane follows one ray: it walks the top
// level with the binary traversal's slab test and, at each instance of a
// leaf, moves the ray into the instance's space (ori' = A (ori - T),
// dir' = A dir with A the stored inverse) and runs the 4-wide traversal of
// that instance's BVH, capped at the closest hit so far.  t is the same
// parameter in both spaces (the direction is not renormalized), so the
// closest hit over all instances is the smallest t.
#include "vr_internal.h"

#include <cstring>
#include <vector>

namespace vr {

namespace {

constexpr int BLOCK = 256;
constexpr int TLAS_STACK = 64;  // deferred top-level nodes; the host checks the depth

// one instance on the device
struct alignas(16) inst_rec {
    float inv[9];  // the inverse 3x3 map, row by row
    float ti[3];   // the negated translation
    uint32_t blas;
    int32_t inst_id;
    uint32_t pad[2];
};
static_assert(sizeof(inst_rec) == 64, "instance record");

struct tlas_args {
    const pair_node* pairs;
    uint32_t num_pairs;
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

__global__ __launch_bounds__(BLOCK) void k_tlas_trace(tlas_args A)
{
    __shared__ uint32_t stack[LDS_STACK * BLOCK];
    uint32_t* st = stack + threadIdx.x;
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= A.n) return;
    const float4* rp = reinterpret_cast<const float4*>(A.rays + i);
    const float4 r0 = rp[0], r1 = rp[1];
    const f3 ori = mk3(r0.x, r0.y, r0.z), dir = mk3(r0.w, r1.x, r1.y);
    const float tmin = r1.z, tmax = r1.w;
    const f3 inv = mk3(dir.x != 0.0f ? 1.0f / dir.x : FLT_MAX, dir.y != 0.0f ? 1.0f / dir.y : FLT_MAX,
                       dir.z != 0.0f ? 1.0f / dir.z : FLT_MAX);
    float best_t = FLT_MAX, best_u = 0.0f, best_v = 0.0f;
    int32_t best_tri = -1, best_inst = -1;
    uint32_t tst[TLAS_STACK];
    int sp = 0;
    uint32_t cur = A.root;
    uint32_t budget = 2u * A.num_slots + 2u;  // each node at most once
    bool done = A.num_slots == 0;
    while (!done) {
        while (!(cur & LEAF_BIT)) {
            if (--budget == 0 || cur >= A.num_pairs) {
                done = true;
                break;
            }
            const float4* p = reinterpret_cast<const float4*>(A.pairs + cur);
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
            const bool h0 = tf0 >= tn0 && tn0 < best_t;
            const bool h1 = tf1 >= tn1 && tn1 < best_t;
            if (h0 && h1) {
                const bool near1 = !(tn0 < tn1);
                if (sp >= TLAS_STACK) {  // cannot happen for trees the host accepted
                    done = true;
                    break;
                }
                tst[sp++] = near1 ? q3.x : q3.y;
                cur = near1 ? q3.y : q3.x;
            } else if (h0) {
                cur = q3.x;
            } else if (h1) {
                cur = q3.y;
            } else if (sp == 0) {
                done = true;
                break;
            } else {
                cur = tst[--sp];
            }
        }
        if (done) break;
        const uint32_t first = cur & LEAF_FIRST_MASK;
        const uint32_t count = ((cur >> 26) & 31u) + 1u;
        if (first + count > A.num_slots) break;  // corrupt entry
        for (uint32_t k = first; k < first + count; ++k) {
            const uint32_t ii = A.slots[k];
            const inst_rec* rec = A.insts + ii;
            float m[9], ti[3];
            for (int q = 0; q < 9; ++q) m[q] = rec->inv[q];
            for (int q = 0; q < 3; ++q) ti[q] = rec->ti[q];
            // index_bvh_inst_t::transform_ray (bvh.h:769-776)
            const f3 o2 = row_mul(m, mk3(ori.x + ti[0], ori.y + ti[1], ori.z + ti[2]));
            const f3 d2 = row_mul(m, dir);
            const hit_t h = traverse<false, false, LDS_STACK, BLOCK>(A.views[rec->blas], o2, d2, tmin,
                                                                     tmin_f(tmax, best_t), st, nullptr);
            // is_closer: strictly nearer, inside [tmin, tmax]
            if (h.tri >= 0 && h.t < best_t && h.t >= tmin && h.t <= tmax) {
                best_t = h.t;
                best_u = h.u;
                best_v = h.v;
                best_tri = h.tri;
                best_inst = (int32_t)ii;
            }
        }
        if (sp == 0) break;
        cur = tst[--sp];
    }
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
    A.hits[i] = out;
}

}  // namespace

}  // namespace vr

using namespace vr;

struct vr_tlas_t {
    int device = 0;
    uint32_t num_instances = 0;
    uint32_t num_pairs = 0;
    uint32_t root = 0;
    pair_node* pairs = nullptr;
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
    void* ptrs[] = {t->pairs, t->slots, t->insts, t->views};
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
        VR_CHECK(s.quads, VR_ERR_INVALID_ARGUMENT, "vr_tlas_build: BVH without traversal nodes");
        views[b] = s.view();
        VR_HIP_CHECK(hipMemcpyAsync(&root_boxes[6 * (size_t)b], s.root_box, 6 * sizeof(float), hipMemcpyDeviceToHost,
                                    st));
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
        rec.pad[0] = rec.pad[1] = 0;
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
            boxes[6 * (size_t)i + k] = blo[k];
            boxes[6 * (size_t)i + 3 + k] = bhi[k];
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
        if (depth > (uint32_t)TLAS_STACK)
            return bail(fail(VR_ERR_UNSUPPORTED, "vr_tlas_build: top level deeper than its traversal stack"));
    }
    t->num_pairs = (uint32_t)pairs.size();
    auto upload = [&](auto** dst, const auto* src, size_t count) -> bool {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(**dst);
        if (hipMalloc(reinterpret_cast<void**>(dst), bytes) != hipSuccess) return false;
        return count == 0 || hipMemcpy(*dst, src, count * sizeof(**dst), hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!upload(&t->pairs, pairs.data(), pairs.size()) || !upload(&t->slots, slots.data(), slots.size()) ||
        !upload(&t->insts, recs.data(), recs.size()) || !upload(&t->views, views.data(), views.size()))
        return bail(fail(VR_ERR_OUT_OF_MEMORY, "vr_tlas_build: device allocation or upload failed"));
    *out = t;
    return VR_OK;
    VR_TRY_END
}

int vr_tlas_closest_hit(vr_tlas t, const vr_ray* d_rays, size_t n, vr_inst_hit* d_hits, void* stream)
{
    VR_CHECK(t, VR_ERR_INVALID_ARGUMENT, "vr_tlas_closest_hit: null handle");
    VR_CHECK(n == 0 || (d_rays && d_hits), VR_ERR_INVALID_ARGUMENT, "vr_tlas_closest_hit: null array");
    VR_CHECK(n <= 0xFFFFFFFFull, VR_ERR_UNSUPPORTED, "vr_tlas_closest_hit: too many rays");
    if (n == 0) return VR_OK;
    int rc = ensure_device();
    if (rc) return rc;
    tlas_args A{t->pairs, t->num_pairs, t->root, t->slots, t->num_instances, t->insts, t->views,
                d_rays, (uint32_t)n, d_hits};
    hipLaunchKernelGGL(k_tlas_trace, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, (hipStream_t)stream, A);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

int vr_tlas_nodes(vr_tlas t, vr_bvh_node* h_nodes, uint32_t* h_indices, size_t* num_nodes)
{
    VR_CHECK(t && num_nodes, VR_ERR_INVALID_ARGUMENT, "vr_tlas_nodes: null argument");
    *num_nodes = t->nodes.size();
    if (h_nodes && !t->nodes.empty()) std::memcpy(h_nodes, t->nodes.data(), t->nodes.size() * sizeof(vr_bvh_node));
    if (h_indices && !t->indices.empty()) std::memcpy(h_indices, t->indices.data(), t->indices.size() * 4);
    return VR_OK;
}
