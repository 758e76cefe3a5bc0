// host_build.cpp -- the reference's two host builders, driven top down like
// build_top_down (detail/bvh/build_top_down.h:27-174): children appended at
// the end of the node array, the right subtree numbered before the left, leaf
// indices appended in visit order.  The recursion is an explicit stack here.
//
//  * binned SAH (binned_sah_builder with object splits, detail/bvh/sah.h:176-807)
//  * CPU LBVH   (lbvh_builder::init/find_split/split, detail/bvh/lbvh.h:500-656):
//    Morton-sorted references split at the highest differing code bit, equal
//    codes at the midpoint, leaves of up to max_leaf_size references.
//
// SAH: the tree is meant to be the reference's tree node for node, so the decisions
// follow the same arithmetic: 16 bins along the largest centroid extent, the
// split cost 1 + (A_l / A) 3 N_l + (A_r / A) 3 N_r over half surface areas,
// ties resolved towards the rightmost bin boundary, a leaf when the best split
// costs more than 3 N, the partition done by std::partition (so the reference
// order within each range is reproduced), children appended at the end of the
// node array and the right subtree numbered before the left.  The recursion is
// an explicit stack here.
#include "vr_internal.h"

#include <algorithm>
#include <array>
#include <vector>

namespace vr {

namespace {

constexpr int NUM_BINS = 16;

struct box {
    f3 lo{FLT_MAX, FLT_MAX, FLT_MAX};
    f3 hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void insert(f3 p)
    {
        lo = min3v(lo, p);
        hi = max3v(hi, p);
    }
    void insert(const box& b)
    {
        lo = min3v(lo, b.lo);
        hi = max3v(hi, b.hi);
    }
    f3 center() const { return muls3(add3(hi, lo), 0.5f); }
    f3 size() const { return sub3(hi, lo); }
};

inline float axis_of(f3 v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; }

// safe_half_surface_area (math/detail/aabb.inl:211-217)
inline float safe_hsa(const box& b)
{
    f3 s = b.size();
    s = f3{tmax_f(0.0f, s.x), tmax_f(0.0f, s.y), tmax_f(0.0f, s.z)};
    return s.x * s.y + s.y * s.z + s.z * s.x;
}

// max_index (math/detail/vector.inl:874-884): ties go to the later axis
inline int max_index3(f3 u)
{
    int i = u.y < u.x ? 0 : 1;
    return axis_of(u, 2) < axis_of(u, i) ? i : 2;
}

struct prim_ref {
    box bounds;
    int index;
};

struct bin {
    box prim_bounds, cent_bounds;
    int enter = 0, leave = 0;
    void merge(const bin& r)
    {
        prim_bounds.insert(r.prim_bounds);
        cent_bounds.insert(r.cent_bounds);
        enter += r.enter;
        leave += r.leave;
    }
};

struct range_info {
    box prim_bounds, cent_bounds;
    int first, last;  // refs[first, last)
    uint32_t node;
};

struct projection {
    float k0, k1;
    int axis;
    projection(const box& b, int a) : k0(axis_of(b.lo, a)), k1((float)NUM_BINS / (axis_of(b.hi, a) - k0)), axis(a) {}
    int project_unsafe(f3 v) const { return (int)(k1 * (axis_of(v, axis) - k0)); }
    int project(f3 v) const { return std::min(std::max(project_unsafe(v), 0), NUM_BINS - 1); }
};

inline void set_node(vr_bvh_node& n, const box& b, uint32_t first, uint32_t count, uint8_t axis, uint8_t sign)
{
    n.lo[0] = b.lo.x; n.lo[1] = b.lo.y; n.lo[2] = b.lo.z;
    n.hi[0] = b.hi.x; n.hi[1] = b.hi.y; n.hi[2] = b.hi.z;
    n.first = first;
    n.num_prims = (uint16_t)count;
    n.axis = axis;
    n.sign = sign;
}

}  // namespace

int sah_build_host(const vr_triangle* prims, size_t n, int max_leaf_size, vr_bvh_node* nodes, size_t* num_nodes,
                   uint32_t* indices)
{
    *num_nodes = 0;
    if (n == 0) return VR_OK;
    VR_CHECK(n < (size_t)1 << 31, VR_ERR_UNSUPPORTED, "vr_sah_build: too many primitives");
    if (max_leaf_size <= 0) max_leaf_size = 4;  // build_top_down.h:153-156
    VR_CHECK(max_leaf_size <= 65535, VR_ERR_INVALID_ARGUMENT, "vr_sah_build: leaf size exceeds bvh_node range");

    std::vector<prim_ref> refs(n);
    range_info root;
    root.first = 0;
    root.last = (int)n;
    root.node = 0;
    for (size_t i = 0; i < n; ++i) {
        const vr_triangle& t = prims[i];
        const f3 v1 = mk3(t.v1[0], t.v1[1], t.v1[2]);
        box b;
        b.insert(v1);
        b.insert(add3(v1, mk3(t.e1[0], t.e1[1], t.e1[2])));
        b.insert(add3(v1, mk3(t.e2[0], t.e2[1], t.e2[2])));
        refs[i].bounds = b;
        refs[i].index = (int)i;
        root.prim_bounds.insert(b);
        root.cent_bounds.insert(b.center());
    }

    size_t node_count = 1, index_count = 0;
    std::vector<range_info> stack{root};
    while (!stack.empty()) {
        const range_info leaf = stack.back();
        stack.pop_back();
        const int size = leaf.last - leaf.first;
        bool do_split = size > max_leaf_size;
        int axis = 0;
        projection pr(leaf.cent_bounds, 0);
        int best = -1;
        bin L, R;
        if (do_split) {
            const f3 cs = leaf.cent_bounds.size();
            axis = max_index3(cs);
            if (!(axis_of(cs, axis) > 0.0f)) do_split = false;
        }
        if (do_split) {
            pr = projection(leaf.cent_bounds, axis);
            std::array<bin, NUM_BINS> bins;
            for (int i = leaf.first; i < leaf.last; ++i) {
                const f3 c = refs[i].bounds.center();
                bin& b = bins[pr.project(c)];
                b.prim_bounds.insert(refs[i].bounds);
                b.cent_bounds.insert(c);
                b.enter++;
                b.leave++;
            }
            // sweeps (find_split, detail/bvh/sah.h:343-400)
            const float hsa_parent = safe_hsa(leaf.prim_bounds);
            std::array<bin, NUM_BINS> acc_l, acc_r;
            acc_l[0] = bins[0];
            for (int i = 1; i < NUM_BINS; ++i) {
                acc_l[i] = acc_l[i - 1];
                acc_l[i].merge(bins[i]);
            }
            acc_r[NUM_BINS - 1] = bins[NUM_BINS - 1];
            float best_cost = FLT_MAX;
            for (int i = NUM_BINS - 1; i > 0; --i) {
                acc_r[i - 1] = acc_r[i];
                acc_r[i - 1].merge(bins[i - 1]);
                const bin& l = acc_l[i - 1];
                const bin& r = acc_r[i];
                const float cost = 1.0f + (safe_hsa(l.prim_bounds) / hsa_parent) * (3.0f * (float)l.enter) +
                                   (safe_hsa(r.prim_bounds) / hsa_parent) * (3.0f * (float)r.leave);
                if (cost < best_cost) {
                    best_cost = cost;
                    best = i;
                }
            }
            if (best < 0 || best_cost > 3.0f * (float)size) {
                do_split = false;
            } else {
                L = acc_l[best - 1];
                R = acc_r[best];
            }
        }
        if (!do_split) {
            // leaf: its references, in range order, go to the index list
            VR_CHECK(size <= 65535, VR_ERR_UNSUPPORTED, "vr_sah_build: leaf larger than bvh_node allows");
            box b = leaf.prim_bounds;
            set_node(nodes[leaf.node], b, (uint32_t)index_count, (uint32_t)size, 0, 0);
            for (int i = leaf.first; i < leaf.last; ++i) indices[index_count++] = (uint32_t)refs[i].index;
            continue;
        }
        const int split_bin = best;
        auto pivot = std::partition(refs.begin() + leaf.first, refs.begin() + leaf.last,
                                    [&](const prim_ref& x) { return pr.project_unsafe(x.bounds.center()) < split_bin; });
        const int mid = (int)(pivot - refs.begin());
        const uint32_t first_child = (uint32_t)node_count;
        VR_CHECK(node_count + 2 <= 2 * n - 1, VR_ERR_INVALID_ARGUMENT, "vr_sah_build: node array overflow");
        const uint8_t sign = axis_of(L.prim_bounds.lo, axis) < axis_of(R.prim_bounds.lo, axis) ? 0 : 1;
        set_node(nodes[leaf.node], leaf.prim_bounds, first_child, 0, (uint8_t)axis, sign);
        node_count += 2;
        range_info lc{L.prim_bounds, L.cent_bounds, leaf.first, mid, first_child};
        range_info rc{R.prim_bounds, R.cent_bounds, mid, leaf.last, first_child + 1};
        // the right subtree is built (and numbered) first
        stack.push_back(lc);
        stack.push_back(rc);
    }
    *num_nodes = node_count;
    return VR_OK;
}

namespace {

// morton_encode3D (morton.h:33-47)
inline uint32_t spread10(uint32_t n)
{
    n &= 0x000003FFu;
    n = (n ^ (n << 16)) & 0xFF0000FFu;
    n = (n ^ (n << 8)) & 0x0300F00Fu;
    n = (n ^ (n << 4)) & 0x030C30C3u;
    n = (n ^ (n << 2)) & 0x09249249u;
    return n;
}

struct code_ref {
    int id;
    uint32_t code;
};

inline int clz32(uint32_t v) { return v == 0 ? 32 : __builtin_clz(v); }

// lbvh_builder::find_split (detail/bvh/lbvh.h:500-532): the last reference
// sharing the first reference's code prefix beyond the range's common prefix
int lbvh_find_split(const std::vector<code_ref>& r, int first, int last)
{
    const uint32_t cf = r[first].code, cl = r[last - 1].code;
    if (cf == cl) return (first + last) / 2;
    const int common = clz32(cf ^ cl);
    int result = first, step = last - first;
    do {
        step = (step + 1) / 2;
        const int next = result + step;
        if (next < last) {
            const uint32_t c = r[next].code;
            if (cf == c || clz32(cf ^ c) > common) result = next;
        }
    } while (step > 1);
    return result;
}

}  // namespace

namespace {

// init (detail/bvh/lbvh.h:545-599) from the primitive boxes: scene and
// centroid boxes, Morton codes, std::stable_sort by code; then the top-down
// splits
int lbvh_build_boxes(const std::vector<box>& pb, int max_leaf_size, vr_bvh_node* nodes, size_t* num_nodes,
                     uint32_t* indices)
{
    const size_t n = pb.size();
    box scene, cent;
    std::vector<f3> cs(n);
    for (size_t i = 0; i < n; ++i) {
        scene.insert(pb[i]);
        cs[i] = pb[i].center();
        cent.insert(cs[i]);
    }
    const f3 center = cent.center(), size = cent.size();
    std::vector<code_ref> refs(n);
    for (size_t i = 0; i < n; ++i) {
        f3 c = sub3(cs[i], center);
        c = add3(c, muls3(size, 0.5f));
        c = f3{c.x / size.x, c.y / size.y, c.z / size.z};
        c = min3v(max3v(muls3(c, 1024.0f), mk3(0.0f, 0.0f, 0.0f)), mk3(1023.0f, 1023.0f, 1023.0f));
        refs[i].id = (int)i;
        refs[i].code = spread10((uint32_t)c.x) | (spread10((uint32_t)c.y) << 1) | (spread10((uint32_t)c.z) << 2);
    }
    std::stable_sort(refs.begin(), refs.end(), [](const code_ref& a, const code_ref& b) { return a.code < b.code; });

    struct item {
        box bounds;
        int first, last;
        uint32_t node;
    };
    size_t node_count = 1, index_count = 0;
    std::vector<item> stack{{scene, 0, (int)n, 0u}};
    while (!stack.empty()) {
        const item it = stack.back();
        stack.pop_back();
        if (it.last - it.first <= max_leaf_size) {
            set_node(nodes[it.node], it.bounds, (uint32_t)index_count, (uint32_t)(it.last - it.first), 0, 0);
            for (int i = it.first; i < it.last; ++i) indices[index_count++] = (uint32_t)refs[i].id;
            continue;
        }
        const int split = lbvh_find_split(refs, it.first, it.last);
        box lb, rb;
        for (int i = it.first; i != split + 1; ++i) lb.insert(pb[refs[i].id]);
        for (int i = split + 1; i != it.last; ++i) rb.insert(pb[refs[i].id]);
        const uint32_t first_child = (uint32_t)node_count;
        VR_CHECK(node_count + 2 <= 2 * n - 1, VR_ERR_INVALID_ARGUMENT, "vr_lbvh_build_host: node array overflow");
        set_node(nodes[it.node], it.bounds, first_child, 0, 0, 0);
        node_count += 2;
        stack.push_back({lb, it.first, split + 1, first_child});
        stack.push_back({rb, split + 1, it.last, first_child + 1});
    }
    *num_nodes = node_count;
    return VR_OK;
}

}  // namespace

int lbvh_build_host(const vr_triangle* prims, size_t n, int max_leaf_size, vr_bvh_node* nodes, size_t* num_nodes,
                    uint32_t* indices)
{
    *num_nodes = 0;
    if (n == 0) return VR_OK;
    VR_CHECK(n < (size_t)1 << 31, VR_ERR_UNSUPPORTED, "vr_lbvh_build_host: too many primitives");
    if (max_leaf_size <= 0) max_leaf_size = 4;
    VR_CHECK(max_leaf_size <= 65535, VR_ERR_INVALID_ARGUMENT, "vr_lbvh_build_host: leaf size exceeds bvh_node range");
    std::vector<box> pb(n);
    for (size_t i = 0; i < n; ++i) {
        // get_bounds(triangle) (math/detail/triangle.inl:41-53)
        const vr_triangle& t = prims[i];
        const f3 v1 = mk3(t.v1[0], t.v1[1], t.v1[2]);
        box b;
        b.insert(v1);
        b.insert(add3(v1, mk3(t.e1[0], t.e1[1], t.e1[2])));
        b.insert(add3(v1, mk3(t.e2[0], t.e2[1], t.e2[2])));
        pb[i] = b;
    }
    return lbvh_build_boxes(pb, max_leaf_size, nodes, num_nodes, indices);
}

int lbvh_build_host_boxes(const float* boxes, size_t n, int max_leaf_size, vr_bvh_node* nodes, size_t* num_nodes,
                          uint32_t* indices)
{
    *num_nodes = 0;
    if (n == 0) return VR_OK;
    VR_CHECK(n < (size_t)1 << 31, VR_ERR_UNSUPPORTED, "lbvh over boxes: too many primitives");
    if (max_leaf_size <= 0) max_leaf_size = 4;
    VR_CHECK(max_leaf_size <= 65535, VR_ERR_INVALID_ARGUMENT, "lbvh over boxes: leaf size exceeds bvh_node range");
    std::vector<box> pb(n);
    for (size_t i = 0; i < n; ++i) {
        pb[i].lo = mk3(boxes[6 * i + 0], boxes[6 * i + 1], boxes[6 * i + 2]);
        pb[i].hi = mk3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]);
    }
    return lbvh_build_boxes(pb, max_leaf_size, nodes, num_nodes, indices);
}

// ---------------------------------------------------------------------------
// tree statistics and rotations over the reference node layout
// ---------------------------------------------------------------------------

namespace {

inline box node_box(const vr_bvh_node& n)
{
    box b;
    b.lo = f3{n.lo[0], n.lo[1], n.lo[2]};
    b.hi = f3{n.hi[0], n.hi[1], n.hi[2]};
    return b;
}

inline void set_box(vr_bvh_node& n, const box& b)
{
    n.lo[0] = b.lo.x, n.lo[1] = b.lo.y, n.lo[2] = b.lo.z;
    n.hi[0] = b.hi.x, n.hi[1] = b.hi.y, n.hi[2] = b.hi.z;
}

// surface_area (math/detail/aabb.inl:199-224): 2 (s.x s.y + s.y s.z + s.z s.x), s = max - min
inline float surface_area(const box& b)
{
    const f3 s = b.size();
    return 2.0f * (s.x * s.y + s.y * s.z + s.z * s.x);
}

inline box joined(box a, const box& b)
{
    a.insert(b);
    return a;
}

}  // namespace

// sah_cost(tree, ci, cl, cp) (detail/bvh/statistics.h:38-68): float sums in
// node order
int sah_cost_host(const vr_bvh_node* nodes, size_t n, float ci, float cl, float cp, float* out)
{
    VR_CHECK(n > 0, VR_ERR_INVALID_ARGUMENT, "sah_cost: empty tree");
    const float a_r = surface_area(node_box(nodes[0]));
    float a_l = 0.0f, a_i = 0.0f, a_l_n = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        const float sa = surface_area(node_box(nodes[i]));
        if (nodes[i].num_prims) {
            a_l += sa;
            a_l_n += sa * (float)nodes[i].num_prims;
        } else {
            a_i += sa;
        }
    }
    *out = ci * (a_i / a_r) + cl * (a_l / a_r) + cp * (a_l_n / a_r);
    return VR_OK;
}

// bvh_optimizer::optimize_tree_rotations (detail/bvh/optimize.h:22-262;
// Kensler, "Tree rotations for improving bounding volume hierarchies"):
// per-node costs C_t + (SA_0 C_0 + SA_1 C_1) / SA (leaves C_t + C_i n),
// assigned bottom up by sweeps over the node array; then one depth-first pass
// from the root that, at every inner node, tries the four exchanges of a
// grandchild with its uncle and applies the cheapest one cheaper than the
// node's cost.  As in the reference, the cost table is not updated after an
// exchange (entries stay with array positions while node records move), and
// the two affected boxes are recomputed.  A tree whose root is a leaf is left
// as it is (the reference's cost sweep would not end); the depth-first stack
// grows as needed (the reference's holds 64 entries).
int optimize_rotations_host(vr_bvh_node* nodes, size_t n, int* count)
{
    VR_CHECK(n > 0, VR_ERR_INVALID_ARGUMENT, "optimize_rotations: empty tree");
    *count = 0;
    if (nodes[0].num_prims) return VR_OK;
    const float C_t = 1.0f, C_i = 1.0f;
    auto child = [&](uint32_t i, int k) -> uint32_t { return nodes[i].first + (uint32_t)k; };
    auto inner = [&](uint32_t i) { return nodes[i].num_prims == 0; };
    for (size_t i = 0; i < n; ++i)
        if (inner((uint32_t)i)) VR_CHECK((size_t)nodes[i].first + 1 < n, VR_ERR_INVALID_ARGUMENT,
                                         "optimize_rotations: child index out of range");
    std::vector<float> costs(n, FLT_MAX);
    for (bool root_done = false; !root_done;) {
        bool progress = false;
        for (size_t i = 0; i < n; ++i) {
            if (costs[i] < FLT_MAX) continue;
            const vr_bvh_node& nd = nodes[i];
            if (nd.num_prims) {
                costs[i] = C_t + C_i * (float)nd.num_prims;
                progress = true;
            } else {
                const uint32_t c0 = nd.first, c1 = nd.first + 1;
                if (costs[c0] < FLT_MAX && costs[c1] < FLT_MAX) {
                    const float sa0 = surface_area(node_box(nodes[c0]));
                    const float sa1 = surface_area(node_box(nodes[c1]));
                    costs[i] = C_t + (sa0 * costs[c0] + sa1 * costs[c1]) / surface_area(node_box(nd));
                    progress = true;
                    if (i == 0) root_done = true;
                }
            }
        }
        VR_CHECK(progress || root_done, VR_ERR_INVALID_ARGUMENT, "optimize_rotations: not a tree (cycle)");
    }
    std::vector<uint32_t> st{0u};
    uint32_t addr = 0;
    int rotations = 0;
    while (!st.empty()) {
        const vr_bvh_node node = nodes[addr];
        if (!node.num_prims) {
            const float C = costs[addr];
            enum { R00, R01, R11, R10 };
            float c[4] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
            const uint32_t k0 = child(addr, 0), k1 = child(addr, 1);
            if (inner(k0)) {
                {  // 0.0 <-> 1
                    const box box00 = node_box(nodes[k1]);
                    const box box01 = node_box(nodes[child(k0, 1)]);
                    const box box0 = joined(box00, box01);
                    const box box1 = node_box(nodes[child(k0, 0)]);
                    const box bx = joined(box0, box1);
                    const float c00 = costs[k1], c01 = costs[child(k0, 1)];
                    const float cc0 = C_t + (surface_area(box00) * c00 + surface_area(box01) * c01) / surface_area(box0);
                    const float cc1 = costs[child(k0, 0)];
                    c[R00] = C_t + (surface_area(box0) * cc0 + surface_area(box1) * cc1) / surface_area(bx);
                }
                {  // 0.1 <-> 1
                    const box box00 = node_box(nodes[child(k0, 0)]);
                    const box box01 = node_box(nodes[k1]);
                    const box box0 = joined(box00, box01);
                    const box box1 = node_box(nodes[child(k0, 1)]);
                    const box bx = joined(box0, box1);
                    const float c00 = costs[child(k0, 0)], c01 = costs[k1];
                    const float cc0 = C_t + (surface_area(box00) * c00 + surface_area(box01) * c01) / surface_area(box0);
                    const float cc1 = costs[child(k0, 1)];
                    c[R01] = C_t + (surface_area(box0) * cc0 + surface_area(box1) * cc1) / surface_area(bx);
                }
            }
            if (inner(k1)) {
                {  // 1.1 <-> 0
                    const box box10 = node_box(nodes[child(k1, 0)]);
                    const box box11 = node_box(nodes[k0]);
                    const box box1 = joined(box10, box11);
                    const box box0 = node_box(nodes[child(k1, 1)]);
                    const box bx = joined(box0, box1);
                    const float c10 = costs[child(k1, 0)], c11 = costs[k0];
                    const float cc0 = costs[child(k1, 1)];
                    const float cc1 = C_t + (surface_area(box10) * c10 + surface_area(box11) * c11) / surface_area(box1);
                    c[R11] = C_t + (surface_area(box0) * cc0 + surface_area(box1) * cc1) / surface_area(bx);
                }
                {  // 1.0 <-> 0
                    const box box10 = node_box(nodes[k0]);
                    const box box11 = node_box(nodes[child(k1, 1)]);
                    const box box1 = joined(box10, box11);
                    const box box0 = node_box(nodes[child(k1, 0)]);
                    const box bx = joined(box0, box1);
                    const float c10 = costs[k0], c11 = costs[child(k1, 1)];
                    const float cc0 = costs[child(k1, 0)];
                    const float cc1 = C_t + (surface_area(box10) * c10 + surface_area(box11) * c11) / surface_area(box1);
                    c[R10] = C_t + (surface_area(box0) * cc0 + surface_area(box1) * cc1) / surface_area(bx);
                }
            }
            int rotation = -1;
            float c_new = C;
            for (int i = 0; i < 4; ++i) {
                if (c[i] < c_new) {
                    rotation = i;
                    c_new = c[i];
                }
            }
            uint32_t n0 = ~0u, n1 = ~0u;
            switch (rotation) {
            case R00: n0 = child(k0, 0); n1 = k1; break;
            case R01: n0 = child(k0, 1); n1 = k1; break;
            case R11: n0 = k0; n1 = child(k1, 1); break;
            case R10: n0 = k0; n1 = child(k1, 0); break;
            default: break;
            }
            if (n0 != ~0u && n1 != ~0u) {
                std::swap(nodes[n0], nodes[n1]);
                if (rotation == R00 || rotation == R01) {
                    const box b0 = joined(node_box(nodes[child(k0, 0)]), node_box(nodes[child(k0, 1)]));
                    set_box(nodes[k0], b0);
                    set_box(nodes[addr], joined(b0, node_box(nodes[k1])));
                } else {
                    const box b1 = joined(node_box(nodes[child(k1, 0)]), node_box(nodes[child(k1, 1)]));
                    set_box(nodes[k1], b1);
                    set_box(nodes[addr], joined(node_box(nodes[k0]), b1));
                }
                ++rotations;
            }
            addr = child(addr, 0);
            st.push_back(k1);
        } else {
            addr = st.back();
            st.pop_back();
        }
    }
    *count = rotations;
    return VR_OK;
}

}  // namespace vr
