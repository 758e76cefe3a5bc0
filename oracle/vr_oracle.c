/*
 * vr_oracle.c -- CPU restatement of Visionaray's ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see vr_oracle.h
 * for who may load it and how each function is pinned.  Build with
 * -ffp-contract=off (oracle/Makefile) so every float operation rounds exactly
 * as written.  Citations are relative to the reference repository root,
 * under include/visionaray/ unless stated otherwise.
 */
#include "vr_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * scalar helpers with the reference's semantics
 * ------------------------------------------------------------------------- */

/* math/detail/math.h:116-128: ternary min/max (NaN in x selects y for min,
 * NaN anywhere falls to the second operand's branch as written). */
static inline float tmin_f(float x, float y) { return x < y ? x : y; }
static inline float tmax_f(float x, float y) { return x < y ? y : x; }

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
static inline v3 add3(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 muls(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 smul(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 div3(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 neg3(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 min3v(v3 a, v3 b) { return mk(tmin_f(a.x, b.x), tmin_f(a.y, b.y), tmin_f(a.z, b.z)); }
static inline v3 max3v(v3 a, v3 b) { return mk(tmax_f(a.x, b.x), tmax_f(a.y, b.y), tmax_f(a.z, b.z)); }

/* math/detail/vector3.inl:292-307 */
static inline v3 cross3(v3 u, v3 v)
{
    return mk(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
static inline float dot3(v3 u, v3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }

/* math/detail/vector3.inl:332-335 with rsqrt = 1/sqrt (math/detail/math.h:596-599) */
static inline v3 normalize3(v3 v) { return muls(v, 1.0f / sqrtf(dot3(v, v))); }

/* math/detail/vector3.inl:432-447 */
static inline float min_el(v3 u)
{
    float r = u.x;
    if (u.y < r) r = u.y;
    if (u.z < r) r = u.z;
    return r;
}
static inline float max_el(v3 u)
{
    float r = u.x;
    if (u.y > r) r = u.y;
    if (u.z > r) r = u.z;
    return r;
}

/* math/detail/vector.inl:679-686 */
static inline v3 faceforward3(v3 n, v3 i, v3 nref) { return dot3(nref, i) < 0.0f ? neg3(n) : n; }

/* math/detail/vector3.inl:356-366 (also brdf.h:62-66) */
static inline void onb(v3 w, v3* u, v3* v)
{
    *v = fabsf(w.x) > fabsf(w.y) ? normalize3(mk(-w.z, 0.0f, w.x)) : normalize3(mk(0.0f, w.z, -w.y));
    *u = cross3(*v, w);
}

/* ---------------------------------------------------------------------------
 * Morton codes (morton.h:33-47)
 * ------------------------------------------------------------------------- */

static inline uint32_t separate_bits3(uint32_t n)
{
    n &= 0x000003FFu;
    n = (n ^ (n << 16)) & 0xFF0000FFu;
    n = (n ^ (n << 8)) & 0x0300F00Fu;
    n = (n ^ (n << 4)) & 0x030C30C3u;
    n = (n ^ (n << 2)) & 0x09249249u;
    return n;
}

uint32_t vro_morton_encode3d(uint32_t x, uint32_t y, uint32_t z)
{
    return separate_bits3(x) | (separate_bits3(y) << 1) | (separate_bits3(z) << 2);
}

/* math/detail/triangle.inl:41-53 with aabb::invalidate/insert
 * (math/detail/aabb.inl:136-155) */
void vro_tri_bounds(const vro_tri* t, float lo[3], float hi[3])
{
    v3 v1 = ld3(t->v1), e1 = ld3(t->e1), e2 = ld3(t->e2);
    v3 mn = mk(FLT_MAX, FLT_MAX, FLT_MAX), mx = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    v3 p[3] = {v1, add3(v1, e1), add3(v1, e2)};
    for (int k = 0; k < 3; ++k) {
        mn = min3v(mn, p[k]);
        mx = max3v(mx, p[k]);
    }
    lo[0] = mn.x; lo[1] = mn.y; lo[2] = mn.z;
    hi[0] = mx.x; hi[1] = mx.y; hi[2] = mx.z;
}

/* aabb::center (math/detail/aabb.inl:56-61) */
static inline v3 tri_centroid(const vro_tri* t)
{
    float lo[3], hi[3];
    vro_tri_bounds(t, lo, hi);
    return muls(add3(ld3(hi), ld3(lo)), 0.5f);
}

/* Centroid bounds as accumulated by compute_bounds_and_centroids
 * (detail/bvh/lbvh.h:216-240): min/max over all centroids, which the float
 * atomics make order independent. */
void vro_centroid_bounds(const vro_tri* tris, size_t n, float lo[3], float hi[3])
{
    v3 mn = mk(FLT_MAX, FLT_MAX, FLT_MAX), mx = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (size_t i = 0; i < n; ++i) {
        v3 c = tri_centroid(&tris[i]);
        mn = min3v(mn, c);
        mx = max3v(mx, c);
    }
    lo[0] = mn.x; lo[1] = mn.y; lo[2] = mn.z;
    hi[0] = mx.x; hi[1] = mx.y; hi[2] = mx.z;
}

/* assign_morton_codes (detail/bvh/lbvh.h:242-269; CPU twin :577-594) */
void vro_morton_codes(const vro_tri* tris, size_t n, uint32_t* codes)
{
    float lo[3], hi[3];
    vro_centroid_bounds(tris, n, lo, hi);
    v3 cb_lo = ld3(lo), cb_hi = ld3(hi);
    v3 center = muls(add3(cb_hi, cb_lo), 0.5f);
    v3 size = sub3(cb_hi, cb_lo);
    for (size_t i = 0; i < n; ++i) {
        v3 c = tri_centroid(&tris[i]);
        c = sub3(c, center);
        c = div3(add3(c, muls(size, 0.5f)), size);
        c = min3v(max3v(muls(c, 1024.0f), mk(0.0f, 0.0f, 0.0f)), mk(1023.0f, 1023.0f, 1023.0f));
        codes[i] = vro_morton_encode3d((uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z);
    }
}

/* Stable sort of prim_ref{id, code} by code (detail/bvh/lbvh.h:740-755 CUB
 * StableSortKeys; CPU std::stable_sort :596).  Input ids are 0..n-1, so the
 * result is the unique ordering by (code, id).  Counting sort on 30 bits in
 * three stable 10-bit passes. */
void vro_sort_prim_refs(const uint32_t* codes, size_t n, uint32_t* sorted_codes, int32_t* sorted_ids)
{
    uint32_t* kbuf = (uint32_t*)malloc(n * sizeof(uint32_t) + 1);
    int32_t* ibuf = (int32_t*)malloc(n * sizeof(int32_t) + 1);
    for (size_t i = 0; i < n; ++i) {
        sorted_codes[i] = codes[i];
        sorted_ids[i] = (int32_t)i;
    }
    uint32_t* ks = sorted_codes; int32_t* is = sorted_ids;
    uint32_t* kd = kbuf; int32_t* id = ibuf;
    for (int pass = 0; pass < 4; ++pass) {
        int shift = pass * 8;
        size_t count[257];
        memset(count, 0, sizeof(count));
        for (size_t i = 0; i < n; ++i) count[((ks[i] >> shift) & 0xFFu) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (size_t i = 0; i < n; ++i) {
            size_t dst = count[(ks[i] >> shift) & 0xFFu]++;
            kd[dst] = ks[i];
            id[dst] = is[i];
        }
        uint32_t* tk = ks; ks = kd; kd = tk;
        int32_t* ti = is; is = id; id = ti;
    }
    /* four passes: data is back in the caller's arrays */
    free(kbuf);
    free(ibuf);
}

/* clz on 32-bit (detail/bvh/lbvh.h:36-46) */
static inline int clz32(uint32_t v) { return v == 0 ? 32 : __builtin_clz(v); }

/* Karras' delta (detail/bvh/lbvh.h:123-145) */
static inline int karras_delta(const uint32_t* k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint32_t x = k[i] ^ k[j];
    if (x == 0) return clz32((uint32_t)i ^ (uint32_t)j) + 32;
    return clz32(x);
}

/* determine_range (detail/bvh/lbvh.h:120-185), restated literally including
 * the float step schedule t = ceil(l / divf), divf = 2, 4, 8, ... which keeps
 * running with t = 1 until divf overflows to infinity. */
void vro_determine_range(const uint32_t* k, int n, int i, int* first, int* last, int* split)
{
    int d = karras_delta(k, n, i, i + 1) >= karras_delta(k, n, i, i - 1) ? 1 : -1;
    int delta_min = karras_delta(k, n, i, i - d);
    int l_max = 2;
    while (karras_delta(k, n, i, i + l_max * d) > delta_min) l_max *= 2;
    int l = 0;
    for (int t = l_max >> 1; t >= 1; t >>= 1) {
        if (karras_delta(k, n, i, i + (l + t) * d) > delta_min) l += t;
    }
    int j = i + l * d;
    int delta_node = karras_delta(k, n, i, j);
    int s = 0;
    float divf = 2.0f;
    int t = (int)ceilf((float)l / divf);
    for (; t >= 1; divf *= 2.0f, t = (int)ceilf((float)l / divf)) {
        if (karras_delta(k, n, i, i + (s + t) * d) > delta_node) s += t;
    }
    *split = i + s * d + (d < 0 ? d : 0);
    if (d == 1) { *first = i; *last = j; }
    else { *first = j; *last = i; }
}

/* The GPU LBVH builder (detail/bvh/lbvh.h:681-844): Morton codes, stable
 * sort, Karras hierarchy (build_hierarchy :281-330), node boxes as the exact
 * union of leaf boxes (assign_node_bounds :332-372) and the collapse to
 * bvh_node[2n-1] (:374-451): root at 0 with first_child 1, the children of
 * Karras inner node i at 1+2i and 2+2i, leaves {prim box, prim id, 1},
 * axis = sign = 0.  indices = 0..n-1 (sequence :453-466).
 * Returns 0 on success, -1 on allocation failure. */
int vro_lbvh_build(const vro_tri* tris, size_t n, vro_node* nodes, uint32_t* indices,
                   uint32_t* sorted_codes, int32_t* sorted_ids)
{
    if (n == 0) return 0;
    uint32_t* codes = (uint32_t*)malloc(n * sizeof(uint32_t));
    int* left = (int*)malloc(n * sizeof(int));
    int* right = (int*)malloc(n * sizeof(int));
    int* parent = (int*)malloc(2 * n * sizeof(int));
    float* box = (float*)malloc(2 * n * 6 * sizeof(float));
    int* order = (int*)malloc(n * sizeof(int));
    if (!codes || !left || !right || !parent || !box || !order) {
        free(codes); free(left); free(right); free(parent); free(box); free(order);
        return -1;
    }
    vro_morton_codes(tris, n, codes);
    vro_sort_prim_refs(codes, n, sorted_codes, sorted_ids);
    for (size_t i = 0; i < n; ++i) indices[i] = (uint32_t)i;

    int nl = (int)n, ni = (int)n - 1;
    for (int i = 0; i < 2 * nl; ++i) parent[i] = -1;
    for (int i = 0; i < ni; ++i) {
        int f, l, s;
        vro_determine_range(sorted_codes, nl, i, &f, &l, &s);
        int lc = s, rc = s + 1;
        if (lc == f) { left[i] = ni + lc; parent[ni + lc] = i; }
        else { left[i] = lc; parent[lc] = i; }
        if (rc == l) { right[i] = ni + rc; parent[ni + rc] = i; }
        else { right[i] = rc; parent[rc] = i; }
    }
    /* leaf boxes, then inner boxes as unions in reverse topological order:
     * an iterative post-order walk from the root */
    for (int j = 0; j < nl; ++j) vro_tri_bounds(&tris[sorted_ids[j]], &box[(ni + j) * 6], &box[(ni + j) * 6 + 3]);
    if (ni > 0) {
        int* stack = (int*)malloc(2 * n * sizeof(int));
        int sp = 0, cnt = 0;
        stack[sp++] = 0;
        while (sp) {
            int x = stack[--sp];
            order[cnt++] = x;
            if (left[x] < ni) stack[sp++] = left[x];
            if (right[x] < ni) stack[sp++] = right[x];
        }
        for (int q = cnt - 1; q >= 0; --q) {
            int x = order[q];
            float* b = &box[x * 6];
            const float* a = &box[left[x] * 6];
            const float* c = &box[right[x] * 6];
            for (int k = 0; k < 3; ++k) {
                b[k] = tmin_f(a[k], c[k]);
                b[3 + k] = tmax_f(a[3 + k], c[3 + k]);
            }
        }
        free(stack);
    }
    /* collapse */
    memset(nodes, 0, (2 * n - 1) * sizeof(vro_node));
    if (ni == 0) {
        memcpy(nodes[0].lo, &box[0], 3 * sizeof(float));
        memcpy(nodes[0].hi, &box[3], 3 * sizeof(float));
        nodes[0].first = (uint32_t)sorted_ids[0];
        nodes[0].num_prims = 1;
    } else {
        memcpy(nodes[0].lo, &box[0], 3 * sizeof(float));
        memcpy(nodes[0].hi, &box[3], 3 * sizeof(float));
        nodes[0].first = 1;
        for (int x = 0; x < ni; ++x) {
            for (int side = 0; side < 2; ++side) {
                int c = side == 0 ? left[x] : right[x];
                vro_node* o = &nodes[1 + 2 * x + side];
                memcpy(o->lo, &box[c * 6], 3 * sizeof(float));
                memcpy(o->hi, &box[c * 6 + 3], 3 * sizeof(float));
                if (c >= ni) { o->first = (uint32_t)sorted_ids[c - ni]; o->num_prims = 1; }
                else { o->first = (uint32_t)(1 + 2 * c); o->num_prims = 0; }
            }
        }
    }
    free(codes); free(left); free(right); free(parent); free(box); free(order);
    return 0;
}

/* Refit (detail/bvh/refit.h:28-114): every node box becomes the exact union of
 * the primitive boxes below it. */
static void refit_rec(vro_node* nodes, uint32_t x, const vro_tri* tris, const uint32_t* indices)
{
    vro_node* nd = &nodes[x];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    if (nd->num_prims) {
        for (uint32_t i = nd->first; i < nd->first + nd->num_prims; ++i) {
            float a[3], b[3];
            vro_tri_bounds(&tris[indices[i]], a, b);
            for (int k = 0; k < 3; ++k) { lo[k] = tmin_f(lo[k], a[k]); hi[k] = tmax_f(hi[k], b[k]); }
        }
    } else {
        for (uint32_t c = nd->first; c < nd->first + 2; ++c) {
            refit_rec(nodes, c, tris, indices);
            for (int k = 0; k < 3; ++k) {
                lo[k] = tmin_f(lo[k], nodes[c].lo[k]);
                hi[k] = tmax_f(hi[k], nodes[c].hi[k]);
            }
        }
    }
    memcpy(nd->lo, lo, sizeof(lo));
    memcpy(nd->hi, hi, sizeof(hi));
}

void vro_refit(vro_node* nodes, size_t num_nodes, const vro_tri* tris, const uint32_t* indices)
{
    if (num_nodes) refit_rec(nodes, 0, tris, indices);
}

/* ---------------------------------------------------------------------------
 * Ray / triangle (math/intersect.h:176-229), Moller-Trumbore, two-sided
 * ------------------------------------------------------------------------- */

void vro_intersect_tri(const vro_ray* r, const vro_tri* t, int* hit, float* tt, float* u, float* v)
{
    v3 ori = ld3(r->ori), dir = ld3(r->dir);
    v3 v1 = ld3(t->v1), e1 = ld3(t->e1), e2 = ld3(t->e2);
    *hit = 0; *tt = -1.0f; *u = 0.0f; *v = 0.0f;
    v3 s1 = cross3(dir, e2);
    float div = dot3(s1, e1);
    if (!(div != 0.0f)) return;
    float inv_div = 1.0f / div;
    v3 d = sub3(ori, v1);
    float b1 = dot3(d, s1) * inv_div;
    if (!(b1 >= 0.0f && b1 <= 1.0f)) return;
    v3 s2 = cross3(d, e1);
    float b2 = dot3(dir, s2) * inv_div;
    if (!(b2 >= 0.0f && b1 + b2 <= 1.0f)) return;
    *hit = 1;
    *tt = dot3(e2, s2) * inv_div;
    *u = b1;
    *v = b2;
}

/* is_closer with tmin/tmax (update_if.h:68-107): strict <, so the first hit
 * found wins a tie. */
static inline int closer(int hit, float t, float ref_t, float tmin, float tmax)
{
    return hit && t < ref_t && t >= tmin && t <= tmax;
}

typedef struct {
    int hit;
    int32_t prim_id;
    uint32_t geom_id;
    uint32_t list_index;
    float t, u, v;
} hitrec;

/* Ray / BVH, device traversal order (detail/bvh/intersect.inl:166-356):
 * the root box is not tested; at an inner node both child boxes are tested
 * with tnear = max(tmin, max_el(min(tlo, thi))), tfar = min(tmax,
 * min_el(max(tlo, thi))), a child is entered when tfar >= tnear and
 * tnear < t_closest; with both entered the nearer (tnear1 < tnear2 ? 0 : 1)
 * goes first.  The reference keeps the deferred children in a 5-entry
 * stack plus a restart trail; a full stack visits the same nodes in the same
 * order (a restart only re-culls nodes whose tnear is no longer < t_closest,
 * whose primitives could not be accepted either). */
static hitrec trace_one(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                        const vro_ray* r, int any_hit, uint64_t* counters)
{
    hitrec res;
    res.hit = 0; res.prim_id = 0; res.geom_id = 0; res.list_index = 0;
    res.t = FLT_MAX; res.u = 0.0f; res.v = 0.0f;
    if (num_prims == 0) return res;
    v3 ori = ld3(r->ori), dir = ld3(r->dir);
    v3 inv = mk(dir.x != 0.0f ? 1.0f / dir.x : FLT_MAX,
                dir.y != 0.0f ? 1.0f / dir.y : FLT_MAX,
                dir.z != 0.0f ? 1.0f / dir.z : FLT_MAX);
    uint32_t stack[256];
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
        const vro_node* node = &nodes[cur];
        int dead = 0;
        while (node->num_prims == 0) {
            const vro_node* c = &nodes[node->first];
            if (counters) counters[0]++;
            float tn[2], tf[2];
            for (int k = 0; k < 2; ++k) {
                v3 tlo = mul3(sub3(ld3(c[k].lo), ori), inv);
                v3 thi = mul3(sub3(ld3(c[k].hi), ori), inv);
                v3 tnr = min3v(tlo, thi);
                v3 tfr = max3v(tlo, thi);
                tn[k] = tmax_f(r->tmin, max_el(tnr));
                tf[k] = tmin_f(r->tmax, min_el(tfr));
            }
            int b1 = tf[0] >= tn[0] && tn[0] < res.t;
            int b2 = tf[1] >= tn[1] && tn[1] < res.t;
            if (b1 && b2) {
                uint32_t nearc = tn[0] < tn[1] ? 0u : 1u;
                stack[sp++] = node->first + (1u - nearc);
                cur = node->first + nearc;
            } else if (b1 || b2) {
                cur = node->first + (b1 ? 0u : 1u);
            } else {
                dead = 1;
                break;
            }
            node = &nodes[cur];
        }
        if (!dead) {
            for (uint32_t i = node->first; i < node->first + node->num_prims; ++i) {
                const vro_tri* p = &tris[indices[i]];
                int h; float t, u, v;
                if (counters) counters[1]++;
                vro_intersect_tri(r, p, &h, &t, &u, &v);
                if (closer(h, t, res.t, r->tmin, r->tmax)) {
                    res.hit = 1; res.t = t; res.u = u; res.v = v;
                    res.prim_id = (int32_t)p->prim_id; res.geom_id = p->geom_id; res.list_index = i;
                    if (any_hit) return res;
                }
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return res;
}

static void store_hit(vro_hit* o, const hitrec* h)
{
    o->prim_id = h->hit ? h->prim_id : -1;
    o->t = h->hit ? h->t : FLT_MAX;
    o->u = h->hit ? h->u : 0.0f;
    o->v = h->hit ? h->v : 0.0f;
}

void vro_trace(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
               const vro_ray* rays, size_t num_rays, int any_hit, vro_hit* hits)
{
    for (size_t i = 0; i < num_rays; ++i) {
        hitrec h = trace_one(nodes, tris, indices, num_prims, &rays[i], any_hit, NULL);
        store_hit(&hits[i], &h);
    }
}

/* closest_hit/any_hit over a plain primitive list
 * (detail/traverse_linear.inl:24-56): lowest index wins a tie. */
void vro_trace_brute(const vro_tri* tris, size_t num_prims, const vro_ray* rays, size_t num_rays,
                     int any_hit, vro_hit* hits)
{
    for (size_t i = 0; i < num_rays; ++i) {
        hitrec res;
        res.hit = 0; res.prim_id = 0; res.t = FLT_MAX; res.u = res.v = 0.0f;
        for (size_t k = 0; k < num_prims; ++k) {
            int h; float t, u, v;
            vro_intersect_tri(&rays[i], &tris[k], &h, &t, &u, &v);
            if (closer(h, t, res.t, rays[i].tmin, rays[i].tmax)) {
                res.hit = 1; res.t = t; res.u = u; res.v = v; res.prim_id = (int32_t)tris[k].prim_id;
                if (any_hit) break;
            }
        }
        store_hit(&hits[i], &res);
    }
}

/* ---------------------------------------------------------------------------
 * Seeds and the CPU random engine
 * ------------------------------------------------------------------------- */

/* make_random_seed.h:23-37 (TEA-style, 4 rounds) */
uint32_t vro_make_random_seed(int32_t val0, int32_t val1)
{
    uint32_t v0 = (uint32_t)val0, v1 = (uint32_t)val1, s0 = 0;
    for (uint32_t n = 0; n < 4; ++n) {
        s0 += 0x9E3779B9u;
        v0 += ((v1 << 4) + 0xA341316Cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xC8013EA4u);
        v1 += ((v0 << 4) + 0xAd90777Du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7E95761Eu);
    }
    return v0;
}

/* random_generator<float> on the CPU (random_generator.h:29-62):
 * std::default_random_engine is libstdc++'s minstd_rand0 (x <- 16807 x mod
 * 2^31-1, a zero seed becomes 1) and uniform_real_distribution<float>(0, 1)
 * draws one engine value through generate_canonical<float, 24>:
 * float(x - 1) / float(2^31 - 2), the divisor rounding to 2^31, with results
 * >= 1 replaced by the largest float below 1. */
typedef struct { uint32_t x; } rng_t;

static inline void rng_seed(rng_t* g, uint32_t seed)
{
    uint32_t x = seed % 2147483647u;
    g->x = x == 0 ? 1u : x;
}

static inline float rng_next(rng_t* g)
{
    g->x = (uint32_t)(((uint64_t)g->x * 16807u) % 2147483647u);
    float r = (float)(g->x - 1u) / 2147483648.0f;
    if (r >= 1.0f) r = 0x1.fffffep-1f;
    return r * (1.0f - 0.0f) + 0.0f;
}

void vro_rng_stream(uint32_t seed, size_t count, float* out)
{
    rng_t g;
    rng_seed(&g, seed);
    for (size_t i = 0; i < count; ++i) out[i] = rng_next(&g);
}

/* two draws for one argument list, in the evaluation order the reference
 * build used */
static inline void rng_pair(rng_t* g, uint32_t arg_order, float* first_arg, float* second_arg)
{
    if (arg_order == 0) {
        *first_arg = rng_next(g);
        *second_arg = rng_next(g);
    } else {
        *second_arg = rng_next(g);
        *first_arg = rng_next(g);
    }
}

/* ---------------------------------------------------------------------------
 * Primary rays: make_primary_ray (jittered) (detail/sched_common.h:209-233)
 * and pinhole_camera::primary_ray (detail/pinhole_camera.inl:103-121)
 * ------------------------------------------------------------------------- */

static vro_ray primary_ray(const vro_camera* cam, rng_t* g, uint32_t arg_order, int x, int y, int w, int h)
{
    float a, b;
    rng_pair(g, arg_order, &a, &b);
    float jx = a - 0.5f, jy = b - 0.5f;
    float fx = (float)x + jx, fy = (float)y + jy;
    float sx = (fx + 0.5f) / (float)w;
    float sy = (fy + 0.5f) / (float)h;
    sx = (1.0f - sx) * 0.0f + sx * 1.0f;
    sy = (1.0f - sy) * 0.0f + sy * 1.0f;
    float u = 2.0f * sx - 1.0f;
    float v = 2.0f * sy - 1.0f;
    v3 d = add3(add3(muls(ld3(cam->U), u), muls(ld3(cam->V), v)), ld3(cam->W));
    d = normalize3(d);
    vro_ray r;
    r.ori[0] = cam->eye[0]; r.ori[1] = cam->eye[1]; r.ori[2] = cam->eye[2];
    r.dir[0] = d.x; r.dir[1] = d.y; r.dir[2] = d.z;
    r.tmin = 0.0f;
    r.tmax = FLT_MAX;
    return r;
}

/* cosine_sample_hemisphere (sampling.h:117-127) */
static inline v3 cosine_hemisphere(float u1, float u2)
{
    float r = sqrtf(u1);
    float theta = 6.28318530717958623199592693709e+00f * u2;
    float x = r * cosf(theta);
    float y = r * sinf(theta);
    float z = sqrtf(1.0f - u1);
    return mk(x, y, z);
}

static inline vro_ray mkray(v3 o, v3 d, float tmin, float tmax)
{
    vro_ray r;
    r.ori[0] = o.x; r.ori[1] = o.y; r.ori[2] = o.z;
    r.dir[0] = d.x; r.dir[1] = d.y; r.dir[2] = d.z;
    r.tmin = tmin; r.tmax = tmax;
    return r;
}

/* pixel_access::blend for RGBA32F accum (detail/pixel_access.h:1040-1131)
 * and the store of the blended accum into the color target
 * (detail/sched_common.h:474-509). */
static void blend_store(float* accum, float* color, size_t px, const float c[4], float sf, float df)
{
    for (int k = 0; k < 4; ++k) {
        float a = c[k] * sf + accum[px * 4 + k] * df;
        accum[px * 4 + k] = a;
        color[px * 4 + k] = a;
    }
}

/* AO kernel of src/examples/ao/main.cpp:192-260 under tiled_sched +
 * jittered_blend (one primary sample per pixel and frame). */
void vro_render_ao(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                   const vro_camera* cam, const vro_render_params* p, float* accum, float* color,
                   int y_begin, int y_end)
{
    int w = (int)p->width, h = (int)p->height;
    for (int y = y_begin; y < y_end; ++y) {
        for (int x = 0; x < w; ++x) {
            rng_t g;
            rng_seed(&g, vro_make_random_seed(y * w + x, (int32_t)p->frame_id));
            float sum[4] = {0, 0, 0, 0};
            for (uint32_t s = 0; s < p->spp; ++s) {
                vro_ray r = primary_ray(cam, &g, p->arg_order, x, y, w, h);
                float c[4] = {p->bg[0], p->bg[1], p->bg[2], 1.0f};
                hitrec hr = trace_one(nodes, tris, indices, num_prims, &r, 0, NULL);
                if (hr.hit) {
                    v3 ori = ld3(r.ori), dir = ld3(r.dir);
                    v3 pos = add3(ori, muls(dir, hr.t));
                    const vro_tri* pr = &tris[indices[hr.list_index]];
                    v3 n = normalize3(cross3(ld3(pr->e1), ld3(pr->e2)));
                    n = faceforward3(n, neg3(dir), n);
                    v3 u, v, wv = n;
                    onb(wv, &u, &v);
                    float ao = 0.0f;
                    float inc = 1.0f / (float)p->ao_samples;
                    for (uint32_t i = 0; i < p->ao_samples; ++i) {
                        float a, b;
                        rng_pair(&g, p->arg_order, &a, &b);
                        v3 sp = cosine_hemisphere(a, b);
                        v3 d = normalize3(add3(add3(smul(sp.x, u), smul(sp.y, v)), smul(sp.z, wv)));
                        vro_ray ar = mkray(add3(pos, muls(d, 1e-3f)), d, 0.0f, p->ao_radius);
                        hitrec ah = trace_one(nodes, tris, indices, num_prims, &ar, 1, NULL);
                        if (ah.hit) ao = ao + inc;
                    }
                    ao = 1.0f - ao;
                    c[0] = ao; c[1] = ao; c[2] = ao; c[3] = 1.0f;
                }
                for (int k = 0; k < 4; ++k) sum[k] += c[k];
            }
            for (int k = 0; k < 4; ++k) sum[k] /= (float)p->spp;
            blend_store(accum, color, (size_t)y * w + x, sum, p->sfactor, p->dfactor);
        }
    }
}

/* area_light<float, basic_triangle>::sample (detail/area_light.inl:31-56):
 * a point by sample_surface (math/detail/triangle.inl:69-82; two draws in
 * statement order), the direction and distance from the reference point,
 * intensity cl * kl, the normal normalize(cross(e1, e2)) (get_normal.h:108-113),
 * the area 0.5 |e1 x e2| (triangle.inl:36-39) and the solid-angle pdf. */
typedef struct {
    v3 dir, intensity, normal;
    float dist, area, pdf;
} light_sample;

static light_sample sample_area_light(const vro_light* l, v3 ref, rng_t* g)
{
    light_sample r;
    const float u1 = rng_next(g);
    const float u2 = rng_next(g);
    const v3 e1 = ld3(l->geometry.e1), e2 = ld3(l->geometry.e2);
    const v3 a = ld3(l->geometry.v1);
    const v3 b = add3(a, e1);
    const v3 c = add3(a, e2);
    const v3 pos = add3(add3(muls(a, 1.0f - sqrtf(u1)), muls(muls(b, sqrtf(u1)), 1.0f - u2)),
                        muls(muls(c, sqrtf(u1)), u2));
    r.dir = sub3(pos, ref);
    r.dist = sqrtf(dot3(r.dir, r.dir));
    r.intensity = muls(ld3(l->cl), l->kl);
    const v3 cr = cross3(e1, e2);
    r.normal = normalize3(cr);
    r.area = 0.5f * sqrtf(dot3(cr, cr));
    const float ldotln = fabsf(dot3(neg3(normalize3(r.dir)), r.normal));
    const float solid_angle = (ldotln * r.area) / (r.dist * r.dist);
    r.pdf = 1.0f / solid_angle;
    return r;
}

/* Optional log of the queries the path tracer traces (diagnostics of the
 * tests): ray, any-hit flag, result, bounce and pixel. */
static vro_query* g_log;
static size_t g_log_cap, g_log_n;
static uint32_t g_log_pixel;

void vro_query_log(vro_query* buf, size_t cap)
{
    g_log = buf;
    g_log_cap = cap;
    g_log_n = 0;
}

size_t vro_query_log_count(void) { return g_log_n; }

static void log_query(const vro_ray* r, int any, const hitrec* h, uint32_t bounce)
{
    if (!g_log || g_log_n >= g_log_cap) return;
    vro_query* q = &g_log[g_log_n++];
    q->ray = *r;
    q->any = any;
    q->hit = h->hit;
    q->prim_id = h->hit ? h->prim_id : -1;
    q->t = h->t;
    q->bounce = bounce;
    q->pixel = g_log_pixel;
}

/* power_heuristic (sampling.h:73-79) */
static inline float power_heuristic(float a, float b) { return (a * a) / (a * a + b * b); }

/* pathtracing::kernel::operator() (detail/pathtracing.inl:43-233) for scalar
 * rays with generic_material<emissive, matte> (detail/material/matte.inl:28-71,
 * detail/material/emissive.inl:14-40, brdf.h:39-84) and geometric normals
 * normalize(cross(e1, e2)) (get_normal.h:108-113, get_surface.h:135-160).
 * Without lights every MIS weight is 1.  With num_lights > 0 (lines 116-196):
 * an emitter hit at bounce > 0 is weighted by power_heuristic(1, pdf_light /
 * num_lights), pdf_light the solid-angle density of the hit triangle (the
 * emissive material's sample() reports pdf 1); and every hit surface samples
 * one light uniformly (light = int(u * num_lights)), traces the shadow ray
 * [pos + L eps, tmin eps, tmax dist - eps] for any hit, and adds
 * power_heuristic(pdf_light / num_lights, pdf_brdf * max(throughput)) *
 * throughput * f * (cos / pdf_light) * num_lights when the surface still
 * scatters, the shadow ray is clear and both cosines are positive.  The
 * light's three draws come after the BRDF sample and before the roulette, on
 * every hit (emitters included).
 * A ray that leaves the path (emission or a zero pdf) before bounce 2 ends the
 * loop at the next closest_hit in the reference; that extra query changes no
 * output and is not performed (or counted) here; neither is a shadow ray
 * whose contribution is masked. */
static void pt_sample(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                      const vro_material* mats, const vro_light* lights, size_t num_lights,
                      const vro_render_params* p, vro_ray ray, rng_t* g, float out[4], uint64_t* counts)
{
    const float inv_pi = 3.18309886183790691216444201928e-01f;
    const float pi = 3.14159265358979323846264338328e+00f;
    float I[3] = {0, 0, 0};
    float tp[3] = {1, 1, 1};
    int result_hit = 0;
    for (uint32_t bounce = 0; bounce < p->num_bounces; ++bounce) {
        hitrec hr = trace_one(nodes, tris, indices, num_prims, &ray, 0, NULL);
        log_query(&ray, 0, &hr, bounce);
        if (counts) counts[bounce == 0 ? 0 : 1]++;
        if (!hr.hit) {
            for (int k = 0; k < 3; ++k) I[k] += p->ambient[k] * tp[k];
            break;
        }
        if (bounce == 0) result_hit = 1;
        v3 dir = ld3(ray.dir), ori = ld3(ray.ori);
        v3 view = neg3(dir);
        v3 pos = add3(ori, muls(dir, hr.t));
        const vro_tri* pr = &tris[indices[hr.list_index]];
        v3 gn = normalize3(cross3(ld3(pr->e1), ld3(pr->e2)));
        v3 sn = gn;
        const vro_material* m = &mats[hr.geom_id];
        v3 refl = mk(0.0f, 0.0f, 0.0f);
        float pdf = 0.0f;
        float src[3];
        int emission = m->type == 1;
        if (emission) {
            for (int k = 0; k < 3; ++k) src[k] = (1.0f * m->ce[k]) * m->ls;
            pdf = 1.0f;
        } else {
            v3 n = faceforward3(sn, view, gn);
            v3 u, v, w = n;
            onb(w, &u, &v);
            float a, b;
            rng_pair(g, p->arg_order, &a, &b);
            v3 sp = cosine_hemisphere(a, b);
            refl = normalize3(add3(add3(smul(sp.x, u), smul(sp.y, v)), smul(sp.z, w)));
            pdf = dot3(n, refl) * inv_pi;
            for (int k = 0; k < 3; ++k) src[k] = 1.0f * ((m->cd[k] * m->kd) * inv_pi);
        }
        int zero_pdf = pdf <= 0.0f;
        int active = 1;
        if (emission) {
            float wgt = 1.0f;
            if (num_lights > 0 && bounce > 0) {
                const v3 ce1 = ld3(pr->e1), ce2 = ld3(pr->e2);
                const v3 cc = cross3(ce1, ce2);
                const float A = 0.5f * sqrtf(dot3(cc, cc));
                const v3 dl = sub3(pos, ori);
                const float ld = sqrtf(dot3(dl, dl));
                const v3 L = normalize3(dl);
                const float ldotln = fabsf(dot3(neg3(L), gn));
                const float solid_angle = (ldotln * A) / (ld * ld);
                const float light_pdf = 1.0f / solid_angle;
                wgt = power_heuristic(pdf, light_pdf / (float)num_lights);
            }
            for (int k = 0; k < 3; ++k) I[k] += (wgt * tp[k]) * src[k];
            active = 0;
        }
        if (zero_pdf) active = 0;
        v3 n = faceforward3(sn, view, gn);
        if (num_lights > 0) {
            const float u = rng_next(g);
            size_t id = (size_t)(int)(u * (float)num_lights);
            if (id >= num_lights) id = num_lights - 1;
            const light_sample ls = sample_area_light(&lights[id], pos, g);
            const float ld = ls.dist;
            const v3 L = normalize3(ls.dir);
            const v3 ln = faceforward3(ls.normal, neg3(L), ls.normal);
            const float ldotn = dot3(L, n);
            const float ldotln = fabsf(dot3(neg3(L), ln));
            if (active && ldotn > 0.0f && ldotln > 0.0f) {
                const vro_ray shadow = mkray(add3(pos, muls(L, p->epsilon)), L, p->epsilon, ld - p->epsilon);
                const hitrec lhr = trace_one(nodes, tris, indices, num_prims, &shadow, 1, NULL);
                log_query(&shadow, 1, &lhr, bounce);
                if (counts) counts[1]++;
                if (!lhr.hit) {
                    /* matte: pdf dot(n, L) / pi; shade ((f pi) E) max(0, dot(n, L)) */
                    float bpdf = dot3(n, L) * inv_pi;
                    float prob = tp[0];
                    if (tp[1] > prob) prob = tp[1];
                    if (tp[2] > prob) prob = tp[2];
                    bpdf *= prob;
                    const float ndotl = tmax_f(0.0f, dot3(n, L));
                    const float E[3] = {ls.intensity.x, ls.intensity.y, ls.intensity.z};
                    const float wgt = power_heuristic(ls.pdf / (float)num_lights, bpdf);
                    for (int k = 0; k < 3; ++k) {
                        const float f = 1.0f * ((m->cd[k] * m->kd) * inv_pi);
                        const float sh = ((f * pi) * E[k]) * ndotl;
                        const float s = (sh * inv_pi) / ldotn;
                        I[k] += (((wgt * tp[k]) * s) * (ldotn / ls.pdf)) * (float)num_lights;
                    }
                }
            }
        }
        float q = dot3(n, refl) / pdf;
        for (int k = 0; k < 3; ++k) tp[k] = tp[k] * (src[k] * q);
        if (zero_pdf) tp[0] = tp[1] = tp[2] = 0.0f;
        if (bounce >= 2) {
            float prob = tp[0];
            if (tp[1] > prob) prob = tp[1];
            if (tp[2] > prob) prob = tp[2];
            int terminate = rng_next(g) > prob;
            if (terminate) active = 0;
            for (int k = 0; k < 3; ++k) tp[k] = tp[k] / prob;
            if (!active) break;
        }
        if (!active) break;
        v3 o2 = add3(pos, muls(refl, p->epsilon));
        ray = mkray(o2, refl, ray.tmin, ray.tmax);
    }
    if (result_hit) { out[0] = I[0]; out[1] = I[1]; out[2] = I[2]; out[3] = 1.0f; }
    else { out[0] = p->bg[0]; out[1] = p->bg[1]; out[2] = p->bg[2]; out[3] = 1.0f; }
}

/* ---------------------------------------------------------------------------
 * Adjoint of the path tracer (no reference counterpart: the reference has no
 * differentiation).  FP64 accumulation of d(loss)/d(albedo cd) and
 * d(loss)/d(emission ce) per material, replaying each sample's path with the
 * forward's random stream.
 *
 * With no light list a path's estimate is I = T_K * E: E the emission
 * ce * ls of the emitter that ends it (or the ambient term on escape) and
 * T_K the throughput, the product of f_j = cd_{m_j} kd / pi * (cos / pdf)
 * over the matte vertices j < K divided by the Russian-roulette
 * probabilities of the vertices j >= 2.  The roulette probabilities are held
 * constant (their decisions and values are not differentiated), which keeps
 * the gradient estimate unbiased for the gradient of the expected image:
 *   d I_c / d cd_{m,c} = E_c * sum_{j : m_j = m} g_j * prod_{k != j} f_{k,c}
 *   d I_c / d ce_{e,c} = T_c * ls_e   for the ending emitter e,
 * with f_{j,c} = cd_{m_j,c} * g_j, g_j = kd / pi * (cos / pdf).  When no
 * factor of channel c is zero the sum is I_c * #{j : m_j = m} / cd_{m,c}.
 * When exactly one factor (vertex j0) is zero -- a zero albedo channel, a
 * zero kd, or an underflow -- only material m_{j0} has a nonzero derivative,
 * g_{j0} * (product of the other factors) * E_c; with two or more zero
 * factors every term holds a zero factor.  A zero albedo channel therefore
 * gets its finite derivative instead of 0 / 0.
 * ------------------------------------------------------------------------- */

/* leave-one-out bookkeeping of a path's per-channel throughput factors */
typedef struct {
    double nzp[3];   /* product of the nonzero factors over the roulette probabilities */
    int zc[3];       /* zero factors so far */
    uint32_t zm[3];  /* material of the (first) zero-factor vertex */
    double zg[3];    /* its d factor / d cd */
} adj_factors;

/* d I / d cd for every material of the path's matte vertices, given the
 * ending term's emission E_c (ambient or ce * ls) and I_c = T_c * E_c */
static void adj_add_cd(const adj_factors* af, const uint32_t* mids, int nm, const vro_material* mats,
                       const double w[3], const double E[3], const double I[3], double* grad_cd)
{
    for (int c = 0; c < 3; ++c) {
        if (af->zc[c] == 0) {
            for (int j = 0; j < nm; ++j) grad_cd[3 * mids[j] + c] += w[c] * I[c] / (double)mats[mids[j]].cd[c];
        } else if (af->zc[c] == 1) {
            grad_cd[3 * af->zm[c] + c] += w[c] * E[c] * af->zg[c] * af->nzp[c];
        }
    }
}
static void pt_sample_adjoint(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                              const vro_material* mats, const vro_render_params* p, vro_ray ray, rng_t* g,
                              const double w[3], double* grad_cd, double* grad_ce)
{
    float tp[3] = {1, 1, 1};
    uint32_t mids[64];
    int nm = 0;
    adj_factors af = {{1.0, 1.0, 1.0}, {0, 0, 0}, {0, 0, 0}, {0.0, 0.0, 0.0}};
    for (uint32_t bounce = 0; bounce < p->num_bounces; ++bounce) {
        hitrec hr = trace_one(nodes, tris, indices, num_prims, &ray, 0, NULL);
        if (!hr.hit) {
            double I[3], E[3];
            for (int k = 0; k < 3; ++k) {
                I[k] = (double)(p->ambient[k] * tp[k]);
                E[k] = (double)p->ambient[k];
            }
            adj_add_cd(&af, mids, nm, mats, w, E, I, grad_cd);
            return;
        }
        v3 dir = ld3(ray.dir), ori = ld3(ray.ori);
        v3 view = neg3(dir);
        v3 pos = add3(ori, muls(dir, hr.t));
        const vro_tri* pr = &tris[indices[hr.list_index]];
        v3 gn = normalize3(cross3(ld3(pr->e1), ld3(pr->e2)));
        uint32_t mid = hr.geom_id;
        const vro_material* m = &mats[mid];
        if (m->type == 1) {
            double I[3], E[3];
            for (int k = 0; k < 3; ++k) {
                E[k] = (double)((1.0f * m->ce[k]) * m->ls);
                I[k] = (double)((1.0f * tp[k]) * ((1.0f * m->ce[k]) * m->ls));
            }
            for (int c = 0; c < 3; ++c) grad_ce[3 * mid + c] += w[c] * (double)tp[c] * (double)m->ls;
            adj_add_cd(&af, mids, nm, mats, w, E, I, grad_cd);
            /* the forward still draws its roulette number here */
            if (bounce >= 2) (void)rng_next(g);
            return;
        }
        v3 n = faceforward3(gn, view, gn);
        v3 u, v, wv = n;
        onb(wv, &u, &v);
        float a, b;
        rng_pair(g, p->arg_order, &a, &b);
        v3 sp = cosine_hemisphere(a, b);
        v3 refl = normalize3(add3(add3(smul(sp.x, u), smul(sp.y, v)), smul(sp.z, wv)));
        float pdf = dot3(n, refl) * 3.18309886183790691216444201928e-01f;
        float src[3];
        for (int k = 0; k < 3; ++k) src[k] = 1.0f * ((m->cd[k] * m->kd) * 3.18309886183790691216444201928e-01f);
        if (pdf <= 0.0f) {
            if (bounce >= 2) (void)rng_next(g);
            return;
        }
        if (nm < 64) mids[nm++] = mid;
        float q = dot3(n, refl) / pdf;
        for (int k = 0; k < 3; ++k) {
            const float f = src[k] * q;
            tp[k] = tp[k] * f;
            if (f == 0.0f) {
                if (af.zc[k]++ == 0) {
                    af.zm[k] = mid;
                    af.zg[k] = (double)m->kd / 3.14159265358979323846 * (double)q;
                }
            } else {
                af.nzp[k] *= (double)f;
            }
        }
        if (bounce >= 2) {
            float prob = tp[0];
            if (tp[1] > prob) prob = tp[1];
            if (tp[2] > prob) prob = tp[2];
            int terminate = rng_next(g) > prob;
            for (int k = 0; k < 3; ++k) {
                tp[k] = tp[k] / prob;
                af.nzp[k] /= (double)prob;
            }
            if (terminate) return;
        }
        v3 o2 = add3(pos, muls(refl, p->epsilon));
        ray = mkray(o2, refl, ray.tmin, ray.tmax);
    }
}

void vro_render_pt_adjoint(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                           const vro_material* mats, const vro_camera* cam, const vro_render_params* p,
                           const float* dl_dimg, double* grad_cd, double* grad_ce, int y_begin, int y_end)
{
    int w = (int)p->width, h = (int)p->height;
    for (int y = y_begin; y < y_end; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t px = (size_t)y * w + x;
            double wgt[3];
            for (int k = 0; k < 3; ++k) wgt[k] = (double)dl_dimg[px * 4 + k] * (double)p->sfactor / (double)p->spp;
            rng_t g;
            rng_seed(&g, vro_make_random_seed(y * w + x, (int32_t)p->frame_id));
            for (uint32_t s = 0; s < p->spp; ++s) {
                vro_ray r = primary_ray(cam, &g, p->arg_order, x, y, w, h);
                pt_sample_adjoint(nodes, tris, indices, num_prims, mats, p, r, &g, wgt, grad_cd, grad_ce);
            }
        }
    }
}

/* sample_pixel_impl with jittered_blend (detail/sched_common.h:427-526) under
 * basic_sched::frame seeding (detail/basic_sched.inl:100-112). */
void vro_render_pt_lights(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                          const vro_material* mats, const vro_light* lights, size_t num_lights,
                          const vro_camera* cam, const vro_render_params* p, float* accum, float* color,
                          int y_begin, int y_end, uint64_t* ray_counts)
{
    int w = (int)p->width, h = (int)p->height;
    for (int y = y_begin; y < y_end; ++y) {
        for (int x = 0; x < w; ++x) {
            rng_t g;
            rng_seed(&g, vro_make_random_seed(y * w + x, (int32_t)p->frame_id));
            g_log_pixel = (uint32_t)(y * w + x);
            float sum[4] = {0, 0, 0, 0};
            for (uint32_t s = 0; s < p->spp; ++s) {
                vro_ray r = primary_ray(cam, &g, p->arg_order, x, y, w, h);
                float c[4];
                pt_sample(nodes, tris, indices, num_prims, mats, lights, num_lights, p, r, &g, c, ray_counts);
                for (int k = 0; k < 4; ++k) sum[k] += c[k];
            }
            for (int k = 0; k < 4; ++k) sum[k] /= (float)p->spp;
            blend_store(accum, color, (size_t)y * w + x, sum, p->sfactor, p->dfactor);
        }
    }
}

void vro_render_pt(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                   const vro_material* mats, const vro_camera* cam, const vro_render_params* p,
                   float* accum, float* color, int y_begin, int y_end, uint64_t* ray_counts)
{
    vro_render_pt_lights(nodes, tris, indices, num_prims, mats, NULL, 0, cam, p, accum, color, y_begin, y_end,
                         ray_counts);
}

/* ---- wide trees: bvh_multi_node<W>, the collapse, bvh_compressed_node<W> ---- */

/* bvh_multi_node::encode_leaf (bvh.h:242-245) */
static int64_t wide_encode_leaf(uint64_t first, uint64_t num)
{
    return (int64_t)~(num << 48 | (first & 0xFFFFFFFFFFFFull));
}

typedef struct {
    float b[6][8]; /* minx, miny, minz, maxx, maxy, maxz of each child slot */
    int64_t c[8];  /* children; INT64_MAX = unused */
} wide_tmp;

/* bvh_multi_node::get_num_children (bvh.h:272-283) */
static int wide_num_children(const wide_tmp* n, int W)
{
    for (int i = 0; i < W; ++i)
        if (n->c[i] == INT64_MAX) return i;
    return W;
}

static void wide_set_box(wide_tmp* n, int slot, const float lo[3], const float hi[3])
{
    for (int k = 0; k < 3; ++k) {
        n->b[k][slot] = lo[k];
        n->b[3 + k][slot] = hi[k];
    }
}

/* bvh_multi_node::init (bvh.h:183-240): the multi-node of binary node id */
static void wide_init(wide_tmp* m, size_t id, const vro_node* nodes, int W)
{
    const vro_node* n = &nodes[id];
    for (int i = 0; i < W; ++i) {
        m->c[i] = INT64_MAX;
        for (int k = 0; k < 3; ++k) {
            m->b[k][i] = FLT_MAX;
            m->b[3 + k][i] = -FLT_MAX;
        }
    }
    if (n->num_prims == 0) {
        for (int i = 0; i < 2; ++i) {
            const vro_node* c = &nodes[n->first + i];
            m->c[i] = (int64_t)(n->first + i);
            wide_set_box(m, i, c->lo, c->hi);
            if (c->num_prims != 0) m->c[i] = wide_encode_leaf(c->first, c->num_prims);
        }
    } else if (id == 0) {
        /* a root that is a leaf (bvh.h:221-234) */
        wide_set_box(m, 0, n->lo, n->hi);
        m->c[0] = wide_encode_leaf(n->first, n->num_prims);
    }
}

/* surface_area (math/detail/aabb.inl:201-224): 2 (sx sy + sy sz + sz sx) */
static float wide_child_area(const wide_tmp* n, int i)
{
    const float sx = n->b[3][i] - n->b[0][i], sy = n->b[4][i] - n->b[1][i], sz = n->b[5][i] - n->b[2][i];
    return 2.0f * (sx * sy + sy * sz + sz * sx);
}

/* bvh_multi_node::collapse_child (bvh.h:256-269) */
static void wide_collapse_child(wide_tmp* n, wide_tmp* child, int dest, int src)
{
    n->c[dest] = child->c[src];
    for (int k = 0; k < 6; ++k) n->b[k][dest] = child->b[k][src];
    if (src != 0) child->c[src] = INT64_MAX;
}

size_t vro_collapse_wide(const vro_node* nodes, size_t num_nodes, int W, float* out_bounds, int64_t* out_children,
                         size_t* max_stack)
{
    if (num_nodes == 0 || (W != 4 && W != 8)) return 0;
    wide_tmp* m = (wide_tmp*)malloc(num_nodes * sizeof(wide_tmp));
    for (size_t i = 0; i < num_nodes; ++i) wide_init(&m[i], i, nodes, W);
    /* depth-first from the root (collapse.h:37-118); the reference's stack
     * holds 64 entries, this one as many as the tree needs (max_stack
     * reports the deepest it got) */
    size_t cap = 64, sp = 0, deepest = 0;
    size_t* st = (size_t*)malloc(cap * sizeof(size_t));
    st[sp++] = 0;
    deepest = 1;
    while (sp > 0) {
        const size_t addr = st[--sp];
        wide_tmp* node = &m[addr];
        while (wide_num_children(node, W) < W) {
            int best = -1;
            float best_sa = 0.0f;
            const int nc = wide_num_children(node, W);
            for (int i = 0; i < nc; ++i) {
                if (node->c[i] < 0) continue;
                const wide_tmp* child = &m[node->c[i]];
                const int cc = wide_num_children(child, W);
                int inner = 0;
                for (int c = 0; c < cc; ++c)
                    if (child->c[c] > 0) ++inner;
                /* leaves are never collapsed into the root */
                if (inner == 0 && addr == 0) continue;
                if (nc - 1 + cc <= W) {
                    const float sa = wide_child_area(node, i);
                    if (sa > best_sa) {
                        best = i;
                        best_sa = sa;
                    }
                }
            }
            if (best == -1) break;
            wide_tmp* bc = &m[node->c[best]];
            wide_collapse_child(node, bc, best, 0);
            int dest = wide_num_children(node, W);
            for (int i = 1; i < wide_num_children(bc, W); ++i) wide_collapse_child(node, bc, dest++, i);
        }
        const int nc = wide_num_children(node, W);
        for (int i = 0; i < nc; ++i) {
            if (node->c[i] > 0) {
                if (sp == cap) {
                    cap *= 2;
                    st = (size_t*)realloc(st, cap * sizeof(size_t));
                }
                st[sp++] = (size_t)node->c[i];
                if (sp > deepest) deepest = sp;
            }
        }
    }
    free(st);
    if (max_stack) *max_stack = deepest;
    /* drop empty nodes, renumber the inner children (collapse.h:120-150) */
    int64_t* prefix = (int64_t*)malloc(num_nodes * sizeof(int64_t));
    prefix[0] = 0;
    for (size_t i = 1; i < num_nodes; ++i) prefix[i] = prefix[i - 1] + (m[i].c[0] == INT64_MAX ? 0 : 1);
    size_t out = 0;
    for (size_t i = 0; i < num_nodes; ++i) {
        if (m[i].c[0] == INT64_MAX) continue;
        const int nc = wide_num_children(&m[i], W);
        for (int c = 0; c < W; ++c) {
            int64_t e = m[i].c[c];
            if (c < nc && e >= 0) e = prefix[e];
            out_children[out * W + c] = e;
            for (int k = 0; k < 6; ++k) out_bounds[(out * 6 + k) * W + c] = m[i].b[k][c];
        }
        ++out;
    }
    free(prefix);
    free(m);
    return out;
}

/* float -> int as x86's truncating conversion gives it (cvttss2si): values
 * outside int32 and NaN give INT32_MIN; the reference stores such results
 * into char / unsigned char, whose low byte this keeps */
static int32_t cvt_trunc_i32(float x)
{
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}

/* bvh_compressed_node<W>::init (bvh.h:352-400) in the reference layout:
 * child_bounds (6 x W bytes), children (W x {int id; short num_prims},
 * packed), origin (3 floats), e (3 chars); W = 4: 64 bytes, W = 8: 112 */
void vro_compress_wide(const float* bounds, const int64_t* children, size_t num, int W, uint8_t* out)
{
    const size_t stride = W == 4 ? 64 : 112;
    for (size_t n = 0; n < num; ++n) {
        const float* b = bounds + n * 6 * (size_t)W;
        const int64_t* c = children + n * (size_t)W;
        uint8_t* o = out + n * stride;
        memset(o, 0, stride);
        /* get_bounds: the union of all W child boxes (aabb::insert) */
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int i = 0; i < W; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = tmin_f(lo[k], b[k * W + i]);
                hi[k] = tmax_f(hi[k], b[(3 + k) * W + i]);
            }
        signed char e[3];
        for (int k = 0; k < 3; ++k)
            e[k] = (signed char)cvt_trunc_i32(ceilf(log2f((hi[k] - lo[k]) / (powf(2.0f, 8.0f) - 1.0f))));
        for (int i = 0; i < W; ++i) {
            for (int k = 0; k < 3; ++k) {
                o[k * W + i] = (uint8_t)cvt_trunc_i32(floorf((b[k * W + i] - lo[k]) / powf(2.0f, (float)e[k])));
                o[(3 + k) * W + i] = (uint8_t)cvt_trunc_i32(ceilf((b[(3 + k) * W + i] - lo[k]) / powf(2.0f, (float)e[k])));
            }
            int32_t id;
            int16_t np;
            if (c[i] < 0) {
                const uint64_t a = ~(uint64_t)c[i];
                id = (int32_t)(a & 0xFFFFFFFFFFFFull);
                np = (int16_t)(a >> 48);
            } else {
                id = (int32_t)c[i];
                np = 0;
            }
            memcpy(o + 6 * W + 6 * i, &id, 4);
            memcpy(o + 6 * W + 6 * i + 4, &np, 2);
        }
        memcpy(o + 12 * W, lo, 12);
        memcpy(o + 12 * W + 12, e, 3);
    }
}
