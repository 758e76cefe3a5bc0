// render.hip -- the path-tracing bounce loop and the AO kernel under the
// jittered-blend pixel sampler.
//
// pathtracing::kernel (detail/pathtracing.inl:43-233) runs on a persistent
// wavefront: each lane owns one pixel at a time and runs its spp samples in
// sequence (the reference draws all samples of a pixel from one generator,
// detail/sched_common.h:427-526, so a pixel cannot be split across lanes
// without changing its random stream).  Each lane holds one ray query in
// flight.  The wave advances its queries an iteration at a time (a node
// visit or a triangle test, whichever more lanes need) until 32 lanes have a
// finished query, then shades those lanes together: each starts its next
// bounce, its next sample, or (pixel done) takes the next pixel, with one
// atomic per wave.  Neither a long path nor a long query keeps the other
// lanes of the wave idle.
#include "vr_internal.h"

#include <algorithm>
#include <cstring>

namespace vr {

namespace {

constexpr int BLOCK = 256;

__device__ inline f3 ldtri3(float4 q) { return mk3(q.x, q.y, q.z); }

__device__ inline uint64_t lane_mask_lt_count(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Everything one scene contributes to a frame (the kernel_params, camera,
// sampler and render target of one reference frame() call).
struct scene_desc {
    bvh_view b;
    const vr_material* mats;
    uint32_t num_mats;
    uint32_t grad_base;  // adjoint: first gradient slot (material) of this scene
    vr_kernel_params kp;
    vr_camera cam;
    vr_sched_params sp;
    float* accum;
    float* color;
    const float* dl;  // adjoint: d loss / d image, RGBA per pixel
    const vr_area_light* lights;  // NEE kernels: the light list
    uint32_t num_lights;
};

struct pt_args {
    scene_desc single;           // one scene (BATCH = false)
    const scene_desc* scenes;    // BATCH: num_scenes descriptors
    const uint32_t* prefix;      // BATCH: first global pixel of every scene, num_scenes + 1 entries
    uint32_t num_scenes;
    uint32_t total_pixels;
    const uint32_t* frame_offset;  // optional device frame counter added to every frame_id
    unsigned long long* ray_counts;
    uint32_t* work;
    uint32_t arg_order;
    uint32_t tile;        // pixel queue order (tile x tile squares; <= 1: scanlines)
    uint32_t grad_slots;  // adjoint: materials covered by the LDS partials
    double* grad;         // adjoint: 6 per material slot (d/d cd rgb, d/d ce rgb), accumulated
    const vr_material* rr_mats;  // PT_FROZEN_RR: materials whose throughput sets the roulette
};

__device__ inline void write_pixel(const scene_desc& S, uint32_t px, float4 c)
{
    // sample average, pixel_access::blend into accum, accum -> color
    // (detail/sched_common.h:466-509, detail/pixel_access.h:1040-1131)
    const float spp = (float)S.sp.spp;
    c.x = c.x / spp; c.y = c.y / spp; c.z = c.z / spp; c.w = c.w / spp;
    float4* acc = reinterpret_cast<float4*>(S.accum) + px;
    const float4 old = *acc;
    const float sf = S.sp.sfactor, df = S.sp.dfactor;
    const float4 r = make_float4(c.x * sf + old.x * df, c.y * sf + old.y * df, c.z * sf + old.z * df,
                                 c.w * sf + old.w * df);
    *acc = r;
    reinterpret_cast<float4*>(S.color)[px] = r;
}

// kernel modes; PT_NEE (a light list: next-event estimation with MIS) is
// combined with PT_RENDER or PT_COUNT
enum pt_mode { PT_RENDER = 0, PT_COUNT = 1, PT_ADJOINT = 2, PT_FROZEN_RR = 4, PT_NEE = 8 };
constexpr int ADJ_MAX_VERTICES = (int)ADJ_MAX_BOUNCES;

// State of one pixel's path (the reference keeps it in the locals of
// pathtracing::kernel and sample_pixel, detail/pathtracing.inl:43-233,
// detail/sched_common.h:427-526).
template <bool ADJ>
struct path_t {
    rng_t g;
    uint32_t s;       // sample
    uint32_t bounce;
    f3 tp, tp0, I;    // throughput, throughput under A.rr_mats (frozen roulette), radiance
    bool rhit;        // the primary ray hit
    float4 sum;       // samples so far
    f3 wgt;           // adjoint: d loss / d sample
    uint32_t mids[ADJ ? ADJ_MAX_VERTICES : 1];
    uint32_t nmid;
    // adjoint, per channel: product of the nonzero vertex factors over the
    // roulette probabilities, count of zero factors, material and
    // d factor / d cd of the first zero-factor vertex
    f3 nzp, zg;
    uint32_t zc[3], zm[3];
};

// Next-event estimation state of a lane: the query in flight is the shadow
// ray of the last vertex (its contribution `add` counts when nothing is hit),
// and the path goes on with (co, cd) afterwards unless it ended there.
template <bool NEE>
struct nee_t {
    static constexpr bool shadow = false;
};
template <>
struct nee_t<true> {
    bool shadow, cont;
    f3 add, co, cd;
};

// power_heuristic (sampling.h:73-79)
__device__ inline float power_heuristic(float a, float b) { return (a * a) / (a * a + b * b); }

template <bool ADJ>
__device__ __forceinline__ void start_sample(path_t<ADJ>& P)
{
    P.tp = mk3(1, 1, 1);
    P.tp0 = mk3(1, 1, 1);
    P.I = mk3(0, 0, 0);
    P.bounce = 0;
    P.rhit = false;
    if (ADJ) {
        P.nmid = 0;
        P.nzp = mk3(1, 1, 1);
        P.zg = mk3(0, 0, 0);
        P.zc[0] = P.zc[1] = P.zc[2] = 0;
    }
}

// Adjoint: d I_c / d cd_{m,c} for the materials of the path's matte
// vertices, given the ending term's emission E (ambient or ce * ls) and its
// contribution e = T * E.  With f_{j,c} = cd_{m_j,c} * g_j (g_j = kd / pi *
// cos / pdf) the derivative is E_c * sum_{j : m_j = m} g_j * prod_{k != j}
// f_{k,c}: e_c * (vertices with m) / cd_{m,c} when no factor of channel c is
// zero; only the zero vertex's material, g * (other factors) * E_c, when
// exactly one is (a zero albedo channel gets its finite derivative, not
// 0 / 0); nothing when two or more are.
template <bool ADJ>
__device__ __forceinline__ void adj_add_cd(const scene_desc& S, const path_t<ADJ>& P, f3 E, f3 e, float* s_grad)
{
    const float w[3] = {P.wgt.x, P.wgt.y, P.wgt.z};
    const float ec[3] = {e.x, e.y, e.z};
    const float Ec[3] = {E.x, E.y, E.z};
    const float nz[3] = {P.nzp.x, P.nzp.y, P.nzp.z};
    const float zg[3] = {P.zg.x, P.zg.y, P.zg.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (P.zc[c] == 0) {
            for (uint32_t j = 0; j < P.nmid; ++j) {
                const vr_material* mj = S.mats + P.mids[j];
                atomicAdd(&s_grad[6 * (S.grad_base + P.mids[j]) + c], w[c] * ec[c] / mj->cd[c]);
            }
        } else if (P.zc[c] == 1) {
            atomicAdd(&s_grad[6 * (S.grad_base + P.zm[c]) + c], w[c] * ((Ec[c] * zg[c]) * nz[c]));
        }
    }
}

// One bounce of pathtracing::kernel for the query (o, d) with result hr:
// emission / ambient contributions, the matte sample, Russian roulette.
// Returns true when the path ended; otherwise (no, nd) is the next ray.
// Adjoint (no counterpart in the reference, which has no differentiation):
// a path's estimate is I = T * E, E the ending emitter's ce * ls or the
// ambient term on escape, T the product over its matte vertices of
// cd kd / pi * cos / pdf divided by the roulette probabilities, which are
// held constant (unbiased for the gradient of the expected image):
//   d I_c / d cd_{m,c} = I_c * (vertices with material m) / cd_{m,c}
//     (leave-one-out products when a factor is zero: adj_add_cd)
//   d I_c / d ce_{e,c} = T_c * ls_e.
// Per-block FP32 partials in LDS; FP64 only in the final reduction.
//
// NEE (a light list, pathtracing.inl:116-196): an emitter hit at bounce > 0
// is weighted by power_heuristic(1, pdf_light / num_lights) (the emissive
// material's sample() reports pdf 1), pdf_light the solid-angle density of
// the hit triangle; every hit draws a light (u), a point on it (u1, u2) after
// the BRDF sample and before the roulette, and a vertex that still scatters
// with both cosines positive gets a shadow ray [pos + L eps, tmin eps,
// tmax dist - eps] (so, sd, smax; Q.shadow) whose contribution Q.add
// = power_heuristic(pdf_light / num_lights, pdf_brdf max(throughput))
//   * throughput * f * (cos / pdf_light) * num_lights
// counts if it reaches the light.  The reference traces the masked shadow
// rays too; they change no output and are not traced here.
template <int MODE>
__device__ __forceinline__ bool shade(const scene_desc& S, const pt_args& A,
                                      path_t<(MODE & ~PT_NEE) == PT_ADJOINT>& P, nee_t<(MODE & PT_NEE) != 0>& Q,
                                      const hit_t& hr, f3 o, f3 d, f3& no, f3& nd, f3& so, f3& sd, float& smax,
                                      float* s_grad)
{
    constexpr bool ADJ = (MODE & ~PT_NEE) == PT_ADJOINT;
    constexpr bool FROZEN = (MODE & ~PT_NEE) == PT_FROZEN_RR;
    constexpr bool NEE = (MODE & PT_NEE) != 0;
    constexpr float INV_PI = 3.18309886183790691216444201928e-01f;
    if (hr.tri < 0) {
        // escaped: ambient light times throughput (pathtracing.inl:69-77)
        const f3 amb = mk3(S.kp.ambient[0], S.kp.ambient[1], S.kp.ambient[2]);
        const f3 e = mul3(amb, P.tp);
        P.I = add3(P.I, e);
        if (ADJ) adj_add_cd(S, P, amb, e, s_grad);
        return true;
    }
    if (P.bounce == 0) P.rhit = true;
    const f3 view = neg3(d);
    const f3 pos = add3(o, muls3(d, hr.t));
    const float4* tq = reinterpret_cast<const float4*>(S.b.tris + hr.tri);
    const float4 tb = tq[1], tcq = tq[2];
    const f3 gn = normalize3(cross3(ldtri3(tb), ldtri3(tcq)));
    uint32_t mid = __float_as_uint(tb.w);
    mid = mid < S.num_mats ? mid : S.num_mats - 1u;
    const vr_material* m = S.mats + mid;
    f3 refl = mk3(0.0f, 0.0f, 0.0f);
    float pdf;
    f3 src, src0 = mk3(0.0f, 0.0f, 0.0f);
    const bool emission = m->type == 1;
    if (emission) {
        // emissive::sample: tex_color * ce * ls, pdf 1 (emissive.inl:31-40)
        const float ls = m->ls;
        src = mk3((1.0f * m->ce[0]) * ls, (1.0f * m->ce[1]) * ls, (1.0f * m->ce[2]) * ls);
        pdf = 1.0f;
    } else {
        // matte::sample -> lambertian::sample_f (matte.inl:43-56, brdf.h:50-84)
        const f3 n = faceforward3(gn, view, gn);
        f3 u, v;
        onb(n, u, v);
        float a1, a2;
        rng_pair(P.g, A.arg_order, a1, a2);
        const f3 sp = cosine_hemisphere(a1, a2);
        refl = normalize3(add3(add3(smul3(sp.x, u), smul3(sp.y, v)), smul3(sp.z, n)));
        pdf = dot3(n, refl) * 3.18309886183790691216444201928e-01f;
        const float kd = m->kd;
        src = mk3(1.0f * ((m->cd[0] * kd) * 3.18309886183790691216444201928e-01f),
                  1.0f * ((m->cd[1] * kd) * 3.18309886183790691216444201928e-01f),
                  1.0f * ((m->cd[2] * kd) * 3.18309886183790691216444201928e-01f));
        if (FROZEN) {
            const vr_material* m0 = A.rr_mats + mid;
            const float kd0 = m0->kd;
            src0 = mk3(1.0f * ((m0->cd[0] * kd0) * 3.18309886183790691216444201928e-01f),
                       1.0f * ((m0->cd[1] * kd0) * 3.18309886183790691216444201928e-01f),
                       1.0f * ((m0->cd[2] * kd0) * 3.18309886183790691216444201928e-01f));
        }
    }
    const bool zero_pdf = pdf <= 0.0f;
    bool active = true;
    if (emission) {
        // MIS weight 1 without a light list (pathtracing.inl:116-146)
        float wgt = 1.0f;
        if (NEE && S.num_lights > 0 && P.bounce > 0) {
            const f3 cc = cross3(ldtri3(tb), ldtri3(tcq));
            const float area = 0.5f * sqrtf(dot3(cc, cc));
            const f3 dl = sub3(pos, o);
            const float ld = sqrtf(dot3(dl, dl));
            const f3 L = normalize3(dl);
            const float ldotln = fabsf(dot3(neg3(L), gn));
            const float solid_angle = (ldotln * area) / (ld * ld);
            wgt = power_heuristic(pdf, (1.0f / solid_angle) / (float)S.num_lights);
        }
        const f3 e = mul3(mk3(wgt * P.tp.x, wgt * P.tp.y, wgt * P.tp.z), src);
        P.I = add3(P.I, e);
        active = false;
        if (ADJ) {
            const float ls = m->ls;
            const uint32_t es = 6 * (S.grad_base + mid);
            atomicAdd(&s_grad[es + 3], P.wgt.x * P.tp.x * ls);
            atomicAdd(&s_grad[es + 4], P.wgt.y * P.tp.y * ls);
            atomicAdd(&s_grad[es + 5], P.wgt.z * P.tp.z * ls);
            adj_add_cd(S, P, src, e, s_grad);
        }
    } else if (ADJ && !zero_pdf && P.nmid < ADJ_MAX_VERTICES) {
        P.mids[P.nmid++] = mid;
    }
    if (zero_pdf) active = false;
    const f3 n = faceforward3(gn, view, gn);
    if constexpr (NEE) {
      if (S.num_lights > 0) {
        // sample_random_light (sampling.h:194-210): light int(u * num_lights);
        // area_light::sample (detail/area_light.inl:31-56) with sample_surface
        // (math/detail/triangle.inl:69-82)
        const float ul = rng_next(P.g);
        uint32_t li = (uint32_t)(int)(ul * (float)S.num_lights);
        li = li < S.num_lights ? li : S.num_lights - 1u;
        const vr_area_light* lt = S.lights + li;
        const float u1 = rng_next(P.g);
        const float u2 = rng_next(P.g);
        const f3 lv1 = mk3(lt->geometry.v1[0], lt->geometry.v1[1], lt->geometry.v1[2]);
        const f3 le1 = mk3(lt->geometry.e1[0], lt->geometry.e1[1], lt->geometry.e1[2]);
        const f3 le2 = mk3(lt->geometry.e2[0], lt->geometry.e2[1], lt->geometry.e2[2]);
        const f3 lv2 = add3(lv1, le1), lv3 = add3(lv1, le2);
        const float s1 = sqrtf(u1);
        const f3 lp = add3(add3(muls3(lv1, 1.0f - s1), muls3(muls3(lv2, s1), 1.0f - u2)), muls3(muls3(lv3, s1), u2));
        const f3 ldir = sub3(lp, pos);
        const float ldist = sqrtf(dot3(ldir, ldir));
        const f3 lc = cross3(le1, le2);
        const f3 lnorm = normalize3(lc);
        const float larea = 0.5f * sqrtf(dot3(lc, lc));
        const float lsolid = (fabsf(dot3(neg3(normalize3(ldir)), lnorm)) * larea) / (ldist * ldist);
        const float lpdf = 1.0f / lsolid;
        const f3 L = normalize3(ldir);
        const f3 ln = faceforward3(lnorm, neg3(L), lnorm);
        const float ldotn = dot3(L, n);
        const float ldotln = fabsf(dot3(neg3(L), ln));
        if (active && ldotn > 0.0f && ldotln > 0.0f) {
            // matte::pdf and matte::shade (matte.inl:28-71): pdf n.L / pi,
            // shade ((f pi) E) max(0, n.L), then * 1/pi / (L.n)
            float bpdf = dot3(n, L) * INV_PI;
            float prob = P.tp.x;
            if (P.tp.y > prob) prob = P.tp.y;
            if (P.tp.z > prob) prob = P.tp.z;
            bpdf *= prob;
            const float ndotl = tmax_f(0.0f, dot3(n, L));
            const float wgt = power_heuristic(lpdf / (float)S.num_lights, bpdf);
            const float nl = (float)S.num_lights;
            const float kd = m->kd;
            const float el[3] = {lt->cl[0] * lt->kl, lt->cl[1] * lt->kl, lt->cl[2] * lt->kl};
            const float tpc[3] = {P.tp.x, P.tp.y, P.tp.z};
            float add[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float f = 1.0f * ((m->cd[k] * kd) * INV_PI);
                const float sh = ((f * 3.14159265358979323846264338328e+00f) * el[k]) * ndotl;
                const float sc = (sh * INV_PI) / ldotn;
                add[k] = (((wgt * tpc[k]) * sc) * (ldotn / lpdf)) * nl;
            }
            Q.add = mk3(add[0], add[1], add[2]);
            Q.shadow = true;
            so = add3(pos, muls3(L, S.kp.epsilon));
            sd = L;
            smax = ldist - S.kp.epsilon;
        }
      }
    }
    const float q = dot3(n, refl) / pdf;
    P.tp = mk3(P.tp.x * (src.x * q), P.tp.y * (src.y * q), P.tp.z * (src.z * q));
    if (zero_pdf) P.tp = mk3(0.0f, 0.0f, 0.0f);
    if (ADJ && !emission && !zero_pdf) {
        // the vertex's factors f_c = src_c * q: zero ones are counted, the
        // others multiplied into nzp
        const float f[3] = {src.x * q, src.y * q, src.z * q};
        const float g = (m->kd * INV_PI) * q;
        float nz[3] = {P.nzp.x, P.nzp.y, P.nzp.z};
        float zg[3] = {P.zg.x, P.zg.y, P.zg.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (f[c] == 0.0f) {
                if (P.zc[c] == 0) {
                    P.zm[c] = mid;
                    zg[c] = g;
                }
                P.zc[c] = P.zc[c] < 2u ? P.zc[c] + 1u : 2u;
            } else {
                nz[c] = nz[c] * f[c];
            }
        }
        P.nzp = mk3(nz[0], nz[1], nz[2]);
        P.zg = mk3(zg[0], zg[1], zg[2]);
    }
    if (FROZEN) {
        P.tp0 = emission ? P.tp : mk3(P.tp0.x * (src0.x * q), P.tp0.y * (src0.y * q), P.tp0.z * (src0.z * q));
        if (zero_pdf) P.tp0 = mk3(0.0f, 0.0f, 0.0f);
    }
    if (P.bounce >= 2) {
        // Russian roulette (pathtracing.inl:201-213)
        const f3 rt = FROZEN ? P.tp0 : P.tp;
        float prob = rt.x;
        if (rt.y > prob) prob = rt.y;
        if (rt.z > prob) prob = rt.z;
        const bool terminate = rng_next(P.g) > prob;
        if (terminate) active = false;
        P.tp = mk3(P.tp.x / prob, P.tp.y / prob, P.tp.z / prob);
        if (FROZEN) P.tp0 = mk3(P.tp0.x / prob, P.tp0.y / prob, P.tp0.z / prob);
        if (ADJ) P.nzp = mk3(P.nzp.x / prob, P.nzp.y / prob, P.nzp.z / prob);
    }
    if (!active) return true;
    no = add3(pos, muls3(refl, S.kp.epsilon));
    nd = refl;
    ++P.bounce;
    return P.bounce >= S.kp.num_bounces;
}

// The pixel's sum of samples: every sample adds alpha 1 (exact in float up
// to 2^24 samples), so the alpha sum is the sample count.
__device__ __forceinline__ float4 sample_sum(const scene_desc& S, float4 sum)
{
    sum.w = (float)S.sp.spp;
    return sum;
}

// A path ended: add its rgb sample (the background on a primary miss);
// returns true when the pixel's last sample is done.
template <bool ADJ>
__device__ __forceinline__ bool end_sample(const scene_desc& S, path_t<ADJ>& P)
{
    if (!ADJ) {
        if (P.rhit) {
            P.sum.x += P.I.x; P.sum.y += P.I.y; P.sum.z += P.I.z;
        } else {
            P.sum.x += S.kp.bg[0]; P.sum.y += S.kp.bg[1]; P.sum.z += S.kp.bg[2];
        }
    }
    ++P.s;
    return P.s >= S.sp.spp;
}

// One kernel for the frame (PT_RENDER), the frame with traversal counters
// (PT_COUNT), the path-replay adjoint (PT_ADJOINT) and the frozen-roulette
// frame (PT_FROZEN_RR): all draw the same random numbers in the same order,
// so the adjoint replays exactly the paths of the forward frame with the
// same seeds, recomputing every hit.  PT_FROZEN_RR takes the roulette
// probabilities from the throughput under A.rr_mats instead of the shading
// materials: with rr_mats = mats it is the frame itself, and for mats near
// rr_mats it is the estimator the adjoint differentiates (central
// differences of it at common random numbers keep every roulette decision).
//
// BATCH: many scenes in one persistent launch.  The global pixel index runs
// over all scenes back to back; a lane finds its scene by binary search over
// the pixel prefix and reads that scene's descriptor.
//
// BLK (threads per block) and LDSD (LDS stack entries per lane) are the
// autotuned launch shape; A.tile orders pixels in tile x tile squares.
// waves per SIMD the register allocation must leave room for: 5 (96 VGPRs)
// for the plain kernels; 4 (128 VGPRs) with NEE, whose shadow-ray state would
// otherwise spill (measured: NEE frame 58.9 -> 54.7 ms on the large scene,
// the plain frame 36.9 -> 38.5 ms at 4)
#ifndef VR_PT_WAVES_PER_EU
#define VR_PT_WAVES_PER_EU 5
#endif
#ifndef VR_PT_NEE_WAVES_PER_EU
#define VR_PT_NEE_WAVES_PER_EU 4
#endif
#define VR_PT_OCCUPANCY \
    __attribute__((amdgpu_waves_per_eu((MODE & PT_NEE) ? VR_PT_NEE_WAVES_PER_EU : VR_PT_WAVES_PER_EU)))

// lanes that must wait for shading before the wave stops advancing its queries
#ifndef VR_PT_REGEN_LANES
#define VR_PT_REGEN_LANES 32
#endif

// Hand-over of samples inside a workgroup once the pixel queue is empty.  A
// pixel's samples run in sequence on one lane (one random stream per pixel),
// so at the end of the frame lanes idle while their waves finish the last
// pixels.  A wave (other than the workgroup's first) with fewer than
// VR_PT_DONATE_LANES lanes holding a pixel then hands each of its pixels over
// at the pixel's next sample boundary: pixel, next sample index, random
// state and sum so far go to the workgroup's pool in LDS, and the wave ends
// once all its lanes are idle.  Idle lanes of the other waves take them and
// go on with the next sample exactly as the giving lane would have (same
// random numbers, same sum), so the image does not change; the work gathers
// in fewer, fuller waves.  The first wave never gives and ends last, after
// every other wave has ended and the pool is empty, so no sample is left
// behind.  0 disables it.
#ifndef VR_PT_DONATE_LANES
#define VR_PT_DONATE_LANES 24
#endif
struct handed_sample {
    uint4 a;     // pixel, next sample, random state, scene
    float4 sum;  // the pixel's samples so far
};

template <int MODE, bool BATCH, int BLK, int LDSD, int NODE = NODE_QUAD>
__global__ __launch_bounds__(BLK) VR_PT_OCCUPANCY void k_pathtrace(pt_args A)
{
    constexpr bool COUNT = (MODE & ~PT_NEE) == PT_COUNT;
    constexpr bool ADJ = (MODE & ~PT_NEE) == PT_ADJOINT;
    constexpr bool NEE = (MODE & PT_NEE) != 0;
    // the adjoint keeps the plain end of frame (its gradient partials are
    // summed per workgroup, which a hand-over would reorder), and so does the
    // NEE frame (measured: no gain, its shadow-ray state leaves no registers)
    constexpr bool DON = !ADJ && !NEE && VR_PT_DONATE_LANES > 0 && BLK > 64;
    // each lane of the waves other than the first gives at most once (a
    // giving wave takes nothing afterwards)
    constexpr uint32_t POOL = DON ? BLK - 64 : 1;
    __shared__ uint32_t stack[LDSD * BLK];
    __shared__ handed_sample s_pool[POOL];
    __shared__ uint32_t s_pool_ready[POOL];
    __shared__ uint32_t s_given, s_taken, s_waves;
    extern __shared__ float s_grad[];  // ADJ: 6 * num_mats
    uint32_t* st = stack + threadIdx.x;
    const uint32_t npix = BATCH ? A.total_pixels : padded_pixels(A.single.sp.width, A.single.sp.height, A.tile);
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const uint32_t frame_add = A.frame_offset ? *A.frame_offset : 0u;
    if (ADJ) {
        for (uint32_t i = threadIdx.x; i < 6 * A.grad_slots; i += BLK) s_grad[i] = 0.0f;
        __syncthreads();
    }
    const uint32_t wid = threadIdx.x / 64u;
    if (DON) {
        for (uint32_t i = threadIdx.x; i < POOL; i += BLK) s_pool_ready[i] = 0u;
        if (threadIdx.x == 0) {
            s_given = 0u;
            s_taken = 0u;
            s_waves = BLK / 64u;
        }
        __syncthreads();
    }
    // wave-uniform: the pixel queue gave out its last pixel to this wave's
    // view; this wave hands its pixels over
    bool exhausted = false, giver = false;
    uint32_t idle_rounds = 0;
    uint32_t sid = 0;  // scene of this lane's pixel (BATCH)
#define SCN (BATCH ? A.scenes[sid] : A.single)

    int32_t px = -1;
    bool done = false;
    path_t<ADJ> P;
    nee_t<NEE> Q{};
    P.g = rng_t{1u};
    P.s = 0;
    P.sum = make_float4(0, 0, 0, 0);
    P.wgt = mk3(0, 0, 0);
    start_sample(P);
    // queries started by the wave (wave-uniform: counted with ballots where
    // queries start, so no lane holds a counter)
    unsigned long long n_primary = 0, n_secondary = 0;
    trace_counters tc{0, 0};
    unsigned long long n_nodes = 0, n_tris = 0;
    // the lane's query: in flight (tracing) or finished and waiting to be
    // shaded (pending)
    trav_state T;
    uint32_t spill[max_stack_of<NODE>() - LDSD];
    bool tracing = false, pending = false;
    bvh_view bv{};  // BATCH: the BVH of the lane's scene, kept in registers

    for (;;) {
        const bool was_tracing = tracing;
        bool give = false;  // DON: this lane hands its pixel's next sample over
        // 1. shade the finished queries; each such lane starts its next query
        //    or finishes its sample or pixel
        if (pending) {
            pending = false;
            const scene_desc& S = SCN;
            f3 no, nd;  // next ray
            if (COUNT) {
                n_nodes += tc.nodes;
                n_tris += tc.tris;
                tc.nodes = 0;
                tc.tris = 0;
            }
            bool ended = false, shadow_now = false;
            if (NEE && Q.shadow) {
                // the vertex's shadow ray is done: the light counts if it is clear
                if constexpr (NEE) {
                    Q.shadow = false;
                    if (T.r.tri < 0) P.I = add3(P.I, Q.add);
                    ended = !Q.cont;
                    no = Q.co;
                    nd = Q.cd;
                }
            } else {
                f3 so, sd;
                float smax = 0.0f;
                ended = shade<MODE>(S, A, P, Q, T.r, T.ori, T.dir, no, nd, so, sd, smax, s_grad);
                if constexpr (NEE) {
                    if (Q.shadow) {
                        Q.cont = !ended;
                        Q.co = no;
                        Q.cd = nd;
                        trav_begin(T, S.b, so, sd, S.kp.epsilon, smax);
                        tracing = true;
                        shadow_now = true;
                    }
                }
            }
            if (!shadow_now) {
                if (ended) {
                    if (end_sample(S, P)) {
                        if (!ADJ) write_pixel(S, (uint32_t)px, sample_sum(S, P.sum));
                        px = -1;
                    } else if (DON && giver) {
                        give = true;
                    } else {
                        const uint32_t w = S.sp.width, h = S.sp.height;
                        const int x = (int)((uint32_t)px % w), y = (int)((uint32_t)px / w);
                        primary_ray(S.cam, P.g, A.arg_order, x, y, (int)w, (int)h, no, nd);
                        start_sample(P);
                    }
                }
                if (px >= 0 && !give) {
                    trav_begin(T, S.b, no, nd, 0.0f, FLT_MAX);
                    tracing = true;
                }
            }
        }

        if (DON) {
            // the samples handed over in this round, one LDS atomic per wave
            const uint64_t gm = __ballot(give);
            if (gm) {
                const int leader = __ffsll((unsigned long long)gm) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(&s_given, (uint32_t)__popcll(gm));
                base = __shfl(base, leader);
                if (give) {
                    const uint32_t i = min(base + (uint32_t)lane_mask_lt_count(gm), POOL - 1u);
                    s_pool[i].a = make_uint4((uint32_t)px, P.s, P.g.x, sid);
                    s_pool[i].sum = P.sum;
                    __hip_atomic_store(&s_pool_ready[i], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    px = -1;
                }
            }
        }

        // 2. lanes without a pixel take the next ones, one atomic per wave
        const bool need = !done && px < 0;
        const uint64_t needm = DON && exhausted ? 0ull : __ballot(need);
        bool dry = false;  // DON: the queue had no pixel left for this lane
        if (needm) {
            const int leader = __ffsll((unsigned long long)needm) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(A.work, (uint32_t)__popcll(needm));
            base = __shfl(base, leader);
            if (need) {
                const uint32_t my = base + (uint32_t)lane_mask_lt_count(needm);
                if (my < npix) {
                    uint32_t lpx = my;
                    if (BATCH) {
                        uint32_t lo = 0, hi = A.num_scenes;
                        while (hi - lo > 1) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (A.prefix[mid] <= my) lo = mid;
                            else hi = mid;
                        }
                        sid = lo;
                        lpx = my - A.prefix[lo];
                        bv = A.scenes[sid].b;
                    }
                    const scene_desc& S = SCN;
                    const uint32_t w = S.sp.width, h = S.sp.height;
                    int x, y;
                    // pixels of the padding of a partial tile are skipped (px stays -1)
                    if (tile_pixel(lpx, w, h, A.tile, x, y)) {
                        lpx = (uint32_t)y * w + (uint32_t)x;
                        px = (int32_t)lpx;
                        rng_seed(P.g, make_random_seed(y * (int)w + x, (int32_t)(S.sp.frame_id + frame_add)));
                        P.s = 0;
                        P.sum = make_float4(0, 0, 0, 0);
                        f3 o, d;
                        primary_ray(S.cam, P.g, A.arg_order, x, y, (int)w, (int)h, o, d);
                        start_sample(P);
                        if (ADJ) {
                            // the image is accum = mean(samples) * sfactor + ...
                            const float4 dl = reinterpret_cast<const float4*>(S.dl)[lpx];
                            const float k = S.sp.sfactor / (float)S.sp.spp;
                            P.wgt = mk3(dl.x * k, dl.y * k, dl.z * k);
                        }
                        trav_begin(T, S.b, o, d, 0.0f, FLT_MAX);
                        tracing = true;
                    }
                } else if (DON) {
                    dry = true;
                } else {
                    done = true;
                }
            }
        }
        if (DON) {
            if (__ballot(dry)) exhausted = true;
            if (exhausted) {
                const uint32_t live = (uint32_t)__popcll(__ballot(px >= 0));
                if (wid != 0 && live < (uint32_t)VR_PT_DONATE_LANES) giver = true;
                if (giver) {
                    // every pixel handed over or finished: this wave ends
                    // (its hand-overs are in the pool before the count drops)
                    if (live == 0) {
                        if (lane == 0) __hip_atomic_fetch_add(&s_waves, -1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        break;
                    }
                } else {
                    // idle lanes take handed-over samples: the leader reserves
                    // up to one per idle lane of those given so far
                    const uint64_t im = __ballot(px < 0);
                    uint32_t got = 0, base = 0;
                    if (im) {
                        const int leader = __ffsll((unsigned long long)im) - 1;
                        if ((int)lane == leader) {
                            for (int tries = 0; tries < 16; ++tries) {
                                uint32_t p = __hip_atomic_load(&s_taken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                const uint32_t q = __hip_atomic_load(&s_given, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                if (p >= q) break;
                                const uint32_t k = min(q - p, (uint32_t)__popcll(im));
                                if (__hip_atomic_compare_exchange_strong(&s_taken, &p, p + k, __ATOMIC_RELAXED,
                                                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                                    base = p;
                                    got = k;
                                    break;
                                }
                            }
                        }
                        got = __shfl(got, leader);
                        base = __shfl(base, leader);
                        const uint32_t r = (uint32_t)lane_mask_lt_count(im);
                        if (px < 0 && r < got) {
                            const uint32_t i = min(base + r, POOL - 1u);
                            // the giver writes a record right after reserving it
                            for (uint32_t wt = 0; wt < (1u << 20); ++wt) {
                                if (__hip_atomic_load(&s_pool_ready[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                                __builtin_amdgcn_s_sleep(1);
                            }
                            const uint4 a = s_pool[i].a;
                            P.sum = s_pool[i].sum;
                            px = (int32_t)a.x;
                            P.s = a.y;
                            P.g.x = a.z;
                            if (BATCH) {
                                sid = a.w;
                                bv = A.scenes[sid].b;
                            }
                            const scene_desc& S = SCN;
                            const uint32_t w = S.sp.width, h = S.sp.height;
                            const int x = (int)(a.x % w), y = (int)(a.x / w);
                            f3 o, d;
                            primary_ray(S.cam, P.g, A.arg_order, x, y, (int)w, (int)h, o, d);
                            start_sample(P);
                            trav_begin(T, S.b, o, d, 0.0f, FLT_MAX);
                            tracing = true;
                        }
                    }
                    if (__ballot(px >= 0) == 0) {
                        // only the first wave gets here with nothing to do
                        // (the others hand over once they run low): it ends
                        // when every other wave has ended and nothing is
                        // left in the pool, and waits otherwise
                        const uint32_t waves = __hip_atomic_load(&s_waves, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        const uint32_t q = __hip_atomic_load(&s_given, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        const uint32_t p = __hip_atomic_load(&s_taken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if ((waves == 1u && p >= q) || ++idle_rounds > (1u << 22)) break;
                        __builtin_amdgcn_s_sleep(32);
                        continue;
                    }
                }
            }
        } else if (__ballot(!done) == 0) {
            break;
        }
        if (A.ray_counts) {
            // every lane that holds a query now and did not before this round
            // started one: primary if its path is at bounce 0 (a shadow ray
            // is secondary)
            const uint64_t started = __ballot(tracing && !was_tracing);
            const uint64_t prim = __ballot(tracing && !was_tracing && P.bounce == 0 && !(NEE && Q.shadow));
            n_primary += (unsigned long long)__popcll(prim);
            n_secondary += (unsigned long long)__popcll(started & ~prim);
        }

        // 3. advance the queries in flight until VR_PT_REGEN_LANES lanes wait
        //    to be shaded (or none is in flight): a lane whose query ends
        //    early does not idle until the wave's longest query is done
        for (;;) {
            if (__ballot(tracing) == 0) break;
            if (__popcll(__ballot(pending)) >= VR_PT_REGEN_LANES) break;
            if (tracing &&
                trav_iter<false, COUNT, LDSD, BLK, TRAV_VOTED, default_isect, NODE>(T, BATCH ? bv : A.single.b, st,
                                                                                    spill, &tc, NEE && Q.shadow)) {
                tracing = false;
                pending = true;
            }
        }
    }
    if (A.ray_counts) {
        // wave reduction, one atomic pair per wave (the query counts are
        // already the wave's)
        if (COUNT) {
            for (int off = 32; off >= 1; off >>= 1) {
                n_nodes += __shfl_xor(n_nodes, off);
                n_tris += __shfl_xor(n_tris, off);
            }
        }
        if (lane == 0) {
            atomicAdd(&A.ray_counts[0], n_primary);
            atomicAdd(&A.ray_counts[1], n_secondary);
            if (COUNT) {
                atomicAdd(&A.ray_counts[2], n_nodes);
                atomicAdd(&A.ray_counts[3], n_tris);
            }
        }
    }
    if (ADJ) {
        // FP32 block partials -> FP64 gradient
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < 6 * A.grad_slots; i += BLK) {
            const float v = s_grad[i];
            if (v != 0.0f) atomicAdd(&A.grad[i], (double)v);
        }
    }
#undef SCN
}

// The frame as the reference's GPU path renders it, for comparisons: the
// render kernel of cuda_sched (detail/cuda_sched.inl:27-70) launches one
// thread per pixel over the image, which runs its samples and bounces in
// sequence (sample_pixel, detail/sched_common.h:427-526), tracing each ray
// with the binary-tree traversal of detail/bvh/intersect.inl:166-356 (the
// pair records, near child first); a thread whose pixel is done idles.  The
// same random numbers as k_pathtrace, so the same image up to ties.
template <int BLK>
__global__ __launch_bounds__(BLK) void k_pathtrace_plain(pt_args A)
{
    __shared__ uint32_t stack[LDS_STACK * BLK];
    uint32_t* st = stack + threadIdx.x;
    const scene_desc& S = A.single;
    const uint32_t w = S.sp.width, h = S.sp.height;
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const uint32_t idx = blockIdx.x * (uint32_t)BLK + threadIdx.x;
    unsigned long long n_primary = 0, n_secondary = 0;
    int x, y;
    if (idx < padded_pixels(w, h, A.tile) && tile_pixel(idx, w, h, A.tile, x, y)) {
        const uint32_t px = (uint32_t)y * w + (uint32_t)x;
        path_t<false> P;
        rng_seed(P.g, make_random_seed(y * (int)w + x, (int32_t)S.sp.frame_id));
        P.sum = make_float4(0, 0, 0, 0);
        for (P.s = 0; P.s < S.sp.spp;) {
            f3 o, d;
            primary_ray(S.cam, P.g, A.arg_order, x, y, (int)w, (int)h, o, d);
            start_sample(P);
            for (;;) {
                if (P.bounce == 0) ++n_primary;
                else ++n_secondary;
                const hit_t hr = traverse_binary<false, LDS_STACK, BLK>(S.b, o, d, 0.0f, FLT_MAX, st);
                f3 no, nd, so, sd;
                float smax;
                nee_t<false> Q;
                if (shade<PT_RENDER>(S, A, P, Q, hr, o, d, no, nd, so, sd, smax, nullptr)) break;
                o = no;
                d = nd;
            }
            end_sample(S, P);
        }
        write_pixel(S, px, sample_sum(S, P.sum));
    }
    if (A.ray_counts) {
        for (int off = 32; off >= 1; off >>= 1) {
            n_primary += __shfl_xor(n_primary, off);
            n_secondary += __shfl_xor(n_secondary, off);
        }
        if (lane == 0) {
            atomicAdd(&A.ray_counts[0], n_primary);
            atomicAdd(&A.ray_counts[1], n_secondary);
        }
    }
}



struct ao_args {
    bvh_view b;
    vr_camera cam;
    vr_sched_params sp;
    uint32_t samples;
    float radius;
    float bg[3];
    float* accum;
    float* color;
    uint32_t arg_order;
};

// AO kernel of src/examples/ao/main.cpp:192-260: primary closest hit, then
// `samples` cosine-distributed any-hit rays of length `radius` from
// P + dir * 1e-3; color = 1 - hits / samples.
__global__ __launch_bounds__(BLOCK) void k_ao(ao_args A)
{
    __shared__ uint32_t stack[LDS_STACK * BLOCK];
    uint32_t* st = stack + threadIdx.x;
    const uint32_t w = A.sp.width, h = A.sp.height;
    const uint32_t px = blockIdx.x * BLOCK + threadIdx.x;
    if (px >= w * h) return;
    const int x = (int)(px % w), y = (int)(px / w);
    rng_t g;
    rng_seed(g, make_random_seed(y * (int)w + x, (int32_t)A.sp.frame_id));
    float4 sum = make_float4(0, 0, 0, 0);
    const float inc = 1.0f / (float)A.samples;
    for (uint32_t s = 0; s < A.sp.spp; ++s) {
        f3 o, d;
        primary_ray(A.cam, g, A.arg_order, x, y, (int)w, (int)h, o, d);
        float4 c = make_float4(A.bg[0], A.bg[1], A.bg[2], 1.0f);
        const hit_t hr = traverse<false, false, LDS_STACK, BLOCK, TRAV_LANE>(A.b, o, d, 0.0f, FLT_MAX, st, nullptr);
        if (hr.tri >= 0) {
            const f3 pos = add3(o, muls3(d, hr.t));
            const float4* tq = reinterpret_cast<const float4*>(A.b.tris + hr.tri);
            f3 n = normalize3(cross3(ldtri3(tq[1]), ldtri3(tq[2])));
            n = faceforward3(n, neg3(d), n);
            f3 u, v;
            onb(n, u, v);
            float ao = 0.0f;
            for (uint32_t i = 0; i < A.samples; ++i) {
                float a1, a2;
                rng_pair(g, A.arg_order, a1, a2);
                const f3 sp = cosine_hemisphere(a1, a2);
                const f3 dir = normalize3(add3(add3(smul3(sp.x, u), smul3(sp.y, v)), smul3(sp.z, n)));
                const f3 ao_o = add3(pos, muls3(dir, 1e-3f));
                const hit_t ah = traverse<true, false, LDS_STACK, BLOCK, TRAV_LANE>(A.b, ao_o, dir, 0.0f, A.radius, st, nullptr);
                if (ah.tri >= 0) ao = ao + inc;
            }
            ao = 1.0f - ao;
            c = make_float4(ao, ao, ao, 1.0f);
        }
        sum.x += c.x; sum.y += c.y; sum.z += c.z; sum.w += c.w;
    }
    const float spp = (float)A.sp.spp;
    sum.x = sum.x / spp; sum.y = sum.y / spp; sum.z = sum.z / spp; sum.w = sum.w / spp;
    float4* acc = reinterpret_cast<float4*>(A.accum) + px;
    const float4 old = *acc;
    const float sf = A.sp.sfactor, df = A.sp.dfactor;
    const float4 r = make_float4(sum.x * sf + old.x * df, sum.y * sf + old.y * df, sum.z * sf + old.z * df,
                                 sum.w * sf + old.w * df);
    *acc = r;
    reinterpret_cast<float4*>(A.color)[px] = r;
}

// One compiled launch shape: a persistent grid of exactly the blocks the
// device holds at once (occupancy x CUs), computed once per shape.
template <int MODE, bool BATCH, int BLK, int LDSD, int NODE = NODE_QUAD>
int launch_variant(const pt_args& A, size_t dyn_lds, hipStream_t st)
{
    static int grid = 0;
    static size_t grid_lds = (size_t)-1;
    if (grid == 0 || grid_lds != dyn_lds) {
        int dev = 0, per_cu = 0, cus = 0;
        VR_HIP_CHECK(hipGetDevice(&dev));
        VR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pathtrace<MODE, BATCH, BLK, LDSD, NODE>,
                                                                 BLK, dyn_lds));
        VR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid = std::max(1, per_cu) * std::max(1, cus);
        grid_lds = dyn_lds;
    }
    hipLaunchKernelGGL((k_pathtrace<MODE, BATCH, BLK, LDSD, NODE>), dim3(grid), dim3(BLK), dyn_lds, st, A);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

// One thread per pixel with the reference's binary traversal (cuda_sched's
// launch): a grid over the pixel order.
template <int BLK>
int launch_plain(const pt_args& A, hipStream_t st)
{
    const uint32_t npix = padded_pixels(A.single.sp.width, A.single.sp.height, A.tile);
    hipLaunchKernelGGL((k_pathtrace_plain<BLK>), dim3(div_up(npix, BLK)), dim3(BLK), 0, st, A);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

// The frame kernel is compiled for 2 block sizes x 3 LDS stack depths; the
// counting, adjoint and frozen-roulette variants use the default shape.
template <int MODE, bool BATCH>
int launch_pt(const pt_args& A, size_t dyn_lds, const vr_launch_shape& sh, hipStream_t st)
{
    if (MODE != PT_RENDER) return launch_variant<MODE, BATCH, 256, 16>(A, dyn_lds, st);
    if (!BATCH && sh.schedule == VR_SCHEDULE_PER_PIXEL)
        return sh.block_size == 128 ? launch_plain<128>(A, st) : launch_plain<256>(A, st);
    switch (sh.block_size * 100 + sh.stack_lds) {
    case 12808: return launch_variant<MODE, BATCH, 128, 8>(A, dyn_lds, st);
    case 12816: return launch_variant<MODE, BATCH, 128, 16>(A, dyn_lds, st);
    case 12832: return launch_variant<MODE, BATCH, 128, 32>(A, dyn_lds, st);
    case 25608: return launch_variant<MODE, BATCH, 256, 8>(A, dyn_lds, st);
    case 25632: return launch_variant<MODE, BATCH, 256, 32>(A, dyn_lds, st);
    default: return launch_variant<MODE, BATCH, 256, 16>(A, dyn_lds, st);
    }
}

__global__ void k_frame_begin(uint32_t* work, uint32_t* frame, int advance)
{
    *work = 0u;
    if (frame && advance) *frame += 1u;
}

pt_args single_args(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_kernel_params& kp,
                    const vr_camera& cam, const vr_sched_params& sp, float* accum, float* color, const float* dl,
                    unsigned long long* ray_counts, uint32_t* work, double* grad, uint32_t arg_order)
{
    pt_args A{};
    A.single.b = bs.view();
    A.single.mats = mats;
    A.single.num_mats = (uint32_t)num_mats;
    A.single.grad_base = 0;
    A.single.kp = kp;
    A.single.cam = cam;
    A.single.sp = sp;
    A.single.accum = accum;
    A.single.color = color;
    A.single.dl = dl;
    A.scenes = nullptr;
    A.prefix = nullptr;
    A.num_scenes = 1;
    A.total_pixels = sp.width * sp.height;
    A.frame_offset = nullptr;
    A.ray_counts = ray_counts;
    A.work = work;
    A.arg_order = arg_order;
    A.grad_slots = grad ? (uint32_t)num_mats : 0u;
    A.grad = grad;
    return A;
}

}  // namespace

// argument evaluation order of the reference host build: GCC evaluates
// f(gen.next(), gen.next()) right to left (1)
uint32_t g_arg_order = REFERENCE_ARG_ORDER;

// current launch shape of the frame kernel (vr_set_launch_shape / autotuner)
vr_launch_shape g_shape = {256, 16, 1, VR_SCHEDULE_PERSISTENT};

bool valid_shape(const vr_launch_shape& s)
{
    const bool b = s.block_size == 128 || s.block_size == 256;
    const bool d = s.stack_lds == 8 || s.stack_lds == 16 || s.stack_lds == 32;
    const bool q = s.schedule == VR_SCHEDULE_PERSISTENT || (s.schedule == VR_SCHEDULE_PER_PIXEL && s.stack_lds == 16);
    return b && d && q && s.tile >= 1 && s.tile <= 64;
}

// largest material table the adjoint keeps in LDS partials
constexpr size_t ADJ_MAX_MATERIALS = ADJ_MAX_MATERIALS_TOTAL;

int render_pt_shaped(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_kernel_params& kp,
                     const vr_camera& cam, const vr_sched_params& sp, float* accum, float* color,
                     unsigned long long* ray_counts, uint32_t* work, bool counted, const vr_launch_shape& shape,
                     hipStream_t st);

int render_pt(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_kernel_params& kp,
              const vr_camera& cam, const vr_sched_params& sp, float* accum, float* color,
              unsigned long long* ray_counts, uint32_t* work, bool counted, hipStream_t st)
{
    if ((size_t)sp.width * sp.height == 0) return VR_OK;
    VR_CHECK(stack_fits(bs), VR_ERR_UNSUPPORTED, "render_pt: BVH deeper than the traversal stack");
    return render_pt_shaped(bs, mats, num_mats, kp, cam, sp, accum, color, ray_counts, work, counted, g_shape, st);
}

int render_pt_shaped(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_kernel_params& kp,
                     const vr_camera& cam, const vr_sched_params& sp, float* accum, float* color,
                     unsigned long long* ray_counts, uint32_t* work, bool counted, const vr_launch_shape& shape,
                     hipStream_t st)
{
    if ((size_t)sp.width * sp.height == 0) return VR_OK;
    VR_CHECK((uint64_t)padded_pixels(sp.width, sp.height, shape.tile) < (1ull << 31), VR_ERR_INVALID_ARGUMENT,
             "render_pt: image too large");
    pt_args A =
        single_args(bs, mats, num_mats, kp, cam, sp, accum, color, nullptr, ray_counts, work, nullptr, g_arg_order);
    A.tile = shape.tile;
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(1), 0, st, work, nullptr, 0);
    // a BVH with 8-wide nodes: the persistent kernel over them (default
    // block size and stack depth)
    if (shape.schedule != VR_SCHEDULE_PER_PIXEL) {
        switch (traversal_kind(bs)) {
        case NODE_OCT:
            return counted ? launch_variant<PT_COUNT, false, 256, 16, NODE_OCT>(A, 0, st)
                           : launch_variant<PT_RENDER, false, 256, 16, NODE_OCT>(A, 0, st);
        case NODE_OCTQ:
            return counted ? launch_variant<PT_COUNT, false, 256, 16, NODE_OCTQ>(A, 0, st)
                           : launch_variant<PT_RENDER, false, 256, 16, NODE_OCTQ>(A, 0, st);
        default: break;
        }
    }
    return counted ? launch_pt<PT_COUNT, false>(A, 0, shape, st) : launch_pt<PT_RENDER, false>(A, 0, shape, st);
}

// the frame with a light list (kernel_params w/ lights, kernels.h:137-187)
int render_pt_lights(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_area_light* lights,
                     size_t num_lights, const vr_kernel_params& kp, const vr_camera& cam, const vr_sched_params& sp,
                     float* accum, float* color, unsigned long long* ray_counts, uint32_t* work, bool counted,
                     hipStream_t st)
{
    if (num_lights == 0) return render_pt(bs, mats, num_mats, kp, cam, sp, accum, color, ray_counts, work, counted, st);
    if ((size_t)sp.width * sp.height == 0) return VR_OK;
    VR_CHECK(stack_fits(bs), VR_ERR_UNSUPPORTED, "render_pt_lights: BVH deeper than the traversal stack");
    VR_CHECK(num_lights < (1ull << 31), VR_ERR_INVALID_ARGUMENT, "render_pt_lights: more than 2^31 lights");
    VR_CHECK((uint64_t)padded_pixels(sp.width, sp.height, g_shape.tile) < (1ull << 31), VR_ERR_INVALID_ARGUMENT,
             "render_pt_lights: image too large");
    pt_args A =
        single_args(bs, mats, num_mats, kp, cam, sp, accum, color, nullptr, ray_counts, work, nullptr, g_arg_order);
    A.tile = g_shape.tile;
    A.single.lights = lights;
    A.single.num_lights = (uint32_t)num_lights;
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(1), 0, st, work, nullptr, 0);
    return counted ? launch_pt<PT_NEE | PT_COUNT, false>(A, 0, g_shape, st)
                   : launch_pt<PT_NEE | PT_RENDER, false>(A, 0, g_shape, st);
}

int render_pt_frozen_rr(const bvh_state& bs, const vr_material* mats, const vr_material* rr_mats, size_t num_mats,
                        const vr_kernel_params& kp, const vr_camera& cam, const vr_sched_params& sp, float* accum,
                        float* color, unsigned long long* ray_counts, uint32_t* work, hipStream_t st)
{
    if ((size_t)sp.width * sp.height == 0) return VR_OK;
    VR_CHECK(stack_fits(bs), VR_ERR_UNSUPPORTED, "render_pt_frozen_rr: BVH deeper than the stack");
    VR_CHECK((uint64_t)padded_pixels(sp.width, sp.height, g_shape.tile) < (1ull << 31), VR_ERR_INVALID_ARGUMENT,
             "render_pt_frozen_rr: image too large");
    pt_args A =
        single_args(bs, mats, num_mats, kp, cam, sp, accum, color, nullptr, ray_counts, work, nullptr, g_arg_order);
    A.tile = g_shape.tile;
    A.rr_mats = rr_mats;
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(1), 0, st, work, nullptr, 0);
    return launch_pt<PT_FROZEN_RR, false>(A, 0, g_shape, st);
}

int render_pt_adjoint(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_kernel_params& kp,
                      const vr_camera& cam, const vr_sched_params& sp, const float* dl, double* grad,
                      unsigned long long* ray_counts, uint32_t* work, hipStream_t st)
{
    if ((size_t)sp.width * sp.height == 0) return VR_OK;
    VR_CHECK(num_mats <= ADJ_MAX_MATERIALS, VR_ERR_UNSUPPORTED, "render_pt_backward: more than 2048 materials");
    VR_CHECK(kp.num_bounces <= (uint32_t)ADJ_MAX_VERTICES, VR_ERR_UNSUPPORTED,
             "render_pt_backward: more than 16 bounces");
    VR_CHECK(stack_fits(bs), VR_ERR_UNSUPPORTED, "render_pt_backward: BVH deeper than the stack");
    pt_args A =
        single_args(bs, mats, num_mats, kp, cam, sp, nullptr, nullptr, dl, ray_counts, work, grad, g_arg_order);
    A.tile = g_shape.tile;
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(1), 0, st, work, nullptr, 0);
    return launch_pt<PT_ADJOINT, false>(A, 6 * num_mats * sizeof(float), g_shape, st);
}

// Many scenes in one persistent launch.  d_scenes / d_prefix are device
// arrays prepared by the batch handle; d_frame (optional) is a device frame
// counter incremented before every frame, so a captured graph renders
// successive frames.  Only kernel launches are enqueued (graph capturable).
int render_pt_batch(const void* d_scenes, const uint32_t* d_prefix, uint32_t num_scenes, uint32_t total_pixels,
                    uint32_t* d_frame, bool advance_frame, unsigned long long* ray_counts, uint32_t* work,
                    uint32_t grad_slots, double* grad, const vr_launch_shape& shape, hipStream_t st)
{
    if (total_pixels == 0) return VR_OK;
    pt_args A{};
    A.tile = shape.tile;
    A.scenes = static_cast<const scene_desc*>(d_scenes);
    A.prefix = d_prefix;
    A.num_scenes = num_scenes;
    A.total_pixels = total_pixels;
    A.frame_offset = d_frame;
    A.ray_counts = ray_counts;
    A.work = work;
    A.arg_order = g_arg_order;
    A.grad_slots = grad ? grad_slots : 0u;
    A.grad = grad;
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(1), 0, st, work, d_frame, advance_frame ? 1 : 0);
    if (grad) {
        VR_CHECK(grad_slots <= ADJ_MAX_MATERIALS, VR_ERR_UNSUPPORTED, "batch adjoint: more than 2048 materials");
        return launch_pt<PT_ADJOINT, true>(A, 6 * (size_t)grad_slots * sizeof(float), shape, st);
    }
    return launch_pt<PT_RENDER, true>(A, 0, shape, st);
}

uint32_t padded_pixel_count(uint32_t w, uint32_t h, uint32_t tile) { return padded_pixels(w, h, tile); }

size_t scene_desc_size() { return sizeof(scene_desc); }

// fill one scene descriptor (host side) for the batch handle
void make_scene_desc(void* out, const bvh_state& bs, const vr_material* mats, uint32_t num_mats, uint32_t grad_base,
                     const vr_kernel_params& kp, const vr_camera& cam, const vr_sched_params& sp, float* accum,
                     float* color, const float* dl)
{
    scene_desc d{};
    d.b = bs.view();
    d.mats = mats;
    d.num_mats = num_mats;
    d.grad_base = grad_base;
    d.kp = kp;
    d.cam = cam;
    d.sp = sp;
    d.accum = accum;
    d.color = color;
    d.dl = dl;
    std::memcpy(out, &d, sizeof(d));
}

// L2 image loss sum_px sum_rgb (I - T)^2 (FP64 reduction) and its image
// gradient 2 (I - T) (alpha 0).
__global__ __launch_bounds__(BLOCK) void k_l2_loss(const float4* __restrict__ img, const float4* __restrict__ tgt,
                                                   size_t npix, float4* __restrict__ dl, double* __restrict__ loss)
{
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)BLOCK + threadIdx.x; i < npix; i += (size_t)gridDim.x * BLOCK) {
        const float4 a = img[i], b = tgt[i];
        const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        if (dl) dl[i] = make_float4(2.0f * dx, 2.0f * dy, 2.0f * dz, 0.0f);
        acc += (double)dx * dx + (double)dy * dy + (double)dz * dz;
    }
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ double s[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < BLOCK / 64; ++q) t += s[q];
        atomicAdd(loss, t);
    }
}

int image_l2_loss(const float* img, const float* tgt, size_t npix, float* dl, double* loss, hipStream_t st)
{
    if (npix == 0) return VR_OK;
    const unsigned grid = std::max(1u, std::min(div_up(npix, BLOCK), 2048u));
    hipLaunchKernelGGL(k_l2_loss, dim3(grid), dim3(BLOCK), 0, st, reinterpret_cast<const float4*>(img),
                       reinterpret_cast<const float4*>(tgt), npix, reinterpret_cast<float4*>(dl), loss);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

int render_ao(const bvh_state& bs, const vr_camera& cam, const vr_sched_params& sp, uint32_t samples, float radius,
              const float bg[3], float* accum, float* color, hipStream_t st)
{
    const size_t npix = (size_t)sp.width * sp.height;
    if (npix == 0) return VR_OK;
    ao_args A;
    A.b = bs.view();
    A.cam = cam;
    A.sp = sp;
    A.samples = samples;
    A.radius = radius;
    A.bg[0] = bg[0]; A.bg[1] = bg[1]; A.bg[2] = bg[2];
    A.accum = accum;
    A.color = color;
    A.arg_order = g_arg_order;
    VR_CHECK(stack_fits(bs), VR_ERR_UNSUPPORTED, "render_ao: BVH deeper than the traversal stack");
    hipLaunchKernelGGL(k_ao, dim3(div_up(npix, BLOCK)), dim3(BLOCK), 0, st, A);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

}  // namespace vr
