// trace.hip -- closest-hit / any-hit queries over a batch of rays
// (closest_hit / any_hit, detail/traverse_linear.inl:149-214, one BVH in the
// list; intersect<Traversal>, detail/bvh/intersect.inl:41-357).
//
// One lane per ray, rays and hits in AoS records loaded as 16-byte vectors,
// the traversal stack in LDS (one column per lane).
#include "vr_internal.h"

namespace vr {

namespace {

constexpr int BLOCK = 256;

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace(bvh_view b, const vr_ray* __restrict__ rays, size_t n,
                                                 vr_hit* __restrict__ hits, uint32_t* __restrict__ counts)
{
    __shared__ uint32_t stack[LDS_STACK * BLOCK];
    const size_t i = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 r0 = rp[0], r1 = rp[1];
    trace_counters c{0, 0};
    const hit_t h = traverse<ANY, COUNT, LDS_STACK, BLOCK>(b, mk3(r0.x, r0.y, r0.z), mk3(r0.w, r1.x, r1.y), r1.z, r1.w,
                                                      stack + threadIdx.x, &c);
    float4 o;
    if (h.tri >= 0) {
        const float prim = b.tris[h.tri].a.w;  // prim_id bits
        o = make_float4(prim, h.t, h.u, h.v);
    } else {
        o = make_float4(__int_as_float(-1), FLT_MAX, 0.0f, 0.0f);
    }
    *reinterpret_cast<float4*>(hits + i) = o;
    if (COUNT) {
        counts[2 * i] = c.nodes;
        counts[2 * i + 1] = c.tris;
    }
}

template <bool ANY, bool COUNT>
int launch_trace_depth(const bvh_state& bs, const vr_ray* rays, size_t n, vr_hit* hits, uint32_t* counts,
                       hipStream_t st)
{
    VR_CHECK(stack_fits(bs), VR_ERR_UNSUPPORTED, "trace: BVH deeper than the traversal stack");
    hipLaunchKernelGGL((k_trace<ANY, COUNT>), dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, st, bs.view(), rays, n, hits,
                       counts);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

}  // namespace

int trace_batch(const bvh_state& bs, bool any_hit, const vr_ray* rays, size_t n, vr_hit* hits, uint32_t* counts,
                hipStream_t st)
{
    if (n == 0) return VR_OK;
    if (counts) {
        return any_hit ? launch_trace_depth<true, true>(bs, rays, n, hits, counts, st)
                       : launch_trace_depth<false, true>(bs, rays, n, hits, counts, st);
    }
    return any_hit ? launch_trace_depth<true, false>(bs, rays, n, hits, nullptr, st)
                   : launch_trace_depth<false, false>(bs, rays, n, hits, nullptr, st);
}

}  // namespace vr
