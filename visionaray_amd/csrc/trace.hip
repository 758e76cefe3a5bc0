// trace.hip -- closest-hit / any-hit queries over a batch of rays
// (closest_hit / any_hit, detail/traverse_linear.inl:149-214, one BVH in the
// list; intersect<Traversal>, detail/bvh/intersect.inl:41-357).
//
// A resident grid of waves; each lane holds one query and advances it one
// traversal iteration (a node visit or a triangle test) per loop iteration.  A lane whose query ends writes its hit
// and takes the next ray at once, so no lane idles until the longest query
// of its wave is done.  A wave reserves ray indices 64 at a time with one
// atomic and hands them to its idle lanes in order (neighbouring rays stay
// in one wave).  Rays and hits are AoS records loaded as 16-byte vectors;
// the traversal stack lives in LDS (one column per lane).
#include "vr_internal.h"

#include <algorithm>
#include <mutex>

namespace vr {

// Scratch of vr_trace_noalloc: the work counter of its resident grid (the
// other entry points take theirs from a counter ring the library allocates
// on a device's first use).  A file-scope device array, so that entry point
// allocates nothing; every call on a device shares that device's copy and
// keeps the last call's contents (the counter is cleared by a fill kernel at
// the start of each call), so calls that overlap on one GPU race.  256 bytes:
// word 0 is the counter, the rest of the line stays unused so the counter
// shares its cache line with nothing.
constexpr uint32_t NOALLOC_SCRATCH_WORDS = 64;
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];

namespace {

// the device address of `scratch` on each device (the array's name is a
// host-side stand-in: only these addresses may reach kernels and copies)
constexpr int NOALLOC_MAX_DEVICES = 64;
uint32_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mutex;

constexpr int BLOCK = 256;
constexpr uint32_t CHUNK = 64;

template <bool ANY, bool COUNT, int NODE>
__global__ __launch_bounds__(BLOCK) void k_trace(bvh_view b, const vr_ray* __restrict__ rays, uint32_t n,
                                                 vr_hit* __restrict__ hits, uint32_t* __restrict__ counts,
                                                 uint32_t* __restrict__ work)
{
    __shared__ uint32_t stack[LDS_STACK * BLOCK];
    uint32_t* st = stack + threadIdx.x;
    uint32_t spill[max_stack_of<NODE>() - LDS_STACK];
    trav_state T;
    trace_counters c{0, 0};
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    uint32_t idx = 0;
    bool tracing = false, done = false;
    uint32_t chunk_base = 0, chunk_left = 0;  // the wave's reserved indices (uniform)
    for (;;) {
        // lanes without a query take the next rays of the wave's chunk, once
        // VR_TRACE_REFILL_LANES of them wait (or no lane has a query left)
        const bool need = !done && !tracing;
        const uint64_t needm = __ballot(need);
        if (needm && (__popcll(needm) >= VR_TRACE_REFILL_LANES || __ballot(tracing) == 0)) {
            const uint32_t k = (uint32_t)__popcll(needm);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(needm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)needm, 0u));
            uint32_t fresh = 0;
            if (chunk_left < k) {
                const int leader = __ffsll((unsigned long long)needm) - 1;
                if ((int)lane == leader) fresh = atomicAdd(work, CHUNK);
                fresh = __shfl(fresh, leader);
            }
            if (need) {
                const uint32_t my = rank < chunk_left ? chunk_base + rank : fresh + (rank - chunk_left);
                if (my < n) {
                    const float4* rp = reinterpret_cast<const float4*>(rays + my);
                    const float4 r0 = rp[0], r1 = rp[1];
                    trav_begin(T, b, mk3(r0.x, r0.y, r0.z), mk3(r0.w, r1.x, r1.y), r1.z, r1.w);
                    idx = my;
                    tracing = true;
                } else {
                    done = true;
                }
            }
            if (chunk_left < k) {
                chunk_base = fresh + (k - chunk_left);
                chunk_left = CHUNK - (k - chunk_left);
            } else {
                chunk_base += k;
                chunk_left -= k;
            }
        }
        if (__ballot(!done) == 0) break;
        if (tracing && trav_iter<ANY, COUNT, LDS_STACK, BLOCK, COUNT ? TRAV_LANE : TRAV_VOTED, paired_isect, NODE>(
                           T, b, st, spill, &c)) {
            float4 o;
            if (T.r.tri >= 0) {
                const float prim = b.tris[T.r.tri].a.w;  // prim_id bits
                o = make_float4(prim, T.r.t, T.r.u, T.r.v);
            } else {
                o = make_float4(__int_as_float(-1), FLT_MAX, 0.0f, 0.0f);
            }
            *reinterpret_cast<float4*>(hits + idx) = o;
            if (COUNT) {
                counts[2 * (size_t)idx] = c.nodes;
                counts[2 * (size_t)idx + 1] = c.tris;
                c.nodes = 0;
                c.tris = 0;
            }
            tracing = false;
        }
    }
}

__global__ void k_reset(uint32_t* work) { *work = 0u; }

// a resident grid: the blocks the device holds at once
template <bool ANY, bool COUNT, int NODE>
int launch_trace(const bvh_state& bs, const vr_ray* rays, uint32_t n, vr_hit* hits, uint32_t* counts,
                 uint32_t* work, hipStream_t st)
{
    static int grid = 0;
    if (grid == 0) {
        int dev = 0, per_cu = 0, cus = 0;
        VR_HIP_CHECK(hipGetDevice(&dev));
        VR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<ANY, COUNT, NODE>, BLOCK, 0));
        VR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid = std::max(1, per_cu) * std::max(1, cus);
    }
    // fewer rays than lanes: no more waves than there are 64-ray chunks
    const unsigned blocks = std::max(1u, std::min((unsigned)grid, div_up(n, BLOCK)));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, st, work);
    hipLaunchKernelGGL((k_trace<ANY, COUNT, NODE>), dim3(blocks), dim3(BLOCK), 0, st, bs.view(), rays, n, hits,
                       counts, work);
    VR_HIP_CHECK(hipGetLastError());
    return VR_OK;
}

}  // namespace

int trace_batch(const bvh_state& bs, bool any_hit, const vr_ray* rays, size_t n, vr_hit* hits, uint32_t* counts,
                uint32_t* work, hipStream_t st);

namespace {
template <int NODE>
int trace_kind(const bvh_state& bs, bool any_hit, const vr_ray* rays, uint32_t m, vr_hit* hits, uint32_t* counts,
               uint32_t* work, hipStream_t st)
{
    if (counts)
        return any_hit ? launch_trace<true, true, NODE>(bs, rays, m, hits, counts, work, st)
                       : launch_trace<false, true, NODE>(bs, rays, m, hits, counts, work, st);
    return any_hit ? launch_trace<true, false, NODE>(bs, rays, m, hits, nullptr, work, st)
                   : launch_trace<false, false, NODE>(bs, rays, m, hits, nullptr, work, st);
}
}  // namespace

// Look up the current device's address of `scratch` once (vr_init, else the
// first vr_trace_noalloc call on the device).
int trace_scratch_setup()
{
    int dev = 0;
    VR_HIP_CHECK(hipGetDevice(&dev));
    VR_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, VR_ERR_UNSUPPORTED, "vr_trace_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev]) return VR_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    VR_CHECK(e == hipSuccess && p != nullptr, VR_ERR_HIP,
             std::string("vr_trace_noalloc: hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
    return VR_OK;
}

// vr_trace with the work counter in `scratch`: no device allocation.
int trace_batch_noalloc(const bvh_state& bs, bool any_hit, const vr_ray* rays, size_t n, vr_hit* hits,
                        hipStream_t st)
{
    VR_CHECK(n <= (1ull << 31), VR_ERR_INVALID_ARGUMENT,
             "vr_trace_noalloc: more than 2^31 rays in one call (the scratch holds one 32-bit work counter)");
    if (n == 0) return VR_OK;
    const int rc = trace_scratch_setup();
    if (rc) return rc;
    int dev = 0;
    VR_HIP_CHECK(hipGetDevice(&dev));
    return trace_batch(bs, any_hit, rays, n, hits, nullptr, g_scratch_ptr[dev], st);
}

int trace_batch(const bvh_state& bs, bool any_hit, const vr_ray* rays, size_t n, vr_hit* hits, uint32_t* counts,
                uint32_t* work, hipStream_t st)
{
    if (n == 0) return VR_OK;
    VR_CHECK(n <= (1ull << 31), VR_ERR_INVALID_ARGUMENT, "trace: more than 2^31 rays in one call");
    VR_CHECK(stack_fits(bs), VR_ERR_UNSUPPORTED, "trace: BVH deeper than the traversal stack");
    const uint32_t m = (uint32_t)n;
    switch (traversal_kind(bs)) {
    case NODE_OCT: return trace_kind<NODE_OCT>(bs, any_hit, rays, m, hits, counts, work, st);
    case NODE_OCTQ: return trace_kind<NODE_OCTQ>(bs, any_hit, rays, m, hits, counts, work, st);
    default: return trace_kind<NODE_QUAD>(bs, any_hit, rays, m, hits, counts, work, st);
    }
}

}  // namespace vr
