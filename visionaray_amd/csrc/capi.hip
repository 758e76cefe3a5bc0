// capi.hip -- the C ABI (include/visionaray_amd.h): status codes, handles,
// device memory helpers and the host-side parts of the path (camera basis,
// conversion of reference-layout BVHs into the traversal layout).
#include "vr_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

namespace vr {

int lbvh_build_into(bvh_state& bs, const vr_triangle* d_prims, size_t n, bool keep_nodes, bool rotate,
                    hipStream_t st);
int lbvh_refit(bvh_state& bs, const vr_triangle* d_prims, hipStream_t st);
int lbvh_morton_only(const vr_triangle* d_prims, size_t n, uint32_t* d_codes, hipStream_t st);
int trace_batch(const bvh_state& bs, bool any_hit, const vr_ray* rays, size_t n, vr_hit* hits, uint32_t* counts,
                uint32_t* work, hipStream_t st);
int trace_scratch_setup();
int trace_batch_noalloc(const bvh_state& bs, bool any_hit, const vr_ray* rays, size_t n, vr_hit* hits,
                        hipStream_t st);
int render_pt(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_kernel_params& kp,
              const vr_camera& cam, const vr_sched_params& sp, float* accum, float* color,
              unsigned long long* ray_counts, uint32_t* work, bool counted, hipStream_t st);
int render_ao(const bvh_state& bs, const vr_camera& cam, const vr_sched_params& sp, uint32_t samples, float radius,
              const float bg[3], float* accum, float* color, hipStream_t st);
int render_pt_lights(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_area_light* lights,
                     size_t num_lights, const vr_kernel_params& kp, const vr_camera& cam, const vr_sched_params& sp,
                     float* accum, float* color, unsigned long long* ray_counts, uint32_t* work, bool counted,
                     hipStream_t st);
int sah_build_host(const vr_triangle* prims, size_t n, int max_leaf_size, vr_bvh_node* nodes, size_t* num_nodes,
                   uint32_t* indices);
int lbvh_build_host(const vr_triangle* prims, size_t n, int max_leaf_size, vr_bvh_node* nodes, size_t* num_nodes,
                    uint32_t* indices);
int device_gather_lines(const void* src, size_t src_bytes, const uint32_t* lines, size_t n, uint32_t* out,
                        hipStream_t st);
int device_copy(void* dst, const void* src, size_t bytes, hipStream_t st);
int render_pt_adjoint(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_kernel_params& kp,
                      const vr_camera& cam, const vr_sched_params& sp, const float* dl, double* grad,
                      unsigned long long* ray_counts, uint32_t* work, hipStream_t st);
int image_l2_loss(const float* img, const float* tgt, size_t npix, float* dl, double* loss, hipStream_t st);
int render_pt_frozen_rr(const bvh_state& bs, const vr_material* mats, const vr_material* rr_mats, size_t num_mats,
                        const vr_kernel_params& kp, const vr_camera& cam, const vr_sched_params& sp, float* accum,
                        float* color, unsigned long long* ray_counts, uint32_t* work, hipStream_t st);
int render_pt_batch(const void* d_scenes, const uint32_t* d_prefix, uint32_t num_scenes, uint32_t total_pixels,
                    uint32_t* d_frame, bool advance_frame, unsigned long long* ray_counts, uint32_t* work,
                    uint32_t grad_slots, double* grad, const vr_launch_shape& shape, hipStream_t st);
int render_pt_shaped(const bvh_state& bs, const vr_material* mats, size_t num_mats, const vr_kernel_params& kp,
                     const vr_camera& cam, const vr_sched_params& sp, float* accum, float* color,
                     unsigned long long* ray_counts, uint32_t* work, bool counted, const vr_launch_shape& shape,
                     hipStream_t st);
uint32_t padded_pixel_count(uint32_t w, uint32_t h, uint32_t tile);
extern vr_launch_shape g_shape;
bool valid_shape(const vr_launch_shape& s);
size_t scene_desc_size();
void make_scene_desc(void* out, const bvh_state& bs, const vr_material* mats, uint32_t num_mats, uint32_t grad_base,
                     const vr_kernel_params& kp, const vr_camera& cam, const vr_sched_params& sp, float* accum,
                     float* color, const float* dl);

namespace {
thread_local std::string g_error;

// Work counters of persistent kernels: every launch takes the next of
// WORK_SLOTS counters (each on its own 128-byte line) and resets it on its
// stream first, so launches in flight on different streams never share one.
constexpr uint32_t WORK_SLOTS = 1024;
struct device_state {
    bool ready = false;
    uint32_t* work = nullptr;  // WORK_SLOTS counters, 32 words apart
    std::atomic<uint32_t> next{0};
    // the library's own non-blocking stream: small read-backs and host-tree
    // uploads that must not wait for the caller's streams
    hipStream_t aux = nullptr;
};
std::mutex g_mutex;
device_state g_dev[64];

device_state* dev_state()
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= 64) return nullptr;
    return &g_dev[d];
}
}  // namespace

void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

int current_device()
{
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

}  // namespace vr

using namespace vr;


struct vr_batch_t {
    int device = 0;
    uint32_t num_scenes = 0;
    uint32_t total_pixels = 0;
    uint32_t total_materials = 0;
    uint32_t max_depth = 0;
    vr_launch_shape shape{256, 16, 1, VR_SCHEDULE_PERSISTENT};  // fixed at creation (the pixel prefix depends on the tile)
    void* d_scenes = nullptr;      // scene descriptors
    uint32_t* d_prefix = nullptr;  // num_scenes + 1
    uint32_t* d_frame = nullptr;   // frame counter (starts at -1, advanced before each frame)
    uint32_t* d_work = nullptr;    // the batch's own work counter (graph replays)
    // every scene has d loss / d image and at most ADJ_MAX_BOUNCES bounces
    // (vr_batch_backward needs both)
    bool adjoint_ok = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int vr::ensure_device()
{
    device_state* ds = dev_state();
    VR_CHECK(ds != nullptr, VR_ERR_HIP, "no current HIP device");
    if (ds->ready) return VR_OK;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ds->ready) {
        VR_HIP_CHECK(hipMalloc(&ds->work, (size_t)WORK_SLOTS * 128));
        // LBVH arrays come from the device's stream-ordered pool; freed blocks
        // stay in the pool, so the next build of a similar size allocates
        // without a driver call
        int dev = 0;
        hipMemPool_t pool = nullptr;
        VR_HIP_CHECK(hipGetDevice(&dev));
        VR_HIP_CHECK(hipDeviceGetDefaultMemPool(&pool, dev));
        uint64_t keep = UINT64_MAX;
        VR_HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
        VR_HIP_CHECK(hipStreamCreateWithFlags(&ds->aux, hipStreamNonBlocking));
        ds->ready = true;
    }
    return VR_OK;
}

// the work counter of the next persistent launch on the current device
static uint32_t* work_counter()
{
    device_state* ds = dev_state();
    return ds->work + 32u * (ds->next.fetch_add(1u, std::memory_order_relaxed) % WORK_SLOTS);
}

uint32_t* vr::next_work_counter() { return work_counter(); }

static hipStream_t aux_stream() { return dev_state()->aux; }

// The handle's last build, rebuild or refit: an event on the caller's stream,
// so a read-back of what the build wrote waits for that work only.
static int mark_built(vr_bvh h, hipStream_t st)
{
    if (!h->built) VR_HIP_CHECK(hipEventCreateWithFlags(&h->built, hipEventDisableTiming));
    VR_HIP_CHECK(hipEventRecord(h->built, st));
    return VR_OK;
}

// copy device -> host after the handle's last build, on the library's own
// stream (no wait for unrelated work)
static int read_after_build(vr_bvh h, void* dst, const void* src, size_t bytes)
{
    if (h->built) VR_HIP_CHECK(hipEventSynchronize(h->built));
    VR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, aux_stream()));
    VR_HIP_CHECK(hipStreamSynchronize(aux_stream()));
    return VR_OK;
}

extern "C" {

int vr_init(int device)
{
    int count = 0;
    VR_HIP_CHECK(hipGetDeviceCount(&count));
    VR_CHECK(device >= 0 && device < count, VR_ERR_INVALID_ARGUMENT, "vr_init: no such device");
    VR_HIP_CHECK(hipSetDevice(device));
    const int rc = ensure_device();
    return rc ? rc : trace_scratch_setup();
}

const char* vr_last_error(void) { return g_error.c_str(); }

int vr_version(int* major, int* minor)
{
    VR_CHECK(major && minor, VR_ERR_INVALID_ARGUMENT, "vr_version: null argument");
    *major = 0;
    *minor = 1;
    return VR_OK;
}

int vr_malloc(void** d_ptr, size_t bytes)
{
    VR_CHECK(d_ptr, VR_ERR_INVALID_ARGUMENT, "vr_malloc: null out pointer");
    *d_ptr = nullptr;
    if (bytes == 0) return VR_OK;
    hipError_t e = hipMalloc(d_ptr, bytes);
    if (e == hipErrorOutOfMemory) return fail(VR_ERR_OUT_OF_MEMORY, "vr_malloc: out of device memory");
    VR_HIP_CHECK(e);
    return VR_OK;
}

int vr_free(void* d_ptr)
{
    if (d_ptr) VR_HIP_CHECK(hipFree(d_ptr));
    return VR_OK;
}

int vr_memcpy_htod(void* d_dst, const void* h_src, size_t bytes)
{
    if (bytes == 0) return VR_OK;
    VR_CHECK(d_dst && h_src, VR_ERR_INVALID_ARGUMENT, "vr_memcpy_htod: null pointer");
    VR_HIP_CHECK(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
    return VR_OK;
}

int vr_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes)
{
    if (bytes == 0) return VR_OK;
    VR_CHECK(h_dst && d_src, VR_ERR_INVALID_ARGUMENT, "vr_memcpy_dtoh: null pointer");
    VR_HIP_CHECK(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    return VR_OK;
}

int vr_memset(void* d_dst, int value, size_t bytes)
{
    if (bytes == 0) return VR_OK;
    VR_CHECK(d_dst, VR_ERR_INVALID_ARGUMENT, "vr_memset: null pointer");
    VR_HIP_CHECK(hipMemset(d_dst, value, bytes));
    return VR_OK;
}

int vr_synchronize(void)
{
    VR_HIP_CHECK(hipDeviceSynchronize());
    return VR_OK;
}

int vr_stream_create(void** stream)
{
    VR_CHECK(stream, VR_ERR_INVALID_ARGUMENT, "vr_stream_create: null out pointer");
    hipStream_t s;
    VR_HIP_CHECK(hipStreamCreate(&s));
    *stream = s;
    return VR_OK;
}

int vr_stream_destroy(void* stream)
{
    if (stream) VR_HIP_CHECK(hipStreamDestroy((hipStream_t)stream));
    return VR_OK;
}

int vr_stream_synchronize(void* stream)
{
    VR_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return VR_OK;
}

int vr_event_create(void** event)
{
    VR_CHECK(event, VR_ERR_INVALID_ARGUMENT, "vr_event_create: null out pointer");
    hipEvent_t e;
    VR_HIP_CHECK(hipEventCreate(&e));
    *event = e;
    return VR_OK;
}

int vr_event_destroy(void* event)
{
    if (event) VR_HIP_CHECK(hipEventDestroy((hipEvent_t)event));
    return VR_OK;
}

int vr_event_record(void* event, void* stream)
{
    VR_CHECK(event, VR_ERR_INVALID_ARGUMENT, "vr_event_record: null event");
    VR_HIP_CHECK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return VR_OK;
}

int vr_event_elapsed_ms(void* start, void* stop, float* ms)
{
    VR_CHECK(start && stop && ms, VR_ERR_INVALID_ARGUMENT, "vr_event_elapsed_ms: null argument");
    VR_HIP_CHECK(hipEventSynchronize((hipEvent_t)stop));
    VR_HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return VR_OK;
}

int vr_device_copy(void* d_dst, const void* d_src, size_t bytes, void* stream)
{
    VR_CHECK(bytes == 0 || (d_dst && d_src), VR_ERR_INVALID_ARGUMENT, "vr_device_copy: null pointer");
    return device_copy(d_dst, d_src, bytes, (hipStream_t)stream);
}

int vr_device_gather_lines(const void* d_src, size_t src_bytes, const uint32_t* d_lines, size_t n, uint32_t* d_out,
                           void* stream)
{
    VR_CHECK(n == 0 || (d_src && d_lines && d_out), VR_ERR_INVALID_ARGUMENT, "vr_device_gather_lines: null pointer");
    return device_gather_lines(d_src, src_bytes, d_lines, n, d_out, (hipStream_t)stream);
}

// pinhole_camera::look_at + perspective + begin_frame
// (detail/pinhole_camera.inl:12-29, 31-64, 87-97), host float arithmetic
int vr_camera_look_at(const float eye[3], const float center[3], const float up[3], float fovy, float aspect,
                      vr_camera* out)
{
    VR_CHECK(eye && center && up && out, VR_ERR_INVALID_ARGUMENT, "vr_camera_look_at: null argument");
    const f3 e = mk3(eye[0], eye[1], eye[2]);
    const f3 c = mk3(center[0], center[1], center[2]);
    const f3 u0 = mk3(up[0], up[1], up[2]);
    const f3 f = normalize3(sub3(e, c));
    const f3 s = normalize3(cross3(u0, f));
    const f3 u = cross3(f, s);
    const float th = std::tan(fovy / 2.0f);
    const f3 U = muls3(muls3(s, th), aspect);
    const f3 V = muls3(u, std::tan(fovy / 2.0f));
    const f3 W = neg3(f);
    const float vals[12] = {e.x, e.y, e.z, U.x, U.y, U.z, V.x, V.y, V.z, W.x, W.y, W.z};
    std::memcpy(out, vals, sizeof(vals));
    return VR_OK;
}

int vr_lbvh_build(const vr_triangle* d_prims, size_t n, uint32_t flags, void* stream, vr_bvh* out)
{
    VR_TRY_BEGIN
    VR_CHECK(out, VR_ERR_INVALID_ARGUMENT, "vr_lbvh_build: null out handle");
    VR_CHECK(n == 0 || d_prims, VR_ERR_INVALID_ARGUMENT, "vr_lbvh_build: null primitives");
    int rc = ensure_device();
    if (rc) return rc;
    vr_bvh h = new vr_bvh_t;
    h->s.device = current_device();
    rc = lbvh_build_into(h->s, d_prims, n, (flags & VR_BVH_KEEP_NODES) != 0, (flags & VR_BVH_ROTATE) != 0,
                         (hipStream_t)stream);
    if (rc == VR_OK) rc = mark_built(h, (hipStream_t)stream);
    if (rc) {
        vr_bvh_destroy(h);
        return rc;
    }
    *out = h;
    return VR_OK;
    VR_TRY_END
}

int vr_lbvh_rebuild(vr_bvh bvh, const vr_triangle* d_prims, size_t n, void* stream)
{
    VR_TRY_BEGIN
    VR_CHECK(bvh, VR_ERR_INVALID_ARGUMENT, "vr_lbvh_rebuild: null handle");
    VR_CHECK(bvh->s.is_lbvh || bvh->s.num_prims == 0, VR_ERR_UNSUPPORTED, "vr_lbvh_rebuild: not an LBVH handle");
    VR_CHECK(n == 0 || d_prims, VR_ERR_INVALID_ARGUMENT, "vr_lbvh_rebuild: null primitives");
    int rc = ensure_device();
    if (rc) return rc;
    rc = lbvh_build_into(bvh->s, d_prims, n, bvh->s.ref_nodes != nullptr, bvh->s.want_rotate, (hipStream_t)stream);
    if (rc == VR_OK) rc = collapse_wide(bvh->s, (hipStream_t)stream);
    return rc ? rc : mark_built(bvh, (hipStream_t)stream);
    VR_TRY_END
}

int vr_lbvh_morton(const vr_triangle* d_prims, size_t n, uint32_t* d_codes, void* stream)
{
    VR_CHECK(n == 0 || (d_prims && d_codes), VR_ERR_INVALID_ARGUMENT, "vr_lbvh_morton: null pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return lbvh_morton_only(d_prims, n, d_codes, (hipStream_t)stream);
}

int vr_bvh_sorted_refs(vr_bvh bvh, uint32_t* d_codes, int32_t* d_ids, void* stream)
{
    VR_CHECK(bvh, VR_ERR_INVALID_ARGUMENT, "vr_bvh_sorted_refs: null handle");
    VR_CHECK(bvh->s.is_lbvh, VR_ERR_UNSUPPORTED, "vr_bvh_sorted_refs: not an LBVH");
    const size_t n = bvh->s.num_prims;
    if (n == 0) return VR_OK;
    if (d_codes)
        VR_HIP_CHECK(hipMemcpyAsync(d_codes, bvh->s.sorted_codes, n * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    if (d_ids)
        VR_HIP_CHECK(hipMemcpyAsync(d_ids, bvh->s.sorted_ids, n * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return VR_OK;
}

// Adopt a reference-layout tree: inner nodes become pair records in
// depth-first order, leaf triangles are copied into traversal order.
int vr_bvh_from_nodes(const vr_triangle* h_prims, size_t num_prims, const vr_bvh_node* h_nodes, size_t num_nodes,
                      const uint32_t* h_indices, vr_bvh* out)
{
    VR_TRY_BEGIN
    VR_CHECK(out, VR_ERR_INVALID_ARGUMENT, "vr_bvh_from_nodes: null out handle");
    VR_CHECK(num_prims == 0 || (h_prims && h_nodes && h_indices && num_nodes > 0), VR_ERR_INVALID_ARGUMENT,
             "vr_bvh_from_nodes: null array");
    VR_CHECK(num_prims <= (size_t)MAX_LEAF_FIRST, VR_ERR_UNSUPPORTED, "vr_bvh_from_nodes: too many primitives");
    // the tree is checked and converted on the host before any device call
    std::vector<pair_node> pairs;
    std::vector<tri_node> tris;
    tris.reserve(num_prims);
    uint32_t depth = 0, max_leaf = 1, root = 0;
    if (num_prims > 0) {
        // iterative DFS over (node, depth); inner nodes get pair indices in visit order
        struct item {
            uint32_t node;
            uint32_t depth;
            int64_t parent_pair;  // pair whose child slot receives this node
            uint32_t side;
        };
        std::vector<item> stack{{0u, 1u, -1, 0u}};
        std::vector<bool> seen(num_nodes, false);
        // append the leaf's triangles in index order (ranges checked by the caller)
        auto add_leaf = [&](const vr_bvh_node& nd) -> uint32_t {
            const uint32_t first = (uint32_t)tris.size();
            for (uint32_t k = nd.first; k < nd.first + nd.num_prims; ++k) {
                const uint32_t id = h_indices[k];
                const vr_triangle& t = h_prims[id];
                float a4, b4, c4;
                std::memcpy(&a4, &t.prim_id, 4);
                std::memcpy(&b4, &t.geom_id, 4);
                std::memcpy(&c4, &id, 4);
                tri_node tn;
                tn.a = make_float4(t.v1[0], t.v1[1], t.v1[2], a4);
                tn.b = make_float4(t.e1[0], t.e1[1], t.e1[2], b4);
                tn.c = make_float4(t.e2[0], t.e2[1], t.e2[2], c4);
                tris.push_back(tn);
            }
            max_leaf = std::max<uint32_t>(max_leaf, nd.num_prims);
            return leaf_entry(first, nd.num_prims);
        };
        while (!stack.empty()) {
            item it = stack.back();
            stack.pop_back();
            VR_CHECK(it.node < num_nodes, VR_ERR_INVALID_ARGUMENT, "vr_bvh_from_nodes: child index out of range");
            VR_CHECK(it.depth <= 4096, VR_ERR_INVALID_ARGUMENT, "vr_bvh_from_nodes: tree too deep (cycle?)");
            // a tree reaches every node once: a node reached twice means shared
            // subtrees (a DAG, which the DFS could expand exponentially) or a cycle
            VR_CHECK(!seen[it.node], VR_ERR_INVALID_ARGUMENT, "vr_bvh_from_nodes: node reached twice (not a tree)");
            seen[it.node] = true;
            const vr_bvh_node& nd = h_nodes[it.node];
            depth = std::max(depth, it.depth);
            uint32_t entry;
            if (nd.num_prims) {
                VR_CHECK(nd.num_prims <= MAX_LEAF_COUNT, VR_ERR_UNSUPPORTED, "vr_bvh_from_nodes: leaf larger than 32");
                VR_CHECK((size_t)nd.first + nd.num_prims <= num_prims, VR_ERR_INVALID_ARGUMENT,
                         "vr_bvh_from_nodes: leaf range out of bounds");
                for (uint32_t k = nd.first; k < nd.first + nd.num_prims; ++k)
                    VR_CHECK(h_indices[k] < num_prims, VR_ERR_INVALID_ARGUMENT, "vr_bvh_from_nodes: bad index");
                entry = add_leaf(nd);
            } else {
                VR_CHECK((size_t)nd.first + 1 < num_nodes, VR_ERR_INVALID_ARGUMENT,
                         "vr_bvh_from_nodes: child index out of range");
                const uint32_t p = (uint32_t)pairs.size();
                pair_node pn;
                std::memset(&pn, 0, sizeof(pn));
                for (int side = 0; side < 2; ++side) {
                    const vr_bvh_node& c = h_nodes[nd.first + side];
                    for (int k = 0; k < 3; ++k) {
                        pn.f[side * 6 + k] = c.lo[k];
                        pn.f[side * 6 + 3 + k] = c.hi[k];
                    }
                }
                pn.e[2] = it.parent_pair < 0 ? 0xFFFFFFFFu : (uint32_t)((it.parent_pair << 1) | it.side);
                pairs.push_back(pn);
                entry = p;
                // push right first so the left subtree is numbered first
                stack.push_back({nd.first + 1, it.depth + 1, (int64_t)p, 1u});
                stack.push_back({nd.first, it.depth + 1, (int64_t)p, 0u});
            }
            if (it.parent_pair < 0) root = entry;
            else pairs[(size_t)it.parent_pair].e[it.side] = entry;
        }
    }
    int rc = ensure_device();
    if (rc) return rc;
    vr_bvh h = new vr_bvh_t;
    bvh_state& s = h->s;
    s.device = current_device();
    s.num_prims = num_prims;
    s.num_pairs = pairs.size();
    s.root = root;
    s.depth = depth;
    s.max_leaf = max_leaf;
    s.is_lbvh = false;
    auto bail = [&](int code) {
        vr_bvh_destroy(h);
        return code;
    };
    if (!stack_fits(s)) return bail(fail(VR_ERR_UNSUPPORTED, "vr_bvh_from_nodes: tree deeper than the traversal stack"));
    if (num_prims > 0) {
        if (hipMalloc(&s.pairs, std::max<size_t>(pairs.size(), 1) * sizeof(pair_node)) != hipSuccess ||
            hipMalloc(&s.tris, tris.size() * sizeof(tri_node)) != hipSuccess ||
            hipMalloc(&s.prims, num_prims * sizeof(vr_triangle)) != hipSuccess ||
            hipMalloc(&s.ref_nodes, num_nodes * sizeof(vr_bvh_node)) != hipSuccess ||
            hipMalloc(&s.indices, num_prims * 4) != hipSuccess)
            return bail(fail(VR_ERR_OUT_OF_MEMORY, "vr_bvh_from_nodes: device allocation failed"));
        s.num_ref_nodes = num_nodes;
        if ((pairs.size() && hipMemcpy(s.pairs, pairs.data(), pairs.size() * sizeof(pair_node), hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(s.tris, tris.data(), tris.size() * sizeof(tri_node), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s.prims, h_prims, num_prims * sizeof(vr_triangle), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s.ref_nodes, h_nodes, num_nodes * sizeof(vr_bvh_node), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s.indices, h_indices, num_prims * 4, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(VR_ERR_HIP, "vr_bvh_from_nodes: upload failed"));
        const int rc = collapse_quads(s, 0u, aux_stream());
        if (rc == VR_OK && hipStreamSynchronize(aux_stream()) != hipSuccess)
            return bail(fail(VR_ERR_HIP, "vr_bvh_from_nodes: collapse failed"));
        if (rc) return bail(rc);
    }
    *out = h;
    return VR_OK;
    VR_TRY_END
}

int vr_bvh_refit(vr_bvh bvh, const vr_triangle* d_prims, void* stream)
{
    VR_CHECK(bvh && (bvh->s.num_prims == 0 || d_prims), VR_ERR_INVALID_ARGUMENT, "vr_bvh_refit: null argument");
    VR_CHECK(bvh->s.is_lbvh, VR_ERR_UNSUPPORTED, "vr_bvh_refit: only LBVHs are refitted on the device");
    int rc = lbvh_refit(bvh->s, d_prims, (hipStream_t)stream);
    if (rc == VR_OK) rc = collapse_wide(bvh->s, (hipStream_t)stream);
    return rc ? rc : mark_built(bvh, (hipStream_t)stream);
}

int vr_bvh_set_width(vr_bvh bvh, int width, uint32_t flags, void* stream)
{
    VR_TRY_BEGIN
    VR_CHECK(bvh, VR_ERR_INVALID_ARGUMENT, "vr_bvh_set_width: null handle");
    VR_CHECK(width == 4 || width == 8, VR_ERR_INVALID_ARGUMENT, "vr_bvh_set_width: width is 4 or 8");
    VR_CHECK((flags & ~(VR_WIDE_COMPRESSED | VR_WIDE_BINARY_LEAVES)) == 0, VR_ERR_INVALID_ARGUMENT,
             "vr_bvh_set_width: unknown flags");
    bvh_state& s = bvh->s;
    int cur = 0;
    VR_HIP_CHECK(hipGetDevice(&cur));
    VR_CHECK(cur == s.device, VR_ERR_INVALID_ARGUMENT, "vr_bvh_set_width: the BVH lives on another device");
    if (width == 4) {
        if (s.wide) VR_HIP_CHECK(hipFreeAsync(s.wide, (hipStream_t)stream));
        s.wide = nullptr;
        s.wide_bytes = 0;
        s.wide_depth = 0;
        s.wide_kind = NODE_QUAD;
        return VR_OK;
    }
    s.wide_kind = (flags & VR_WIDE_COMPRESSED) ? NODE_OCTQ : NODE_OCT;
    s.wide_binary_leaves = (flags & VR_WIDE_BINARY_LEAVES) != 0;
    int rc = collapse_wide(s, (hipStream_t)stream);
    if (rc == VR_OK) rc = mark_built(bvh, (hipStream_t)stream);
    return rc;
    VR_TRY_END
}

int vr_bvh_download_wide(vr_bvh bvh, void* h_out, size_t capacity, size_t* bytes, uint32_t* depth)
{
    VR_TRY_BEGIN
    VR_CHECK(bvh && bytes, VR_ERR_INVALID_ARGUMENT, "vr_bvh_download_wide: null argument");
    const bvh_state& s = bvh->s;
    const size_t node_bytes = s.wide_kind == NODE_OCT ? sizeof(oct_node) : sizeof(octq_node);
    *bytes = s.wide ? s.num_pairs * node_bytes : 0;
    if (depth) *depth = s.wide_depth;
    if (!h_out || *bytes == 0) return VR_OK;
    VR_CHECK(capacity >= *bytes, VR_ERR_INVALID_ARGUMENT, "vr_bvh_download_wide: buffer too small");
    return read_after_build(bvh, h_out, s.wide, *bytes);
    VR_TRY_END
}

int vr_bvh_get_info(vr_bvh bvh, vr_bvh_info* info)
{
    VR_CHECK(bvh && info, VR_ERR_INVALID_ARGUMENT, "vr_bvh_get_info: null argument");
    const bvh_state& s = bvh->s;
    info->num_prims = s.num_prims;
    info->num_nodes = s.is_lbvh ? (s.num_prims ? 2 * s.num_prims - 1 : 0) : s.num_ref_nodes;
    info->max_leaf_size = s.max_leaf;
    info->depth = s.depth;
    if (s.is_lbvh && s.num_prims > 1) {
        // written by the build: wait for the build's stream up to the build
        uint32_t d = 0;
        const int rc = read_after_build(bvh, &d, s.root_box + 6, 4);
        if (rc) return rc;
        info->depth = d;
    }
    return VR_OK;
}

int vr_bvh_get_ref(vr_bvh bvh, vr_bvh_ref* out)
{
    VR_CHECK(bvh && out, VR_ERR_INVALID_ARGUMENT, "vr_bvh_get_ref: null argument");
    const bvh_state& s = bvh->s;
    VR_CHECK(stack_fits(s), VR_ERR_UNSUPPORTED, "vr_bvh_get_ref: BVH deeper than the traversal stack");
    const bvh_view v = s.view();
    out->nodes = v.quads;
    out->prims = v.tris;
    out->pairs = v.pairs;
    out->num_prims = v.num_prims;
    out->root = v.root;
    out->num_nodes = v.num_quads;
    out->reserved = 0;
    return VR_OK;
}

int vr_bvh_nodes(vr_bvh bvh, vr_bvh_node* d_nodes, void* stream)
{
    VR_CHECK(bvh && d_nodes, VR_ERR_INVALID_ARGUMENT, "vr_bvh_nodes: null argument");
    const bvh_state& s = bvh->s;
    if (s.num_prims == 0) return VR_OK;
    VR_CHECK(s.ref_nodes, VR_ERR_UNSUPPORTED, "vr_bvh_nodes: build with VR_BVH_KEEP_NODES to keep the node layout");
    const size_t nn = s.is_lbvh ? 2 * s.num_prims - 1 : s.num_ref_nodes;
    VR_HIP_CHECK(hipMemcpyAsync(d_nodes, s.ref_nodes, nn * sizeof(vr_bvh_node), hipMemcpyDeviceToDevice,
                                (hipStream_t)stream));
    return VR_OK;
}

int vr_bvh_indices(vr_bvh bvh, uint32_t* d_indices, void* stream)
{
    VR_CHECK(bvh && d_indices, VR_ERR_INVALID_ARGUMENT, "vr_bvh_indices: null argument");
    const bvh_state& s = bvh->s;
    if (s.num_prims == 0) return VR_OK;
    if (s.is_lbvh) return lbvh_indices(d_indices, s.num_prims, (hipStream_t)stream);
    VR_HIP_CHECK(hipMemcpyAsync(d_indices, s.indices, s.num_prims * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return VR_OK;
}

int vr_bvh_destroy_async(vr_bvh bvh, void* stream)
{
    if (!bvh) return VR_OK;
    bvh_state& s = bvh->s;
    void* ptrs[] = {s.pairs, s.quads, s.tris, s.prims, s.sorted_codes, s.sorted_ids, s.parents,
                    s.flags, s.heights, s.ref_nodes, s.indices, s.root_box, s.scratch};
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(s.device);
    if (s.wide) (void)hipFreeAsync(s.wide, (hipStream_t)stream);
    if (s.pooled) {
        // stream-ordered release: work enqueued on `stream` before this call
        // still sees the arrays; the caller orders other streams' use before it
        if (s.arena) {
            (void)hipFreeAsync(s.arena, (hipStream_t)stream);
        } else {
            for (void* p : ptrs)
                if (p) (void)hipFreeAsync(p, (hipStream_t)stream);
        }
    } else {
        // plain allocations: wait for `stream` only, then free
        (void)hipStreamSynchronize((hipStream_t)stream);
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    if (bvh->built) (void)hipEventDestroy(bvh->built);
    (void)hipSetDevice(cur);
    delete bvh;
    return VR_OK;
}

int vr_bvh_destroy(vr_bvh bvh)
{
    if (!bvh) return VR_OK;
    bvh_state& s = bvh->s;
    void* ptrs[] = {s.pairs, s.quads, s.tris, s.prims, s.sorted_codes, s.sorted_ids, s.parents,
                    s.flags, s.heights, s.ref_nodes, s.indices, s.root_box, s.scratch};
    if (s.pooled) {
        // back to the pool once the device is done with them: the caller may
        // still have work on the BVH in flight on any stream
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(s.device);
        (void)hipDeviceSynchronize();
        if (s.wide) (void)hipFreeAsync(s.wide, nullptr);
        if (s.arena) {
            (void)hipFreeAsync(s.arena, nullptr);
        } else {
            for (void* p : ptrs)
                if (p) (void)hipFreeAsync(p, nullptr);
        }
        (void)hipSetDevice(cur);
    } else {
        if (s.wide) {
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(s.device);
            (void)hipDeviceSynchronize();
            (void)hipFreeAsync(s.wide, nullptr);
            (void)hipSetDevice(cur);
        }
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    if (bvh->built) (void)hipEventDestroy(bvh->built);
    delete bvh;
    return VR_OK;
}

int vr_sah_build_device(const vr_triangle* d_prims, size_t n, int max_leaf_size, void* stream, vr_bvh* out)
{
    VR_TRY_BEGIN
    VR_CHECK(out, VR_ERR_INVALID_ARGUMENT, "vr_sah_build_device: null out handle");
    VR_CHECK(n == 0 || d_prims, VR_ERR_INVALID_ARGUMENT, "vr_sah_build_device: null primitives");
    int rc = ensure_device();
    if (rc) return rc;
    vr_bvh h = new vr_bvh_t;
    h->s.device = current_device();
    rc = sah_build_device(d_prims, n, max_leaf_size, h->s, (hipStream_t)stream);
    if (rc == VR_OK) rc = mark_built(h, (hipStream_t)stream);
    if (rc) {
        vr_bvh_destroy(h);
        return rc;
    }
    *out = h;
    return VR_OK;
    VR_TRY_END
}

int vr_sah_build(const vr_triangle* h_prims, size_t n, int max_leaf_size, vr_bvh_node* h_nodes, size_t* num_nodes,
                 uint32_t* h_indices)
{
    VR_TRY_BEGIN
    VR_CHECK(num_nodes && (n == 0 || (h_prims && h_nodes && h_indices)), VR_ERR_INVALID_ARGUMENT,
             "vr_sah_build: null argument");
    return sah_build_host(h_prims, n, max_leaf_size, h_nodes, num_nodes, h_indices);
    VR_TRY_END
}

int vr_lbvh_build_host(const vr_triangle* h_prims, size_t n, int max_leaf_size, vr_bvh_node* h_nodes,
                       size_t* num_nodes, uint32_t* h_indices)
{
    VR_TRY_BEGIN
    VR_CHECK(num_nodes && (n == 0 || (h_prims && h_nodes && h_indices)), VR_ERR_INVALID_ARGUMENT,
             "vr_lbvh_build_host: null argument");
    return lbvh_build_host(h_prims, n, max_leaf_size, h_nodes, num_nodes, h_indices);
    VR_TRY_END
}

int vr_bvh_sah_cost_host(const vr_bvh_node* h_nodes, size_t num_nodes, float ci, float cl, float cp, float* cost)
{
    VR_TRY_BEGIN
    VR_CHECK(h_nodes && cost && num_nodes > 0, VR_ERR_INVALID_ARGUMENT, "vr_bvh_sah_cost_host: null argument");
    return sah_cost_host(h_nodes, num_nodes, ci, cl, cp, cost);
    VR_TRY_END
}

int vr_bvh_sah_cost(vr_bvh bvh, float ci, float cl, float cp, double* cost, void* stream)
{
    VR_TRY_BEGIN
    VR_CHECK(bvh && cost, VR_ERR_INVALID_ARGUMENT, "vr_bvh_sah_cost: null argument");
    int rc = ensure_device();
    if (rc) return rc;
    return sah_cost_device(bvh->s, ci, cl, cp, cost, (hipStream_t)stream);
    VR_TRY_END
}

int vr_bvh_optimize_rotations_host(vr_bvh_node* h_nodes, size_t num_nodes, int* num_rotations)
{
    VR_TRY_BEGIN
    VR_CHECK(h_nodes && num_rotations && num_nodes > 0, VR_ERR_INVALID_ARGUMENT,
             "vr_bvh_optimize_rotations_host: null argument");
    return optimize_rotations_host(h_nodes, num_nodes, num_rotations);
    VR_TRY_END
}

int vr_trace(vr_bvh bvh, int any_hit, const vr_ray* d_rays, size_t n, vr_hit* d_hits, void* stream)
{
    VR_CHECK(bvh && (n == 0 || (d_rays && d_hits)), VR_ERR_INVALID_ARGUMENT, "vr_trace: null argument");
    int rc = ensure_device();
    if (rc) return rc;
    return trace_batch(bvh->s, any_hit != 0, d_rays, n, d_hits, nullptr, work_counter(), (hipStream_t)stream);
}

int vr_trace_noalloc(vr_bvh bvh, int any_hit, const vr_ray* d_rays, size_t n, vr_hit* d_hits, void* stream)
{
    VR_CHECK(bvh && (n == 0 || (d_rays && d_hits)), VR_ERR_INVALID_ARGUMENT, "vr_trace_noalloc: null argument");
    return trace_batch_noalloc(bvh->s, any_hit != 0, d_rays, n, d_hits, (hipStream_t)stream);
}

int vr_trace_counted(vr_bvh bvh, int any_hit, const vr_ray* d_rays, size_t n, vr_hit* d_hits, uint32_t* d_counts,
                     void* stream)
{
    VR_CHECK(bvh && (n == 0 || (d_rays && d_hits && d_counts)), VR_ERR_INVALID_ARGUMENT,
             "vr_trace_counted: null argument");
    int rc = ensure_device();
    if (rc) return rc;
    return trace_batch(bvh->s, any_hit != 0, d_rays, n, d_hits, d_counts, work_counter(), (hipStream_t)stream);
}

int vr_render_ao(vr_bvh bvh, const vr_camera* cam, const vr_sched_params* sp, uint32_t ao_samples, float ao_radius,
                 const float bg[3], float* d_accum, float* d_color, void* stream)
{
    VR_CHECK(bvh && cam && sp && bg && d_accum && d_color, VR_ERR_INVALID_ARGUMENT, "vr_render_ao: null argument");
    VR_CHECK(sp->spp >= 1 && ao_samples >= 1, VR_ERR_INVALID_ARGUMENT, "vr_render_ao: spp and samples must be >= 1");
    int rc = ensure_device();
    if (rc) return rc;
    return render_ao(bvh->s, *cam, *sp, ao_samples, ao_radius, bg, d_accum, d_color, (hipStream_t)stream);
}

int vr_render_pt(vr_bvh bvh, const vr_material* d_materials, size_t num_materials, const vr_kernel_params* kp,
                 const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color,
                 uint64_t* d_ray_counts, void* stream)
{
    VR_CHECK(bvh && kp && cam && sp && d_accum && d_color, VR_ERR_INVALID_ARGUMENT, "vr_render_pt: null argument");
    VR_CHECK(bvh->s.num_prims == 0 || (d_materials && num_materials > 0), VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt: no materials");
    VR_CHECK(sp->spp >= 1, VR_ERR_INVALID_ARGUMENT, "vr_render_pt: spp must be >= 1");
    VR_CHECK((uint64_t)sp->width * sp->height < (1ull << 31), VR_ERR_INVALID_ARGUMENT, "vr_render_pt: image too large");
    int rc = ensure_device();
    if (rc) return rc;
    return render_pt(bvh->s, d_materials, num_materials, *kp, *cam, *sp, d_accum, d_color,
                     reinterpret_cast<unsigned long long*>(d_ray_counts), work_counter(), false,
                     (hipStream_t)stream);
}

int vr_render_pt_lights(vr_bvh bvh, const vr_material* d_materials, size_t num_materials,
                        const vr_area_light* d_lights, size_t num_lights, const vr_kernel_params* kp,
                        const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color,
                        uint64_t* d_stats, int counted, void* stream)
{
    VR_CHECK(bvh && kp && cam && sp && d_accum && d_color && (d_stats || !counted), VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_lights: null argument");
    VR_CHECK(num_lights == 0 || d_lights, VR_ERR_INVALID_ARGUMENT, "vr_render_pt_lights: null light list");
    VR_CHECK(bvh->s.num_prims == 0 || (d_materials && num_materials > 0), VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_lights: no materials");
    VR_CHECK(sp->spp >= 1, VR_ERR_INVALID_ARGUMENT, "vr_render_pt_lights: spp must be >= 1");
    VR_CHECK((uint64_t)sp->width * sp->height < (1ull << 31), VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_lights: image too large");
    int rc = ensure_device();
    if (rc) return rc;
    return render_pt_lights(bvh->s, d_materials, num_materials, d_lights, num_lights, *kp, *cam, *sp, d_accum,
                            d_color, reinterpret_cast<unsigned long long*>(d_stats), work_counter(), counted != 0,
                            (hipStream_t)stream);
}

int vr_render_pt_counted(vr_bvh bvh, const vr_material* d_materials, size_t num_materials, const vr_kernel_params* kp,
                         const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color,
                         uint64_t* d_stats, void* stream)
{
    VR_CHECK(bvh && kp && cam && sp && d_accum && d_color && d_stats, VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_counted: null argument");
    VR_CHECK(bvh->s.num_prims == 0 || (d_materials && num_materials > 0), VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_counted: no materials");
    VR_CHECK(sp->spp >= 1, VR_ERR_INVALID_ARGUMENT, "vr_render_pt_counted: spp must be >= 1");
    VR_CHECK((uint64_t)sp->width * sp->height < (1ull << 31), VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_counted: image too large");
    int rc = ensure_device();
    if (rc) return rc;
    return render_pt(bvh->s, d_materials, num_materials, *kp, *cam, *sp, d_accum, d_color,
                     reinterpret_cast<unsigned long long*>(d_stats), work_counter(), true, (hipStream_t)stream);
}

int vr_image_l2_loss(const float* d_image, const float* d_target, size_t num_pixels, float* d_dl_dimage,
                     double* d_loss, void* stream)
{
    VR_CHECK(num_pixels == 0 || (d_image && d_target && d_loss), VR_ERR_INVALID_ARGUMENT,
             "vr_image_l2_loss: null argument");
    return image_l2_loss(d_image, d_target, num_pixels, d_dl_dimage, d_loss, (hipStream_t)stream);
}

int vr_render_pt_backward(vr_bvh bvh, const vr_material* d_materials, size_t num_materials,
                          const vr_kernel_params* kp, const vr_camera* cam, const vr_sched_params* sp,
                          const float* d_dl_dimage, double* d_grad, uint64_t* d_ray_counts, void* stream)
{
    VR_CHECK(bvh && kp && cam && sp && d_dl_dimage && d_grad, VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_backward: null argument");
    VR_CHECK(d_materials && num_materials > 0, VR_ERR_INVALID_ARGUMENT, "vr_render_pt_backward: no materials");
    VR_CHECK(sp->spp >= 1, VR_ERR_INVALID_ARGUMENT, "vr_render_pt_backward: spp must be >= 1");
    VR_CHECK((uint64_t)sp->width * sp->height < (1ull << 31), VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_backward: image too large");
    int rc = ensure_device();
    if (rc) return rc;
    return render_pt_adjoint(bvh->s, d_materials, num_materials, *kp, *cam, *sp, d_dl_dimage, d_grad,
                             reinterpret_cast<unsigned long long*>(d_ray_counts), work_counter(),
                             (hipStream_t)stream);
}

int vr_render_pt_frozen_rr(vr_bvh bvh, const vr_material* d_materials, const vr_material* d_rr_materials,
                           size_t num_materials, const vr_kernel_params* kp, const vr_camera* cam,
                           const vr_sched_params* sp, float* d_accum, float* d_color, uint64_t* d_ray_counts,
                           void* stream)
{
    VR_CHECK(bvh && kp && cam && sp && d_accum && d_color, VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_frozen_rr: null argument");
    VR_CHECK(bvh->s.num_prims == 0 || (d_materials && d_rr_materials && num_materials > 0), VR_ERR_INVALID_ARGUMENT,
             "vr_render_pt_frozen_rr: no materials");
    VR_CHECK(sp->spp >= 1, VR_ERR_INVALID_ARGUMENT, "vr_render_pt_frozen_rr: spp must be >= 1");
    int rc = ensure_device();
    if (rc) return rc;
    return render_pt_frozen_rr(bvh->s, d_materials, d_rr_materials, num_materials, *kp, *cam, *sp, d_accum, d_color,
                               reinterpret_cast<unsigned long long*>(d_ray_counts), work_counter(),
                               (hipStream_t)stream);
}

int vr_batch_create(const vr_scene* h_scenes, size_t num_scenes, vr_batch* out)
{
    VR_TRY_BEGIN
    VR_CHECK(out && (num_scenes == 0 || h_scenes), VR_ERR_INVALID_ARGUMENT, "vr_batch_create: null argument");
    VR_CHECK(num_scenes < (1u << 20), VR_ERR_INVALID_ARGUMENT, "vr_batch_create: too many scenes");
    int rc = ensure_device();
    if (rc) return rc;
    const size_t dsz = scene_desc_size();
    std::vector<unsigned char> descs(std::max<size_t>(num_scenes, 1) * dsz);
    std::vector<uint32_t> prefix(num_scenes + 1, 0);
    uint64_t total = 0, mats = 0;
    uint32_t depth = 0;
    bool adjoint_ok = true;
    for (size_t i = 0; i < num_scenes; ++i) {
        const vr_scene& sc = h_scenes[i];
        VR_CHECK(sc.bvh && sc.d_accum && sc.d_color, VR_ERR_INVALID_ARGUMENT, "vr_batch_create: scene without bvh/target");
        VR_CHECK(sc.bvh->s.num_prims == 0 || (sc.d_materials && sc.num_materials > 0), VR_ERR_INVALID_ARGUMENT,
                 "vr_batch_create: scene without materials");
        VR_CHECK(sc.sp.spp >= 1, VR_ERR_INVALID_ARGUMENT, "vr_batch_create: spp must be >= 1");
        make_scene_desc(&descs[i * dsz], sc.bvh->s, sc.d_materials, (uint32_t)sc.num_materials, (uint32_t)mats,
                        sc.kp, sc.cam, sc.sp, sc.d_accum, sc.d_color, sc.d_dl_dimage);
        prefix[i] = (uint32_t)total;
        total += padded_pixel_count(sc.sp.width, sc.sp.height, g_shape.tile);
        mats += sc.num_materials;
        depth = std::max(depth, sc.bvh->s.depth);
        adjoint_ok = adjoint_ok && sc.d_dl_dimage != nullptr && sc.kp.num_bounces <= ADJ_MAX_BOUNCES;
    }
    VR_CHECK(total < (1ull << 31), VR_ERR_INVALID_ARGUMENT, "vr_batch_create: more than 2^31 pixels");
    VR_CHECK(3u * depth + 1u <= (uint32_t)MAX_STACK, VR_ERR_UNSUPPORTED,
             "vr_batch_create: BVH deeper than the traversal stack");
    prefix[num_scenes] = (uint32_t)total;
    vr_batch b = new vr_batch_t;
    b->device = current_device();
    b->num_scenes = (uint32_t)num_scenes;
    b->total_pixels = (uint32_t)total;
    b->total_materials = (uint32_t)std::min<uint64_t>(mats, 0xFFFFFFFFu);
    b->max_depth = depth;
    b->adjoint_ok = adjoint_ok;
    b->shape = g_shape;
    b->shape.schedule = VR_SCHEDULE_PERSISTENT;  // a batch is one persistent launch
    const uint32_t start = 0xFFFFFFFFu;
    if (hipMalloc(&b->d_scenes, descs.size()) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->d_prefix), prefix.size() * 4) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->d_frame), 256) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->d_work), 256) != hipSuccess ||
        hipMemcpy(b->d_scenes, descs.data(), descs.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_prefix, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_frame, &start, 4, hipMemcpyHostToDevice) != hipSuccess) {
        vr_batch_destroy(b);
        return fail(VR_ERR_OUT_OF_MEMORY, "vr_batch_create: device allocation or upload failed");
    }
    *out = b;
    return VR_OK;
    VR_TRY_END
}

int vr_batch_render(vr_batch batch, uint64_t* d_ray_counts, void* stream)
{
    VR_CHECK(batch, VR_ERR_INVALID_ARGUMENT, "vr_batch_render: null handle");
    return render_pt_batch(batch->d_scenes, batch->d_prefix, batch->num_scenes, batch->total_pixels, batch->d_frame,
                           true, reinterpret_cast<unsigned long long*>(d_ray_counts), batch->d_work, 0, nullptr,
                           batch->shape, (hipStream_t)stream);
}

int vr_batch_backward(vr_batch batch, double* d_grad, uint64_t* d_ray_counts, void* stream)
{
    VR_CHECK(batch && d_grad, VR_ERR_INVALID_ARGUMENT, "vr_batch_backward: null argument");
    VR_CHECK(batch->adjoint_ok, VR_ERR_INVALID_ARGUMENT,
             "vr_batch_backward: every scene needs d_dl_dimage and at most 16 bounces");
    VR_CHECK(batch->total_materials <= ADJ_MAX_MATERIALS_TOTAL, VR_ERR_UNSUPPORTED,
             "vr_batch_backward: more than 2048 materials over all scenes");
    return render_pt_batch(batch->d_scenes, batch->d_prefix, batch->num_scenes, batch->total_pixels, batch->d_frame,
                           false, reinterpret_cast<unsigned long long*>(d_ray_counts), batch->d_work,
                           batch->total_materials, d_grad, batch->shape, (hipStream_t)stream);
}

int vr_set_launch_shape(const vr_launch_shape* shape)
{
    VR_CHECK(shape, VR_ERR_INVALID_ARGUMENT, "vr_set_launch_shape: null argument");
    VR_CHECK(valid_shape(*shape), VR_ERR_INVALID_ARGUMENT,
             "vr_set_launch_shape: block_size 128|256, stack_lds 8|16|32, tile 1..64, schedule persistent|per-pixel "
             "(per-pixel with stack_lds 16)");
    g_shape = *shape;
    return VR_OK;
}

int vr_get_launch_shape(vr_launch_shape* shape)
{
    VR_CHECK(shape, VR_ERR_INVALID_ARGUMENT, "vr_get_launch_shape: null argument");
    *shape = g_shape;
    return VR_OK;
}

int vr_autotune_pt(vr_bvh bvh, const vr_material* d_materials, size_t num_materials, const vr_kernel_params* kp,
                   const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color, int reps,
                   vr_launch_shape* best, float* best_ms, void* stream)
{
    VR_CHECK(bvh && kp && cam && sp && d_accum && d_color && best, VR_ERR_INVALID_ARGUMENT,
             "vr_autotune_pt: null argument");
    VR_CHECK(bvh->s.num_prims == 0 || (d_materials && num_materials > 0), VR_ERR_INVALID_ARGUMENT,
             "vr_autotune_pt: no materials");
    VR_CHECK(sp->spp >= 1, VR_ERR_INVALID_ARGUMENT, "vr_autotune_pt: spp must be >= 1");
    int rc = ensure_device();
    if (rc) return rc;
    reps = std::max(1, reps);
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    VR_HIP_CHECK(hipEventCreate(&e0));
    VR_HIP_CHECK(hipEventCreate(&e1));
    float best_t = 1e30f;
    vr_launch_shape best_s = g_shape;
    const uint32_t blocks[] = {128, 256}, stacks[] = {8, 16, 32}, tiles[] = {1, 8, 16};
    for (uint32_t b : blocks)
        for (uint32_t d : stacks)
            for (uint32_t t : tiles) {
                const vr_launch_shape shp{b, d, t, VR_SCHEDULE_PERSISTENT};
                rc = render_pt_shaped(bvh->s, d_materials, num_materials, *kp, *cam, *sp, d_accum, d_color, nullptr,
                                      work_counter(), false, shp, st);
                if (rc) break;
                VR_HIP_CHECK(hipEventRecord(e0, st));
                for (int r = 0; r < reps && rc == 0; ++r)
                    rc = render_pt_shaped(bvh->s, d_materials, num_materials, *kp, *cam, *sp, d_accum, d_color,
                                          nullptr, work_counter(), false, shp, st);
                VR_HIP_CHECK(hipEventRecord(e1, st));
                VR_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                VR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                ms /= (float)reps;
                if (rc == 0 && ms < best_t) {
                    best_t = ms;
                    best_s = shp;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    g_shape = best_s;
    *best = best_s;
    if (best_ms) *best_ms = best_t;
    return VR_OK;
}

int vr_batch_capture_graph(vr_batch batch, uint64_t* d_ray_counts, void* stream)
{
    VR_CHECK(batch, VR_ERR_INVALID_ARGUMENT, "vr_batch_capture_graph: null handle");
    VR_CHECK(stream != nullptr, VR_ERR_INVALID_ARGUMENT, "vr_batch_capture_graph: capture needs a created stream");
    hipStream_t st = (hipStream_t)stream;
    if (batch->exec) {
        (void)hipGraphExecDestroy(batch->exec);
        batch->exec = nullptr;
    }
    if (batch->graph) {
        (void)hipGraphDestroy(batch->graph);
        batch->graph = nullptr;
    }
    VR_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = vr_batch_render(batch, d_ray_counts, stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    VR_HIP_CHECK(e);
    batch->graph = g;
    VR_HIP_CHECK(hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0));
    return VR_OK;
}

int vr_batch_launch_graph(vr_batch batch, void* stream)
{
    VR_CHECK(batch && batch->exec, VR_ERR_INVALID_ARGUMENT, "vr_batch_launch_graph: no captured graph");
    VR_HIP_CHECK(hipGraphLaunch(batch->exec, (hipStream_t)stream));
    return VR_OK;
}

int vr_batch_destroy(vr_batch batch)
{
    if (!batch) return VR_OK;
    if (batch->exec) (void)hipGraphExecDestroy(batch->exec);
    if (batch->graph) (void)hipGraphDestroy(batch->graph);
    void* ptrs[] = {batch->d_scenes, batch->d_prefix, batch->d_frame, batch->d_work};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete batch;
    return VR_OK;
}

// ---- the minimum export set named for the drop-in boundary ----

struct vr_context_t {
    int device = 0;
};

int vr_create_context(int device, vr_context* out)
{
    VR_TRY_BEGIN
    VR_CHECK(out, VR_ERR_INVALID_ARGUMENT, "vr_create_context: null out handle");
    const int rc = vr_init(device);
    if (rc) return rc;
    vr_context c = new vr_context_t;
    c->device = device;
    *out = c;
    return VR_OK;
    VR_TRY_END
}

int vr_destroy_context(vr_context ctx)
{
    delete ctx;
    return VR_OK;
}

int vr_scene_upload(const vr_triangle* h_prims, size_t n, vr_triangle** d_out, void* stream)
{
    VR_CHECK(d_out && (n == 0 || h_prims), VR_ERR_INVALID_ARGUMENT, "vr_scene_upload: null argument");
    *d_out = nullptr;
    if (n == 0) return VR_OK;
    int rc = ensure_device();
    if (rc) return rc;
    void* d = nullptr;
    VR_HIP_CHECK(hipMalloc(&d, n * sizeof(vr_triangle)));
    if (hipMemcpyAsync(d, h_prims, n * sizeof(vr_triangle), hipMemcpyHostToDevice, (hipStream_t)stream) != hipSuccess ||
        hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
        (void)hipFree(d);
        return fail(VR_ERR_HIP, "vr_scene_upload: copy failed");
    }
    *d_out = static_cast<vr_triangle*>(d);
    return VR_OK;
}

int vr_bvh_download_nodes(vr_bvh bvh, vr_bvh_node* h_nodes, size_t* num_nodes)
{
    VR_CHECK(bvh && num_nodes, VR_ERR_INVALID_ARGUMENT, "vr_bvh_download_nodes: null argument");
    const bvh_state& s = bvh->s;
    const size_t nn = s.num_prims == 0 ? 0 : (s.is_lbvh ? 2 * s.num_prims - 1 : s.num_ref_nodes);
    if (!h_nodes) {
        *num_nodes = nn;
        return VR_OK;
    }
    VR_CHECK(*num_nodes >= nn, VR_ERR_INVALID_ARGUMENT, "vr_bvh_download_nodes: buffer too small");
    *num_nodes = nn;
    if (nn == 0) return VR_OK;
    VR_CHECK(s.ref_nodes, VR_ERR_UNSUPPORTED,
             "vr_bvh_download_nodes: build with VR_BVH_KEEP_NODES to keep the node layout");
    return read_after_build(bvh, h_nodes, s.ref_nodes, nn * sizeof(vr_bvh_node));
}

int vr_trace_closest(vr_bvh bvh, const vr_ray* d_rays, size_t n, vr_hit* d_hits, void* stream)
{
    return vr_trace(bvh, 0, d_rays, n, d_hits, stream);
}

int vr_trace_any(vr_bvh bvh, const vr_ray* d_rays, size_t n, vr_hit* d_hits, void* stream)
{
    return vr_trace(bvh, 1, d_rays, n, d_hits, stream);
}

int vr_render_pt_forward(vr_bvh bvh, const vr_material* d_materials, size_t num_materials, const vr_kernel_params* kp,
                         const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color,
                         uint64_t* d_ray_counts, void* stream)
{
    return vr_render_pt(bvh, d_materials, num_materials, kp, cam, sp, d_accum, d_color, d_ray_counts, stream);
}

int vr_autotune(vr_bvh bvh, const vr_material* d_materials, size_t num_materials, const vr_kernel_params* kp,
                const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color, int reps,
                vr_launch_shape* best, float* best_ms, void* stream)
{
    return vr_autotune_pt(bvh, d_materials, num_materials, kp, cam, sp, d_accum, d_color, reps, best, best_ms, stream);
}

int vr_destroy_bvh(vr_bvh bvh) { return vr_bvh_destroy(bvh); }

int vr_destroy_batch(vr_batch batch) { return vr_batch_destroy(batch); }

const char* vr_last_error_string(void) { return vr_last_error(); }

}  // extern "C"
