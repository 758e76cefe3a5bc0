// mesh.hip -- triangles of a loaded model built on the device from its
// indexed form: the host hands over positions (12 B per vertex), indices and
// material ids (16 B per triangle); one streaming kernel writes the
// reference's 44-byte triangle records (basic_triangle, math/triangle.h) and
// the geometric normals the loaders compute (obj_loader.cpp:506-511,
// ply_loader.cpp:169).
//
// A block owns 256 consecutive triangles.  Their indices are read with
// coalesced 16-byte loads into LDS, each lane gathers its three vertices and
// assembles its record in LDS (an 11-dword stride, odd, so no bank
// conflicts), and the block then writes its 256 x 44 bytes of records and
// 256 x 12 bytes of normals as contiguous 16-byte stores (dwords for a
// partial last block or unaligned arrays).  Bound by HBM: 16 B read and
// 56 B written per triangle, plus the vertex gathers (12 B per vertex once,
// the rest from L2).
#include "mesh.h"

#include <algorithm>

namespace vr {

namespace {

constexpr int BLOCK = 256;
constexpr int TRI_DW = 11;  // dwords of one vr_triangle

static_assert(sizeof(vr_triangle) == TRI_DW * 4, "triangle record is 11 dwords");

__global__ __launch_bounds__(BLOCK) void k_assemble(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                                                    const uint32_t* __restrict__ geom, uint32_t first, uint32_t n,
                                                    int normalized, int vec, uint32_t* __restrict__ out,
                                                    float* __restrict__ normals)
{
    __shared__ uint4 s_idx4[3 * BLOCK / 4];
    __shared__ uint4 s_tri4[TRI_DW * BLOCK / 4];
    __shared__ uint4 s_n4[3 * BLOCK / 4];
    uint32_t* s_idx = reinterpret_cast<uint32_t*>(s_idx4);
    uint32_t* s_tri = reinterpret_cast<uint32_t*>(s_tri4);
    float* s_n = reinterpret_cast<float*>(s_n4);
    const uint32_t base = first + blockIdx.x * BLOCK;  // first triangle of the block
    const uint32_t cnt = min((uint32_t)BLOCK, first + n - base);
    const uint32_t t = threadIdx.x;
    // a full block with 16-byte aligned rows moves 16-byte vectors; the last
    // (partial) block, or unaligned arrays, move dwords
    const bool v4 = vec && cnt == BLOCK;
    if (v4) {
        const uint4* src = reinterpret_cast<const uint4*>(idx + 3 * (size_t)base);
        for (uint32_t k = t; k < 3 * BLOCK / 4; k += BLOCK) s_idx4[k] = src[k];
    } else {
        for (uint32_t k = t; k < 3 * cnt; k += BLOCK) s_idx[k] = idx[3 * (size_t)base + k];
    }
    __syncthreads();
    if (t < cnt) {
        const uint32_t a = s_idx[3 * t], b = s_idx[3 * t + 1], c = s_idx[3 * t + 2];
        const f3 v1 = mk3(pos[3 * (size_t)a], pos[3 * (size_t)a + 1], pos[3 * (size_t)a + 2]);
        const f3 v2 = mk3(pos[3 * (size_t)b], pos[3 * (size_t)b + 1], pos[3 * (size_t)b + 2]);
        const f3 v3 = mk3(pos[3 * (size_t)c], pos[3 * (size_t)c + 1], pos[3 * (size_t)c + 2]);
        const f3 e1 = sub3(v2, v1);
        const f3 e2 = sub3(v3, v1);
        uint32_t* r = s_tri + TRI_DW * t;
        r[0] = geom[base + t];
        r[1] = base + t;  // prim_id: the triangle's index in the model
        r[2] = __float_as_uint(v1.x);
        r[3] = __float_as_uint(v1.y);
        r[4] = __float_as_uint(v1.z);
        r[5] = __float_as_uint(e1.x);
        r[6] = __float_as_uint(e1.y);
        r[7] = __float_as_uint(e1.z);
        r[8] = __float_as_uint(e2.x);
        r[9] = __float_as_uint(e2.y);
        r[10] = __float_as_uint(e2.z);
        if (normals) {
            const f3 cr = cross3(e1, e2);
            const f3 nn = normalized ? normalize3(cr) : cr;
            s_n[3 * t] = nn.x;
            s_n[3 * t + 1] = nn.y;
            s_n[3 * t + 2] = nn.z;
        }
    }
    __syncthreads();
    if (v4) {
        uint4* o = reinterpret_cast<uint4*>(out + (size_t)TRI_DW * base);
        for (uint32_t k = t; k < TRI_DW * BLOCK / 4; k += BLOCK) o[k] = s_tri4[k];
        if (normals) {
            uint4* on = reinterpret_cast<uint4*>(normals + 3 * (size_t)base);
            for (uint32_t k = t; k < 3 * BLOCK / 4; k += BLOCK) on[k] = s_n4[k];
        }
    } else {
        uint32_t* o = out + (size_t)TRI_DW * base;
        for (uint32_t k = t; k < TRI_DW * cnt; k += BLOCK) o[k] = s_tri[k];
        if (normals) {
            float* on = normals + 3 * (size_t)base;
            for (uint32_t k = t; k < 3 * cnt; k += BLOCK) on[k] = s_n[k];
        }
    }
}

}  // namespace

int assemble_triangles(mesh_model& m, vr_triangle* d_prims, float* d_normals, hipStream_t st)
{
    const size_t nt = m.geom.size();
    if (nt == 0) return VR_OK;
    int dev = 0;
    VR_HIP_CHECK(hipGetDevice(&dev));
    if (!m.d_pos) {
        VR_HIP_CHECK(hipMalloc(&m.d_pos, m.pos.size() * sizeof(float)));
        VR_HIP_CHECK(hipMalloc(&m.d_idx, m.idx.size() * sizeof(uint32_t)));
        VR_HIP_CHECK(hipMalloc(&m.d_geom, m.geom.size() * sizeof(uint32_t)));
        m.device = dev;
        VR_HIP_CHECK(hipMemcpyAsync(m.d_pos, m.pos.data(), m.pos.size() * sizeof(float), hipMemcpyHostToDevice, st));
        VR_HIP_CHECK(hipMemcpyAsync(m.d_idx, m.idx.data(), m.idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        VR_HIP_CHECK(
            hipMemcpyAsync(m.d_geom, m.geom.data(), m.geom.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
    VR_CHECK(m.device == dev, VR_ERR_INVALID_ARGUMENT, "vr_mesh_upload: the mesh was uploaded to another device");
    for (const mesh_span& s : m.spans) {
        const unsigned blocks = (unsigned)div_up(s.count, BLOCK);
        // 16-byte rows: the span starts at a multiple of 4 triangles (44 and
        // 12 bytes each) and the output arrays are 16-byte aligned
        const bool vec = s.first % 4 == 0 && ((uintptr_t)d_prims % 16) == 0 && ((uintptr_t)d_normals % 16) == 0;
        hipLaunchKernelGGL(k_assemble, dim3(blocks), dim3(BLOCK), 0, st, m.d_pos, m.d_idx, m.d_geom,
                           (uint32_t)s.first, (uint32_t)s.count, s.normalized ? 1 : 0, vec ? 1 : 0,
                           reinterpret_cast<uint32_t*>(d_prims), d_normals);
        VR_HIP_CHECK(hipGetLastError());
    }
    return VR_OK;
}

}  // namespace vr
