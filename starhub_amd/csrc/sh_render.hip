// Solver maps without OpenGL: the reference renders its model surfels as
// one-pixel points with a depth test (Renderer::drawSolverMaps,
// libs/star/geometry/render/RendererDraw.cu:104-155; shaders/solver_map.vert:
// 21-53, solver_map.frag:1-17) and maps the colour attachments to CUDA.  Here:
//   k_zclear    the 64-bit z-buffer to all ones (cleared depth 1, no point);
//   k_splat     one thread per surfel: transform, cull, project, and an
//               atomicMin of (depth bits << 32 | surfel id) on its pixel;
//   k_resolve   one thread per pixel: the winner's reference vertex
//               (w = 1), normal (w = 0) and index, or the clear values.
// The depth is positive, so its float bits order as unsigned integers, and
// the minimum key is the nearest point with equal depths going to the lower
// id: the result of GL_LESS with points processed in order, independent of
// the order the atomics land in.
//
// Contraction is off: the transform and projection are rounded as written,
// as in the restatement they are checked against.
#include "sh_common.hpp"
#include "sh_kernels.hpp"

#pragma clang fp contract(off)

namespace sh
{
namespace
{
constexpr unsigned long long kEmptyKey = ~0ull;

__global__ void k_zclear(unsigned long long *__restrict__ zbuf, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        zbuf[i] = kEmptyKey;
}

struct RenderParams
{
    float m[12]; // world2camera rows 0..2, row-major
    float fx, fy, cx, cy;
    float max_depth, current_time, time_threshold;
    int width, height;
};

__global__ void k_splat(const float4 *__restrict__ vertex, const float4 *__restrict__ color_time, uint32_t n,
                        RenderParams rp, unsigned long long *__restrict__ zbuf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float4 v = vertex[i];
    const float px = ((rp.m[0] * v.x + rp.m[1] * v.y) + rp.m[2] * v.z) + rp.m[3];
    const float py = ((rp.m[4] * v.x + rp.m[5] * v.y) + rp.m[6] * v.z) + rp.m[7];
    const float pz = ((rp.m[8] * v.x + rp.m[9] * v.y) + rp.m[10] * v.z) + rp.m[11];
    if (pz > rp.max_depth || pz <= 0.f)
        return;
    // low-confidence surfels count only while recently observed (solver_map.vert:32-33)
    if (color_time && v.w < 1.f && fabsf(color_time[i].z - rp.current_time) >= rp.time_threshold)
        return;
    const float wx = ((rp.fx * px) / pz + rp.cx) + 0.5f;
    const float wy = ((rp.fy * py) / pz + rp.cy) + 0.5f;
    const float uf = floorf(wx), vf = floorf(wy);
    if (!(uf >= 0.f && uf < float(rp.width) && vf >= 0.f && vf < float(rp.height)))
        return;
    const float depth = (pz / rp.max_depth) * 0.5f + 0.5f;
    if (!(depth < 1.f))
        return; // depth 1 fails GL_LESS against the cleared depth
    const unsigned long long key = (static_cast<unsigned long long>(__float_as_uint(depth)) << 32) | i;
    atomicMin(zbuf + size_t(int(vf)) * rp.width + int(uf), key);
}

__global__ void k_resolve(const unsigned long long *__restrict__ zbuf, uint32_t n,
                          const float4 *__restrict__ vertex, const float4 *__restrict__ normal,
                          float4 *__restrict__ out_vertex, float4 *__restrict__ out_normal,
                          uint32_t *__restrict__ out_index)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const unsigned long long key = zbuf[p];
    if (key == kEmptyKey)
    {
        out_vertex[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        out_normal[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        out_index[p] = 0xffffffffu;
        return;
    }
    const uint32_t id = uint32_t(key & 0xffffffffull);
    const float4 v = vertex[id], nr = normal[id];
    out_vertex[p] = make_float4(v.x, v.y, v.z, 1.f);
    out_normal[p] = make_float4(nr.x, nr.y, nr.z, 0.f);
    out_index[p] = id;
}
} // namespace

hipError_t launch_render_solver_maps(const float *vertex4, const float *normal4, const float *color_time4,
                                     uint32_t num_surfels, const float world2camera[16], int width, int height,
                                     float fx, float fy, float cx, float cy, float max_depth, float current_time,
                                     bool with_recent_observation, unsigned long long *zbuf, float *out_vertex4,
                                     float *out_normal4, uint32_t *out_index, hipStream_t st)
{
    const uint32_t npix = uint32_t(width) * uint32_t(height);
    if (npix == 0)
        return hipSuccess;
    RenderParams rp{};
    for (int k = 0; k < 12; ++k)
        rp.m[k] = world2camera[k];
    rp.fx = fx;
    rp.fy = fy;
    rp.cx = cx;
    rp.cy = cy;
    rp.max_depth = max_depth;
    rp.current_time = current_time;
    rp.time_threshold = with_recent_observation ? 3.f : -1.f; // d_rendering_recent_time_threshold, constants.h:18
    rp.width = width;
    rp.height = height;
    hipLaunchKernelGGL(k_zclear, dim3((npix + 255) / 256), dim3(256), 0, st, zbuf, npix);
    if (num_surfels)
        hipLaunchKernelGGL(k_splat, dim3((num_surfels + 255) / 256), dim3(256), 0, st,
                           reinterpret_cast<const float4 *>(vertex4), reinterpret_cast<const float4 *>(color_time4),
                           num_surfels, rp, zbuf);
    hipLaunchKernelGGL(k_resolve, dim3((npix + 255) / 256), dim3(256), 0, st, zbuf, npix,
                       reinterpret_cast<const float4 *>(vertex4), reinterpret_cast<const float4 *>(normal4),
                       reinterpret_cast<float4 *>(out_vertex4), reinterpret_cast<float4 *>(out_normal4), out_index);
    return hipGetLastError();
}
} // namespace sh
