// Depth image -> the solver's measurement maps (the step before the solve):
//   k_filter_depth      filterUnreliableDepthKernel (libs/star/img_proc/surfel_reliable_filter.cu:37-154)
//   k_scaled_vertex     ComputeScaledVertexFromDepthKernel (scale 1) + filterAndScaleVertexKernel
//                       (libs/star/geometry/geometry_map/SurfelMapInitializer.cu:83-137)
//   k_normal_radius     createNormalRadiusMapKernel (libs/star/img_proc/normal_map.cu:32-117)
//   k_vertex_confid     createVertexConfidMapKernel (libs/star/img_proc/vertex_map.cu:13-33)
// The reference reads these images through textures with point sampling and
// zero border; here they are linear row-major arrays, and every 7x7 window
// is staged once per 16x16 tile in LDS (22x22 texels with the 3-texel apron).
// One thread per output pixel; the maps are a few hundred thousand pixels,
// so each kernel is a single pass over its input and output.
//
// Floating-point contraction is off in this file: every product and sum is
// rounded as written, as in the CPU restatement these kernels are checked
// against.
#include <cfloat>

#include "sh_common.hpp"
#include "sh_kernels.hpp"

#pragma clang fp contract(off)

namespace sh
{
namespace
{
constexpr int kMTile = 16;
constexpr int kApron = 3;
constexpr int kMSpan = kMTile + 2 * kApron; // 22

__device__ __forceinline__ bool zero_vertex(float x, float y, float z)
{
    // is_zero_vertex (vector_ops.hpp:122-126): compared with the double literal
    return fabs(double(x)) < 1e-3 && fabs(double(y)) < 1e-3 && fabs(double(z)) < 1e-3;
}

__device__ __forceinline__ float3 cross3(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// ---- smallest eigenvector of a packed PSD 3x3 (00 01 02 11 12 22): the
// closed-form solver of eigen33 (libs/star/math/eigen33.hpp): scale into
// [-1, 1], roots of the characteristic cubic by the trigonometric formula,
// eigenvectors as the longest row cross product of (A - lambda I).
__device__ void roots2(float b, float c, float3 &r)
{
    r.x = 0.f;
    float d = b * b - 4.f * c;
    if (d < 0.f)
        d = 0.f;
    const float sd = sqrtf(d);
    r.z = 0.5f * (b + sd);
    r.y = 0.5f * (b - sd);
}

__device__ void roots3(float c0, float c1, float c2, float3 &r)
{
    if (fabsf(c0) < FLT_EPSILON)
    {
        roots2(c2, c1, r);
        return;
    }
    const float inv3 = 1.f / 3.f;
    const float sqrt3 = sqrtf(3.f);
    const float c2_3 = c2 * inv3;
    float a_3 = (c1 - c2 * c2_3) * inv3;
    if (a_3 > 0.f)
        a_3 = 0.f;
    const float half_b = 0.5f * (c0 + c2_3 * (2.f * c2_3 * c2_3 - c1));
    float q = half_b * half_b + a_3 * a_3 * a_3;
    if (q > 0.f)
        q = 0.f;
    const float rho = sqrtf(-a_3);
    const float theta = atan2f(sqrtf(-q), half_b) * inv3;
    const float ct = cosf(theta), st = sinf(theta);
    r.x = c2_3 + 2.f * rho * ct;
    r.y = c2_3 - rho * (ct + sqrt3 * st);
    r.z = c2_3 - rho * (ct - sqrt3 * st);
    float t;
    if (r.x >= r.y)
    {
        t = r.x;
        r.x = r.y;
        r.y = t;
    }
    if (r.y >= r.z)
    {
        t = r.y;
        r.y = r.z;
        r.z = t;
        if (r.x >= r.y)
        {
            t = r.x;
            r.x = r.y;
            r.y = t;
        }
    }
    if (r.x <= 0.f)
        roots2(c2, c1, r);
}

__device__ __forceinline__ bool much_smaller(float x, float y)
{
    return x * x <= FLT_EPSILON * FLT_EPSILON * y * y;
}

__device__ float3 unit_orthogonal(float3 s)
{
    if (!much_smaller(s.x, s.z) || !much_smaller(s.y, s.z))
    {
        const float inv = rsqrtf(s.x * s.x + s.y * s.y);
        return make_float3(-s.y * inv, s.x * inv, 0.f);
    }
    const float inv = rsqrtf(s.z * s.z + s.y * s.y);
    return make_float3(0.f, -s.z * inv, s.y * inv);
}

__device__ __forceinline__ float3 normalized3(float3 v)
{
    const float inv = rsqrtf(dot3(v, v));
    return make_float3(v.x * inv, v.y * inv, v.z * inv);
}

// unit null vector of (A - lambda I); chosen: its squared length before
// normalisation; third: the third cross product's (read by the bookkeeping)
__device__ float3 null_vector(const float m[6], float lambda, float &chosen, float &third)
{
    const float3 r0 = make_float3(m[0] - lambda, m[1], m[2]);
    const float3 r1 = make_float3(m[1], m[3] - lambda, m[4]);
    const float3 r2 = make_float3(m[2], m[4], m[5] - lambda);
    const float3 v0 = cross3(r0, r1), v1 = cross3(r0, r2), v2 = cross3(r1, r2);
    const float l1 = dot3(v0, v0), l2 = dot3(v1, v1), l3 = dot3(v2, v2);
    third = l3;
    float3 v;
    if (l1 >= l2 && l1 >= l3)
    {
        chosen = l1;
        v = v0;
    }
    else if (l2 >= l1 && l2 >= l3)
    {
        chosen = l2;
        v = v1;
    }
    else
    {
        chosen = l3;
        v = v2;
    }
    const float inv = rsqrtf(chosen);
    return make_float3(v.x * inv, v.y * inv, v.z * inv);
}

__device__ float3 smallest_eigvec(const float cov[6])
{
    float scale = 0.f;
#pragma unroll
    for (int i = 0; i < 6; ++i)
        scale = fmaxf(scale, fabsf(cov[i]));
    if (scale <= FLT_MIN)
        scale = 1.f;
    float m[6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
        m[i] = cov[i] / scale;
    const float c0 =
        m[0] * m[3] * m[5] + 2.f * m[1] * m[2] * m[4] - m[0] * m[4] * m[4] - m[3] * m[2] * m[2] - m[5] * m[1] * m[1];
    const float c1 = m[0] * m[3] - m[1] * m[1] + m[0] * m[5] - m[2] * m[2] + m[3] * m[5] - m[4] * m[4];
    const float c2 = m[0] + m[3] + m[5];
    float3 ev;
    roots3(c0, c1, c2, ev);
    float chosen, third;
    if (ev.z - ev.x <= FLT_EPSILON)
        return make_float3(1.f, 0.f, 0.f);
    if (ev.y - ev.x <= FLT_EPSILON)
    {
        const float3 e2 = null_vector(m, ev.z, chosen, third);
        const float3 e1 = unit_orthogonal(e2);
        return cross3(e1, e2);
    }
    if (ev.z - ev.y <= FLT_EPSILON)
        return null_vector(m, ev.x, chosen, third);
    float3 e[3];
    float mmax[3];
    e[2] = null_vector(m, ev.z, mmax[2], third);
    e[1] = null_vector(m, ev.y, mmax[1], third);
    unsigned min_el = 2, max_el = 2;
    min_el = mmax[1] <= mmax[min_el] ? 1u : min_el;
    max_el = mmax[1] > mmax[max_el] ? 1u : max_el;
    e[0] = null_vector(m, ev.x, mmax[0], third);
    min_el = third <= mmax[min_el] ? 0u : min_el; // the third product, in every branch
    max_el = third > mmax[max_el] ? 0u : max_el;
    const unsigned mid_el = 3u - min_el - max_el;
    e[min_el] = normalized3(cross3(e[(min_el + 1) % 3], e[(min_el + 2) % 3]));
    e[mid_el] = normalized3(cross3(e[(mid_el + 1) % 3], e[(mid_el + 2) % 3]));
    return e[0];
}

// The 7x7 window normal around LDS cell (lx, ly) of a kMSpan-wide tile of
// float4 vertices (w unused): centroid of the valid points of the 7x7
// window, covariance over the 6x6 window [-3, +2]^2, smallest eigenvector
// oriented against the centre.  false: at most 24 valid points.
__device__ bool window_normal(const float4 *tile, int lx, int ly, float3 center, float3 &normal)
{
    float cx = 0.f, cy = 0.f, cz = 0.f;
    int counter = 0;
    for (int dy = -kApron; dy <= kApron; ++dy)
#pragma unroll
        for (int dx = -kApron; dx <= kApron; ++dx)
        {
            const float4 p = tile[(ly + dy) * kMSpan + lx + dx];
            if (!zero_vertex(p.x, p.y, p.z))
            {
                cx += p.x;
                cy += p.y;
                cz += p.z;
                ++counter;
            }
        }
    if (counter <= 24)
        return false;
    const float inv = 1.f / float(counter);
    cx *= inv;
    cy *= inv;
    cz *= inv;
    float cov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int dy = -kApron; dy < kApron; ++dy)
#pragma unroll
        for (int dx = -kApron; dx < kApron; ++dx)
        {
            const float4 p = tile[(ly + dy) * kMSpan + lx + dx];
            if (!zero_vertex(p.x, p.y, p.z))
            {
                const float ux = p.x - cx, uy = p.y - cy, uz = p.z - cz;
                cov[0] += ux * ux;
                cov[1] += ux * uy;
                cov[2] += ux * uz;
                cov[3] += uy * uy;
                cov[4] += uy * uz;
                cov[5] += uz * uz;
            }
        }
    float3 n = smallest_eigvec(cov);
    if (dot3(n, center) >= 0.f)
        n = make_float3(-n.x, -n.y, -n.z);
    normal = n;
    return true;
}

// ------------------------------------------------------------ kernels
__global__ __launch_bounds__(kMTile *kMTile) void k_filter_depth(const uint16_t *__restrict__ raw, int W, int H,
                                                                float cx_, float cy_, float inv_fx, float inv_fy,
                                                                float clip_near, float clip_far,
                                                                float *__restrict__ out)
{
    __shared__ float4 tile[kMSpan * kMSpan];
    const int x0 = blockIdx.x * kMTile - kApron, y0 = blockIdx.y * kMTile - kApron;
    const int tid = threadIdx.y * kMTile + threadIdx.x;
    // the raw vertices of the tile and its apron (get_vertex_confid, :14-27), zero outside
    for (int i = tid; i < kMSpan * kMSpan; i += kMTile * kMTile)
    {
        const int gx = x0 + i % kMSpan, gy = y0 + i / kMSpan;
        const uint32_t d = (gx >= 0 && gy >= 0 && gx < W && gy < H) ? raw[size_t(gy) * W + gx] : 0u;
        const float z = float(d) / 1000.f;
        tile[i] = make_float4((float(gx) - cx_) * inv_fx * z, (float(gy) - cy_) * inv_fy * z, z, 0.f);
    }
    __syncthreads();
    const int x = blockIdx.x * kMTile + threadIdx.x, y = blockIdx.y * kMTile + threadIdx.y;
    if (x >= W || y >= H)
        return;
    const int lx = threadIdx.x + kApron, ly = threadIdx.y + kApron;
    const uint32_t rd = raw[size_t(y) * W + x];
    const float dm = float(rd) / 1000.f;
    bool reliable = true, background = false;
    if (fabsf(dm) < clip_near)
        reliable = false;
    if (fabsf(dm) > clip_far)
        background = true;
    else
    {
        const float4 c4 = tile[ly * kMSpan + lx];
        const float3 vc = make_float3(c4.x, c4.y, c4.z);
        float3 nc = make_float3(0.f, 0.f, 0.f);
        if (!zero_vertex(vc.x, vc.y, vc.z))
        {
            float3 n;
            if (window_normal(tile, lx, ly, vc, n))
                nc = n;
        }
        const float ci = 1e-6f + sqrtf(dot3(vc, vc));
        const float ni = 1e-6f + sqrtf(dot3(nc, nc));
        const float3 cd = make_float3(vc.x / ci, vc.y / ci, vc.z / ci);
        const float3 nd = make_float3(nc.x / ni, nc.y / ni, nc.z / ni);
        if (dot3(cd, nd) > -0.1f)
            reliable = false;
    }
    out[size_t(y) * W + x] = !reliable ? 0.f : (background ? 10.f : float(rd) / 1000.f);
}

__global__ void k_scaled_vertex(const float *__restrict__ filtered, int W, int H, int w, int h, float sinv,
                                float cx_, float cy_, float fx, float fy, float clip_near, float clip_far,
                                float4 *__restrict__ vertex)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h)
        return;
    const int rx = int(float(x) * sinv), ry = int(float(y) * sinv);
    const float d = (rx < W && ry < H) ? filtered[size_t(ry) * W + rx] : 0.f;
    float4 v;
    if (fabsf(d) > clip_far || fabsf(d) < clip_near)
        v = make_float4(0.f, 0.f, 0.f, 0.f);
    else if (fabsf(d) < 1e-6f || fabsf(d) > 10.f)
        v = make_float4(0.f, 0.f, 0.f, 1.f);
    else
        v = make_float4((float(rx) - cx_) / fx * d, (float(ry) - cy_) / fy * d, d, 1.f);
    vertex[size_t(y) * w + x] = v;
}

__global__ __launch_bounds__(kMTile *kMTile) void k_normal_radius(const float4 *__restrict__ vertex, int w, int h,
                                                                 float radius_scale, float4 *__restrict__ out)
{
    __shared__ float4 tile[kMSpan * kMSpan];
    const int x0 = blockIdx.x * kMTile - kApron, y0 = blockIdx.y * kMTile - kApron;
    const int tid = threadIdx.y * kMTile + threadIdx.x;
    for (int i = tid; i < kMSpan * kMSpan; i += kMTile * kMTile)
    {
        const int gx = x0 + i % kMSpan, gy = y0 + i / kMSpan;
        tile[i] = (gx >= 0 && gy >= 0 && gx < w && gy < h) ? vertex[size_t(gy) * w + gx]
                                                          : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    const int x = blockIdx.x * kMTile + threadIdx.x, y = blockIdx.y * kMTile + threadIdx.y;
    if (x >= w || y >= h)
        return;
    const int lx = threadIdx.x + kApron, ly = threadIdx.y + kApron;
    const float4 c4 = tile[ly * kMSpan + lx];
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    float3 n;
    if (!zero_vertex(c4.x, c4.y, c4.z) && window_normal(tile, lx, ly, make_float3(c4.x, c4.y, c4.z), n))
    {
        // compute_radius (normal_map.cu:9-22), [mm]
        const float radius = c4.z * (1000 * 1.414f / 570.0f) * radius_scale;
        const float nz = fabsf(n.z);
        const float rn = nz > 0.5f ? radius / nz : 2.f * radius;
        o = make_float4(n.x, n.y, n.z, rn);
    }
    out[size_t(y) * w + x] = o;
}

__global__ void k_vertex_confid(const uint16_t *__restrict__ raw, int W, int H, float cx_, float cy_, float ifx,
                                float ify, float4 *__restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H)
        return;
    const float z = float(raw[size_t(y) * W + x]) / 1000.f;
    out[size_t(y) * W + x] = make_float4((float(x) - cx_) * ifx * z, (float(y) - cy_) * ify * z, z, 1.f);
}
} // namespace

hipError_t launch_filter_depth(const uint16_t *raw, int W, int H, float fx, float fy, float cx, float cy,
                               float clip_near, float clip_far, float *out, hipStream_t st)
{
    if (W <= 0 || H <= 0)
        return hipSuccess;
    // the filter's intrinsic inverse (surfel_reliable_filter.cu:172-176)
    const float inv_fx = 1.f / fx, inv_fy = 1.f / fy;
    hipLaunchKernelGGL(k_filter_depth, dim3((W + kMTile - 1) / kMTile, (H + kMTile - 1) / kMTile),
                       dim3(kMTile, kMTile), 0, st, raw, W, H, cx, cy, inv_fx, inv_fy, clip_near, clip_far, out);
    return hipGetLastError();
}

hipError_t launch_scaled_vertex(const float *filtered, int W, int H, int w, int h, float scale, float fx, float fy,
                                float cx, float cy, float clip_near, float clip_far, float *vertex4, hipStream_t st)
{
    if (w <= 0 || h <= 0)
        return hipSuccess;
    const float sinv = 1.f / scale;
    hipLaunchKernelGGL(k_scaled_vertex, dim3((w + 15) / 16, (h + 15) / 16), dim3(16, 16), 0, st, filtered, W, H, w, h,
                       sinv, cx, cy, fx, fy, clip_near, clip_far, reinterpret_cast<float4 *>(vertex4));
    return hipGetLastError();
}

hipError_t launch_normal_radius(const float *vertex4, int w, int h, float radius_scale, float *normal4,
                                hipStream_t st)
{
    if (w <= 0 || h <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_normal_radius, dim3((w + kMTile - 1) / kMTile, (h + kMTile - 1) / kMTile),
                       dim3(kMTile, kMTile), 0, st, reinterpret_cast<const float4 *>(vertex4), w, h, radius_scale,
                       reinterpret_cast<float4 *>(normal4));
    return hipGetLastError();
}

hipError_t launch_vertex_confid(const uint16_t *raw, int W, int H, float fx, float fy, float cx, float cy,
                                float *vertex4, hipStream_t st)
{
    if (W <= 0 || H <= 0)
        return hipSuccess;
    // IntrinsicInverse as common_utils.h:51-58 builds it: 1 / double(focal)
    const float ifx = float(1.0 / double(fx)), ify = float(1.0 / double(fy));
    hipLaunchKernelGGL(k_vertex_confid, dim3((W + 15) / 16, (H + 15) / 16), dim3(16, 16), 0, st, raw, W, H, cx, cy,
                       ifx, ify, reinterpret_cast<float4 *>(vertex4));
    return hipGetLastError();
}
} // namespace sh
