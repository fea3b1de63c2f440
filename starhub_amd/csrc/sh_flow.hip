// Optical-flow input and the node-motion prior (the inputs of the solve's flow
// and node-translation terms):
//   k_flow_from_of16     VolumeDeformFileFetch::FetchOFImage's decode
//                        (libs/star/io/VolumeDeformFileFetch.cpp:39-52) of the
//                        .of.png encoding u16 = (flow + 1000) * 10
//                        (tools/transfer_hdf5.py:42-52), fused with
//                        createScaledOpticalFlowMap (libs/star/img_proc/
//                        optical_flow_map.cu:5-57): only the sampled raw pixels
//                        are decoded.
//   k_node_motion_sum    AtomicSumIdealNodeMotionFromOpticalFlowKernel
//                        (libs/star/geometry/node_graph/node_graph_opticalflow.cu:
//                        14-75): per previous-frame pixel, the camera-space motion
//                        of its surfel along the flow, weighted, rotated to world
//                        and added to the surfel's nearest node.
//   k_node_motion_mean   AverageNodeMotionKernel (:77-92).
// The reference adds floats atomically, so its sums depend on the order the
// atomics land in.  Here every contribution is quantised to an integer
// number of 2^-31 units and added with 64-bit integer atomics: the sums are
// exact, so the result is the same whatever the order.
#include <mutex>

#include "sh_common.hpp"
#include "sh_kernels.hpp"

#pragma clang fp contract(off)

namespace sh
{
// Node sums of the allocation-free entry point (sh_node_motion_from_flow_noalloc):
// kNoAllocMaxNodes nodes x 4 64-bit words = 128 KiB, one copy per GPU, held
// until the process ends.  Host code never uses this name as an address: each
// device's address comes from hipGetSymbolAddress (noalloc_sums below).
static __device__ __align__(256) unsigned long long scratch[kNoAllocMaxNodes * 4];

namespace
{
constexpr float kFixScale = 2147483648.f; // 2^31

__global__ void k_flow_from_of16(const uint16_t *__restrict__ of16, int W, int H, int w, int h, float window,
                                 float scale, float2 *__restrict__ flow)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h)
        return;
    // __float2int_rn of x * window: round half to even
    const int rx = int(rintf(float(x) * window)), ry = int(rintf(float(y) * window));
    float2 f = make_float2(0.f, 0.f);
    if (rx >= 0 && ry >= 0 && rx < W && ry < H)
    {
        const uint16_t *p = of16 + (size_t(ry) * W + rx) * 3;
        // convertTo(CV_32FC2, 1/10, -1000): the double product and sum, rounded once to float
        f = make_float2(float(double(p[0]) * 0.1 - 1000.0), float(double(p[1]) * 0.1 - 1000.0));
    }
    flow[size_t(y) * w + x] = make_float2(f.x * scale, f.y * scale);
}

struct MotionParams
{
    float r[9]; // cam2world rotation, row-major
    float fx, fy, cx, cy;
    int width, height, K;
    uint32_t num_surfels, num_nodes;
};

__device__ __forceinline__ float4 texel4(const float4 *__restrict__ m, int W, int H, int x, int y)
{
    return (x >= 0 && y >= 0 && x < W && y < H) ? m[size_t(y) * W + x] : make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_node_motion_sum(const float4 *__restrict__ rgbd_prev, const float4 *__restrict__ rgbd,
                                  const float2 *__restrict__ flow, const uint32_t *__restrict__ index_prev,
                                  const uint16_t *__restrict__ surfel_knn, const float *__restrict__ surfel_spatial,
                                  MotionParams mp, unsigned long long *__restrict__ sums)
{
    const int xp = blockIdx.x * blockDim.x + threadIdx.x, yp = blockIdx.y * blockDim.y + threadIdx.y;
    if (xp >= mp.width || yp >= mp.height)
        return;
    const size_t pp = size_t(yp) * mp.width + xp;
    const float2 of = flow[pp];
    const float xf = float(xp) + of.x, yf = float(yp) + of.y;
    // point sampling at a float coordinate: the texel floor(x), zero outside
    const float4 c = texel4(rgbd, mp.width, mp.height, int(floorf(xf)), int(floorf(yf)));
    const uint32_t s = index_prev[pp];
    const float4 cp = rgbd_prev[pp];
    if (s == 0xffffffffu || c.w == 0.f)
        return;
    if (s >= mp.num_surfels)
        return; // not a surfel of this geometry: contributes nothing
    const uint32_t node = surfel_knn[size_t(s) * mp.K];
    if (node >= mp.num_nodes)
        return;
    const float wgt = surfel_spatial[size_t(s) * mp.K];
    const float depth = 1.f / c.w, depth_prev = 1.f / cp.w;
    const float vx = (xf - mp.cx) / mp.fx * depth, vy = (yf - mp.cy) / mp.fy * depth, vz = depth;
    const float ux = (float(xp) - mp.cx) / mp.fx * depth_prev, uy = (float(yp) - mp.cy) / mp.fy * depth_prev;
    float mx = vx - ux, my = vy - uy, mz = vz - depth_prev;
    if (fabsf(mz) > 0.05f) // optical_flow_z_shift, node_graph_opticalflow.cu:10
        return;
    mx *= wgt;
    my *= wgt;
    mz *= wgt;
    const float wx = (mp.r[0] * mx + mp.r[1] * my) + mp.r[2] * mz;
    const float wy = (mp.r[3] * mx + mp.r[4] * my) + mp.r[5] * mz;
    const float wz = (mp.r[6] * mx + mp.r[7] * my) + mp.r[8] * mz;
    // exact integer sums (two's complement wraps identically in any order)
    unsigned long long *o = sums + size_t(node) * 4;
    atomicAdd(o + 0, static_cast<unsigned long long>(llrintf(wx * kFixScale)));
    atomicAdd(o + 1, static_cast<unsigned long long>(llrintf(wy * kFixScale)));
    atomicAdd(o + 2, static_cast<unsigned long long>(llrintf(wz * kFixScale)));
    atomicAdd(o + 3, static_cast<unsigned long long>(llrintf(wgt * kFixScale)));
}

__global__ void k_zero_u64(unsigned long long *__restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0ull;
}

__global__ void k_node_motion_mean(const unsigned long long *__restrict__ sums, uint32_t n,
                                   float4 *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const long long *s = reinterpret_cast<const long long *>(sums) + size_t(i) * 4;
    const double inv = 1.0 / 2147483648.0;
    float4 m = make_float4(float(double(s[0]) * inv), float(double(s[1]) * inv), float(double(s[2]) * inv),
                           float(double(s[3]) * inv));
    if (m.w != 0.f)
    {
        const float k = 1.f / m.w;
        m = make_float4(m.x * k, m.y * k, m.z * k, m.w * k);
    }
    out[i] = m;
}
} // namespace

hipError_t launch_flow_from_of16(const uint16_t *of16, int W, int H, float scale, float *flow2, hipStream_t st)
{
    // createScaledOpticalFlowMap's size: floor of the float product
    const int w = int(floorf(float(W) * scale)), h = int(floorf(float(H) * scale));
    if (w <= 0 || h <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_flow_from_of16, dim3((w + 15) / 16, (h + 15) / 16), dim3(16, 16), 0, st, of16, W, H, w, h,
                       1.f / scale, scale, reinterpret_cast<float2 *>(flow2));
    return hipGetLastError();
}

hipError_t noalloc_sums(unsigned long long **out)
{
    // the device address of `scratch` on the current device, looked up once per
    // device (each GPU holds its own copy of the array at its own address)
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static unsigned long long *addr[kMaxDevices] = {};
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev])
    {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        if (!p)
            return hipErrorInvalidSymbol;
        addr[dev] = static_cast<unsigned long long *>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

hipError_t launch_node_motion(const float *rgbd_prev4, const float *rgbd4, const float *flow2,
                              const uint32_t *index_prev, int width, int height, const uint16_t *surfel_knn,
                              const float *surfel_spatial, uint32_t num_surfels, int K, const float cam2world[16],
                              float fx, float fy, float cx, float cy, uint32_t num_nodes, unsigned long long *sums,
                              float *out4, hipStream_t st)
{
    if (num_nodes == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3((num_nodes * 4 + 255) / 256), dim3(256), 0, st, sums, num_nodes * 4);
    MotionParams mp{};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            mp.r[3 * a + b] = cam2world[4 * a + b];
    mp.fx = fx;
    mp.fy = fy;
    mp.cx = cx;
    mp.cy = cy;
    mp.width = width;
    mp.height = height;
    mp.K = K;
    mp.num_surfels = num_surfels;
    mp.num_nodes = num_nodes;
    if (width > 0 && height > 0)
        hipLaunchKernelGGL(k_node_motion_sum, dim3((width + 15) / 16, (height + 15) / 16), dim3(16, 16), 0, st,
                           reinterpret_cast<const float4 *>(rgbd_prev4), reinterpret_cast<const float4 *>(rgbd4),
                           reinterpret_cast<const float2 *>(flow2), index_prev, surfel_knn, surfel_spatial, mp, sums);
    hipLaunchKernelGGL(k_node_motion_mean, dim3((num_nodes + 255) / 256), dim3(256), 0, st, sums, num_nodes,
                       reinterpret_cast<float4 *>(out4));
    return hipGetLastError();
}
} // namespace sh
