// Forward warp of surfels and nodes by a solved warp field: the step after
// the solve (SurfelNodeDeformer::ForwardWarpSurfelsAndNodes,
// libs/star/geometry/surfel/SurfelNodeDeformer.cu:13-80).
//
// Per surfel (and per node, with the node's own KNN): the K node dual
// quaternions, each normalised first, blended with weights spatial x connect
// (averageDualQuaternion, DualQuaternion.hpp:289-306 -- unlike the solver's
// blend, which does not normalise per node), then se3_matrix() of the blend
// (normalise, rotation, translation 2 q1 q0*, DualQuaternion.hpp:182-189):
// v' = R v + t, n' = R n, the w components kept.  One thread per element,
// surfels then nodes; coalesced float4 streams.
#include "sh_common.hpp"
#include "sh_kernels.hpp"

namespace sh
{
template <int K>
__device__ __forceinline__ rt34 blend_normalised(const float4 *__restrict__ warp, const uint16_t *__restrict__ knn,
                                                 const float *__restrict__ spatial, const float *__restrict__ connect,
                                                 size_t i)
{
    dq8 avg;
    avg.q0 = make_float4(0.f, 0.f, 0.f, 0.f);
    avg.q1 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        const uint32_t n = knn[i * K + k];
        const float w = spatial[i * K + k] * connect[i * K + k];
        dq8 q;
        q.q0 = warp[2 * size_t(n)];
        q.q1 = warp[2 * size_t(n) + 1];
        q = dq_normalize(q);
        avg.q0.x += w * q.q0.x;
        avg.q0.y += w * q.q0.y;
        avg.q0.z += w * q.q0.z;
        avg.q0.w += w * q.q0.w;
        avg.q1.x += w * q.q1.x;
        avg.q1.y += w * q.q1.y;
        avg.q1.z += w * q.q1.z;
        avg.q1.w += w * q.q1.w;
    }
    return dq_se3(avg);
}

template <int K>
__global__ __launch_bounds__(256) void k_forward_warp(const float4 *__restrict__ warp, const float4 *__restrict__ ref_v,
                                                      const float4 *__restrict__ ref_n,
                                                      const uint16_t *__restrict__ s_knn,
                                                      const float *__restrict__ s_spatial,
                                                      const float *__restrict__ s_connect, uint32_t num_surfels,
                                                      const float4 *__restrict__ ref_node,
                                                      const uint16_t *__restrict__ n_knn,
                                                      const float *__restrict__ n_spatial,
                                                      const float *__restrict__ n_connect, uint32_t num_nodes,
                                                      float4 *__restrict__ live_v, float4 *__restrict__ live_n,
                                                      float4 *__restrict__ live_node)
{
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < num_surfels)
    {
        const rt34 T = blend_normalised<K>(warp, s_knn, s_spatial, s_connect, idx);
        const float4 v = ref_v[idx], n = ref_n[idx];
        const f3 v3 = rot_apply(T.r, mk3(v.x, v.y, v.z)) + T.t;
        const f3 n3 = rot_apply(T.r, mk3(n.x, n.y, n.z));
        live_v[idx] = make_float4(v3.x, v3.y, v3.z, v.w);
        live_n[idx] = make_float4(n3.x, n3.y, n3.z, n.w);
    }
    else if (idx < num_surfels + num_nodes)
    {
        const uint32_t m = idx - num_surfels;
        const rt34 T = blend_normalised<K>(warp, n_knn, n_spatial, n_connect, m);
        const float4 v = ref_node[m];
        const f3 v3 = rot_apply(T.r, mk3(v.x, v.y, v.z)) + T.t;
        live_node[m] = make_float4(v3.x, v3.y, v3.z, v.w);
    }
}

hipError_t launch_forward_warp(uint32_t K, const float4 *warp, const float4 *ref_v, const float4 *ref_n,
                               const uint16_t *s_knn, const float *s_spatial, const float *s_connect,
                               uint32_t num_surfels, const float4 *ref_node, const uint16_t *n_knn,
                               const float *n_spatial, const float *n_connect, uint32_t num_nodes, float4 *live_v,
                               float4 *live_n, float4 *live_node, hipStream_t st)
{
    const uint32_t total = num_surfels + num_nodes;
    if (total == 0)
        return hipSuccess;
    const dim3 grid((total + 255) / 256);
    if (K == 8)
        hipLaunchKernelGGL(k_forward_warp<8>, grid, dim3(256), 0, st, warp, ref_v, ref_n, s_knn, s_spatial, s_connect,
                           num_surfels, ref_node, n_knn, n_spatial, n_connect, num_nodes, live_v, live_n, live_node);
    else
        hipLaunchKernelGGL(k_forward_warp<4>, grid, dim3(256), 0, st, warp, ref_v, ref_n, s_knn, s_spatial, s_connect,
                           num_surfels, ref_node, n_knn, n_spatial, n_connect, num_nodes, live_v, live_n, live_node);
    return hipGetLastError();
}
} // namespace sh
