// Device helpers shared by the GN kernels and the pipelined PCG (which applies
// the warp update to its own nodes when it ends).
#pragma once
#include "sh_common.hpp"
#include "sh_desc.hpp"

namespace sh
{
__device__ __forceinline__ dq8 load_dq(const dq8 *__restrict__ a, uint32_t i)
{
    const float4 *p = reinterpret_cast<const float4 *>(a + i);
    dq8 r;
    r.q0 = p[0];
    r.q1 = p[1];
    return r;
}

// Total weighted cost (dense + ARAP) from the work-group partials of
// the dense and ARAP parts of k_terms, FP64, fixed order; one wave.
__device__ inline void reduce_cost(const SolveDesc &d, int slot)
{
    const uint32_t n = (uint32_t(d.P) + 63u) / 64u + (uint32_t(d.E + d.N) + 255u) / 256u;
    double v = 0.0;
    for (uint32_t i = threadIdx.x % 64u; i < n; i += 64)
        v += d.cost_part[i];
    v = wave_sum(v);
    if (threadIdx.x % 64u == 0)
        d.cost[slot] = v;
}

// T <- exp(xi) T for node n, xi = xv[0..5] (FP32, as the PCG stores x), stored
// as a dual quaternion (ApplyWarpFieldUpdateKernel, SolverIterationData.cu:19-74).
// Records whether the rotation was dropped (|w|_1 < 1e-4, :42) for this GN
// iteration, and the |w|_1 it was decided on, so a check can replay the same
// decisions and compare the increments.
__device__ inline void warp_update_node(const SolveDesc &d, uint32_t n, int parity, const float *xv)
{
    const f3 w = mk3(xv[0], xv[1], xv[2]);
    const f3 t = mk3(xv[3], xv[4], xv[5]);
    rt34 S;
    const float l1 = fabsf(w.x) + fabsf(w.y) + fabsf(w.z);
    const bool cut = l1 < 1e-4f;
    const size_t rec = size_t(d.gn_iter + parity) * uint32_t(d.N) + n;
    d.rot_cut[rec] = cut ? 1u : 0u;
    d.twist_l1[rec] = l1;
    if (cut)
    {
        S.r[0] = 1.f;
        S.r[1] = 0.f;
        S.r[2] = 0.f;
        S.r[3] = 0.f;
        S.r[4] = 1.f;
        S.r[5] = 0.f;
        S.r[6] = 0.f;
        S.r[7] = 0.f;
        S.r[8] = 1.f;
    }
    else
    {
        const float angle = sqrtf(dot3(w, w));
        const f3 ax = (1.0f / angle) * w;
        const float c = cosf(angle), s = sinf(angle), tt = 1.0f - c;
        S.r[0] = tt * ax.x * ax.x + c;
        S.r[1] = tt * ax.x * ax.y - ax.z * s;
        S.r[2] = tt * ax.x * ax.z + ax.y * s;
        S.r[3] = tt * ax.x * ax.y + ax.z * s;
        S.r[4] = tt * ax.y * ax.y + c;
        S.r[5] = tt * ax.y * ax.z - ax.x * s;
        S.r[6] = tt * ax.x * ax.z - ax.y * s;
        S.r[7] = tt * ax.y * ax.z + ax.x * s;
        S.r[8] = tt * ax.z * ax.z + c;
    }
    S.t = t;
    const rt34 prev = dq_to_mat34(load_dq(pick2(d.dq, (d.gn_iter + parity) & 1), n));
    rt34 nw;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            nw.r[3 * i + j] = S.r[3 * i] * prev.r[j] + S.r[3 * i + 1] * prev.r[3 + j] + S.r[3 * i + 2] * prev.r[6 + j];
    nw.t = rot_apply(S.r, prev.t) + S.t;
    const dq8 out = dq_from_mat34(nw);
    float4 *o = reinterpret_cast<float4 *>(pick2(d.dq, ((d.gn_iter + parity) & 1) ^ 1) + n);
    o[0] = out.q0;
    o[1] = out.q1;
}

} // namespace sh
