// The per-pixel dense term of the GN phase, shared by the term kernel and
// the tile kernel that builds the records itself (sh_kernels.hip, sh_tiles.hip).
#pragma once
#include "sh_common.hpp"
#include "sh_desc.hpp"
#include "sh_update.hpp"

namespace sh
{
template <int K>
__device__ __forceinline__ void load_knn(const SolveDesc &d, uint32_t p, uint16_t (&kn)[K], float (&w)[K])
{
    if constexpr (K == 8)
    {
        const uint4 kk = *reinterpret_cast<const uint4 *>(d.gpknn + size_t(p) * 8);
        const uint32_t v[4] = {kk.x, kk.y, kk.z, kk.w};
#pragma unroll
        for (int a = 0; a < 4; ++a)
        {
            kn[2 * a] = uint16_t(v[a] & 0xffffu);
            kn[2 * a + 1] = uint16_t(v[a] >> 16);
        }
        const float4 w0 = *reinterpret_cast<const float4 *>(d.gpw + size_t(p) * 8);
        const float4 w1 = *reinterpret_cast<const float4 *>(d.gpw + size_t(p) * 8 + 4);
        w[0] = w0.x;
        w[1] = w0.y;
        w[2] = w0.z;
        w[3] = w0.w;
        w[4] = w1.x;
        w[5] = w1.y;
        w[6] = w1.z;
        w[7] = w1.w;
    }
    else
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            kn[k] = d.gpknn[size_t(p) * K + k];
            w[k] = d.gpw[size_t(p) * K + k];
        }
    }
}

// One pixel of the GN phase's order: residual and Jacobian of the depth and
// flow channels, folded into the H/g record hg; returns its cost
// sum_c lam_c^2 r_c^2 (computeDenseImageJacobianKernel,
// DenseImageJacobianResidual.cu:164-310; forms :12-157).
template <int K>
__device__ __forceinline__ double dense_pixel_hg(const SolveDesc &d, uint32_t p, int parity, float (&hg)[kHG])
{
    const dq8 *__restrict__ dq = pick2(d.dq, (d.gn_iter + parity) & 1);
    uint16_t kn[K];
    float w[K];
    load_knn<K>(d, p, kn, w);
    dq8 avg;
    avg.q0 = make_float4(0.f, 0.f, 0.f, 0.f);
    avg.q1 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        const dq8 n = load_dq(dq, kn[k]);
        avg.q0.x += w[k] * n.q0.x;
        avg.q0.y += w[k] * n.q0.y;
        avg.q0.z += w[k] * n.q0.z;
        avg.q0.w += w[k] * n.q0.w;
        avg.q1.x += w[k] * n.q1.x;
        avg.q1.y += w[k] * n.q1.y;
        avg.q1.z += w[k] * n.q1.z;
        avg.q1.w += w[k] * n.q1.w;
    }
    const rt34 T = dq_se3(avg);
    const ushort4 px = d.gpix[p];
    const uint32_t src = uint32_t(px.y) * uint32_t(d.W) + px.x;
    const float4 cv = d.ref_v[src];
    const float4 cn = d.ref_n[src];
    const f3 vw = rot_apply(T.r, mk3(cv.x, cv.y, cv.z)) + T.t;
    const f3 nw = rot_apply(T.r, mk3(cn.x, cn.y, cn.z));
    const f3 vc = rot_apply(d.w2c_r, vw) + mk3(d.w2c_t[0], d.w2c_t[1], d.w2c_t[2]);
    const f3 nc = rot_apply(d.w2c_r, nw);
    float out[24]; // J[c][0..5] at 6c, r[c] at 18 + c
#pragma unroll
    for (int a = 0; a < 24; ++a)
        out[a] = 0.f;
    if (px.z < d.W && px.w < d.H)
    {
        const uint32_t tgt = uint32_t(px.w) * uint32_t(d.W) + px.z;
        const float4 tv4 = d.meas_v[tgt];
        const float4 tn4 = d.meas_n[tgt];
        const f3 tv = mk3(tv4.x, tv4.y, tv4.z), tn = mk3(tn4.x, tn4.y, tn4.z);
        bool ok = true;
        if (fabsf(tv4.x) < 1e-3f && fabsf(tv4.y) < 1e-3f && fabsf(tv4.z) < 1e-3f) // is_zero_vertex
            ok = false;
        if (dot3(tn, nc) < 0.7f)
            ok = false;
        if (fabsf(tv.z - vc.z) > 0.03f)
            ok = false;
        if (ok)
        {
            const f3 wt = mk3(d.w2c_t[0], d.w2c_t[1], d.w2c_t[2]);
            const f3 tvw = rot_apply_t(d.w2c_r, tv - wt);
            const f3 tnw = rot_apply_t(d.w2c_r, tn);
            const f3 diff = vw - tvw;
            out[18] = dot3(tnw, diff);
            const f3 j0 = cross3(vw, tnw);
            out[0] = j0.x;
            out[1] = j0.y;
            out[2] = j0.z;
            out[3] = tnw.x;
            out[4] = tnw.y;
            out[5] = tnw.z;
            const f3 pr0 = mk3(d.w2c_r[0], d.w2c_r[1], d.w2c_r[2]);
            const f3 pr1 = mk3(d.w2c_r[3], d.w2c_r[4], d.w2c_r[5]);
            const float ex = dot3(pr0, diff), ey = dot3(pr1, diff);
            if (!(sqrtf(ex * ex + ey * ey) < 1e-6f))
            {
                out[19] = ex;
                out[20] = ey;
                const f3 mv = mk3(-vw.x, -vw.y, -vw.z);
                const f3 rows[2] = {pr0, pr1};
#pragma unroll
                for (int c = 0; c < 2; ++c)
                {
                    const f3 a = rows[c];
                    out[6 * (c + 1) + 0] = a.y * mv.z + a.z * (-mv.y);
                    out[6 * (c + 1) + 1] = a.x * (-mv.z) + a.z * mv.x;
                    out[6 * (c + 1) + 2] = a.x * mv.y + a.y * (-mv.x);
                    out[6 * (c + 1) + 3] = a.x;
                    out[6 * (c + 1) + 4] = a.y;
                    out[6 * (c + 1) + 5] = a.z;
                }
            }
        }
    }
    // fold the channels: H = sum_c lam_c^2 J_c^T J_c, g = sum_c lam_c^2 r_c J_c
    const float lam2[3] = {d.lam2_depth, d.lam2_flow, d.lam2_flow};
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = a; b < 6; ++b)
        {
            float h = 0.f;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                h += lam2[c] * out[6 * c + a] * out[6 * c + b];
            hg[sym_index(a, b)] = h;
        }
#pragma unroll
    for (int q = 0; q < 6; ++q)
    {
        float g = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            g += lam2[c] * out[18 + c] * out[6 * c + q];
        hg[kHGg + q] = g;
    }
    hg[kHG - 1] = 0.f;
    return double(d.lam2_depth) * double(out[18]) * double(out[18]) +
           double(d.lam2_flow) * (double(out[19]) * double(out[19]) + double(out[20]) * double(out[20]));
}

// The same, with the record stored at position p (FP32 or FP16 block-scaled).
template <int K, class JT>
__device__ __forceinline__ double dense_pixel(const SolveDesc &d, uint32_t p, int parity, int store)
{
    float hg[kHG];
    const double cost = dense_pixel_hg<K>(d, p, parity, hg);
    if (store)
        store_hg<JT>(d.jr, p, hg);
    return cost;
}

} // namespace sh
