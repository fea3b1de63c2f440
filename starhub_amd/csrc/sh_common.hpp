// Device-side math and data layout shared by the warp-field solver kernels.
//
// gfx950 (CDNA4): wave64, 160 KB LDS per CU.  All kernels here are written for
// 64-lane wavefronts; reductions use cross-lane DPP/permute through the
// compiler's __shfl_xor (ds_swizzle / DPP) and never float atomics, so every
// result is deterministic for a given input.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../../include/sh/warp_solver.h"

namespace sh
{
constexpr int kWave = 64;
constexpr uint32_t kInvalidIndex = 0xffffffffu;
constexpr uint32_t kUninitKey = 0xffffffffu;
constexpr uint32_t kPairBase = 4096u; // node-pair key row*4096+col (solver_encode.h:14-18)
constexpr int kMaxK = 8;

// --------------------------------------------------------------- vector math
struct f3
{
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Row-major 3x3 rotation + translation.
struct rt34
{
    float r[9];
    f3 t;
};
__device__ __forceinline__ f3 rot_apply(const float *r, f3 v)
{
    return {r[0] * v.x + r[1] * v.y + r[2] * v.z, r[3] * v.x + r[4] * v.y + r[5] * v.z,
            r[6] * v.x + r[7] * v.y + r[8] * v.z};
}
__device__ __forceinline__ f3 rot_apply_t(const float *r, f3 v)
{
    return {r[0] * v.x + r[3] * v.y + r[6] * v.z, r[1] * v.x + r[4] * v.y + r[7] * v.z,
            r[2] * v.x + r[5] * v.y + r[8] * v.z};
}

// Quaternion as float4 (x, y, z, w) -- the reference's storage order
// (libs/star/math/Quaternion.hpp:115: float4 q0 with w in .w).
__device__ __forceinline__ float4 qmul(float4 a, float4 b)
{
    float4 q;
    q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    q.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    q.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return q;
}
__device__ __forceinline__ float qdot(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
__device__ __forceinline__ float4 qscale(float s, float4 a) { return make_float4(s * a.x, s * a.y, s * a.z, s * a.w); }

// Unit quaternion -> rotation matrix (Quaternion.hpp:96-111, no renormalisation).
__device__ __forceinline__ void quat_to_rot(float4 q, float *r)
{
    r[0] = 1.f - 2.f * q.y * q.y - 2.f * q.z * q.z;
    r[1] = 2.f * q.x * q.y - 2.f * q.z * q.w;
    r[2] = 2.f * q.x * q.z + 2.f * q.y * q.w;
    r[3] = 2.f * q.x * q.y + 2.f * q.z * q.w;
    r[4] = 1.f - 2.f * q.x * q.x - 2.f * q.z * q.z;
    r[5] = 2.f * q.y * q.z - 2.f * q.x * q.w;
    r[6] = 2.f * q.x * q.z - 2.f * q.y * q.w;
    r[7] = 2.f * q.y * q.z + 2.f * q.x * q.w;
    r[8] = 1.f - 2.f * q.x * q.x - 2.f * q.y * q.y;
}

// Rotation matrix -> quaternion, trace-branched (Quaternion.hpp:17-52).
__device__ __forceinline__ float4 rot_to_quat(const float *r)
{
    const float m00 = r[0], m01 = r[1], m02 = r[2], m10 = r[3], m11 = r[4], m12 = r[5], m20 = r[6], m21 = r[7],
                m22 = r[8];
    const float tr = m00 + m11 + m22;
    float4 q;
    if (tr > 0.f)
    {
        const float s = sqrtf(tr + 1.0f) * 2.f;
        q.w = s * 0.25f;
        q.x = (m21 - m12) / s;
        q.y = (m02 - m20) / s;
        q.z = (m10 - m01) / s;
    }
    else if ((m00 > m11) && (m00 > m22))
    {
        const float s = sqrtf(1.0f + m00 - m11 - m22) * 2.f;
        q.w = (m21 - m12) / s;
        q.x = 0.25f * s;
        q.y = (m01 + m10) / s;
        q.z = (m02 + m20) / s;
    }
    else if (m11 > m22)
    {
        const float s = sqrtf(1.0f + m11 - m00 - m22) * 2.f;
        q.w = (m02 - m20) / s;
        q.x = (m01 + m10) / s;
        q.y = 0.25f * s;
        q.z = (m12 + m21) / s;
    }
    else
    {
        const float s = sqrtf(1.0f + m22 - m00 - m11) * 2.f;
        q.w = (m10 - m01) / s;
        q.x = (m02 + m20) / s;
        q.y = (m12 + m21) / s;
        q.z = 0.25f * s;
    }
    return q;
}

struct dq8
{
    float4 q0, q1;
};

// DualQuaternion::normalize (DualQuaternion.hpp:150-156).
__device__ __forceinline__ dq8 dq_normalize(dq8 d)
{
    const float inv = rsqrtf(qdot(d.q0, d.q0));
    d.q0 = qscale(inv, d.q0);
    d.q1 = qscale(inv, d.q1);
    const float c = qdot(d.q0, d.q1);
    d.q1 = make_float4(d.q1.x - c * d.q0.x, d.q1.y - c * d.q0.y, d.q1.z - c * d.q0.z, d.q1.w - c * d.q0.w);
    return d;
}

// se3_matrix(): normalise, R(q0), t = vec(2 q1 q0*)  (DualQuaternion.hpp:182-189).
__device__ __forceinline__ rt34 dq_se3(dq8 d)
{
    d = dq_normalize(d);
    rt34 m;
    quat_to_rot(d.q0, m.r);
    const float4 conj = make_float4(-d.q0.x, -d.q0.y, -d.q0.z, d.q0.w);
    const float4 tp = qmul(qscale(2.f, d.q1), conj);
    m.t = mk3(tp.x, tp.y, tp.z);
    return m;
}

// operator mat34(): normalised copy, q0 re-normalised inside matrix()
// (DualQuaternion.hpp:169-179, Quaternion.hpp:77-94).
__device__ __forceinline__ rt34 dq_to_mat34(dq8 d)
{
    d = dq_normalize(d);
    const float inv = rsqrtf(qdot(d.q0, d.q0));
    rt34 m;
    quat_to_rot(qscale(inv, d.q0), m.r);
    const float4 conj = make_float4(-d.q0.x, -d.q0.y, -d.q0.z, d.q0.w);
    const float4 tp = qmul(qscale(2.f, d.q1), conj);
    m.t = mk3(tp.x, tp.y, tp.z);
    return m;
}

// DualQuaternion(const mat34&): (1 + e t/2) * (q_r + e 0)  (DualQuaternion.hpp:59-66).
__device__ __forceinline__ dq8 dq_from_mat34(const rt34 &m)
{
    const float4 qr = rot_to_quat(m.r);
    const float4 tv = make_float4(0.5f * m.t.x, 0.5f * m.t.y, 0.5f * m.t.z, 0.f);
    dq8 d;
    d.q0 = qr;
    d.q1 = qmul(tv, qr);
    return d;
}

// ------------------------------------------------------------ wave reductions
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}

// FP16 stream helpers
__device__ __forceinline__ float ld_stream(const float *p, size_t i) { return p[i]; }
__device__ __forceinline__ float ld_stream(const __half *p, size_t i) { return __half2float(p[i]); }
__device__ __forceinline__ void st_stream(float *p, size_t i, float v) { p[i] = v; }
__device__ __forceinline__ void st_stream(__half *p, size_t i, float v) { p[i] = __float2half_rn(v); }

// ------------------------------------------------------------ term layout
// Per-pixel normal-equation record (28 scalars), written once per GN iteration
// by the dense-term kernel: the 21 upper-triangle entries of
// H = sum_c lambda_c^2 J_c^T J_c (row-major a <= b) followed by
// g = sum_c lambda_c^2 r_c J_c (6) and one pad.  Node k of the pixel then
// contributes w_i w_k H to block (i, k) and w_i^2 g to node i's J^T W r
// (PreconditionerRhsBuilder.cu:38-108, JtDotResidualIndexed.cu:40-150,
// JtJMaterializer.cu:31-63 with the linear-blend Jacobian w_k J).
constexpr int kHG = 28;
constexpr int kHGg = 21;

// index of (a, b) in the packed upper triangle
__host__ __device__ constexpr int sym_index(int a, int b)
{
    return (a <= b) ? (a * 6 - a * (a - 1) / 2 + (b - a)) : (b * 6 - b * (b - 1) / 2 + (a - b));
}
// Per-edge ARAP record (8 floats): Ti_xj (3), Tj_xj (3), connect weight, valid.
constexpr int kReg = 8;


// The FP16 record is block-scaled: the 27 values are stored as v * 2^-s, with
// s chosen so that the largest magnitude stays below 2^15, and s itself (an
// integer, exact in FP16) is kept in the pad slot.  The products of a
// flow-weighted Jacobian (entries of order f / z) would otherwise overflow
// FP16's range of 65504.
__device__ __forceinline__ int hg_shift(const float (&v)[kHG])
{
    float m = 0.f;
#pragma unroll
    for (int e = 0; e < kHG - 1; ++e)
        m = fmaxf(m, fabsf(v[e]));
    if (!(m > 0.f) || !isfinite(m))
        return 0;
    int x;
    (void)frexpf(m, &x); // m < 2^x
    return x - 15;
}

template <class JT>
__device__ __forceinline__ void store_hg(void *base, uint32_t p, const float (&v)[kHG])
{
    if constexpr (sizeof(JT) == 4)
    {
        float4 *o = reinterpret_cast<float4 *>(reinterpret_cast<float *>(base) + size_t(p) * kHG);
#pragma unroll
        for (int a = 0; a < kHG / 4; ++a)
            o[a] = make_float4(v[4 * a], v[4 * a + 1], v[4 * a + 2], v[4 * a + 3]);
    }
    else
    {
        const int sh = hg_shift(v);
        const float inv = ldexpf(1.f, -sh);
        uint2 *o = reinterpret_cast<uint2 *>(reinterpret_cast<__half *>(base) + size_t(p) * kHG);
#pragma unroll
        for (int a = 0; a < kHG / 4; ++a)
        {
            uint32_t w[2];
#pragma unroll
            for (int b = 0; b < 2; ++b)
            {
                const int e0 = 4 * a + 2 * b, e1 = e0 + 1;
                const __half lo = __float2half_rn(v[e0] * inv);
                const __half hi = (e1 == kHG - 1) ? __float2half_rn(float(sh)) : __float2half_rn(v[e1] * inv);
                w[b] = uint32_t(__half_as_ushort(lo)) | (uint32_t(__half_as_ushort(hi)) << 16);
            }
            o[a] = make_uint2(w[0], w[1]);
        }
    }
}

// Loads the 27 values (pad slot returned as 0).
template <class JT>
__device__ __forceinline__ void load_hg(const void *base, uint32_t p, float (&v)[kHG])
{
    if constexpr (sizeof(JT) == 4)
    {
        const float4 *o = reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(base) + size_t(p) * kHG);
#pragma unroll
        for (int a = 0; a < kHG / 4; ++a)
        {
            const float4 f = o[a];
            v[4 * a] = f.x;
            v[4 * a + 1] = f.y;
            v[4 * a + 2] = f.z;
            v[4 * a + 3] = f.w;
        }
    }
    else
    {
        const uint2 *o = reinterpret_cast<const uint2 *>(reinterpret_cast<const __half *>(base) + size_t(p) * kHG);
        float raw[kHG];
#pragma unroll
        for (int a = 0; a < kHG / 4; ++a)
        {
            const uint2 u = o[a];
            const uint32_t w[2] = {u.x, u.y};
#pragma unroll
            for (int b = 0; b < 2; ++b)
            {
                raw[4 * a + 2 * b] = __half2float(__ushort_as_half(uint16_t(w[b] & 0xffffu)));
                raw[4 * a + 2 * b + 1] = __half2float(__ushort_as_half(uint16_t(w[b] >> 16)));
            }
        }
        const float scale = ldexpf(1.f, int(raw[kHG - 1]));
#pragma unroll
        for (int e = 0; e < kHG - 1; ++e)
            v[e] = raw[e] * scale;
        v[kHG - 1] = 0.f;
    }
}

// One value of the record (e < kHG - 1).
template <class JT>
__device__ __forceinline__ float ld_hg(const void *base, uint32_t p, int e)
{
    if constexpr (sizeof(JT) == 4)
        return ld_stream(reinterpret_cast<const JT *>(base), size_t(p) * kHG + e);
    else
    {
        const __half *h = reinterpret_cast<const __half *>(base) + size_t(p) * kHG;
        return __half2float(h[e]) * ldexpf(1.f, int(__half2float(h[kHG - 1])));
    }
}

// Gauss-Jordan 6x6 inverse with partial pivoting in FP64
// (DenseGaussian<6>::Inverse, libs/star/math/DenseGaussian.hpp:34-113).
__device__ inline void invert6_fp64(const float *A, double *out)
{
    double m[6][12];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j)
        {
            m[i][j] = double(A[6 * i + j]);
            m[i][6 + j] = (i == j) ? 1.0 : 0.0;
        }
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        int piv = i;
        double pv = fabs(m[i][i]);
#pragma unroll
        for (int j = i + 1; j < 6; ++j)
            if (pv < fabs(m[j][i]))
            {
                pv = fabs(m[j][i]);
                piv = j;
            }
        if (piv != i)
        {
#pragma unroll
            for (int j = i + 1; j < 6; ++j)
                if (j == piv)
#pragma unroll
                    for (int k = 0; k < 12; ++k)
                    {
                        const double t = m[i][k];
                        m[i][k] = m[j][k];
                        m[j][k] = t;
                    }
        }
        const double inv = 1.0 / m[i][i];
#pragma unroll
        for (int k = 0; k < 12; ++k)
            m[i][k] *= inv;
#pragma unroll
        for (int r = 0; r < 6; ++r)
            if (r != i)
            {
                const double f = m[r][i];
#pragma unroll
                for (int k = 0; k < 12; ++k)
                    m[r][k] -= f * m[i][k];
            }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j)
            out[6 * i + j] = m[i][6 + j];
}

} // namespace sh
