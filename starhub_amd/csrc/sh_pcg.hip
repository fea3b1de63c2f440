// Persistent block-Jacobi PCG: the whole fixed-iteration loop of
// BlockPCG<6>::SolveNoConvergeCheck (libs/star/pcg_solver/BlockPCG.cuh:325-483)
// in one cooperative launch covering every solve of a batch.
//
// Per iteration two grid barriers replace the reference's 7 launches
// (SpMV, 3 x cublasSdot, two update kernels, swap): phase A forms
// d_k = s + beta d_{k-1} and q = A d_k for the group's own node rows (the
// neighbours' d_k are recomputed from s and d_{k-1}, which no group writes
// during phase A) and publishes the partial d.q; phase B updates x, r, s and
// publishes the partial r.s.  x and r of a row stay in registers; FP64 for
// every dot product, alpha and beta; fixed-order reductions.  All blocks of a
// row are read at once (one lane per block), so a row's SpMV waits on one
// round of loads.
#include <hip/hip_cooperative_groups.h>

#include "sh_common.hpp"
#include "sh_desc.hpp"
#include "sh_kernels.hpp"

namespace sh
{
namespace cg = cooperative_groups;

// One wave per node row, 4 rows per work-group of 256 threads.  Lane b takes
// block b of the row (and b + 64, ...): its 36 entries and the neighbour's 6
// direction entries arrive in one round of loads, and the 6 row sums are
// reduced across the wave in a fixed xor tree.  Lanes 0..5 own the row's
// x, r, s and d entries (a = lane).
constexpr int kCoopNodes = int(kCoopNodesPerWg);
constexpr int kCoopThreads = 64 * kCoopNodes;

__device__ __forceinline__ int find_solve_p(const uint32_t *__restrict__ begin, int n, uint32_t wg)
{
    int lo = 0, hi = n;
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (begin[mid] <= wg)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Sum over the work-group, fixed order: wave trees, then the waves in order.
__device__ __forceinline__ double group_sum(double v, double *sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int w = 1; w < kCoopNodes; ++w)
        r += sh[w];
    __syncthreads();
    return r;
}

// Sum of the work-group partials of one solve, fixed order.
__device__ __forceinline__ double partial_total(const double *__restrict__ part, uint32_t b0, uint32_t b1, double *sh)
{
    double v = 0.0;
    if (threadIdx.x < 64)
    {
        for (uint32_t i = b0 + threadIdx.x; i < b1; i += 64)
            v += part[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        if (threadIdx.x == 0)
            sh[kCoopNodes] = v;
    }
    __syncthreads();
    const double r = sh[kCoopNodes];
    __syncthreads();
    return r;
}

__device__ __forceinline__ float next_dir(double beta, float dold, float s)
{
    return float(fma(beta, double(dold), double(s)));
}

__global__ __launch_bounds__(kCoopThreads) void k_pcg_coop(const SolveDesc *__restrict__ sds,
                                                           const uint32_t *__restrict__ wg_begin, int n_solve,
                                                           int iters, double *__restrict__ part_rs,
                                                           double *__restrict__ part_dq)
{
    cg::grid_group grid = cg::this_grid();
    __shared__ double sh[kCoopNodes + 1];
    const int si = find_solve_p(wg_begin, n_solve, blockIdx.x);
    const SolveDesc &d = sds[si];
    const uint32_t b0 = wg_begin[si], b1 = wg_begin[si + 1];
    const int lane = threadIdx.x & 63;
    const uint32_t node = (blockIdx.x - b0) * kCoopNodes + (threadIdx.x >> 6);
    const bool live = node < uint32_t(d.N); // uniform over the wave
    const bool owner = live && lane < 6;
    const int a = owner ? lane : 0;
    const size_t row = size_t(node) * 6 + a;
    double M[6];
#pragma unroll
    for (int c = 0; c < 6; ++c)
        M[c] = owner ? d.dinv[size_t(node) * 36 + 6 * a + c] : 0.0;
    uint32_t rb0 = 0, rb1 = 0;
    if (live)
    {
        rb0 = d.row_ptr[node] + node;
        rb1 = d.row_ptr[node + 1] + node + 1;
    }

    // r = b, s = M^-1 b, x = 0, d_{-1} = 0  (blockPCGZeroInitKernel, BlockPCG.cuh:68-116)
    float r = owner ? d.rhs[row] : 0.f;
    float x = 0.f, dk = 0.f, s = 0.f;
    {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            acc += M[c] * double(__shfl(r, c, 64));
        if (owner)
        {
            s = float(acc);
            d.s[row] = s;
            d.dv[0][row] = 0.f;
        }
    }
    double tot = group_sum(owner ? double(r) * double(s) : 0.0, sh);
    if (threadIdx.x == 0)
        part_rs[blockIdx.x] = tot;
    grid.sync();

    double delta_prev = 0.0;
    for (int k = 0; k < iters; ++k)
    {
        // ---- phase A: d_k = s + beta d_{k-1}, q = A d_k, partial d.q
        const double delta = partial_total(part_rs, b0, b1, sh);
        const double beta = (k > 0 && delta_prev != 0.0) ? delta / delta_prev : 0.0;
        if (threadIdx.x == 0 && blockIdx.x == b0)
            d.delta[k] = delta;
        const float *__restrict__ dold = d.dv[k & 1];
        float *__restrict__ dnew = d.dv[(k + 1) & 1];
        const float dn = next_dir(beta, dk, s);
        if (owner)
            dnew[row] = dn;
        float qa[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (uint32_t bb = rb0 + lane; bb < rb1; bb += 64)
        {
            const uint32_t col = d.blk_col[bb];
            const float4 *A = reinterpret_cast<const float4 *>(d.blk + size_t(bb) * 36);
            const float2 *po = reinterpret_cast<const float2 *>(dold + size_t(col) * 6);
            const float2 *ps = reinterpret_cast<const float2 *>(d.s + size_t(col) * 6);
            float p[6];
#pragma unroll
            for (int c = 0; c < 3; ++c)
            {
                const float2 ov = po[c], sv = ps[c];
                p[2 * c] = next_dir(beta, ov.x, sv.x);
                p[2 * c + 1] = next_dir(beta, ov.y, sv.y);
            }
            float Av[36];
#pragma unroll
            for (int v = 0; v < 9; ++v)
            {
                const float4 f = A[v];
                Av[4 * v] = f.x;
                Av[4 * v + 1] = f.y;
                Av[4 * v + 2] = f.z;
                Av[4 * v + 3] = f.w;
            }
#pragma unroll
            for (int r6 = 0; r6 < 6; ++r6)
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    qa[r6] = fmaf(Av[6 * r6 + c], p[c], qa[r6]);
        }
        float q = 0.f;
#pragma unroll
        for (int r6 = 0; r6 < 6; ++r6)
        {
            const float v = wave_sum(qa[r6]);
            q = (a == r6) ? v : q;
        }
        tot = group_sum(owner ? double(dn) * double(q) : 0.0, sh);
        if (threadIdx.x == 0)
            part_dq[blockIdx.x] = tot;
        grid.sync();

        // ---- phase B: alpha, x, r, s = M^-1 r, partial r.s
        const double dq = partial_total(part_dq, b0, b1, sh);
        const double alpha = (dq != 0.0) ? delta / dq : 0.0;
        if (owner)
        {
            x = float(fma(alpha, double(dn), double(x)));
            r = float(fma(-alpha, double(q), double(r)));
        }
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            acc += M[c] * double(__shfl(r, c, 64));
        if (owner)
        {
            s = float(acc);
            d.s[row] = s;
        }
        tot = group_sum(owner ? double(r) * double(s) : 0.0, sh);
        if (threadIdx.x == 0)
            part_rs[blockIdx.x] = tot;
        dk = dn;
        delta_prev = delta;
        grid.sync();
    }
    if (owner)
    {
        d.x[row] = x;
        d.r[row] = r;
    }
}

hipError_t launch_pcg_coop(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int iters,
                           double *part_rs, double *part_dq, hipStream_t st)
{
    void *args[] = {(void *)&sds, (void *)&wmap, (void *)&n_solve, (void *)&iters, (void *)&part_rs, (void *)&part_dq};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_pcg_coop), dim3(n_wg), dim3(kCoopThreads), args,
                                      0, st);
}

int pcg_coop_max_groups()
{
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_pcg_coop),
                                                     kCoopThreads, 0) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}
} // namespace sh
