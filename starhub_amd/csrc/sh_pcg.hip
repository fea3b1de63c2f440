// Persistent block-Jacobi PCG: the whole fixed-iteration loop of
// BlockPCG<6>::SolveNoConvergeCheck (libs/star/pcg_solver/BlockPCG.cuh:325-483)
// in one launch covering every solve of a batch.  The groups must be
// co-resident: the host keeps each grid within the kernel's occupancy bound
// (work-groups per CU x CUs), and every barrier wait is capped.
//
// Per iteration two barriers among the groups of one solve replace the
// reference's 7 launches (SpMV, 3 x cublasSdot, two update kernels, swap):
// phase A forms d_k = s + beta d_{k-1} and q = A d_k for the group's rows (the
// neighbours' d_k are recomputed from s and d_{k-1}, which no group writes
// during phase A) and sums d.q; phase B updates x, r, s and sums r.s.  FP64
// for every dot product, alpha and beta; fixed-order reductions.  The
// barriers are per solve (a batch's solves never wait for each other) and
// carry the dot products, so a barrier costs one publish and one poll.

#include "sh_common.hpp"
#include "sh_desc.hpp"
#include "sh_kernels.hpp"
#include "sh_update.hpp"

namespace sh
{

constexpr int kCoopThreads = 512;
constexpr int kCoopWaves = kCoopThreads / 64;
constexpr unsigned kSpinCap = 1u << 22; // default polls per barrier wait

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

// First row whose first block is at or after block index t (rows are whole).
__device__ __forceinline__ uint32_t row_at_block(const uint32_t *__restrict__ row_ptr, uint32_t n, uint32_t t)
{
    uint32_t lo = 0, hi = n; // answer in [0, n]
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (row_ptr[mid] + mid >= t)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ float next_dir(double beta, float dold, float s)
{
    return float(fma(beta, double(dold), double(s)));
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
    for (int w = 1; w < kCoopWaves; ++w)
        r += sh[w];
    __syncthreads();
    return r;
}

// Barrier of the G groups of one solve, fused with the sum of their partials.
// Group g publishes (value, epoch) in its own slot; wave 0 of every group waits
// for all G epochs and sums the G values in slot order, so every group gets
// the same total, bit for bit.  Values alternate between two entries by
// epoch parity: a group can only overwrite an entry two barriers later, after
// every group has read it.  A wait that exceeds the spin cap sets *err and
// ends every later wait at once, so each wave always exits.
__device__ double solve_barrier_sum(double part, PcgSlot *__restrict__ slots, uint32_t b0, uint32_t G, uint32_t ep,
                                    uint32_t *__restrict__ err, uint32_t cap, double *sh)
{
    if (threadIdx.x == 0)
    {
        PcgSlot &me = slots[blockIdx.x];
        me.v[ep & 1u] = part;
        __hip_atomic_store(&me.epoch, ep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x < 64)
    {
        double v = 0.0;
        for (uint32_t i = threadIdx.x; i < G; i += 64)
        {
            PcgSlot &o = slots[b0 + i];
            unsigned spins = 0;
            // relaxed polling (no cache invalidation per poll), one acquire
            // fence once the epoch is seen
            while (__hip_atomic_load(&o.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < ep)
            {
                __builtin_amdgcn_s_sleep(1);
                if (++spins >= cap || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                {
                    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            v += __hip_atomic_load(&o.v[ep & 1u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        if (threadIdx.x == 0)
            sh[kCoopWaves] = v;
    }
    __syncthreads();
    const double r = sh[kCoopWaves];
    __syncthreads();
    return r;
}

// Persistent PCG.  Each solve is split into G groups over contiguous node
// rows holding about equal numbers of blocks.  Phase A: every thread takes
// blocks of the group (2 in flight), forms d_k of the block's column
// (s + beta d_{k-1}, from the previous iteration's vectors) and the 6 row
// sums A_b d_k into LDS; one thread per row then adds its blocks in order,
// writes d_k and q, and the groups sum d.q through the solve barrier.
// Phase B: one thread per row updates x, r, s = M^-1 r and the groups sum r.s.
__global__ __launch_bounds__(kCoopThreads) void k_pcg_coop(const SolveDesc *__restrict__ sds,
                                                           const uint32_t *__restrict__ wg_begin, int n_solve,
                                                           int iters, PcgSlot *__restrict__ slots,
                                                           uint32_t *__restrict__ epoch_base,
                                                           uint32_t *__restrict__ err, uint32_t cap, uint32_t tile)
{
    extern __shared__ float qblk[]; // tile x 6
    __shared__ double sh[kCoopWaves + 1];
    const int si = find_solve_p(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t b0 = wg_begin[si], G = wg_begin[si + 1] - b0, g = blockIdx.x - b0;
    const uint32_t N = uint32_t(d.N);
    const uint32_t NB = d.row_ptr[N] + N;
    const uint32_t r0 = row_at_block(d.row_ptr, N, uint32_t((uint64_t(g) * NB + G - 1) / G));
    const uint32_t r1 = (g + 1 == G) ? N : row_at_block(d.row_ptr, N, uint32_t((uint64_t(g + 1) * NB + G - 1) / G));
    const uint32_t B0 = d.row_ptr[r0] + r0, B1 = d.row_ptr[r1] + r1;
    const bool fits = (B1 - B0) <= tile; // true by construction of the partition
    const uint32_t ep0 = __hip_atomic_load(epoch_base, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t ep = ep0;
    const uint32_t tid = threadIdx.x;
    if (!fits && tid == 0)
        __hip_atomic_store(err, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    // r = b, s = M^-1 b, x = 0, d_{-1} = 0  (blockPCGZeroInitKernel, BlockPCG.cuh:68-116)
    double part = 0.0;
    for (uint32_t r = r0 + tid; r < r1; r += kCoopThreads)
    {
        float rv[6];
#pragma unroll
        for (int a = 0; a < 6; ++a)
            rv[a] = d.rhs[size_t(r) * 6 + a];
        const double *M = d.dinv + size_t(r) * 36;
#pragma unroll
        for (int a = 0; a < 6; ++a)
        {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c)
                acc += M[6 * a + c] * double(rv[c]);
            const float sv = float(acc);
            d.r[size_t(r) * 6 + a] = rv[a];
            d.s[size_t(r) * 6 + a] = sv;
            d.x[size_t(r) * 6 + a] = 0.f;
            d.dv[0][size_t(r) * 6 + a] = 0.f;
            part += double(rv[a]) * double(sv);
        }
    }
    double delta = solve_barrier_sum(group_sum(part, sh), slots, b0, G, ++ep, err, cap, sh);

    double delta_prev = 0.0;
    for (int k = 0; k < iters; ++k)
    {
        // ---- phase A: d_k = s + beta d_{k-1}, q = A d_k, partial d.q
        const double beta = (k > 0 && delta_prev != 0.0) ? delta / delta_prev : 0.0;
        if (tid == 0 && g == 0)
            d.delta[k] = delta;
        const float *__restrict__ dold = pick2(d.dv, k & 1);
        float *__restrict__ dnew = pick2(d.dv, (k + 1) & 1);
        if (fits)
        {
            for (uint32_t bb0 = B0 + tid; bb0 < B1; bb0 += 2 * kCoopThreads)
            {
                uint32_t bbs[2];
                float4 A[2][9];
                float2 po[2][3], ps[2][3];
#pragma unroll
                for (int j = 0; j < 2; ++j)
                {
                    bbs[j] = bb0 + j * kCoopThreads;
                    const uint32_t bb = bbs[j] < B1 ? bbs[j] : B0;
                    const uint32_t col = d.blk_col[bb];
                    const float4 *Ab = reinterpret_cast<const float4 *>(d.blk + size_t(bb) * 36);
#pragma unroll
                    for (int v = 0; v < 9; ++v)
                        A[j][v] = Ab[v];
                    const float2 *pov = reinterpret_cast<const float2 *>(dold + size_t(col) * 6);
                    const float2 *psv = reinterpret_cast<const float2 *>(d.s + size_t(col) * 6);
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                    {
                        po[j][c] = pov[c];
                        ps[j][c] = psv[c];
                    }
                }
#pragma unroll
                for (int j = 0; j < 2; ++j)
                {
                    if (bbs[j] >= B1)
                        continue;
                    float p[6];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                    {
                        p[2 * c] = next_dir(beta, po[j][c].x, ps[j][c].x);
                        p[2 * c + 1] = next_dir(beta, po[j][c].y, ps[j][c].y);
                    }
                    const float *Af = reinterpret_cast<const float *>(A[j]);
                    float *out = qblk + size_t(bbs[j] - B0) * 6;
#pragma unroll
                    for (int a = 0; a < 6; ++a)
                    {
                        float q = 0.f;
#pragma unroll
                        for (int c = 0; c < 6; ++c)
                            q = fmaf(Af[6 * a + c], p[c], q);
                        out[a] = q;
                    }
                }
            }
        }
        __syncthreads();
        part = 0.0;
        for (uint32_t r = r0 + tid; r < r1; r += kCoopThreads)
        {
            const uint32_t e0 = d.row_ptr[r] + r - B0, e1 = d.row_ptr[r + 1] + r + 1 - B0;
#pragma unroll
            for (int a = 0; a < 6; ++a)
            {
                const size_t i6 = size_t(r) * 6 + a;
                const float dn = next_dir(beta, dold[i6], d.s[i6]);
                float q = 0.f;
                if (fits)
                    for (uint32_t e = e0; e < e1; ++e)
                        q += qblk[e * 6 + a];
                dnew[i6] = dn;
                d.q[i6] = q;
                part += double(dn) * double(q);
            }
        }
        const double dq = solve_barrier_sum(group_sum(part, sh), slots, b0, G, ++ep, err, cap, sh);

        // ---- phase B: alpha, x, r, s = M^-1 r, partial r.s
        const double alpha = (dq != 0.0) ? delta / dq : 0.0;
        part = 0.0;
        for (uint32_t r = r0 + tid; r < r1; r += kCoopThreads)
        {
            float rv[6];
#pragma unroll
            for (int a = 0; a < 6; ++a)
            {
                const size_t i6 = size_t(r) * 6 + a;
                const float dn = dnew[i6], q = d.q[i6];
                d.x[i6] = float(fma(alpha, double(dn), double(d.x[i6])));
                rv[a] = float(fma(-alpha, double(q), double(d.r[i6])));
                d.r[i6] = rv[a];
            }
            const double *M = d.dinv + size_t(r) * 36;
#pragma unroll
            for (int a = 0; a < 6; ++a)
            {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    acc += M[6 * a + c] * double(rv[c]);
                const float sv = float(acc);
                d.s[size_t(r) * 6 + a] = sv;
                part += double(rv[a]) * double(sv);
            }
        }
        delta_prev = delta;
        delta = solve_barrier_sum(group_sum(part, sh), slots, b0, G, ++ep, err, cap, sh);
    }
    // next launch starts above every epoch used here; every group has read
    // epoch_base before its first barrier, and this group has passed them all
    if (blockIdx.x == 0 && tid == 0)
        __hip_atomic_store(epoch_base, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Resident variant: the solve's rows are split evenly by count into G groups
// of at most kResRows rows and kResThreads * kResBlocksPerThread blocks (the
// host picks G so both hold).  For the whole launch every thread keeps its
// blocks (column and 36 entries) and one owned vector entry (row, a): its
// row of M^-1 and x, r, s, d, q in registers.  Per iteration the only global
// traffic left is the neighbours' s and d_{k-1} of each block, the owned s
// and d_k going out, and the two solve barriers.
constexpr int kResThreads = 512;
constexpr int kResBlocksPerThread = 3;
constexpr int kResWaves = kResThreads / 64;

__global__ __launch_bounds__(kResThreads) void k_pcg_resident(const SolveDesc *__restrict__ sds,
                                                              const uint32_t *__restrict__ wg_begin, int n_solve,
                                                              int iters, PcgSlot *__restrict__ slots,
                                                              uint32_t *__restrict__ epoch_base,
                                                              uint32_t *__restrict__ err, uint32_t cap)
{
    __shared__ float qblk[kResThreads * kResBlocksPerThread * 6];
    __shared__ float rsh[kResThreads];
    __shared__ double sh[kCoopWaves + 1];
    const int si = find_solve_p(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t b0 = wg_begin[si], G = wg_begin[si + 1] - b0, g = blockIdx.x - b0;
    const uint32_t N = uint32_t(d.N);
    const uint32_t r0 = uint32_t(uint64_t(g) * N / G), r1 = uint32_t(uint64_t(g + 1) * N / G);
    const uint32_t B0 = d.row_ptr[r0] + r0, B1 = d.row_ptr[r1] + r1;
    const uint32_t tid = threadIdx.x;
    const bool fits = (r1 - r0) * 6 <= uint32_t(kResThreads) && (B1 - B0) <= uint32_t(kResThreads * kResBlocksPerThread);
    if (!fits && tid == 0)
        __hip_atomic_store(err, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t ep0 = __hip_atomic_load(epoch_base, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t ep = ep0;

    // resident blocks
    uint32_t col[kResBlocksPerThread];
    float A[kResBlocksPerThread][36];
#pragma unroll
    for (int j = 0; j < kResBlocksPerThread; ++j)
    {
        const uint32_t bb = B0 + tid + j * kResThreads;
        const bool on = fits && bb < B1;
        col[j] = on ? d.blk_col[bb] : 0xffffffffu;
        const float4 *Ab = reinterpret_cast<const float4 *>(d.blk + size_t(on ? bb : 0) * 36);
#pragma unroll
        for (int v = 0; v < 9; ++v)
        {
            const float4 f = on ? Ab[v] : make_float4(0.f, 0.f, 0.f, 0.f);
            A[j][4 * v] = f.x;
            A[j][4 * v + 1] = f.y;
            A[j][4 * v + 2] = f.z;
            A[j][4 * v + 3] = f.w;
        }
    }
    // owned entry
    const bool own = fits && tid < (r1 - r0) * 6;
    const uint32_t row = r0 + tid / 6;
    const int a = int(tid % 6);
    const size_t i6 = size_t(row) * 6 + a;
    const uint32_t lrow = (tid / 6) * 6; // first rsh entry of the row
    double M[6];
    uint32_t e0 = 0, e1 = 0;
#pragma unroll
    for (int c = 0; c < 6; ++c)
        M[c] = own ? d.dinv[size_t(row) * 36 + 6 * a + c] : 0.0;
    if (own)
    {
        e0 = d.row_ptr[row] + row - B0;
        e1 = d.row_ptr[row + 1] + row + 1 - B0;
    }

    // r = b, s = M^-1 b, x = 0, d_{-1} = 0  (blockPCGZeroInitKernel, BlockPCG.cuh:68-116)
    float r = own ? d.rhs[i6] : 0.f;
    float x = 0.f, dk = 0.f, s = 0.f;
    rsh[tid] = r;
    __syncthreads();
    if (own)
    {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            acc += M[c] * double(rsh[lrow + c]);
        s = float(acc);
        d.s[i6] = s;
        d.dv[0][i6] = 0.f;
    }
    double delta = solve_barrier_sum(group_sum(own ? double(r) * double(s) : 0.0, sh), slots, b0, G, ++ep, err, cap, sh);

    double delta_prev = 0.0;
    for (int k = 0; k < iters; ++k)
    {
        // ---- phase A: d_k = s + beta d_{k-1}, q = A d_k, partial d.q
        const double beta = (k > 0 && delta_prev != 0.0) ? delta / delta_prev : 0.0;
        if (tid == 0 && g == 0)
            d.delta[k] = delta;
        const float *__restrict__ dold = pick2(d.dv, k & 1);
        float *__restrict__ dnew = pick2(d.dv, (k + 1) & 1);
        float2 po[kResBlocksPerThread][3], ps[kResBlocksPerThread][3];
#pragma unroll
        for (int j = 0; j < kResBlocksPerThread; ++j)
        {
            const size_t c6 = size_t(col[j] != 0xffffffffu ? col[j] : 0) * 6;
            const float2 *pov = reinterpret_cast<const float2 *>(dold + c6);
            const float2 *psv = reinterpret_cast<const float2 *>(d.s + c6);
#pragma unroll
            for (int c = 0; c < 3; ++c)
            {
                po[j][c] = pov[c];
                ps[j][c] = psv[c];
            }
        }
        const float dn = next_dir(beta, dk, s);
        if (own)
            dnew[i6] = dn;
#pragma unroll
        for (int j = 0; j < kResBlocksPerThread; ++j)
        {
            if (col[j] == 0xffffffffu)
                continue;
            float p[6];
#pragma unroll
            for (int c = 0; c < 3; ++c)
            {
                p[2 * c] = next_dir(beta, po[j][c].x, ps[j][c].x);
                p[2 * c + 1] = next_dir(beta, po[j][c].y, ps[j][c].y);
            }
            float *out = qblk + size_t(tid + j * kResThreads) * 6;
#pragma unroll
            for (int aa = 0; aa < 6; ++aa)
            {
                float q = 0.f;
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    q = fmaf(A[j][6 * aa + c], p[c], q);
                out[aa] = q;
            }
        }
        __syncthreads();
        float q = 0.f;
        for (uint32_t e = e0; e < e1; ++e)
            q += qblk[e * 6 + a];
        const double dq =
            solve_barrier_sum(group_sum(own ? double(dn) * double(q) : 0.0, sh), slots, b0, G, ++ep, err, cap, sh);

        // ---- phase B: alpha, x, r, s = M^-1 r, partial r.s
        const double alpha = (dq != 0.0) ? delta / dq : 0.0;
        x = float(fma(alpha, double(dn), double(x)));
        r = float(fma(-alpha, double(q), double(r)));
        rsh[tid] = r;
        __syncthreads();
        if (own)
        {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c)
                acc += M[c] * double(rsh[lrow + c]);
            s = float(acc);
            d.s[i6] = s;
        }
        dk = dn;
        delta_prev = delta;
        delta = solve_barrier_sum(group_sum(own ? double(r) * double(s) : 0.0, sh), slots, b0, G, ++ep, err, cap, sh);
    }
    if (own)
    {
        d.x[i6] = x;
        d.r[i6] = r;
    }
    if (blockIdx.x == 0 && tid == 0)
        __hip_atomic_store(epoch_base, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pipelined variant (Ghysels and Vanroose, "Hiding global synchronization
// latency in the preconditioned conjugate gradient algorithm", Parallel
// Computing 40(7), 2014, Alg. 4): the same Krylov iterates as the reference's
// PCG in exact arithmetic, with ONE solve barrier per iteration instead of
// two.  Per iteration the two dot products gamma = r.u and delta = w.u
// (u = M^-1 r, w = A u) are local, m = M^-1 w is local, and the barrier both
// sums the dots and publishes m; n = A m then reads the neighbours' m:
//   beta = gamma / gamma_prev, alpha = gamma / (delta - beta gamma / alpha_prev)
//   z = n + beta z, q = m + beta q, s = w + beta s, p = u + beta p
//   x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z.
// The recurrences drift from the directly computed vectors in FP32, so every
// vector is FP64 here (registers; the exchanged m as FP64 in memory) and the
// blocks stay FP32 in registers, converted per product.  gamma_k is the
// reference's delta_k = r_k . M^-1 r_k.  Same partition and register layout
// as the resident kernel.
// distinct block columns of a group kept in LDS (more: read from memory per block)
constexpr int kPipeCols = 768;
constexpr int kPipeFetch = (kPipeCols * 6 + kResThreads - 1) / kResThreads;

struct PipeGroupSum
{
    double a, b;
};

// Tagged publication.  A value published at sequence number ep goes to
// buffer ep & 1 with ((ep + 1) >> 1) & 1 in the lowest bit of its mantissa
// (one ulp, the same value for every reader).  The buffer's previous value,
// from ep - 2, carries the other tag, and older values can no longer be seen:
// every thread waits for its stores to complete (s_waitcnt vmcnt(0)) before
// its group publishes the next sequence number, which a reader of ep has
// waited for.  The kernel before each launch zeroes the launch's slots and
// vectors (tag 0), and a launch starts its sequence numbers after a multiple
// of 4, so its first two carry tag 1 and both slot entries are written by
// then.  No group runs two sequence numbers ahead of another, because each
// needs the others' values of the previous one.  A reader polls each value
// until it carries the expected tag, so every value says by itself that it is
// fresh: no epoch word, no cache flush or invalidation, and the solve's sums
// and the halo arrive in one round trip.  Every value is one 8-byte atomic
// access at agent scope.
__device__ __forceinline__ uint64_t pub_tag(uint32_t ep)
{
    return uint64_t(((ep + 1u) >> 1) & 1u);
}

__device__ __forceinline__ double with_tag(double v, uint64_t t)
{
    return __longlong_as_double(
        static_cast<long long>((static_cast<uint64_t>(__double_as_longlong(v)) & ~1ull) | t));
}

__device__ __forceinline__ bool has_tag(double v, uint64_t t)
{
    return (static_cast<uint64_t>(__double_as_longlong(v)) & 1ull) == t;
}

__device__ __forceinline__ void store_pub(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double load_pub(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Re-reads *p until it carries tag t (v is the first read).  A wait that
// exceeds the spin cap sets *err and ends every later wait at once, so each
// wave always exits.
__device__ __forceinline__ double await_pub(const double *p, double v, uint64_t t, uint32_t *__restrict__ err,
                                            uint32_t cap)
{
    unsigned spins = 0;
    while (!has_tag(v, t))
    {
        __builtin_amdgcn_s_sleep(1);
        if (++spins >= cap || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        {
            __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        v = load_pub(p);
    }
    return v;
}

// The solve's two sums of sequence number ep: the group's tagged partials go
// to its slot, and wave 0 reads every group's in slot order and sums them in a
// fixed tree, so every group gets the same totals, bit for bit; thread 0
// leaves them in out[0..1] (read after the caller's next group sync).
__device__ void publish_sum2(PipeGroupSum part, PcgSlot2 *__restrict__ slots, uint32_t self, uint32_t b0, uint32_t G,
                             uint32_t ep, uint32_t *__restrict__ err, uint32_t cap, double *out)
{
    const uint64_t t = pub_tag(ep);
    if (threadIdx.x == 0)
    {
        PcgSlot2 &me = slots[self];
        store_pub(&me.v[ep & 1u][0], with_tag(part.a, t));
        store_pub(&me.v[ep & 1u][1], with_tag(part.b, t));
    }
    if (threadIdx.x < 64)
    {
        // one slot per lane, in rounds of 64 when the solve has more groups
        // (registers for one round only: the loop below is uniform)
        double va = 0.0, vb = 0.0;
        for (uint32_t c = 0; c < G; c += 64u)
        {
            const uint32_t i = c + threadIdx.x;
            if (i < G)
            {
                const double *pa = &slots[b0 + i].v[ep & 1u][0];
                const double a = load_pub(pa), b = load_pub(pa + 1);
                va += await_pub(pa, a, t, err, cap);
                vb += await_pub(pa + 1, b, t, err, cap);
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
        {
            va += __shfl_xor(va, o, 64);
            vb += __shfl_xor(vb, o, 64);
        }
        if (threadIdx.x == 0)
        {
            out[0] = va;
            out[1] = vb;
        }
    }
}

__global__ __launch_bounds__(kResThreads) void k_pcg_pipe(const SolveDesc *__restrict__ sds,
                                                          const uint32_t *__restrict__ wg_begin, int n_solve, int iters,
                                                          PcgSlot2 *__restrict__ slots,
                                                          uint32_t *__restrict__ epoch_base,
                                                          uint32_t *__restrict__ err, uint32_t cap,
                                                          int update_parity)
{
    __shared__ double qblk[kResThreads * kResBlocksPerThread * 6];
    __shared__ double rsh[kResThreads];
    __shared__ double sh[2 * kResWaves];
    __shared__ double sh2[2];
    __shared__ double halo[kPipeCols * 6];
    __shared__ uint16_t clist[kPipeCols];
    __shared__ uint32_t cbits[kBitmapWords];
    __shared__ uint32_t cpre[kBitmapWords];
    __shared__ uint32_t ncols_sh;
    // a launch may cover a subset of the solves (wg_begin points into the
    // whole batch's map): its groups are numbered from wg_begin[0]
    const uint32_t bid = blockIdx.x + wg_begin[0];
    const int si = find_solve_p(wg_begin, n_solve, bid);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t b0 = wg_begin[si], G = wg_begin[si + 1] - b0, g = bid - b0;
    const uint32_t N = uint32_t(d.N);
    const uint32_t r0 = uint32_t(uint64_t(g) * N / G), r1 = uint32_t(uint64_t(g + 1) * N / G);
    const uint32_t B0 = d.row_ptr[r0] + r0, B1 = d.row_ptr[r1] + r1;
    const uint32_t tid = threadIdx.x;
    const bool fits = (r1 - r0) * 6 <= uint32_t(kResThreads) && (B1 - B0) <= uint32_t(kResThreads * kResBlocksPerThread);
    if (!fits && tid == 0)
        __hip_atomic_store(err, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // publications of this launch start after a multiple of 4 (tag 1 twice)
    uint32_t ep = (__hip_atomic_load(epoch_base, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 3u) & ~3u;

    uint32_t col[kResBlocksPerThread];
    float A[kResBlocksPerThread][36];
#pragma unroll
    for (int j = 0; j < kResBlocksPerThread; ++j)
    {
        const uint32_t bb = B0 + tid + j * kResThreads;
        const bool on = fits && bb < B1;
        col[j] = on ? d.blk_col[bb] : 0xffffffffu;
        const float4 *Ab = reinterpret_cast<const float4 *>(d.blk + size_t(on ? bb : 0) * 36);
#pragma unroll
        for (int v = 0; v < 9; ++v)
        {
            const float4 f = on ? Ab[v] : make_float4(0.f, 0.f, 0.f, 0.f);
            A[j][4 * v] = f.x;
            A[j][4 * v + 1] = f.y;
            A[j][4 * v + 2] = f.z;
            A[j][4 * v + 3] = f.w;
        }
    }
    // the group's distinct block columns (ascending), and each block's rank
    // among them: the exchanged vector is fetched once per distinct node
    for (uint32_t w = tid; w < kBitmapWords; w += kResThreads)
        cbits[w] = 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kResBlocksPerThread; ++j)
        if (col[j] != 0xffffffffu)
            atomicOr(&cbits[col[j] >> 5], 1u << (col[j] & 31)); // set union: order independent
    __syncthreads();
    if (tid < 64)
    {
        const uint32_t c0 = __popc(cbits[2 * tid]), c1 = __popc(cbits[2 * tid + 1]);
        uint32_t v = c0 + c1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t u = __shfl_up(v, o, 64);
            if (int(tid) >= o)
                v += u;
        }
        cpre[2 * tid] = v - c0 - c1;
        cpre[2 * tid + 1] = v - c1;
        if (tid == 63)
            ncols_sh = v;
    }
    __syncthreads();
    const uint32_t ncols = ncols_sh;
    const bool lds_halo = ncols <= uint32_t(kPipeCols); // uniform over the group
    uint32_t lcol[kResBlocksPerThread];
#pragma unroll
    for (int j = 0; j < kResBlocksPerThread; ++j)
    {
        lcol[j] = 0;
        if (col[j] != 0xffffffffu)
        {
            const uint32_t w = col[j] >> 5;
            lcol[j] = cpre[w] + __popc(cbits[w] & ((1u << (col[j] & 31)) - 1u));
            if (lds_halo)
                clist[lcol[j]] = uint16_t(col[j]); // every writer of a rank writes the same id
        }
    }
    __syncthreads();

    const bool own = fits && tid < (r1 - r0) * 6;
    const uint32_t row = r0 + tid / 6;
    const int a = int(tid % 6);
    const size_t i6 = size_t(row) * 6 + a;
    const uint32_t lrow = (tid / 6) * 6;
    double M[6];
    uint32_t e0 = 0, e1 = 0;
#pragma unroll
    for (int c = 0; c < 6; ++c)
        M[c] = own ? d.dinv[size_t(row) * 36 + 6 * a + c] : 0.0;
    if (own)
    {
        e0 = d.row_ptr[row] + row - B0;
        e1 = d.row_ptr[row + 1] + row + 1 - B0;
    }
    // (A v)_a of the owned entry, v published in vb with tag t by every group
    // of the solve (each value is awaited as it is read); two group syncs
    auto apply = [&](const double *__restrict__ vb, uint64_t t) -> double {
        double2 pv[kResBlocksPerThread][3];
        // the blocks look modified here, so their FP64 conversions stay inside
        // the loop instead of being hoisted into twice the registers
#pragma unroll
        for (int j = 0; j < kResBlocksPerThread; ++j)
#pragma unroll
            for (int e = 0; e < 36; ++e)
                __asm__ volatile("" : "+v"(A[j][e]));
        if (lds_halo)
        {
            // the distinct nodes' six entries, in 48-byte runs of ascending nodes
            const uint32_t n6 = ncols * 6;
            double hv[kPipeFetch];
#pragma unroll
            for (int f = 0; f < kPipeFetch; ++f)
            {
                const uint32_t i = tid + uint32_t(f) * kResThreads;
                hv[f] = i < n6 ? load_pub(vb + size_t(clist[i / 6]) * 6 + i % 6) : 0.0;
            }
#pragma unroll
            for (int f = 0; f < kPipeFetch; ++f)
            {
                const uint32_t i = tid + uint32_t(f) * kResThreads;
                if (i < n6)
                    halo[i] = await_pub(vb + size_t(clist[i / 6]) * 6 + i % 6, hv[f], t, err, cap);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kResBlocksPerThread; ++j)
            {
                const double2 *src = reinterpret_cast<const double2 *>(halo + size_t(lcol[j]) * 6);
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    pv[j][c] = src[c];
            }
        }
        else
        {
#pragma unroll
            for (int j = 0; j < kResBlocksPerThread; ++j)
            {
                double v6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                if (col[j] != 0xffffffffu)
                {
                    const double *src = vb + size_t(col[j]) * 6;
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        v6[c] = load_pub(src + c);
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        v6[c] = await_pub(src + c, v6[c], t, err, cap);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    pv[j][c] = make_double2(v6[2 * c], v6[2 * c + 1]);
            }
            __syncthreads(); // keeps the two paths' sync counts equal
        }
#pragma unroll
        for (int j = 0; j < kResBlocksPerThread; ++j)
        {
            if (col[j] == 0xffffffffu)
                continue;
            const double p[6] = {pv[j][0].x, pv[j][0].y, pv[j][1].x, pv[j][1].y, pv[j][2].x, pv[j][2].y};
            double *out = qblk + size_t(tid + j * kResThreads) * 6;
#pragma unroll
            for (int aa = 0; aa < 6; ++aa)
            {
                double qq = 0.0;
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    qq = fma(double(A[j][6 * aa + c]), p[c], qq);
                out[aa] = qq;
            }
        }
        __syncthreads();
        // the row's block products: four interleaved partial sums, then a fixed tree
        double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
        uint32_t e = e0;
        for (; e + 4 <= e1; e += 4)
        {
            q0 += qblk[e * 6 + a];
            q1 += qblk[(e + 1) * 6 + a];
            q2 += qblk[(e + 2) * 6 + a];
            q3 += qblk[(e + 3) * 6 + a];
        }
        for (; e < e1; ++e)
            q0 += qblk[e * 6 + a];
        return (q0 + q1) + (q2 + q3);
    };
    // Stage 1 of an iteration: v's row entries and the wave partials of two
    // dot products go to LDS, one sync, then (M^-1 v)_a and the group sums.
    auto stage = [&](double v, double da, double db, double &mv, PipeGroupSum &gs) {
        __builtin_amdgcn_s_waitcnt(0x0F70); // vmcnt(0): this thread's publications are complete
        rsh[tid] = v;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
        {
            da += __shfl_xor(da, o, 64);
            db += __shfl_xor(db, o, 64);
        }
        if ((tid & 63) == 0)
        {
            sh[2 * (tid >> 6)] = da;
            sh[2 * (tid >> 6) + 1] = db;
        }
        __syncthreads();
        double acc = 0.0;
        if (own)
#pragma unroll
            for (int c = 0; c < 6; ++c)
                acc += M[c] * rsh[lrow + c];
        mv = acc;
        gs = PipeGroupSum{sh[0], sh[1]};
#pragma unroll
        for (int w = 1; w < kResWaves; ++w)
        {
            gs.a += sh[2 * w];
            gs.b += sh[2 * w + 1];
        }
    };

    // r = b, u = M^-1 r, w = A u (x = 0: blockPCGZeroInitKernel, BlockPCG.cuh:68-116)
    double r = own ? double(d.rhs[i6]) : 0.0;
    double u = 0.0;
    PipeGroupSum gs{};
    stage(r, 0.0, 0.0, u, gs);
    // every group uses the published (tagged) u and m, its own entries too
    ++ep;
    u = with_tag(u, pub_tag(ep));
    if (own)
        store_pub(pick2(d.mv, ep & 1u) + i6, u);
    if (tid == 0) // the slot entry this number selects is written too (no sums yet)
    {
        store_pub(&slots[bid].v[ep & 1u][0], with_tag(0.0, pub_tag(ep)));
        store_pub(&slots[bid].v[ep & 1u][1], with_tag(0.0, pub_tag(ep)));
    }
    double w = apply(pick2(d.mv, ep & 1u), pub_tag(ep));
    double x = 0.0, z = 0.0, q = 0.0, s = 0.0, p = 0.0, gprev = 0.0, aprev = 0.0;
    for (int k = 0; k < iters; ++k)
    {
        double m = 0.0;
        stage(w, own ? r * u : 0.0, own ? w * u : 0.0, m, gs);
        ++ep;
        const uint64_t t = pub_tag(ep);
        double *__restrict__ mb = pick2(d.mv, ep & 1u);
        m = with_tag(m, t);
        if (own)
            store_pub(mb + i6, m);
        publish_sum2(gs, slots, bid, b0, G, ep, err, cap, sh2);
        const double n = apply(mb, t); // its group syncs order sh2 for every thread
        const PipeGroupSum gd{sh2[0], sh2[1]};
        const double gam = gd.a, del = gd.b;
        if (tid == 0 && g == 0)
            d.delta[k] = gam;
        const double beta = (k > 0 && gprev != 0.0) ? gam / gprev : 0.0;
        const double den = del - ((k > 0 && aprev != 0.0) ? beta * gam / aprev : 0.0);
        const double alpha = (den != 0.0) ? gam / den : 0.0;
        z = fma(beta, z, n);
        q = fma(beta, q, m);
        s = fma(beta, s, w);
        p = fma(beta, p, u);
        x = fma(alpha, p, x);
        r = fma(-alpha, s, r);
        u = fma(-alpha, q, u);
        w = fma(-alpha, z, w);
        gprev = gam;
        aprev = alpha;
    }
    if (own)
    {
        d.x[i6] = float(x);
        d.r[i6] = float(r);
    }
    if (update_parity >= 0)
    {
        // the warp update of the group's own nodes, from the FP32 x as stored
        // (k_update's work, with its cost reduction on each solve's group 0)
        rsh[tid] = own ? double(float(x)) : 0.0;
        __syncthreads();
        if (g == 0 && tid < 64)
            reduce_cost(d, d.gn_iter + update_parity);
        if (tid < r1 - r0)
        {
            float xv[6];
#pragma unroll
            for (int c = 0; c < 6; ++c)
                xv[c] = float(rsh[tid * 6 + c]);
            warp_update_node(d, r0 + tid, update_parity, xv);
        }
    }
    if (blockIdx.x == 0 && tid == 0)
        __hip_atomic_store(epoch_base, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

hipError_t launch_pcg_pipe(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int iters,
                           PcgSlot2 *slots, uint32_t *epoch_base, uint32_t *err, uint32_t wait_limit,
                           int update_parity, hipStream_t st)
{
    // A plain launch: the kernel needs its groups co-resident (the host keeps
    // n_wg within the occupancy bound, pcg_pipe_max_groups) but no grid-wide
    // sync, and every barrier wait is capped, so a group that is placed late
    // only delays the others.  The cooperative launch API adds a per-launch
    // cost of about 20 us on this device.
    hipLaunchKernelGGL(k_pcg_pipe, dim3(n_wg), dim3(kResThreads), 0, st, sds, wmap, n_solve, iters, slots, epoch_base,
                       err, wait_limit ? wait_limit : kSpinCap, update_parity);
    return hipGetLastError();
}

int pcg_pipe_max_groups()
{
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_pcg_pipe), kResThreads,
                                                     0) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}

hipError_t launch_pcg_resident(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int iters,
                               PcgSlot *slots, uint32_t *epoch_base, uint32_t *err, uint32_t wait_limit,
                               hipStream_t st)
{
    // plain launch, as for k_pcg_pipe (n_wg within pcg_resident_max_groups)
    hipLaunchKernelGGL(k_pcg_resident, dim3(n_wg), dim3(kResThreads), 0, st, sds, wmap, n_solve, iters, slots,
                       epoch_base, err, wait_limit ? wait_limit : kSpinCap);
    return hipGetLastError();
}

int pcg_resident_max_groups()
{
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_pcg_resident),
                                                     kResThreads, 0) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}

hipError_t launch_pcg_coop(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int iters,
                           PcgSlot *slots, uint32_t *epoch_base, uint32_t *err, uint32_t wait_limit, uint32_t tile,
                           hipStream_t st)
{
    const size_t lds = size_t(tile) * 6 * sizeof(float);
    if (lds > 65536)
    {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_pcg_coop),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess)
            return e;
    }
    // plain launch, as for k_pcg_pipe (n_wg within pcg_coop_max_groups(lds))
    hipLaunchKernelGGL(k_pcg_coop, dim3(n_wg), dim3(kCoopThreads), lds, st, sds, wmap, n_solve, iters, slots, epoch_base,
                       err, wait_limit ? wait_limit : kSpinCap, tile);
    return hipGetLastError();
}

int pcg_coop_max_groups(size_t lds)
{
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_pcg_coop),
                                                     kCoopThreads, lds) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}
} // namespace sh
