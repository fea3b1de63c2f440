// The reference's linear-solver plug-in: BlockPCG<6> over an abstract
// ApplySpMVBase<6> operator (libs/star/pcg_solver/BlockPCG.h:14-115,
// BlockPCG.cuh:16-483, ApplySpMVBase.h:9-47), for callers that bring their
// own system.  The warp solver itself runs its PCG inside the persistent
// kernels of sh_pcg.hip; this is the standalone form with the reference's
// call sequence: SetSolverInput(inv_diag, spmv, b, x) and
// Solve(iterations, zero_init).
//
// Per iteration: the operator's q = A d (a host callback that enqueues on the
// solve's stream, or the built-in block-CSR product), then three kernels:
//   k_bp_dot     partial sums of d.q, one per work-group (FP64)
//   k_bp_update  alpha = delta / (d.q); x += alpha d; t = r - alpha q;
//                s = M^-1 t; partial sums of t.s
//   k_bp_dir     delta' = t.s; beta = delta' / delta; d = s + beta d
// Every work-group reduces the same partials in the same order, so alpha and
// beta are the same bits everywhere; vectors are FP32 as in the reference,
// dot products, alpha and beta FP64 (cuBLAS Sdot in the reference,
// BlockPCG.cuh:418-471).  A zero system gives alpha = beta = 0 instead of 0/0.
#include <algorithm>
#include <cmath>
#include <utility>

#include "../../include/sh/warp_solver.h"
#include "sh_common.hpp"
#include "sh_engine.hpp"

#pragma clang fp contract(off)

namespace sh
{
namespace
{
constexpr int kBpThreads = 256; // one 6-row block per thread

// Sum of `n` FP64 partials in a fixed order (lane j takes j, j+64, ...; then a
// fixed shuffle tree); every wave that calls it gets the same bits.
__device__ inline double sum_partials(const double *__restrict__ part, uint32_t n)
{
    double v = 0.0;
    for (uint32_t i = threadIdx.x % 64u; i < n; i += 64)
        v += part[i];
    return wave_sum(v);
}

// Work-group sum of one value per thread, in a fixed order: wave sums, then
// the waves' sums in wave order.  Written by thread 0 to out.
__device__ inline void block_sum_store(double v, double *__restrict__ out)
{
    __shared__ double ws[kBpThreads / 64];
    v = wave_sum(v);
    if (threadIdx.x % 64u == 0)
        ws[threadIdx.x / 64u] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double s = 0.0;
        for (int w = 0; w < kBpThreads / 64; ++w)
            s += ws[w];
        *out = s;
    }
}

// y = M^-1 v for 6-row block b (M^-1: 36 FP32 per block, row-major), FP64 accumulation
__device__ inline void precond6(const float *__restrict__ minv, const float *__restrict__ v, uint32_t b, float *y)
{
    float vb[6];
#pragma unroll
    for (int j = 0; j < 6; ++j)
        vb[j] = v[6 * size_t(b) + j];
    const float *m = minv + 36 * size_t(b);
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j)
            acc += double(m[6 * i + j]) * double(vb[j]);
        y[i] = float(acc);
    }
}

__device__ inline double dot6(const float *a, const float *__restrict__ v, uint32_t b)
{
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j)
        acc += double(a[j]) * double(v[6 * size_t(b) + j]);
    return acc;
}

// r = b, x = 0 (zero_init) or r given (InitResidual already ran); d = M^-1 r;
// partials of r.d  (blockPCGZeroInitKernel / Initialize, BlockPCG.cuh:395-440)
__global__ void k_bp_init(const float *__restrict__ b, const float *__restrict__ minv, float *__restrict__ r,
                          float *__restrict__ d, float *__restrict__ x, uint32_t nblk, int zero_init,
                          double *__restrict__ part)
{
    const uint32_t blk = blockIdx.x * kBpThreads + threadIdx.x;
    double v = 0.0;
    if (blk < nblk)
    {
        if (zero_init)
            for (int j = 0; j < 6; ++j)
            {
                r[6 * size_t(blk) + j] = b[6 * size_t(blk) + j];
                x[6 * size_t(blk) + j] = 0.f;
            }
        float y[6];
        precond6(minv, r, blk, y);
        for (int j = 0; j < 6; ++j)
            d[6 * size_t(blk) + j] = y[j];
        v = dot6(y, r, blk);
    }
    block_sum_store(v, part + blockIdx.x);
}

// delta[0] = sum of the init partials (one wave)
__global__ void k_bp_delta0(const double *__restrict__ part, uint32_t n, double *__restrict__ delta)
{
    const double s = sum_partials(part, n);
    if (threadIdx.x == 0)
        delta[0] = s;
}

// r = b - y  (InitResidual through ApplySpMV when the operator has no InitResidual)
__global__ void k_bp_sub(const float *__restrict__ b, const float *__restrict__ y, float *__restrict__ r, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBpThreads + threadIdx.x;
    if (i < n)
        r[i] = b[i] - y[i];
}

__global__ void k_bp_dot(const float *__restrict__ d, const float *__restrict__ q, uint32_t nblk,
                         double *__restrict__ part)
{
    const uint32_t blk = blockIdx.x * kBpThreads + threadIdx.x;
    double v = 0.0;
    if (blk < nblk)
    {
        float db[6];
        for (int j = 0; j < 6; ++j)
            db[j] = d[6 * size_t(blk) + j];
        v = dot6(db, q, blk);
    }
    block_sum_store(v, part + blockIdx.x);
}

// the second kernel of the reference's iteration (BlockPCG.cuh:442-472)
__global__ void k_bp_update(float *__restrict__ x, const float *__restrict__ r, const float *__restrict__ d,
                            const float *__restrict__ q, const float *__restrict__ minv, float *__restrict__ t,
                            float *__restrict__ s, uint32_t nblk, const double *__restrict__ part_dq,
                            const double *__restrict__ delta, uint32_t it, double *__restrict__ part_ts)
{
    __shared__ double alpha_sh;
    if (threadIdx.x < 64)
    {
        const double dq = sum_partials(part_dq, gridDim.x);
        if (threadIdx.x == 0)
            alpha_sh = dq != 0.0 ? delta[it] / dq : 0.0;
    }
    __syncthreads();
    const double alpha = alpha_sh;
    const uint32_t blk = blockIdx.x * kBpThreads + threadIdx.x;
    double v = 0.0;
    if (blk < nblk)
    {
        const float a = float(alpha);
        for (int j = 0; j < 6; ++j)
        {
            const size_t i = 6 * size_t(blk) + j;
            x[i] = x[i] + a * d[i];
            t[i] = r[i] - a * q[i];
        }
        float y[6];
        precond6(minv, t, blk, y);
        for (int j = 0; j < 6; ++j)
            s[6 * size_t(blk) + j] = y[j];
        v = dot6(y, t, blk);
    }
    block_sum_store(v, part_ts + blockIdx.x);
}

// the third kernel (BlockPCG.cuh:474-483): d = s + beta d; delta[it+1] = t.s
__global__ void k_bp_dir(const float *__restrict__ s, float *__restrict__ d, uint32_t nblk,
                         const double *__restrict__ part_ts, double *__restrict__ delta, uint32_t it)
{
    __shared__ double beta_sh;
    if (threadIdx.x < 64)
    {
        const double dn = sum_partials(part_ts, gridDim.x);
        if (threadIdx.x == 0)
        {
            const double dold = delta[it];
            beta_sh = dold != 0.0 ? dn / dold : 0.0;
            if (blockIdx.x == 0)
                delta[it + 1] = dn;
        }
    }
    __syncthreads();
    const float beta = float(beta_sh);
    const uint32_t blk = blockIdx.x * kBpThreads + threadIdx.x;
    if (blk < nblk)
        for (int j = 0; j < 6; ++j)
        {
            const size_t i = 6 * size_t(blk) + j;
            d[i] = s[i] + beta * d[i];
        }
}

// y = A x over a 6x6 block-CSR (row-major blocks), one thread per scalar row,
// blocks of a row in storage order (performSparseMVKernel, ApplySpMVBinBlockCSR.cuh:13-60)
__global__ void k_bcsr_spmv(const uint32_t *__restrict__ row_ptr, const uint32_t *__restrict__ col,
                            const float *__restrict__ blocks, const float *__restrict__ x, const float *__restrict__ b,
                            float *__restrict__ y, uint32_t nrows)
{
    const uint32_t i = blockIdx.x * kBpThreads + threadIdx.x;
    if (i >= 6 * nrows)
        return;
    const uint32_t row = i / 6u, a = i % 6u;
    float acc = 0.f;
    for (uint32_t k = row_ptr[row]; k < row_ptr[row + 1]; ++k)
    {
        const float *blk = blocks + 36 * size_t(k) + 6 * a;
        const float *xv = x + 6 * size_t(col[k]);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            acc += blk[c] * xv[c];
    }
    y[i] = b ? b[i] - acc : acc; // b: InitResidual's r = b - A x
}

__global__ void k_bcsr_check(const uint32_t *__restrict__ row_ptr, const uint32_t *__restrict__ col, uint32_t nrows,
                             uint32_t nnz, uint32_t *__restrict__ bad)
{
    const uint32_t row = blockIdx.x * kBpThreads + threadIdx.x;
    if (row >= nrows)
        return;
    const uint32_t b0 = row_ptr[row], b1 = row_ptr[row + 1];
    if (b0 > b1 || b1 > nnz || (row == nrows - 1 && b1 != nnz) || (row == 0 && b0 != 0))
    {
        *bad = 1u;
        return;
    }
    for (uint32_t k = b0; k < b1; ++k)
        if (col[k] >= nrows)
            *bad = 1u;
}

inline uint32_t grid_of(uint32_t n) { return (n + kBpThreads - 1) / kBpThreads; }

int status_of(hipError_t e)
{
    if (e == hipSuccess)
        return SH_OK;
    last_hip_error = e;
    return SH_E_HIP;
}
} // namespace
} // namespace sh

struct sh_spmv
{
    uint32_t rows = 0, nnz = 0;
    const uint32_t *row_ptr = nullptr, *col = nullptr;
    const float *blocks = nullptr;
};

struct sh_block_pcg
{
    uint32_t cap = 0; // max_matrix_size
    float *r = nullptr, *t = nullptr, *d = nullptr, *q = nullptr, *s = nullptr;
    double *part_a = nullptr, *part_b = nullptr, *delta = nullptr;
    uint32_t max_iters = 0;  // delta holds max_iters + 1 entries
    uint32_t last_iters = 0; // iterations the last solve ran
    // SetSolverInput
    const float *inv_diag = nullptr, *b = nullptr;
    float *x = nullptr;
    uint32_t n = 0;
    sh_spmv_ops ops{};
    bool has_input = false;
};

namespace
{
void bpcg_free(sh_block_pcg *p)
{
    for (void *a : {static_cast<void *>(p->r), static_cast<void *>(p->t), static_cast<void *>(p->d),
                    static_cast<void *>(p->q), static_cast<void *>(p->s), static_cast<void *>(p->part_a),
                    static_cast<void *>(p->part_b), static_cast<void *>(p->delta)})
        if (a)
            (void)hipFree(a);
    delete p;
}

int ensure_delta(sh_block_pcg *p, uint32_t iters)
{
    if (iters <= p->max_iters && p->delta)
        return SH_OK;
    double *nd = nullptr;
    if (hipMalloc(&nd, size_t(iters + 1) * 8) != hipSuccess)
        return SH_E_NOMEM;
    if (p->delta)
        (void)hipFree(p->delta);
    p->delta = nd;
    p->max_iters = iters;
    return SH_OK;
}
} // namespace

extern "C"
{
    int sh_spmv_create_block_csr(uint32_t num_block_rows, uint32_t num_blocks, const uint32_t *row_ptr,
                                 const uint32_t *block_col, const float *blocks, sh_spmv **out)
    {
        if (!out || num_block_rows == 0 || num_blocks == 0 || !row_ptr || !block_col || !blocks)
            return SH_E_ARG;
        if (uint64_t(num_block_rows) * 6u > 0xffffffffull)
            return SH_E_ARG;
        *out = new sh_spmv{num_block_rows, num_blocks, row_ptr, block_col, blocks};
        return SH_OK;
    }

    int sh_spmv_destroy(sh_spmv *spmv)
    {
        delete spmv;
        return SH_OK;
    }

    int sh_spmv_check(const sh_spmv *spmv, sh_stream_t stream)
    {
        if (!spmv)
            return SH_E_ARG;
        uint32_t *bad = nullptr;
        if (hipMalloc(&bad, 4) != hipSuccess)
            return SH_E_NOMEM;
        const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        uint32_t h = 0;
        int rc = sh::status_of(hipMemsetAsync(bad, 0, 4, st));
        if (rc == SH_OK)
        {
            hipLaunchKernelGGL(sh::k_bcsr_check, dim3(sh::grid_of(spmv->rows)), dim3(sh::kBpThreads), 0, st,
                               spmv->row_ptr, spmv->col, spmv->rows, spmv->nnz, bad);
            rc = sh::status_of(hipGetLastError());
        }
        if (rc == SH_OK)
            rc = sh::status_of(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, st));
        if (rc == SH_OK)
            rc = sh::status_of(hipStreamSynchronize(st));
        (void)hipFree(bad);
        if (rc != SH_OK)
            return rc;
        return h ? SH_E_ARG : SH_OK;
    }

    int sh_spmv_matrix_size(const sh_spmv *spmv, uint32_t *size)
    {
        if (!spmv || !size)
            return SH_E_ARG;
        *size = 6u * spmv->rows;
        return SH_OK;
    }

    int sh_spmv_apply(sh_spmv *spmv, const float *x, float *y, sh_stream_t stream)
    {
        if (!spmv || !x || !y)
            return SH_E_ARG;
        hipLaunchKernelGGL(sh::k_bcsr_spmv, dim3(sh::grid_of(6u * spmv->rows)), dim3(sh::kBpThreads), 0,
                           reinterpret_cast<hipStream_t>(stream), spmv->row_ptr, spmv->col, spmv->blocks, x,
                           static_cast<const float *>(nullptr), y, spmv->rows);
        return sh::status_of(hipGetLastError());
    }

    int sh_spmv_init_residual(sh_spmv *spmv, const float *x_init, const float *b, float *residual,
                              sh_stream_t stream)
    {
        if (!spmv || !x_init || !b || !residual)
            return SH_E_ARG;
        hipLaunchKernelGGL(sh::k_bcsr_spmv, dim3(sh::grid_of(6u * spmv->rows)), dim3(sh::kBpThreads), 0,
                           reinterpret_cast<hipStream_t>(stream), spmv->row_ptr, spmv->col, spmv->blocks, x_init, b,
                           residual, spmv->rows);
        return sh::status_of(hipGetLastError());
    }

    static int bcsr_apply_cb(void *user, const float *x, float *y, sh_stream_t stream)
    {
        return sh_spmv_apply(static_cast<sh_spmv *>(user), x, y, stream);
    }

    static int bcsr_init_cb(void *user, const float *x, const float *b, float *r, sh_stream_t stream)
    {
        return sh_spmv_init_residual(static_cast<sh_spmv *>(user), x, b, r, stream);
    }

    int sh_spmv_ops_of(sh_spmv *spmv, sh_spmv_ops *ops)
    {
        if (!spmv || !ops)
            return SH_E_ARG;
        ops->user = spmv;
        ops->matrix_size = 6u * spmv->rows;
        ops->apply_spmv = bcsr_apply_cb;
        ops->init_residual = bcsr_init_cb;
        return SH_OK;
    }

    int sh_block_pcg_create(uint32_t max_matrix_size, sh_block_pcg **out)
    {
        if (!out || max_matrix_size == 0 || max_matrix_size % 6u != 0)
            return SH_E_ARG;
        sh_block_pcg *p = new sh_block_pcg();
        p->cap = max_matrix_size;
        const size_t vb = size_t(max_matrix_size) * 4;
        const size_t pb = size_t(sh::grid_of(max_matrix_size / 6u)) * 8;
        bool ok = hipMalloc(&p->r, vb) == hipSuccess && hipMalloc(&p->t, vb) == hipSuccess &&
                  hipMalloc(&p->d, vb) == hipSuccess && hipMalloc(&p->q, vb) == hipSuccess &&
                  hipMalloc(&p->s, vb) == hipSuccess && hipMalloc(&p->part_a, pb) == hipSuccess &&
                  hipMalloc(&p->part_b, pb) == hipSuccess && ensure_delta(p, 16) == SH_OK;
        if (!ok)
        {
            bpcg_free(p);
            return SH_E_NOMEM;
        }
        *out = p;
        return SH_OK;
    }

    int sh_block_pcg_destroy(sh_block_pcg *pcg)
    {
        if (pcg)
            bpcg_free(pcg);
        return SH_OK;
    }

    int sh_block_pcg_set_input(sh_block_pcg *pcg, const float *inv_diag_blocks, const sh_spmv_ops *spmv,
                               const float *b, float *x, uint32_t actual_size)
    {
        if (!pcg || !inv_diag_blocks || !spmv || !spmv->apply_spmv || !b || !x)
            return SH_E_ARG;
        const uint32_t n = actual_size ? actual_size : spmv->matrix_size;
        if (n == 0 || n % 6u != 0 || spmv->matrix_size != n)
            return SH_E_ARG;
        if (n > pcg->cap)
            return SH_E_CAPACITY;
        pcg->inv_diag = inv_diag_blocks;
        pcg->ops = *spmv;
        pcg->b = b;
        pcg->x = x;
        pcg->n = n;
        pcg->has_input = true;
        return SH_OK;
    }

    // Solve / SolveNoConvergeCheck and SolveConvergeChecked (BlockPCG.cuh:325-393):
    // epsilon > 0 stops after the first iteration with
    // |delta| < epsilon^2 |delta_0| (one host readback per iteration, as there)
    static int bpcg_run(sh_block_pcg *pcg, uint32_t max_iterations, int zero_init, float epsilon,
                        sh_stream_t stream, uint32_t *iterations)
    {
        if (!pcg)
            return SH_E_ARG;
        if (!pcg->has_input)
            return SH_E_STATE;
        int rc = ensure_delta(pcg, max_iterations);
        if (rc != SH_OK)
            return rc;
        const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        const uint32_t nblk = pcg->n / 6u, g = sh::grid_of(nblk);
        pcg->last_iters = 0;
        if (!zero_init)
        {
            // r = b - A x_init (ApplySpMVBase::InitResidual), or through ApplySpMV
            if (pcg->ops.init_residual)
                rc = pcg->ops.init_residual(pcg->ops.user, pcg->x, pcg->b, pcg->r, stream);
            else if ((rc = pcg->ops.apply_spmv(pcg->ops.user, pcg->x, pcg->q, stream)) == SH_OK)
            {
                hipLaunchKernelGGL(sh::k_bp_sub, dim3(sh::grid_of(pcg->n)), dim3(sh::kBpThreads), 0, st, pcg->b,
                                   pcg->q, pcg->r, pcg->n);
                rc = sh::status_of(hipGetLastError());
            }
            if (rc != SH_OK)
                return rc;
        }
        hipLaunchKernelGGL(sh::k_bp_init, dim3(g), dim3(sh::kBpThreads), 0, st, pcg->b, pcg->inv_diag, pcg->r,
                           pcg->d, pcg->x, nblk, zero_init ? 1 : 0, pcg->part_a);
        hipLaunchKernelGGL(sh::k_bp_delta0, dim3(1), dim3(64), 0, st, pcg->part_a, g, pcg->delta);
        if ((rc = sh::status_of(hipGetLastError())) != SH_OK)
            return rc;
        double delta0 = 0.0;
        if (epsilon > 0.f &&
            (rc = sh::status_of(hipMemcpyAsync(&delta0, pcg->delta, 8, hipMemcpyDeviceToHost, st))) != SH_OK)
            return rc;
        float *r = pcg->r, *t = pcg->t;
        for (uint32_t it = 0; it < max_iterations; ++it)
        {
            if ((rc = pcg->ops.apply_spmv(pcg->ops.user, pcg->d, pcg->q, stream)) != SH_OK)
                return rc;
            hipLaunchKernelGGL(sh::k_bp_dot, dim3(g), dim3(sh::kBpThreads), 0, st, pcg->d, pcg->q, nblk,
                               pcg->part_a);
            hipLaunchKernelGGL(sh::k_bp_update, dim3(g), dim3(sh::kBpThreads), 0, st, pcg->x, r, pcg->d, pcg->q,
                               pcg->inv_diag, t, pcg->s, nblk, pcg->part_a, pcg->delta, it, pcg->part_b);
            hipLaunchKernelGGL(sh::k_bp_dir, dim3(g), dim3(sh::kBpThreads), 0, st, pcg->s, pcg->d, nblk,
                               pcg->part_b, pcg->delta, it);
            if ((rc = sh::status_of(hipGetLastError())) != SH_OK)
                return rc;
            std::swap(r, t);
            pcg->last_iters = it + 1;
            if (epsilon > 0.f)
            {
                double delta = 0.0;
                if ((rc = sh::status_of(hipMemcpyAsync(&delta, pcg->delta + it + 1, 8, hipMemcpyDeviceToHost, st))) !=
                        SH_OK ||
                    (rc = sh::status_of(hipStreamSynchronize(st))) != SH_OK)
                    return rc;
                if (std::fabs(delta) < double(epsilon) * double(epsilon) * std::fabs(delta0))
                    break;
            }
        }
        if (iterations)
            *iterations = pcg->last_iters;
        return sh::status_of(hipStreamSynchronize(st));
    }

    int sh_block_pcg_solve(sh_block_pcg *pcg, uint32_t max_iterations, int zero_init, sh_stream_t stream)
    {
        return bpcg_run(pcg, max_iterations, zero_init, 0.f, stream, nullptr);
    }

    int sh_block_pcg_solve_converge_checked(sh_block_pcg *pcg, uint32_t max_iterations, int zero_init,
                                            float epsilon, sh_stream_t stream, uint32_t *iterations)
    {
        if (!(epsilon > 0.f))
            return SH_E_ARG;
        return bpcg_run(pcg, max_iterations, zero_init, epsilon, stream, iterations);
    }

    int sh_block_pcg_delta_history(const sh_block_pcg *pcg, double *out, uint32_t capacity, uint32_t *count)
    {
        if (!pcg || !count)
            return SH_E_ARG;
        if (!pcg->has_input)
            return SH_E_STATE;
        *count = pcg->last_iters + 1;
        if (!out)
            return SH_OK;
        const uint32_t n = std::min(capacity, pcg->last_iters + 1);
        return sh::status_of(hipMemcpy(out, pcg->delta, size_t(n) * 8, hipMemcpyDeviceToHost));
    }
}
