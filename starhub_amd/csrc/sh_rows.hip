// Row-major normal equations: block-row structure from the node->term lists
// and one-pass materialisation of every 6x6 block of a node's row.
//
// The reference builds the off-diagonal pattern with a 24-bit radix sort of
// K(K-1)/2 pair keys per pixel (NodePair2TermsIndex.cu:11-193,591-691) and then
// gathers each pair's terms again per GN iteration (JtJMaterializer.cu:66-175).
// Here the block-row of node i is the set of nodes sharing a term with i --
// exactly the reference's symmetric pair set (NodePair2TermsIndex.cu:171-193)
// -- collected into an LDS bitmap from i's node->term list, and the GN-phase
// kernel walks that same list once, accumulating all of row i's blocks in LDS
// in a fixed order (no atomics: results are identical alone or in a batch).
#include "sh_common.hpp"
#include "sh_desc.hpp"
#include "sh_kernels.hpp"

namespace sh
{
__device__ __forceinline__ int find_solve_r(const uint32_t *__restrict__ begin, int n, uint32_t wg)
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

constexpr int kRowThreads = 256;

// Entry (c, q) of the x-side ARAP row [-[p]x, I]:  q < 3 -> -skew(p)[c][q],
// q >= 3 -> delta(c, q - 3).  The y side is its negative with p = T_j x_j.
__device__ __forceinline__ float reg_elem(f3 p, int c, int q)
{
    float v = (q - 3 == c) ? 1.f : 0.f;
    v = (c == 0 && q == 1) ? p.z : v;
    v = (c == 0 && q == 2) ? -p.y : v;
    v = (c == 1 && q == 0) ? -p.z : v;
    v = (c == 1 && q == 2) ? p.x : v;
    v = (c == 2 && q == 0) ? p.y : v;
    v = (c == 2 && q == 1) ? -p.x : v;
    return v;
}
constexpr int kMaxWords = kPairBase / 32; // bitmap words for N <= 4096

// One work-group per node: OR the nodes of every term of node i into an LDS
// bitmap, drop i itself, count.  Also counts i's dense terms (they precede the
// ARAP and node-translation terms in the sorted list).
template <int K>
__global__ __launch_bounds__(kRowThreads) void k_row_bitmap(SolveDesc d, uint32_t *__restrict__ bitmap,
                                                            uint32_t *__restrict__ row_len,
                                                            uint32_t *__restrict__ meta)
{
    __shared__ uint32_t bits[kMaxWords];
    __shared__ uint32_t cnt[2];
    const uint32_t i = blockIdx.x;
    const uint32_t words = (uint32_t(d.N) + 31) / 32;
    for (uint32_t w = threadIdx.x; w < words; w += blockDim.x)
        bits[w] = 0u;
    if (threadIdx.x < 2)
        cnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t beg = d.n2t_off[i], end = d.n2t_off[i + 1];
    const uint32_t P = d.P, E = d.E;
    uint32_t dense = 0;
    for (uint32_t it = beg + threadIdx.x; it < end; it += blockDim.x)
    {
        const uint32_t t = d.n2t_terms[it];
        if (t < P)
        {
            ++dense;
#pragma unroll
            for (int k = 0; k < K; ++k)
            {
                const uint32_t n = d.pknn[size_t(t) * K + k];
                atomicOr(&bits[n >> 5], 1u << (n & 31));
            }
        }
        else if (t < P + E)
        {
            const uint32_t e = t - P;
            const uint32_t a = d.edges[3 * e], b = d.edges[3 * e + 1];
            atomicOr(&bits[a >> 5], 1u << (a & 31));
            atomicOr(&bits[b >> 5], 1u << (b & 31));
        }
    }
    // integer sums: order independent
    atomicAdd(&cnt[0], dense);
    __syncthreads();
    if (threadIdx.x == 0)
        bits[i >> 5] &= ~(1u << (i & 31));
    __syncthreads();
    uint32_t pc = 0;
    for (uint32_t w = threadIdx.x; w < words; w += blockDim.x)
    {
        pc += __popc(bits[w]);
        bitmap[size_t(i) * kMaxWords + w] = bits[w];
    }
    atomicAdd(&cnt[1], pc);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        row_len[i] = cnt[1];
        d.dense_cnt[i] = cnt[0];
        atomicMax(&meta[0], cnt[1]);
    }
}

// Expand row i's bitmap into ascending block columns: diagonal block first,
// then the neighbours (the order of the reference's sorted symmetric keys).
__global__ __launch_bounds__(kRowThreads) void k_row_cols(SolveDesc d, const uint32_t *__restrict__ bitmap)
{
    __shared__ uint32_t pre[kMaxWords];
    const uint32_t i = blockIdx.x;
    const uint32_t words = (uint32_t(d.N) + 31) / 32;
    const uint32_t base = d.row_ptr[i] + i;
    // exclusive prefix of the word popcounts (words <= 128, one pass by wave 0)
    if (threadIdx.x < 64)
    {
        const uint32_t w0 = 2 * threadIdx.x, w1 = w0 + 1;
        const uint32_t b0 = (w0 < words) ? bitmap[size_t(i) * kMaxWords + w0] : 0u;
        const uint32_t b1 = (w1 < words) ? bitmap[size_t(i) * kMaxWords + w1] : 0u;
        const uint32_t c0 = __popc(b0), c1 = __popc(b1);
        uint32_t v = c0 + c1;
        // inclusive wave scan
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t u = __shfl_up(v, o, 64);
            if (int(threadIdx.x) >= o)
                v += u;
        }
        const uint32_t ex = v - c0 - c1;
        pre[w0] = ex;
        if (w1 < kMaxWords)
            pre[w1] = ex + c0;
    }
    __syncthreads();
    // rows beyond the block capacity are left unwritten; the host rejects
    // the solve (capacity) after reading U back
    if (threadIdx.x == 0 && base < d.blk_cap)
        d.blk_col[base] = i;
    for (uint32_t w = threadIdx.x; w < words; w += blockDim.x)
    {
        uint32_t b = bitmap[size_t(i) * kMaxWords + w];
        uint32_t o = base + 1 + pre[w];
        while (b)
        {
            const uint32_t bit = __ffs(b) - 1;
            if (o < d.blk_cap)
                d.blk_col[o] = w * 32 + bit;
            ++o;
            b &= b - 1;
        }
    }
}

// U = row_ptr[N] / 2 and the largest row into meta.
__global__ void k_row_meta(SolveDesc d, uint32_t *__restrict__ meta)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        meta[1] = d.row_ptr[d.N] / 2;
}

// One work-group (4 waves) per node row.  Dense terms of the row are split in
// four contiguous chunks, one per wave; per term the wave's 64 lanes cover
// (target slot k, packed H entry e) so each lane adds w_i w_k H_e into a
// private per-wave LDS block accumulator (distinct addresses within a term,
// program order across terms).  Wave 0 then adds the ARAP and node-translation
// terms.  The four wave accumulators are summed in a fixed order.
template <int K, class JT>
__global__ __launch_bounds__(kRowThreads) void k_rows(const SolveDesc *__restrict__ sds,
                                                      const uint32_t *__restrict__ wg_begin, int n_solve)
{
    extern __shared__ float smem[];
    constexpr int LPT = kWave / K;            // lanes per target slot
    constexpr int EPL = (21 + LPT - 1) / LPT; // H entries per lane
    const int si = find_solve_r(wg_begin, n_solve, blockIdx.x);
    const SolveDesc &d = sds[si];
    const uint32_t i = blockIdx.x - wg_begin[si];
    if (i >= uint32_t(d.N))
        return; // uniform over the group
    const uint32_t base = d.row_ptr[i] + i;
    const uint32_t L1 = d.row_ptr[i + 1] - d.row_ptr[i] + 1; // blocks in the row, diagonal included
    uint16_t *colmap = reinterpret_cast<uint16_t *>(smem);
    float *acc = smem + (d.N + 1) / 2;       // 4 x L1 x 21
    float *regblk = acc + 4 * L1 * 21;       // L1 x 36
    float *rhs_sh = regblk + L1 * 36;        // 4 x 8
    for (uint32_t t = threadIdx.x; t < L1; t += blockDim.x)
        colmap[d.blk_col[base + t]] = uint16_t(t);
    for (uint32_t t = threadIdx.x; t < 4 * L1 * 21 + L1 * 36; t += blockDim.x)
        acc[t] = 0.f;
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = lane / LPT, e0 = lane % LPT;
    float *my = acc + wave * L1 * 21;
    float g = 0.f; // lane q < 6 accumulates rhs component q
    const uint32_t beg = d.n2t_off[i];
    const uint32_t dend = beg + d.dense_cnt[i];
    const uint32_t end = d.n2t_off[i + 1];
    const uint32_t nd = dend - beg;
    const uint32_t chunk = (nd + 3) / 4;
    const uint32_t cb = beg + min(nd, chunk * wave), ce = beg + min(nd, chunk * (wave + 1));
    for (uint32_t it = cb; it < ce; ++it)
    {
        const uint32_t t = d.n2t_terms[it];
        uint16_t kn[K];
        float w[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
        {
            kn[kk] = d.pknn[size_t(t) * K + kk];
            w[kk] = d.pw[size_t(t) * K + kk];
        }
        float wi = 0.f;
        uint32_t mk = 0;
        float mw = 0.f;
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
        {
            if (kn[kk] == i)
                wi = w[kk]; // last matching slot (dense_image_jacobian.cuh:17-23)
            if (kk == k)
            {
                mk = kn[kk];
                mw = w[kk];
            }
        }
        bool last = true;
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
            if (kk > k && kn[kk] == mk)
                last = false;
        if (last)
        {
            const uint32_t c = colmap[mk];
            const float coef = wi * mw;
            float *dst = my + c * 21;
#pragma unroll
            for (int m = 0; m < EPL; ++m)
            {
                const int e = e0 + LPT * m;
                if (e < 21)
                    dst[e] += coef * ld_hg<JT>(d.jr, t, e);
            }
        }
        if (lane < 6)
            g += (wi * wi) * ld_hg<JT>(d.jr, t, kHGg + lane);
    }
    if (wave == 0)
    {
        // ARAP terms (node_graph_reg_jacobian.cuh:166-358), sequential in list order
        const uint32_t P = d.P, E = d.E;
        for (uint32_t it = dend; it < end; ++it)
        {
            const uint32_t t = d.n2t_terms[it];
            if (t < P + E)
            {
                const uint32_t e = t - P;
                const float4 *rec = reinterpret_cast<const float4 *>(d.reg + size_t(e) * kReg);
                const float4 u0 = rec[0], u1 = rec[1];
                if (u1.w == 0.f)
                    continue;
                const f3 a = mk3(u0.x, u0.y, u0.z), b = mk3(u0.w, u1.x, u1.y);
                const float cw = u1.z;
                const uint32_t ex = d.edges[3 * e], ey = d.edges[3 * e + 1];
                const bool is_x = (ex == i);
                const uint32_t other = is_x ? ey : ex;
                const float ws = d.lam2_reg * (cw * cw);
                // entries of the unweighted ARAP rows: x side [-[a]x, I], y side [[b]x, -I]
                // (node_graph_reg_jacobian.cuh:250-275)
                const f3 pm = is_x ? a : b;  // this node's warped point
                const f3 po = is_x ? b : a;  // the other node's
                const float sm = is_x ? 1.f : -1.f;
                if (lane < 36)
                {
                    const int aa = lane / 6, bb = lane % 6;
                    float vd = 0.f, vo = 0.f;
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch)
                    {
                        const float ga = sm * reg_elem(pm, ch, aa);
                        const float gb = sm * reg_elem(pm, ch, bb);
                        const float ob = -sm * reg_elem(po, ch, bb);
                        vd += ga * gb;
                        vo += ga * ob;
                    }
                    if (aa <= bb)
                        my[sym_index(aa, bb)] += ws * vd;
                    if (other != i)
                        regblk[colmap[other] * 36 + lane] += ws * vo;
                }
                if (lane < 6)
                {
                    const float res[3] = {(a.x - b.x), (a.y - b.y), (a.z - b.z)};
                    float v = 0.f;
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch)
                        v += res[ch] * (sm * reg_elem(pm, ch, lane));
                    g += ws * v;
                }
            }
            else
            {
                // node translation (node_motion_jacobian.cuh:7-35)
                const uint32_t m = t - P - E;
                const float4 pred = d.node_pred ? d.node_pred[m] : make_float4(0.f, 0.f, 0.f, 0.f);
                const float4 tt = d.ttrans[m];
                const float c2 = pred.w * pred.w;
                if (lane >= 3 && lane < 6)
                {
                    my[sym_index(lane, lane)] += d.lam2_nt * c2;
                    const float tv = (lane == 3) ? tt.x - pred.x : (lane == 4) ? tt.y - pred.y : tt.z - pred.z;
                    g += d.lam2_nt * tv * c2;
                }
            }
        }
    }
    if (lane < 6)
        rhs_sh[wave * 8 + lane] = g;
    __syncthreads();
    const size_t out0 = size_t(base) * 36;
    for (uint32_t t = threadIdx.x; t < L1 * 36; t += blockDim.x)
    {
        const uint32_t c = t / 36, ab = t % 36;
        const int a = int(ab / 6), b = int(ab % 6);
        const int e = sym_index(a, b);
        float v = acc[c * 21 + e];
        v += acc[L1 * 21 + c * 21 + e];
        v += acc[2 * L1 * 21 + c * 21 + e];
        v += acc[3 * L1 * 21 + c * 21 + e];
        v += regblk[c * 36 + ab];
        if (c == 0 && a == b)
            v += d.damping;
        d.blk[out0 + t] = v;
    }
    if (threadIdx.x < 6)
    {
        const float s = rhs_sh[threadIdx.x] + rhs_sh[8 + threadIdx.x] + rhs_sh[16 + threadIdx.x] +
                        rhs_sh[24 + threadIdx.x];
        d.rhs[size_t(i) * 6 + threadIdx.x] = -s;
    }
}

hipError_t launch_row_bitmap(const SolveDesc &d, uint32_t *bitmap, uint32_t *row_len, uint32_t *meta, hipStream_t st)
{
    if (d.K == 8)
        hipLaunchKernelGGL(k_row_bitmap<8>, dim3(d.N), dim3(kRowThreads), 0, st, d, bitmap, row_len, meta);
    else
        hipLaunchKernelGGL(k_row_bitmap<4>, dim3(d.N), dim3(kRowThreads), 0, st, d, bitmap, row_len, meta);
    return hipGetLastError();
}

hipError_t launch_row_cols(const SolveDesc &d, const uint32_t *bitmap, uint32_t *meta, hipStream_t st)
{
    hipLaunchKernelGGL(k_row_cols, dim3(d.N), dim3(kRowThreads), 0, st, d, bitmap);
    hipLaunchKernelGGL(k_row_meta, dim3(1), dim3(64), 0, st, d, meta);
    return hipGetLastError();
}

size_t rows_lds_bytes(uint32_t max_nodes, uint32_t max_row_len)
{
    const size_t L1 = size_t(max_row_len) + 1;
    return sizeof(float) * ((max_nodes + 1) / 2 + 4 * L1 * 21 + L1 * 36 + 32);
}

template <int K, class JT>
static hipError_t allow_lds(size_t lds)
{
    if (lds <= 65536)
        return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(k_rows<K, JT>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
}

hipError_t launch_rows(int K, bool fp16, const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg,
                       size_t lds, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipError_t e = (K == 8) ? (fp16 ? allow_lds<8, __half>(lds) : allow_lds<8, float>(lds))
                            : (fp16 ? allow_lds<4, __half>(lds) : allow_lds<4, float>(lds));
    if (e != hipSuccess)
        return e;
    if (K == 8)
    {
        if (fp16)
            hipLaunchKernelGGL((k_rows<8, __half>), dim3(n_wg), dim3(kRowThreads), lds, st, sds, wmap, n_solve);
        else
            hipLaunchKernelGGL((k_rows<8, float>), dim3(n_wg), dim3(kRowThreads), lds, st, sds, wmap, n_solve);
    }
    else
    {
        if (fp16)
            hipLaunchKernelGGL((k_rows<4, __half>), dim3(n_wg), dim3(kRowThreads), lds, st, sds, wmap, n_solve);
        else
            hipLaunchKernelGGL((k_rows<4, float>), dim3(n_wg), dim3(kRowThreads), lds, st, sds, wmap, n_solve);
    }
    return hipGetLastError();
}
} // namespace sh
