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

#include <algorithm>

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

// TPN threads per node (a wave or the whole work-group), kRowThreads / TPN
// nodes a work-group: OR the nodes of every term of node i into the node's LDS
// bitmap, drop i itself, count.  Also counts i's dense terms (they precede the
// ARAP and node-translation terms in the sorted list) and the segments of
// kSegTerms dense terms they split into.  A node's list is a few hundred
// terms: a wave per node keeps four times as many nodes in flight, which pays
// when the group has enough nodes to fill the device.
template <int K, int TPN>
__global__ __launch_bounds__(kRowThreads) void k_row_bitmap(IndexGroup g, const uint32_t *__restrict__ wmap,
                                                            IndexWs ws)
{
    static_assert(TPN % 64 == 0 && kRowThreads % TPN == 0, "whole waves per node");
    constexpr int kNodes = kRowThreads / TPN;
    __shared__ uint32_t bits_all[kNodes][kMaxWords];
    __shared__ uint32_t cnt[kNodes][2];
    const int it = find_solve_r(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc d = global_desc(g.sds[item.slot]);
    const uint32_t slot = threadIdx.x / TPN, lane = threadIdx.x % TPN;
    const uint32_t i = (blockIdx.x - wmap[it]) * kNodes + slot;
    const uint32_t N = uint32_t(d.N);
    const bool live = i < N;
    uint32_t *bits = bits_all[slot];
    const uint32_t words = (N + 31) / 32;
    for (uint32_t w = lane; w < words; w += TPN)
        bits[w] = 0u;
    if (lane < 2)
        cnt[slot][lane] = 0u;
    __syncthreads();
    const uint32_t beg = live ? d.n2t_off[i] : 0u, end = live ? d.n2t_off[i + 1] : 0u;
    const uint32_t P = d.P, E = d.E;
    uint32_t dense = 0;
    // kU consecutive terms of the list per thread and round: their term ids,
    // then their node lists, are loaded before any is used, and a dense term
    // whose node list equals the previous one's (neighbouring pixels mostly
    // share their surfels' nodes) sets no bits
    constexpr int kU = 4;
    for (uint32_t t0 = beg + lane * kU; t0 < end; t0 += kU * TPN)
    {
        uint32_t tt[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
        {
            const uint32_t t_i = t0 + u;
            tt[u] = t_i < end ? d.n2t_terms[t_i] : 0xffffffffu;
        }
        uint32_t nw[kU][K / 2]; // the node ids of a dense term, two per word
#pragma unroll
        for (int u = 0; u < kU; ++u)
        {
#pragma unroll
            for (int q = 0; q < K / 2; ++q)
                nw[u][q] = 0u;
            if (tt[u] < P)
            {
                if constexpr (K == 8)
                {
                    const uint4 v = *reinterpret_cast<const uint4 *>(d.pknn + size_t(tt[u]) * 8);
                    nw[u][0] = v.x, nw[u][1] = v.y, nw[u][2] = v.z, nw[u][3] = v.w;
                }
                else
                {
                    const uint2 v = *reinterpret_cast<const uint2 *>(d.pknn + size_t(tt[u]) * 4);
                    nw[u][0] = v.x, nw[u][1] = v.y;
                }
            }
            else if (tt[u] < P + E)
            {
                const uint32_t e = tt[u] - P;
                nw[u][0] = uint32_t(d.edges[3 * e]) | (uint32_t(d.edges[3 * e + 1]) << 16);
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
        {
            if (tt[u] < P)
            {
                ++dense;
                bool same = u > 0 && tt[u - 1] < P;
#pragma unroll
                for (int q = 0; q < K / 2; ++q)
                    same = same && nw[u][q] == nw[u > 0 ? u - 1 : 0][q];
                if (!same)
#pragma unroll
                    for (int k = 0; k < K; ++k)
                    {
                        const uint32_t n = (nw[u][k >> 1] >> (16 * (k & 1))) & 0xffffu;
                        atomicOr(&bits[n >> 5], 1u << (n & 31));
                    }
            }
            else if (tt[u] < P + E)
            {
                const uint32_t a = nw[u][0] & 0xffffu, b = nw[u][0] >> 16;
                atomicOr(&bits[a >> 5], 1u << (a & 31));
                atomicOr(&bits[b >> 5], 1u << (b & 31));
            }
        }
    }
    __syncthreads();
    uint32_t pc = 0;
    if (live)
    {
        uint32_t *__restrict__ row_bits = ws.bitmap + size_t(item.node0 + i) * kMaxWords;
        for (uint32_t w = lane; w < words; w += TPN)
        {
            uint32_t b = bits[w];
            if (w == (i >> 5))
                b &= ~(1u << (i & 31));
            pc += __popc(b);
            row_bits[w] = b;
        }
    }
    // integer sums over each wave, then over the node's waves: order independent
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
        pc += __shfl_xor(pc, o, 64);
        dense += __shfl_xor(dense, o, 64);
    }
    if ((threadIdx.x & 63u) == 0)
    {
        atomicAdd(&cnt[slot][0], dense);
        atomicAdd(&cnt[slot][1], pc);
    }
    __syncthreads();
    if (live && lane == 0)
    {
        const uint32_t n_dense = cnt[slot][0], len = cnt[slot][1];
        ws.rowlen[item.node0 + i] = len;
        d.dense_cnt[i] = n_dense;
        ws.segcnt[item.node0 + i] = (n_dense + kSegTerms - 1) / kSegTerms;
        atomicMax(&ws.meta[size_t(it) * 16], len);
    }
}

// One wave per node, kRowNodes nodes a work-group: row i's bitmap words with
// the neighbour count of the words before each (bitpre, for the block lookups
// of the tile build and the transpose map) and, with cols, its ascending block
// columns: diagonal block first, then the neighbours (the order of the
// reference's sorted symmetric keys).  A row has at most 128 words: two a lane.
__global__ __launch_bounds__(kRowThreads) void k_row_cols(IndexGroup g, const uint32_t *__restrict__ wmap, IndexWs ws,
                                                          int cols)
{
    static_assert(kMaxWords == 128 && kRowThreads == 64 * kRowNodes, "two bitmap words a lane, a wave per node");
    const int it = find_solve_r(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc d = global_desc(g.sds[item.slot]);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t i = (blockIdx.x - wmap[it]) * kRowNodes + (threadIdx.x >> 6);
    if (i >= uint32_t(d.N))
        return; // the whole wave
    const uint32_t words = (uint32_t(d.N) + 31) / 32;
    const uint32_t *__restrict__ row_bits = ws.bitmap + size_t(item.node0 + i) * kMaxWords;
    uint2 *__restrict__ row_pre = ws.bitpre + size_t(item.node0 + i) * kMaxWords;
    const uint32_t w0 = 2 * lane, w1 = w0 + 1;
    const uint32_t b0 = (w0 < words) ? row_bits[w0] : 0u;
    const uint32_t b1 = (w1 < words) ? row_bits[w1] : 0u;
    const uint32_t c0 = __popc(b0), c1 = __popc(b1);
    // exclusive prefix of the word popcounts: an inclusive wave scan of the
    // lane pairs' counts
    uint32_t v = c0 + c1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const uint32_t u = __shfl_up(v, o, 64);
        if (int(lane) >= o)
            v += u;
    }
    const uint32_t ex = v - c0 - c1;
    if (w0 < words)
        row_pre[w0] = make_uint2(b0, ex);
    if (w1 < words)
        row_pre[w1] = make_uint2(b1, ex + c0);
    if (!cols)
        return; // uniform
    // rows beyond the block capacity are left unwritten; the host rejects
    // the solve (capacity) after reading U back
    const uint32_t base = d.row_ptr[i] + i;
    if (lane == 0 && base < d.blk_cap)
        d.blk_col[base] = i;
    auto emit = [&](uint32_t b, uint32_t w, uint32_t o) {
        while (b)
        {
            const uint32_t bit = __ffs(b) - 1;
            if (o < d.blk_cap)
                d.blk_col[o] = w * 32 + bit;
            ++o;
            b &= b - 1;
        }
    };
    emit(b0, w0, base + 1 + ex);
    emit(b1, w1, base + 1 + ex + c0);
}

// (a, b) of packed upper-triangle entry e of a symmetric 6x6 block.
__device__ __forceinline__ void sym_pair(int e, int &a, int &b)
{
    a = (e >= 6) + (e >= 11) + (e >= 15) + (e >= 18) + (e >= 20);
    b = e - (a * 6 - a * (a - 1) / 2) + a;
}

constexpr int kRowTasks = 4;  // (column, 4-entry group) tasks per thread and pass
constexpr int kHs = 24;       // staged H stride (21 entries padded to 6 float4)

// Row i of segment `seg`: the last row whose first segment is <= seg.
__device__ __forceinline__ uint32_t seg_row(const uint32_t *__restrict__ seg_ptr, uint32_t n, uint32_t seg)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg_ptr[mid] <= seg)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Dense part of the normal equations, one work-group per segment of up to
// kSegTerms dense terms of one node row i (the terms of a node are very
// uneven: a node facing the camera carries thousands, one behind it none).
// The segment is processed in chunks: thread j stages term j -- its H record
// (21 entries of sum_c lam_c^2 J_c^T J_c) into LDS and, for every slot k, the
// coefficient w_i w_k into a dense chunk x L1 LDS matrix at the column of
// node k (later slots overwrite earlier ones: the reference's "last matching
// slot", dense_image_jacobian.cuh:17-23,92-104).  Every thread then owns
// (column, group of 4 entries) accumulators in registers and runs through the
// chunk in term order: partial block(i, c) = sum_t C[t][c] H_t, a small dense
// product, no atomics.  The partial rhs sum_t w_i^2 g_t is summed by the
// staging threads and reduced in a fixed tree.
template <int K, class JT>
__global__ __launch_bounds__(kRowThreads) void k_rows_seg(const SolveDesc *__restrict__ sds,
                                                          const uint32_t *__restrict__ wg_begin, int n_solve,
                                                          uint32_t chunk)
{
    extern __shared__ float smem[];
    const int si = find_solve_r(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t seg = blockIdx.x - wg_begin[si];
    if (seg >= uint32_t(d.n_seg))
        return; // uniform over the group
    const uint32_t tid = threadIdx.x;
    const uint32_t i = seg_row(d.seg_ptr, uint32_t(d.N), seg);
    const uint32_t beg = d.n2t_off[i] + (seg - d.seg_ptr[i]) * kSegTerms;
    const uint32_t dend = min(beg + kSegTerms, d.n2t_off[i] + d.dense_cnt[i]);
    const uint32_t base = d.row_ptr[i] + i;
    const uint32_t L1 = d.row_ptr[i + 1] - d.row_ptr[i] + 1; // blocks in the row, diagonal first
    // LDS: Hs[chunk][24] | Cm[chunk][L1] | colmap u16[N] | red[4][8]
    float *Hs = smem;
    float *Cm = Hs + size_t(chunk) * kHs;
    uint16_t *colmap = reinterpret_cast<uint16_t *>(Cm + size_t(chunk) * L1);
    float *red = reinterpret_cast<float *>(colmap + ((d.N + 1) & ~1));
    for (uint32_t t = tid; t < L1; t += kRowThreads)
        colmap[d.blk_col[base + t]] = uint16_t(t);
    for (uint32_t t = tid; t < size_t(chunk) * L1; t += kRowThreads)
        Cm[t] = 0.f;
    __syncthreads();

    float *out = d.part + size_t(seg) * d.part_stride;
    const uint32_t n_tasks = L1 * 6;
    float gsum[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint32_t tp = 0; tp < n_tasks; tp += kRowThreads * kRowTasks)
    {
        float acc[kRowTasks][4];
#pragma unroll
        for (int s = 0; s < kRowTasks; ++s)
            acc[s][0] = acc[s][1] = acc[s][2] = acc[s][3] = 0.f;
        for (uint32_t c0 = beg; c0 < dend; c0 += chunk)
        {
            const uint32_t n = min(chunk, dend - c0);
            // ---- stage: thread j < n takes term c0 + j
            uint16_t cols[K];
            const bool stager = tid < n;
            if (stager)
            {
                const uint32_t t = d.n2t_terms[c0 + tid];
                uint16_t kn[K];
                float w[K];
#pragma unroll
                for (int k = 0; k < K; ++k)
                {
                    kn[k] = d.gpknn[size_t(t) * K + k];
                    w[k] = d.gpw[size_t(t) * K + k];
                }
                float hg[kHG];
                load_hg<JT>(d.jr, t, hg);
                float wi = 0.f;
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (kn[k] == i)
                        wi = w[k];
#pragma unroll
                for (int k = 0; k < K; ++k)
                {
                    cols[k] = colmap[kn[k]];
                    Cm[tid * L1 + cols[k]] = wi * w[k];
                }
                float4 *h4 = reinterpret_cast<float4 *>(Hs + tid * kHs);
                h4[0] = make_float4(hg[0], hg[1], hg[2], hg[3]);
                h4[1] = make_float4(hg[4], hg[5], hg[6], hg[7]);
                h4[2] = make_float4(hg[8], hg[9], hg[10], hg[11]);
                h4[3] = make_float4(hg[12], hg[13], hg[14], hg[15]);
                h4[4] = make_float4(hg[16], hg[17], hg[18], hg[19]);
                h4[5] = make_float4(hg[20], 0.f, 0.f, 0.f);
                if (tp == 0)
                {
                    const float wi2 = wi * wi;
#pragma unroll
                    for (int q = 0; q < 6; ++q)
                        gsum[q] += wi2 * hg[kHGg + q];
                }
            }
            __syncthreads();
            // ---- accumulate: thread owns tasks (column c, entries 4g..4g+3)
#pragma unroll
            for (int s = 0; s < kRowTasks; ++s)
            {
                const uint32_t task = tp + tid + s * kRowThreads;
                if (task < n_tasks)
                {
                    const uint32_t c = task / 6, g = task % 6;
                    const float *cp = Cm + c;
                    const float4 *hp = reinterpret_cast<const float4 *>(Hs) + g;
                    float a0 = acc[s][0], a1 = acc[s][1], a2 = acc[s][2], a3 = acc[s][3];
                    for (uint32_t j = 0; j < n; ++j)
                    {
                        const float cv = cp[j * L1];
                        const float4 h = hp[j * (kHs / 4)];
                        const bool on = cv != 0.f;
                        a0 = on ? fmaf(cv, h.x, a0) : a0;
                        a1 = on ? fmaf(cv, h.y, a1) : a1;
                        a2 = on ? fmaf(cv, h.z, a2) : a2;
                        a3 = on ? fmaf(cv, h.w, a3) : a3;
                    }
                    acc[s][0] = a0;
                    acc[s][1] = a1;
                    acc[s][2] = a2;
                    acc[s][3] = a3;
                }
            }
            __syncthreads();
            // ---- clear this term's coefficients for the next chunk
            if (stager)
            {
#pragma unroll
                for (int k = 0; k < K; ++k)
                    Cm[tid * L1 + cols[k]] = 0.f;
            }
        }
        // ---- the pass's partial blocks, packed upper triangles
#pragma unroll
        for (int s = 0; s < kRowTasks; ++s)
        {
            const uint32_t task = tp + tid + s * kRowThreads;
            if (task < n_tasks)
            {
                const uint32_t c = task / 6, g = task % 6;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (int(g) * 4 + q < 21)
                        out[8 + c * 21 + g * 4 + q] = acc[s][q];
            }
        }
    }
    // ---- partial rhs: fixed-order reduction of the staging threads' sums
#pragma unroll
    for (int q = 0; q < 6; ++q)
    {
        float v = gsum[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        gsum[q] = v;
    }
    if ((tid & 63) == 0)
    {
#pragma unroll
        for (int q = 0; q < 6; ++q)
            red[(tid >> 6) * 8 + q] = gsum[q];
    }
    __syncthreads();
    if (tid < 6)
        out[tid] = (red[tid] + red[8 + tid]) + (red[16 + tid] + red[24 + tid]);
}

// One work-group per node row: the row's segment partials summed in segment
// order, plus the ARAP terms (node_graph_reg_jacobian.cuh:166-358) and the
// node-translation term (node_motion_jacobian.cuh:7-35) of the row, added by
// wave 0 in list order, plus the damping on the diagonal block.  Writes the
// full 6x6 blocks of the row and rhs_i = -(sum of the partial g).
__global__ __launch_bounds__(kRowThreads) void k_rows_combine(const SolveDesc *__restrict__ sds,
                                                              const uint32_t *__restrict__ wg_begin, int n_solve)
{
    extern __shared__ float smem[];
    const int si = find_solve_r(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t i = blockIdx.x - wg_begin[si];
    if (i >= uint32_t(d.N))
        return; // uniform over the group
    const uint32_t tid = threadIdx.x;
    const uint32_t base = d.row_ptr[i] + i;
    const uint32_t L1 = d.row_ptr[i + 1] - d.row_ptr[i] + 1;
    // LDS: regblk[L1][36] | colmap u16[N] | red[8]
    float *regblk = smem;
    uint16_t *colmap = reinterpret_cast<uint16_t *>(regblk + size_t(L1) * 36);
    float *red = reinterpret_cast<float *>(colmap + ((d.N + 1) & ~1));
    for (uint32_t t = tid; t < L1; t += kRowThreads)
        colmap[d.blk_col[base + t]] = uint16_t(t);
    for (uint32_t t = tid; t < L1 * 36; t += kRowThreads)
        regblk[t] = 0.f;
    __syncthreads();

    const uint32_t dend = d.n2t_off[i] + d.dense_cnt[i];
    const uint32_t end = d.n2t_off[i + 1];
    if (tid < 64)
    {
        const int lane = int(tid);
        float gA = 0.f;
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
                // unweighted ARAP rows: x side [-[a]x, I], y side [[b]x, -I]
                // (node_graph_reg_jacobian.cuh:250-275)
                const f3 pm = is_x ? a : b; // this node's warped point
                const f3 po = is_x ? b : a; // the other node's
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
                    regblk[lane] += ws * vd;
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
                    gA += ws * v;
                }
            }
            else
            {
                const uint32_t m = t - P - E;
                const float4 pred = d.node_pred ? d.node_pred[m] : make_float4(0.f, 0.f, 0.f, 0.f);
                const float4 tt = d.ttrans[m];
                const float c2 = pred.w * pred.w;
                if (lane >= 3 && lane < 6)
                {
                    regblk[lane * 7] += d.lam2_nt * c2;
                    const float tv = (lane == 3) ? tt.x - pred.x : (lane == 4) ? tt.y - pred.y : tt.z - pred.z;
                    gA += d.lam2_nt * tv * c2;
                }
            }
        }
        if (lane < 6)
            red[lane] = gA;
    }
    __syncthreads();
    const uint32_t s0 = d.seg_ptr[i], s1 = d.seg_ptr[i + 1];
    const float *part = d.part;
    const uint32_t stride = d.part_stride;
    const size_t out0 = size_t(base) * 36;
    for (uint32_t t = tid; t < L1 * 21; t += kRowThreads)
    {
        const uint32_t c = t / 21;
        const int e = int(t % 21);
        float v = 0.f;
        for (uint32_t sg = s0; sg < s1; ++sg)
            v += part[size_t(sg) * stride + 8 + t];
        int a, b;
        sym_pair(e, a, b);
        float u = v + regblk[c * 36 + a * 6 + b];
        if (c == 0 && a == b)
            u += d.damping;
        d.blk[out0 + size_t(c) * 36 + a * 6 + b] = u;
        if (a != b)
            d.blk[out0 + size_t(c) * 36 + b * 6 + a] = v + regblk[c * 36 + b * 6 + a];
    }
    if (tid < 6)
    {
        float g = 0.f;
        for (uint32_t sg = s0; sg < s1; ++sg)
            g += part[size_t(sg) * stride + tid];
        d.rhs[size_t(i) * 6 + tid] = -(g + red[tid]);
    }
}

hipError_t launch_row_bitmap(int K, uint32_t nodes_per_wg, const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg,
                             const IndexWs &ws, hipStream_t st)
{
    static_assert(kRowThreads == 256 && kRowNodes == 4, "launch shapes");
    if (n_wg == 0)
        return hipSuccess;
    if (nodes_per_wg != 1 && nodes_per_wg != kRowNodes)
        return hipErrorInvalidValue;
    const bool wave = nodes_per_wg == kRowNodes;
    if (K == 8 && wave)
        hipLaunchKernelGGL((k_row_bitmap<8, 64>), dim3(n_wg), dim3(kRowThreads), 0, st, g, wmap, ws);
    else if (K == 8)
        hipLaunchKernelGGL((k_row_bitmap<8, 256>), dim3(n_wg), dim3(kRowThreads), 0, st, g, wmap, ws);
    else if (wave)
        hipLaunchKernelGGL((k_row_bitmap<4, 64>), dim3(n_wg), dim3(kRowThreads), 0, st, g, wmap, ws);
    else
        hipLaunchKernelGGL((k_row_bitmap<4, 256>), dim3(n_wg), dim3(kRowThreads), 0, st, g, wmap, ws);
    return hipGetLastError();
}

hipError_t launch_row_cols(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, const IndexWs &ws, bool cols,
                           hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_row_cols, dim3(n_wg), dim3(kRowThreads), 0, st, g, wmap, ws, cols ? 1 : 0);
    return hipGetLastError();
}

RowsPlan rows_plan(uint32_t max_nodes, uint32_t max_row_len)
{
    // Segment kernel: the chunk of terms staged per round is as large as a
    // 40 KB budget allows (several work-groups share a CU), 32..kSegTerms.
    RowsPlan p;
    const size_t L1 = size_t(max_row_len) + 1;
    const size_t colmap = sizeof(uint16_t) * ((max_nodes + 1) & ~1u);
    const size_t fixed = colmap + sizeof(float) * 32;
    const size_t per_term = sizeof(float) * (kHs + L1);
    const size_t budget = 40 * 1024;
    size_t chunk = budget > fixed ? (budget - fixed) / per_term : 0;
    chunk = std::min<size_t>(kSegTerms, std::max<size_t>(32, chunk & ~size_t(31)));
    p.chunk = uint32_t(chunk);
    p.lds = fixed + chunk * per_term;
    p.lds_combine = sizeof(float) * (L1 * 36 + 8) + colmap;
    return p;
}

template <int K, class JT>
static hipError_t allow_lds(size_t lds)
{
    if (lds <= 65536)
        return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(k_rows_seg<K, JT>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
}

hipError_t launch_rows_seg(int K, bool fp16, const SolveDesc *sds, const uint32_t *seg_map, int n_solve,
                           uint32_t n_seg_wg, const RowsPlan &plan, hipStream_t st)
{
    const size_t lds = plan.lds;
    const uint32_t chunk = plan.chunk;
    if (n_seg_wg == 0)
        return hipSuccess;
    hipError_t e = (K == 8) ? (fp16 ? allow_lds<8, __half>(lds) : allow_lds<8, float>(lds))
                            : (fp16 ? allow_lds<4, __half>(lds) : allow_lds<4, float>(lds));
    if (e != hipSuccess)
        return e;
    if (K == 8)
    {
        if (fp16)
            hipLaunchKernelGGL((k_rows_seg<8, __half>), dim3(n_seg_wg), dim3(kRowThreads), lds, st, sds, seg_map,
                               n_solve, chunk);
        else
            hipLaunchKernelGGL((k_rows_seg<8, float>), dim3(n_seg_wg), dim3(kRowThreads), lds, st, sds, seg_map,
                               n_solve, chunk);
    }
    else
    {
        if (fp16)
            hipLaunchKernelGGL((k_rows_seg<4, __half>), dim3(n_seg_wg), dim3(kRowThreads), lds, st, sds, seg_map,
                               n_solve, chunk);
        else
            hipLaunchKernelGGL((k_rows_seg<4, float>), dim3(n_seg_wg), dim3(kRowThreads), lds, st, sds, seg_map,
                               n_solve, chunk);
    }
    return hipGetLastError();
}

hipError_t launch_rows_combine(const SolveDesc *sds, const uint32_t *row_map, int n_solve, uint32_t n_row_wg,
                               const RowsPlan &plan, hipStream_t st)
{
    if (n_row_wg == 0)
        return hipSuccess;
    if (plan.lds_combine > 65536)
    {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_rows_combine),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(plan.lds_combine));
        if (e != hipSuccess)
            return e;
    }
    hipLaunchKernelGGL(k_rows_combine, dim3(n_row_wg), dim3(kRowThreads), plan.lds_combine, st, sds, row_map,
                       n_solve);
    return hipGetLastError();
}
} // namespace sh
