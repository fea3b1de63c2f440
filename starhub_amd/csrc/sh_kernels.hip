// HIP kernels of the warp-field Gauss-Newton solve, hand-written for gfx950.
//
// Launch model: every GN-phase kernel takes the device array of SolveDesc and a
// work map `wg_begin` (n_solve + 1 prefix offsets of work-groups per solve), so
// one launch covers a whole batch of uneven solves.  Work-groups never straddle
// two solves.  No float atomics anywhere: all reductions are fixed-order wave
// (DPP / permute) and work-group trees, so results are deterministic and a
// solve gives identical bits alone or inside a batch.
#include "sh_common.hpp"
#include "sh_desc.hpp"
#include "sh_kernels.hpp"
#include "sh_terms.hpp"
#include "sh_update.hpp"

namespace sh
{
__device__ __forceinline__ int find_solve(const uint32_t *__restrict__ begin, int n, uint32_t wg)
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

// __float2uint_rn with saturation, NaN -> 0 (ImageTermKNNFetcher.cu:164-165)
__device__ __forceinline__ uint32_t f2u_rn_sat(float v)
{
    if (!(v > 0.f))
        return 0u;
    if (v >= 4294967296.0f)
        return 0xffffffffu;
    return uint32_t(rintf(v));
}

// ============================================================ index build
// Tile order of a pixel: 32x32-pixel blocks in raster order (nbx blocks per
// block row), Morton order inside a block.  Indexes in [0, nbx * nby * 1024).
__device__ __forceinline__ uint32_t spread5(uint32_t v)
{
    v &= 31u;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ uint32_t squeeze5(uint32_t v)
{
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v & 31u;
}

__device__ __forceinline__ uint32_t tile_order(uint32_t x, uint32_t y, uint32_t nbx)
{
    return (((y >> 5) * nbx + (x >> 5)) << 10) | spread5(x) | (spread5(y) << 1);
}

// Valid-pixel flags, and the range checks every later index kernel relies on:
// a valid pixel names a surfel below S whose K node ids are below N, and every
// ARAP edge names nodes below N.  A pixel that fails is flagged 0 and sets the
// item's error word; the host rejects the frame before anything reads the
// ids.  Tiled, the threads walk the tile-order index space (nbx blocks per
// block row) and also write each flag at its tile-order index (0 for the
// padding outside the image), for the tile-order scan.
__global__ __launch_bounds__(256) void k_pixel_flags(IndexGroup g, const uint32_t *__restrict__ wmap, bool tiled,
                                                     IndexWs ws)
{
    const int it = find_solve(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc &d = g.sds[item.slot];
    const uint32_t W = uint32_t(d.W), H = uint32_t(d.H), N = uint32_t(d.N), E = uint32_t(d.E);
    const uint32_t n = W * H, nbx = (W + 31u) / 32u;
    const uint32_t nh = tiled ? nbx * ((H + 31u) / 32u) * 1024u : 0u;
    uint32_t *__restrict__ flags = ws.flags + item.m0;
    uint32_t *__restrict__ tflags = ws.tflags + item.h0;
    // kPixPer indices a thread, a work-group's indices contiguous
    const uint32_t hb = (blockIdx.x - wmap[it]) * blockDim.x * kPixPer + threadIdx.x;
    bool bad = false;
#pragma unroll
    for (uint32_t u = 0; u < kPixPer; ++u)
    {
        const uint32_t h = hb + u * blockDim.x;
        uint32_t i = h; // raster index of the pixel this thread decides
        bool inside = h < n;
        if (tiled)
        {
            inside = false;
            if (h < nh)
            {
                const uint32_t b = h >> 10, m = h & 1023u;
                const uint32_t x = (b % nbx) * 32u + squeeze5(m), y = (b / nbx) * 32u + squeeze5(m >> 1);
                inside = x < W && y < H;
                i = y * W + x;
                if (!inside)
                    tflags[h] = 0u;
            }
        }
        if (inside)
        {
            const uint32_t s = d.index_map[i];
            bool valid = s != kInvalidIndex;
            if (valid && s >= d.S)
                valid = false, bad = true;
            if (valid)
            {
                // the largest of the surfel's K node ids (one load when aligned)
                const uint16_t *kn = d.surfel_knn + size_t(s) * d.K;
                uint32_t mx = 0;
                if (d.K == 8 && (reinterpret_cast<uintptr_t>(d.surfel_knn) & 15u) == 0)
                {
                    const uint4 v = *reinterpret_cast<const uint4 *>(kn);
                    mx = max(max(max(v.x & 0xffffu, v.x >> 16), max(v.y & 0xffffu, v.y >> 16)),
                             max(max(v.z & 0xffffu, v.z >> 16), max(v.w & 0xffffu, v.w >> 16)));
                }
                else
                    for (int k = 0; k < d.K; ++k)
                        mx = max(mx, uint32_t(kn[k]));
                if (mx >= N)
                    valid = false, bad = true;
            }
            flags[i] = valid ? 1u : 0u;
            if (tiled)
                tflags[h] = valid ? 1u : 0u;
        }
        if (h < E && (d.edges[3 * size_t(h)] >= N || d.edges[3 * size_t(h) + 1] >= N))
            bad = true;
    }
    if (bad)
        atomicOr(ws.errs + it, 1u);
}

// Stream-compact the valid pixels in row-major order; gathers surfel KNN and
// folds spatial*connect into one weight per slot (queryPixelKNNKernel,
// WarpSolver.cu:7-48, + compactPontentialImageTermPixelsKernel,
// ImageTermKNNFetcher.cu:56-88).  The positions are the group-wide scans less
// the item's first entry.
template <int K>
__global__ __launch_bounds__(256) void k_compact_pixels(IndexGroup g, const uint32_t *__restrict__ wmap, bool tiled,
                                                        IndexWs ws, uint32_t *__restrict__ kv_keys,
                                                        uint32_t *__restrict__ kv_vals)
{
    const int it = find_solve(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc d = global_desc(g.sds[item.slot]);
    const uint32_t blk = blockIdx.x - wmap[it];
    const uint32_t W = uint32_t(d.W), H = uint32_t(d.H), n = W * H;
    // tiled, the threads walk the tile-order index space: a work-group's 256
    // indices are an 8-row band of one 32x32 block, so its tile-order records
    // are one contiguous run and its raster records eight runs of a row each
    const uint32_t nbx = (W + 31u) / 32u;
    const uint32_t n_walk = tiled ? nbx * ((H + 31u) / 32u) * 1024u : n;
    const uint32_t n_pix_wg = max(1u, (n_walk + 255u) / 256u);
    const uint32_t P = d.P, E = d.E, N = d.N;
    if (blk >= n_pix_wg)
    {
        // the ARAP and node-translation terms' (node key, term) pairs: 2 per
        // edge (a self edge's keys sort after the nodes), 1 per node
        // (Node2TermsIndex.cu:6-76)
        const uint32_t t = P + (blk - n_pix_wg) * blockDim.x + threadIdx.x;
        if (t < P + E)
        {
            const uint32_t e = t - P;
            const uint32_t a = d.edges[3 * e], b = d.edges[3 * e + 1];
            const size_t o = item.kv0 + size_t(P) * K + 2 * size_t(e);
            kv_keys[o] = item.node0 + ((a != b) ? a : N);
            kv_keys[o + 1] = item.node0 + ((a != b) ? b : N);
            kv_vals[o] = t;
            kv_vals[o + 1] = t;
        }
        else if (t < P + E + N)
        {
            const uint32_t m = t - P - E;
            const size_t o = item.kv0 + size_t(P) * K + 2 * size_t(E) + m;
            kv_keys[o] = item.node0 + m;
            kv_vals[o] = t;
        }
        return;
    }
    if (blk == 0)
    {
        // the words the later stages accumulate into
        const uint32_t t = threadIdx.x, nt = uint32_t(d.n_tiles);
        if (t < 16)
            ws.meta[size_t(it) * 16 + t] = 0u;
        else if (t == 17)
            ws.rowlen[item.node0 + N] = 0u;
        else if (t == 18)
            ws.segcnt[item.node0 + N] = 0u;
        else if (t == 19 || t == 20)
            ws.tnslot[item.tc0 + nt + (t - 19)] = 0u;
        else if (t == 21)
            ws.tncontrib[item.tc0 + nt] = 0u;
    }
    const uint32_t h = blk * blockDim.x + threadIdx.x; // walk index
    uint32_t i = h, x, y;                              // raster index and pixel
    if (tiled)
    {
        if (h >= n_walk)
            return;
        const uint32_t b = h >> 10, m = h & 1023u;
        x = (b % nbx) * 32u + squeeze5(m);
        y = (b / nbx) * 32u + squeeze5(m >> 1);
        if (x >= W || y >= H)
            return;
        i = y * W + x;
    }
    else
    {
        if (h >= n)
            return;
        x = i % W;
        y = i / W;
    }
    if (ws.flags[item.m0 + i] == 0u)
        return;
    const uint32_t o = ws.pos[item.m0 + i] - ws.pos[item.m0];
    const uint32_t s = d.index_map[i];
    float fx = 0.f, fy = 0.f;
    if (d.flow)
    {
        const float2 f = d.flow[i];
        fx = f.x;
        fy = f.y;
    }
    const uint32_t tx = f2u_rn_sat(fx + float(x));
    const uint32_t ty = f2u_rn_sat(fy + float(y));
    const ushort4 px = make_ushort4(uint16_t(x), uint16_t(y), uint16_t(tx), uint16_t(ty));
    // the surfel's K node ids and folded weights: vector loads when the
    // caller's arrays are aligned for them, vector stores into ours
    uint32_t kw[K / 2]; // node ids, two per word
    float w[K];
    const bool aligned = ((reinterpret_cast<uintptr_t>(d.surfel_knn) | reinterpret_cast<uintptr_t>(d.surfel_spatial) |
                           reinterpret_cast<uintptr_t>(d.surfel_connect)) &
                          15u) == 0;
    if (aligned)
    {
        const uint32_t *kk = reinterpret_cast<const uint32_t *>(d.surfel_knn + size_t(s) * K);
        const float4 *sp = reinterpret_cast<const float4 *>(d.surfel_spatial + size_t(s) * K);
        const float4 *cn = reinterpret_cast<const float4 *>(d.surfel_connect + size_t(s) * K);
        if constexpr (K == 8)
        {
            const uint4 v = *reinterpret_cast<const uint4 *>(kk);
            kw[0] = v.x, kw[1] = v.y, kw[2] = v.z, kw[3] = v.w;
        }
        else
        {
            const uint2 v = *reinterpret_cast<const uint2 *>(kk);
            kw[0] = v.x, kw[1] = v.y;
        }
#pragma unroll
        for (int q = 0; q < K / 4; ++q)
        {
            const float4 a = sp[q], c = cn[q];
            w[4 * q] = a.x * c.x;
            w[4 * q + 1] = a.y * c.y;
            w[4 * q + 2] = a.z * c.z;
            w[4 * q + 3] = a.w * c.w;
        }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < K; k += 2)
            kw[k / 2] = uint32_t(d.surfel_knn[size_t(s) * K + k]) |
                        (uint32_t(d.surfel_knn[size_t(s) * K + k + 1]) << 16);
#pragma unroll
        for (int k = 0; k < K; ++k)
            w[k] = d.surfel_spatial[size_t(s) * K + k] * d.surfel_connect[size_t(s) * K + k];
    }
    uint16_t kn[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
        kn[k] = uint16_t((kw[k >> 1] >> (16 * (k & 1))) & 0xffffu);
    // stores of K ids (2K bytes) and K weights (4K bytes) at o: aligned
    auto put = [&](uint16_t *pk, float *pw, uint32_t at) {
        if constexpr (K == 8)
            *reinterpret_cast<uint4 *>(pk + size_t(at) * 8) = make_uint4(kw[0], kw[1], kw[2], kw[3]);
        else
            *reinterpret_cast<uint2 *>(pk + size_t(at) * 4) = make_uint2(kw[0], kw[1]);
#pragma unroll
        for (int q = 0; q < K / 4; ++q)
            *reinterpret_cast<float4 *>(pw + size_t(at) * K + 4 * q) =
                make_float4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    };
    d.pix[o] = px;
    put(d.pknn, d.pw, o);
    // the pixel term's (node key, term) pairs, K of them (Node2TermsIndex.cu:6-76)
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        kv_keys[item.kv0 + size_t(o) * K + k] = item.node0 + kn[k];
        kv_vals[item.kv0 + size_t(o) * K + k] = o;
    }
    if (tiled)
    {
        // the GN phase's copy, at the pixel's tile-order position
        const uint32_t ot = ws.tpos[item.h0 + h] - ws.tpos[item.h0]; // h = tile_order(x, y, nbx)
        d.gpix[ot] = px;
        put(d.gpknn, d.gpw, ot);
    }
}

// After the stable node sort of the group's pairs: each solve's sorted term
// ids, and offset[v] = the first sorted position whose key is >= v
// (computeTermOffsetKernel, Node2TermsIndex.cu:78-122), v <= N: a thread per
// sorted position j writes offset[v] = j for the keys v after the one below
// it up to its own (the first position of the solve starts at node 0, the
// last also writes the ones after its key, up to N, as n_kv).
__global__ __launch_bounds__(256) void k_n2t_split(IndexGroup g, const uint32_t *__restrict__ wmap,
                                                   const uint32_t *__restrict__ skeys,
                                                   const uint32_t *__restrict__ svals)
{
    const int it = find_solve(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc d = global_desc(g.sds[item.slot]);
    const uint32_t N = uint32_t(d.N);
    const uint32_t n_kv = uint32_t(d.P) * uint32_t(d.K) + 2u * uint32_t(d.E) + N;
    // kSplitPer positions a thread, a work-group's positions contiguous (a
    // work-group's fixed cost, finding its solve and descriptor, is spread
    // over more work)
    const uint32_t j0 = (blockIdx.x - wmap[it]) * blockDim.x * kSplitPer + threadIdx.x;
#pragma unroll
    for (uint32_t u = 0; u < kSplitPer; ++u)
    {
        const uint32_t j = j0 + u * blockDim.x;
        if (j >= n_kv)
            break;
        d.n2t_terms[j] = svals[item.kv0 + j];
        const uint32_t key = skeys[item.kv0 + j] - item.node0; // <= N
        const uint32_t first = j == 0 ? 0u : skeys[item.kv0 + j - 1] - item.node0 + 1u;
        for (uint32_t v = first; v <= key; ++v)
            d.n2t_off[v] = j;
        if (j + 1 == n_kv)
            for (uint32_t v = key + 1; v <= N; ++v)
                d.n2t_off[v] = n_kv;
    }
}

// The counts the host reads between the stages (see launch_index_counts).
__global__ void k_index_counts(IndexGroup g, int stage, IndexWs ws, uint32_t *host)
{
    const int it = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (it >= g.n)
        return;
    const IndexItem item = g.items[it];
    const SolveDesc &d = g.sds[item.slot];
    uint32_t *o = host + size_t(it) * 8;
    if (stage == 1)
    {
        const uint32_t n = uint32_t(d.W) * uint32_t(d.H);
        o[0] = n ? ws.pos[item.m0 + n - 1] + ws.flags[item.m0 + n - 1] - ws.pos[item.m0] : 0u;
        o[1] = ws.errs[it];
        ws.errs[it] = 0u; // for the next build
    }
    else
    {
        const uint32_t N = uint32_t(d.N), nt = uint32_t(d.n_tiles);
        o[0] = d.seg_ptr[N];
        o[1] = ws.meta[size_t(it) * 16];
        o[2] = d.row_ptr[N] / 2u;
        o[3] = d.tile_sbase[nt];
        o[4] = d.tile_cbase[nt];
        o[5] = ws.tnslot[item.tc0 + nt + 1];
    }
}

// (i<j) pair keys i*4096+j per term, padded with the uninitialised key:
// K(K-1)/2 per pixel, 1 per edge.  NodePair2TermsIndex.cu:11-92
template <int K>
__global__ void k_pair_kv(SolveDesc d, uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    constexpr int NP = K * (K - 1) / 2;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t P = d.P, E = d.E;
    if (t < P)
    {
        uint16_t kn[K];
#pragma unroll
        for (int k = 0; k < K; ++k)
            kn[k] = d.pknn[size_t(t) * K + k];
        const size_t base = size_t(t) * NP;
        // at most NP strictly ordered pairs among K slots; fewer with duplicate ids
        uint32_t cnt = 0;
#pragma unroll
        for (int a = 0; a < K; ++a)
#pragma unroll
            for (int c = 0; c < K; ++c)
                if (kn[a] < kn[c])
                {
                    keys[base + cnt] = uint32_t(kn[a]) * kPairBase + kn[c];
                    ++cnt;
                }
        for (uint32_t a = cnt; a < uint32_t(NP); ++a)
            keys[base + a] = kUninitKey;
#pragma unroll
        for (int a = 0; a < NP; ++a)
            vals[base + a] = t;
    }
    else if (t < P + E)
    {
        const uint32_t e = t - P;
        const uint32_t i = d.edges[3 * e], j = d.edges[3 * e + 1];
        uint32_t key = kUninitKey;
        if (i < j)
            key = i * kPairBase + j;
        else if (j < i)
            key = j * kPairBase + i;
        const size_t o = size_t(P) * NP + e;
        keys[o] = key;
        vals[o] = t;
    }
}

// Segment heads of the sorted pair keys (segmentNodePairKernel,
// NodePair2TermsIndex.cu:94-128); `meta[0]` receives the number of valid
// (initialised) keys.
__global__ void k_pair_heads(const uint32_t *__restrict__ keys, uint32_t n, uint32_t *__restrict__ head,
                             uint32_t *__restrict__ meta)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t k = keys[i];
    const bool valid = (k != kUninitKey);
    head[i] = (valid && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
    if (valid && (i + 1 == n || keys[i + 1] == kUninitKey))
        meta[0] = i + 1;
    if (i == 0 && !valid)
        meta[0] = 0;
}

// compactNodePairKeyKernel (NodePair2TermsIndex.cu:130-169)
__global__ void k_pair_compact(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ head,
                               const uint32_t *__restrict__ hpos, uint32_t n, const uint32_t *__restrict__ meta,
                               uint32_t *__restrict__ half_keys, uint32_t *__restrict__ half_off,
                               uint32_t *__restrict__ meta_out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && head[i])
    {
        half_keys[hpos[i]] = keys[i];
        half_off[hpos[i]] = i;
    }
    if (i == 0)
    {
        const uint32_t U = (n > 0) ? hpos[n - 1] + head[n - 1] : 0u;
        half_off[U] = meta[0];
        meta_out[0] = U;
        meta_out[1] = meta[0];
    }
}

// Packed (pair, term) entries: dense -> pixel | slot_i << 25 | slot_j << 28,
// ARAP -> 0x80000000 | (row node == edge.x) << 30 | edge.
// Slots use the last matching position (dense_image_jacobian.cuh:92-104).
template <int K>
__global__ void k_pair_entries(SolveDesc d, const uint32_t *__restrict__ sorted_vals)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= uint32_t(d.n_pair_entries))
        return;
    // h = last half pair with half_off[h] <= e
    uint32_t lo = 0, hi = uint32_t(d.U);
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (d.half_off[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t key = d.half_keys[lo];
    const uint32_t ni = key / kPairBase, nj = key % kPairBase;
    const uint32_t t = sorted_vals[e];
    uint32_t packed;
    if (t < uint32_t(d.P))
    {
        uint32_t si = 0, sj = 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            const uint32_t v = d.pknn[size_t(t) * K + k];
            if (v == ni)
                si = k;
            if (v == nj)
                sj = k;
        }
        packed = t | (si << 25) | (sj << 28);
    }
    else
    {
        const uint32_t edge = t - uint32_t(d.P);
        const uint32_t row_is_x = (uint32_t(d.edges[3 * edge]) == ni) ? 1u : 0u;
        packed = 0x80000000u | (row_is_x << 30) | edge;
    }
    d.pair_entries[e] = packed;
}

// Symmetric keys (computeSymmetricNodePairKernel, NodePair2TermsIndex.cu:171-193);
// values carry (half pair << 1 | mirrored).
__global__ void k_sym_kv(SolveDesc d, uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= uint32_t(d.U))
        return;
    const uint32_t k = d.half_keys[u];
    const uint32_t i = k / kPairBase, j = k % kPairBase;
    keys[2 * u] = k;
    keys[2 * u + 1] = j * kPairBase + i;
    vals[2 * u] = 2 * u;
    vals[2 * u + 1] = 2 * u + 1;
}

__global__ void k_sym_scatter(SolveDesc d, const uint32_t *__restrict__ skeys, const uint32_t *__restrict__ svals)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= 2u * uint32_t(d.U))
        return;
    const uint32_t v = svals[p];
    if (v & 1u)
        d.lower_pos[v >> 1] = p;
    else
        d.upper_pos[v >> 1] = p;
    const uint32_t row = skeys[p] / kPairBase;
    d.blk_col[p + row + 1] = skeys[p] % kPairBase;
}

// Block-row pointers over the sorted symmetric keys (computeRowOffsetKernel,
// NodePairRowColumnIndex.cu:12-36) and the diagonal block columns.
__global__ void k_row_ptr(SolveDesc d, const uint32_t *__restrict__ skeys)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n > uint32_t(d.N))
        return;
    const uint32_t target = n * kPairBase;
    uint32_t lo = 0, hi = 2u * uint32_t(d.U);
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (skeys[mid] < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    d.row_ptr[n] = lo;
    if (n < uint32_t(d.N))
        d.blk_col[lo + n] = n;
}

// ============================================================ GN terms
// (load_knn and the per-pixel term live in sh_terms.hpp)

// Sum over a 256-thread work-group in FP64, fixed order (wave trees, then waves 0..3).
__device__ __forceinline__ double block_sum_256(double v)
{
    __shared__ double part[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = (part[0] + part[1]) + (part[2] + part[3]);
    __syncthreads();
    return r;
}


// Per-pixel residual / Jacobian (computeDenseImageJacobianKernel,
// DenseImageJacobianResidual.cu:164-310; forms :12-157).
// Work-group wg of the pixel map.
template <int K, class JT>
__device__ __forceinline__ void dense_terms_wg(const SolveDesc *__restrict__ sds, const uint32_t *__restrict__ wg_begin,
                                               int n_solve, uint32_t wg, int parity, int store)
{
    const int si = find_solve(wg_begin, n_solve, wg);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t p = (wg - wg_begin[si]) * blockDim.x + threadIdx.x;
    double pix_cost = 0.0;
    if (p < uint32_t(d.P))
        pix_cost = dense_pixel<K, JT>(d, p, parity, store);
    // FP64 cost partial of each 64-pixel group (a wave, fixed order); the
    // tile kernel that builds the records itself writes the same partials
    const double wc = wave_sum(pix_cost);
    if ((threadIdx.x & 63u) == 0 && p < uint32_t(d.P))
        d.cost_part[p / 64u] = wc;
}

// ARAP forward warp per edge + node translations (forwardWarpSmootherNodeKernel,
// NodeGraphHandler.cu:7-46; ComputeNodeMotionKernel, NodeMotionHandler.cu:9-21).
// Work-group wg of the edge map.
__device__ __forceinline__ void reg_terms_wg(const SolveDesc *__restrict__ sds, const uint32_t *__restrict__ wg_begin,
                                             int n_solve, uint32_t wg, int parity, int store)
{
    const int si = find_solve(wg_begin, n_solve, wg);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t item = (wg - wg_begin[si]) * blockDim.x + threadIdx.x;
    const dq8 *__restrict__ dq = pick2(d.dq, (d.gn_iter + parity) & 1);
    double edge_cost = 0.0;
    if (item < uint32_t(d.E))
    {
        const uint32_t i = d.edges[3 * item], j = d.edges[3 * item + 1], k = d.edges[3 * item + 2];
        const float4 xi = d.node_coords[i], xj = d.node_coords[j];
        const rt34 Ti = dq_se3(load_dq(dq, i));
        const rt34 Tj = dq_se3(load_dq(dq, j));
        const f3 vj = mk3(xj.x, xj.y, xj.z);
        const f3 a = rot_apply(Ti.r, vj) + Ti.t;
        const f3 b = rot_apply(Tj.r, vj) + Tj.t;
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float valid = (dx * dx + dy * dy + dz * dz > 64.f * d.node_radius_square) ? 0.f : 1.f;
        const float c = d.node_connect[size_t(i) * d.K + k];
        if (store)
        {
            float4 *o = reinterpret_cast<float4 *>(d.reg + size_t(item) * kReg);
            o[0] = make_float4(a.x, a.y, a.z, b.x);
            o[1] = make_float4(b.y, b.z, c, valid);
        }
        if (valid != 0.f)
        {
            const double ex = double(a.x) - double(b.x), ey = double(a.y) - double(b.y),
                         ez = double(a.z) - double(b.z);
            edge_cost = double(d.lam2_reg) * double(c) * double(c) * (ex * ex + ey * ey + ez * ez);
        }
    }
    else if (item < uint32_t(d.E + d.N))
    {
        const uint32_t n = item - uint32_t(d.E);
        const rt34 T = dq_se3(load_dq(dq, n));
        if (store)
            d.ttrans[n] = make_float4(T.t.x, T.t.y, T.t.z, 0.f);
    }
    const double wc = block_sum_256(edge_cost);
    if (threadIdx.x == 0)
        d.cost_part[(uint32_t(d.P) + 63u) / 64u + (wg - wg_begin[si])] = wc;
}

// Both term kernels as one launch: the pixel map's work-groups, then the edge
// map's (independent work, so the two overlap on the device).
template <int K, class JT>
__global__ __launch_bounds__(256) void k_terms(const SolveDesc *__restrict__ sds, const uint32_t *__restrict__ pix_map,
                                               uint32_t n_pix_wg, const uint32_t *__restrict__ reg_map, int n_solve,
                                               int parity, int store)
{
    if (blockIdx.x < n_pix_wg) // uniform over the group
        dense_terms_wg<K, JT>(sds, pix_map, n_solve, blockIdx.x, parity, store);
    else
        reg_terms_wg(sds, reg_map, n_solve, blockIdx.x - n_pix_wg, parity, store);
}

// ARAP Jacobian rows of one side, scaled by c (node_graph_reg_jacobian.cuh:250-275).
__device__ __forceinline__ void reg_rows(f3 a, f3 s, float c, bool is_i, float (&g)[18])
{
    if (is_i)
    {
        const float r[18] = {0.f, a.z, -a.y, 1.f, 0.f, 0.f, -a.z, 0.f, a.x, 0.f, 1.f, 0.f, a.y, -a.x, 0.f, 0.f, 0.f, 1.f};
#pragma unroll
        for (int q = 0; q < 18; ++q)
            g[q] = r[q] * c;
    }
    else
    {
        const float r[18] = {0.f, -s.z, s.y, -1.f, 0.f, 0.f, s.z, 0.f, -s.x, 0.f, -1.f, 0.f, -s.y, s.x, 0.f, 0.f, 0.f, -1.f};
#pragma unroll
        for (int q = 0; q < 18; ++q)
            g[q] = r[q] * c;
    }
}

__device__ __forceinline__ void load_reg(const SolveDesc &d, uint32_t e, f3 &a, f3 &b, float &cw, float &valid)
{
    const float4 *o = reinterpret_cast<const float4 *>(d.reg + size_t(e) * kReg);
    const float4 u0 = o[0], u1 = o[1];
    a = mk3(u0.x, u0.y, u0.z);
    b = mk3(u0.w, u1.x, u1.y);
    cw = u1.z;
    valid = u1.w;
}

constexpr int kSymIdx[6][6] = {{0, 1, 2, 3, 4, 5},     {1, 6, 7, 8, 9, 10},    {2, 7, 11, 12, 13, 14},
                               {3, 8, 12, 15, 16, 17}, {4, 9, 13, 16, 18, 19}, {5, 10, 14, 17, 19, 20}};

// Diagonal block + damping and -J^T W r per node, one wave per node, lanes
// stride the node's term list (computeBlockDiagonalPreconditionerKernel,
// PreconditionerRhsBuilder.cu:110-262; computeJtResidualWithIndexKernel,
// JtDotResidualIndexed.cu:40-150), then the FP64 block inverse.
template <int K, class JT>
__global__ __launch_bounds__(256) void k_diag_rhs(const SolveDesc *__restrict__ sds,
                                                  const uint32_t *__restrict__ wg_begin, int n_solve)
{
    __shared__ float stage[4][48];
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t n = (blockIdx.x - wg_begin[si]) * 4 + wave;
    if (n >= uint32_t(d.N))
        return;
    float acc[21], g[6];
#pragma unroll
    for (int a = 0; a < 21; ++a)
        acc[a] = 0.f;
#pragma unroll
    for (int a = 0; a < 6; ++a)
        g[a] = 0.f;
    const uint32_t P = d.P, E = d.E;
    const uint32_t beg = d.n2t_off[n], end = d.n2t_off[n + 1];
    for (uint32_t it = beg + lane; it < end; it += kWave)
    {
        const uint32_t t = d.n2t_terms[it];
        if (t < P)
        {
            uint16_t kn[K];
            float w8[K];
            load_knn<K>(d, t, kn, w8);
            float w = 0.f;
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (kn[k] == n)
                    w = w8[k];
            float hg[kHG];
            load_hg<JT>(d.jr, t, hg);
            const float w2 = w * w;
#pragma unroll
            for (int e = 0; e < 21; ++e)
                acc[e] += w2 * hg[e];
#pragma unroll
            for (int q = 0; q < 6; ++q)
                g[q] += w2 * hg[kHGg + q];
        }
        else if (t < P + E)
        {
            const uint32_t e = t - P;
            f3 a, b;
            float cw, valid;
            load_reg(d, e, a, b, cw, valid);
            if (valid != 0.f)
            {
                const bool is_i = (uint32_t(d.edges[3 * e]) == n);
                float gr[18];
                reg_rows(a, b, cw, is_i, gr);
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int aa = 0; aa < 6; ++aa)
#pragma unroll
                        for (int bb = aa; bb < 6; ++bb)
                            acc[kSymIdx[aa][bb]] += d.lam2_reg * gr[6 * c + bb] * gr[6 * c + aa];
                float g1[18];
                reg_rows(a, b, 1.f, is_i, g1);
                const float c2 = cw * cw;
                const float res[3] = {(a.x - b.x) * c2, (a.y - b.y) * c2, (a.z - b.z) * c2};
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int q = 0; q < 6; ++q)
                        g[q] += d.lam2_reg * res[c] * g1[6 * c + q];
            }
        }
        else
        {
            // node-translation term (node_motion_jacobian.cuh:7-35)
            const uint32_t m = t - P - E;
            const float4 pred = d.node_pred ? d.node_pred[m] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float4 tt = d.ttrans[m];
            const float c2 = pred.w * pred.w;
            acc[kSymIdx[3][3]] += d.lam2_nt * c2;
            acc[kSymIdx[4][4]] += d.lam2_nt * c2;
            acc[kSymIdx[5][5]] += d.lam2_nt * c2;
            g[3] += d.lam2_nt * (tt.x - pred.x) * c2;
            g[4] += d.lam2_nt * (tt.y - pred.y) * c2;
            g[5] += d.lam2_nt * (tt.z - pred.z) * c2;
        }
    }
#pragma unroll
    for (int a = 0; a < 21; ++a)
        acc[a] = wave_sum(acc[a]);
#pragma unroll
    for (int a = 0; a < 6; ++a)
        g[a] = wave_sum(g[a]);
    if (lane == 0)
    {
        float D[36];
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int b = 0; b < 6; ++b)
                D[6 * a + b] = acc[kSymIdx[a][b]];
#pragma unroll
        for (int a = 0; a < 6; ++a)
            D[7 * a] += d.damping;
#pragma unroll
        for (int a = 0; a < 36; ++a)
            stage[wave][a] = D[a];
#pragma unroll
        for (int a = 0; a < 6; ++a)
            stage[wave][36 + a] = -g[a];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const uint32_t diag_blk = d.row_ptr[n] + n;
    if (lane < 36)
        d.blk[size_t(diag_blk) * 36 + lane] = stage[wave][lane];
    else if (lane < 42)
        d.rhs[size_t(n) * 6 + (lane - 36)] = stage[wave][lane];
}

// FP64 inverse of each damped diagonal block, one thread per node
// (blockDiagonalInverseKernel, BlockDiagonalPreconditionerInverse.cuh:15-73).
__global__ __launch_bounds__(256) void k_dinv(const SolveDesc *__restrict__ sds, const uint32_t *__restrict__ wg_begin,
                                              int n_solve, PcgSlot2 *__restrict__ slots2, uint32_t n_slots2)
{
    // the pipelined PCG that follows publishes into zeroed slots and vectors
    // (tag 0, which its launch's first publications do not use); the grid has
    // a thread for every one of its groups
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (slots2 && gid < n_slots2)
#pragma unroll
        for (int a = 0; a < 4; ++a)
            slots2[gid].v[a >> 1][a & 1] = 0.0;
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t n = (blockIdx.x - wg_begin[si]) * blockDim.x + threadIdx.x;
    if (n >= uint32_t(d.N))
        return;
    const float *D = d.blk + size_t(d.row_ptr[n] + n) * 36;
    float A[36];
#pragma unroll
    for (int a = 0; a < 36; ++a)
        A[a] = D[a];
    double inv[36];
    invert6_fp64(A, inv);
    double *o = d.dinv + size_t(n) * 36;
#pragma unroll
    for (int a = 0; a < 36; ++a)
        o[a] = inv[a];
#pragma unroll
    for (int a = 0; a < 6; ++a)
    {
        d.mv[0][size_t(n) * 6 + a] = 0.0;
        d.mv[1][size_t(n) * 6 + a] = 0.0;
    }
}

// Off-diagonal blocks, one wave per unique pair (i<j): lanes stride the pair's
// dense entries, ARAP entries (at most two, at the end of the sorted list) are
// added by lane 0; block (i,j) and its transpose (j,i) are written
// (computeJtJNonDiagonalBlockKernel, JtJMaterializer.cu:66-175 with the
// transposed assembly of JtJMaterializerBinBlockCSR.cu:37-39).
template <int K, class JT>
__global__ __launch_bounds__(256) void k_offdiag(const SolveDesc *__restrict__ sds, const uint32_t *__restrict__ wg_begin,
                                                 int n_solve)
{
    __shared__ float stage[4][36];
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t u = (blockIdx.x - wg_begin[si]) * 4 + wave;
    if (u >= uint32_t(d.U))
        return;
    const uint32_t beg = d.half_off[u], end = d.half_off[u + 1];
    float acc[21];
#pragma unroll
    for (int a = 0; a < 21; ++a)
        acc[a] = 0.f;
    for (uint32_t e = beg + lane; e < end; e += kWave)
    {
        const uint32_t pk = d.pair_entries[e];
        if (pk & 0x80000000u)
            continue;
        const uint32_t p = pk & 0x1ffffffu;
        const uint32_t sa = (pk >> 25) & 7u, sb = (pk >> 28) & 7u;
        const float ww = d.gpw[size_t(p) * K + sa] * d.gpw[size_t(p) * K + sb];
        float hg[kHG];
        load_hg<JT>(d.jr, p, hg);
#pragma unroll
        for (int e = 0; e < 21; ++e)
            acc[e] += ww * hg[e];
    }
#pragma unroll
    for (int a = 0; a < 21; ++a)
        acc[a] = wave_sum(acc[a]);
    const uint32_t key = d.half_keys[u];
    const uint32_t ni = key / kPairBase, nj = key % kPairBase;
    if (lane == 0)
    {
        float B[36];
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int b = 0; b < 6; ++b)
                B[6 * a + b] = acc[kSymIdx[a][b]];
        for (uint32_t e = end; e > beg; --e)
        {
            const uint32_t pk = d.pair_entries[e - 1];
            if (!(pk & 0x80000000u))
                break;
            const uint32_t edge = pk & 0x3fffffffu;
            const bool row_is_x = (pk >> 30) & 1u;
            f3 a, b;
            float cw, valid;
            load_reg(d, edge, a, b, cw, valid);
            if (valid == 0.f)
                continue;
            float gi[18], gj[18];
            reg_rows(a, b, 1.f, true, gi);
            reg_rows(a, b, 1.f, false, gj);
            const float ws = d.lam2_reg * (cw * cw);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int aa = 0; aa < 6; ++aa)
#pragma unroll
                    for (int bb = 0; bb < 6; ++bb)
                    {
                        const float gr = row_is_x ? gi[6 * c + aa] : gj[6 * c + aa];
                        const float gc = row_is_x ? gj[6 * c + bb] : gi[6 * c + bb];
                        B[6 * aa + bb] += ws * gr * gc;
                    }
        }
#pragma unroll
        for (int a = 0; a < 36; ++a)
            stage[wave][a] = B[a];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const size_t up = size_t(d.upper_pos[u]) + ni + 1;
    const size_t lo = size_t(d.lower_pos[u]) + nj + 1;
    if (lane < 36)
    {
        d.blk[up * 36 + lane] = stage[wave][lane];
        const int a = lane / 6, b = lane % 6;
        d.blk[lo * 36 + lane] = stage[wave][6 * b + a];
    }
}

// ============================================================ PCG
// Work-groups of 192 threads = 32 node blocks x 6 rows; one solve per group.
constexpr int kPcgNodes = 32;
constexpr int kPcgThreads = kPcgNodes * 6;

__device__ __forceinline__ double block_sum_192(double v, double *sh)
{
    v = wave_sum(v);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        sh[wave] = v;
    __syncthreads();
    const double r = sh[0] + sh[1] + sh[2];
    __syncthreads();
    return r;
}

// Sum of the per-group partials of one solve, fixed order, broadcast to the group.
__device__ __forceinline__ double reduce_partials(const double *__restrict__ part, uint32_t b0, uint32_t b1, double *sh)
{
    double v = 0.0;
    if (threadIdx.x < 64)
    {
        for (uint32_t i = b0 + threadIdx.x; i < b1; i += 64)
            v += part[i];
        v = wave_sum(v);
        if (threadIdx.x == 0)
            sh[3] = v;
    }
    __syncthreads();
    const double r = sh[3];
    __syncthreads();
    return r;
}

// r = b, s = M^-1 b, x = 0, d_prev = 0, partial r.s   (blockPCGZeroInitKernel, BlockPCG.cuh:68-116)
__global__ __launch_bounds__(kPcgThreads) void k_pcg_init(const SolveDesc *__restrict__ sds,
                                                          const uint32_t *__restrict__ wg_begin, int n_solve,
                                                          double *__restrict__ part_rs)
{
    __shared__ double sh[4];
    __shared__ float rb[kPcgThreads];
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t node = (blockIdx.x - wg_begin[si]) * kPcgNodes + threadIdx.x / 6;
    const int a = threadIdx.x % 6;
    const bool live = node < uint32_t(d.N);
    const size_t row = size_t(node) * 6 + a;
    const float r = live ? d.rhs[row] : 0.f;
    rb[threadIdx.x] = r;
    __syncthreads();
    double prod = 0.0;
    if (live)
    {
        const double *Mi = d.dinv + size_t(node) * 36 + 6 * a;
        const int b0 = threadIdx.x - a;
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            s += Mi[c] * double(rb[b0 + c]);
        const float sf = float(s);
        d.r[row] = r;
        d.s[row] = sf;
        d.x[row] = 0.f;
        d.dv[0][row] = 0.f;
        prod = double(r) * double(sf);
    }
    const double tot = block_sum_192(prod, sh);
    if (threadIdx.x == 0)
        part_rs[blockIdx.x] = tot;
}

__device__ __forceinline__ float d_next(double beta, float dold, float s)
{
    return float(fma(beta, double(dold), double(s)));
}

// d_k = s + beta d_{k-1}, q = A d_k (materialised blocks), partial d.q
// (performSparseMVKernel + blockPCGThirdUpdateKernel, BlockPCG.cuh:175-189,
// BinBlockCSR.cuh:7-35).
__global__ __launch_bounds__(kPcgThreads) void k_pcg_spmv(const SolveDesc *__restrict__ sds,
                                                          const uint32_t *__restrict__ wg_begin, int n_solve, int k,
                                                          const double *__restrict__ part_rs,
                                                          double *__restrict__ part_dq)
{
    __shared__ double sh[4];
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const double delta = reduce_partials(part_rs, wg_begin[si], wg_begin[si + 1], sh);
    double beta = 0.0;
    if (k > 0 && d.delta[k - 1] != 0.0)
        beta = delta / d.delta[k - 1];
    if (threadIdx.x == 0 && blockIdx.x == wg_begin[si])
        d.delta[k] = delta;
    const float *__restrict__ dold = pick2(d.dv, k & 1);
    float *__restrict__ dnew = pick2(d.dv, (k + 1) & 1);
    const uint32_t node = (blockIdx.x - wg_begin[si]) * kPcgNodes + threadIdx.x / 6;
    const int a = threadIdx.x % 6;
    double prod = 0.0;
    if (node < uint32_t(d.N))
    {
        const size_t row = size_t(node) * 6 + a;
        const float dn = d_next(beta, dold[row], d.s[row]);
        dnew[row] = dn;
        const uint32_t b0 = d.row_ptr[node] + node;
        const uint32_t b1 = d.row_ptr[node + 1] + node + 1;
        float acc = 0.f;
        // four blocks per round: their loads are issued before their products,
        // which are summed in block order (the same sums as one block a round)
        constexpr int kU = 4;
        uint32_t bb = b0;
        for (; bb + kU <= b1; bb += kU)
        {
            uint32_t col[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u)
                col[u] = d.blk_col[bb + u];
            float av[kU][6], xv[kU][6];
#pragma unroll
            for (int u = 0; u < kU; ++u)
            {
                const float *A = d.blk + size_t(bb + u) * 36 + 6 * a;
                const size_t cr = size_t(col[u]) * 6;
#pragma unroll
                for (int c = 0; c < 6; ++c)
                {
                    av[u][c] = A[c];
                    xv[u][c] = d_next(beta, dold[cr + c], d.s[cr + c]);
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    acc += av[u][c] * xv[u][c];
        }
        for (; bb < b1; ++bb)
        {
            const uint32_t col = d.blk_col[bb];
            const float *A = d.blk + size_t(bb) * 36 + 6 * a;
            const size_t cr = size_t(col) * 6;
#pragma unroll
            for (int c = 0; c < 6; ++c)
                acc += A[c] * d_next(beta, dold[cr + c], d.s[cr + c]);
        }
        d.q[row] = acc;
        prod = double(dn) * double(acc);
    }
    const double tot = block_sum_192(prod, sh);
    if (threadIdx.x == 0)
        part_dq[blockIdx.x] = tot;
}

// d_k = s + beta d_{k-1} only (the matrix-free loop applies A separately).
__global__ __launch_bounds__(kPcgThreads) void k_pcg_dir(const SolveDesc *__restrict__ sds,
                                                         const uint32_t *__restrict__ wg_begin, int n_solve, int k,
                                                         const double *__restrict__ part_rs)
{
    __shared__ double sh[4];
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const double delta = reduce_partials(part_rs, wg_begin[si], wg_begin[si + 1], sh);
    double beta = 0.0;
    if (k > 0 && d.delta[k - 1] != 0.0)
        beta = delta / d.delta[k - 1];
    if (threadIdx.x == 0 && blockIdx.x == wg_begin[si])
        d.delta[k] = delta;
    const uint32_t node = (blockIdx.x - wg_begin[si]) * kPcgNodes + threadIdx.x / 6;
    if (node < uint32_t(d.N))
    {
        const size_t row = size_t(node) * 6 + threadIdx.x % 6;
        pick2(d.dv, (k + 1) & 1)[row] = d_next(beta, pick2(d.dv, k & 1)[row], d.s[row]);
    }
}

// partial d_k . q
__global__ __launch_bounds__(kPcgThreads) void k_pcg_dot(const SolveDesc *__restrict__ sds,
                                                         const uint32_t *__restrict__ wg_begin, int n_solve, int k,
                                                         double *__restrict__ part_dq)
{
    __shared__ double sh[4];
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t node = (blockIdx.x - wg_begin[si]) * kPcgNodes + threadIdx.x / 6;
    double prod = 0.0;
    if (node < uint32_t(d.N))
    {
        const size_t row = size_t(node) * 6 + threadIdx.x % 6;
        prod = double(pick2(d.dv, (k + 1) & 1)[row]) * double(d.q[row]);
    }
    const double tot = block_sum_192(prod, sh);
    if (threadIdx.x == 0)
        part_dq[blockIdx.x] = tot;
}

// alpha = delta / d.q; x += alpha d; r -= alpha q; s = M^-1 r; partial r.s
// (blockPCGSecondUpdateKernel, BlockPCG.cuh:125-169)
__global__ __launch_bounds__(kPcgThreads) void k_pcg_update(const SolveDesc *__restrict__ sds,
                                                            const uint32_t *__restrict__ wg_begin, int n_solve, int k,
                                                            const double *__restrict__ part_dq,
                                                            double *__restrict__ part_rs)
{
    __shared__ double sh[4];
    __shared__ float rb[kPcgThreads];
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const double dq = reduce_partials(part_dq, wg_begin[si], wg_begin[si + 1], sh);
    const double alpha = (dq != 0.0) ? d.delta[k] / dq : 0.0;
    const float *__restrict__ dnew = pick2(d.dv, (k + 1) & 1);
    const uint32_t node = (blockIdx.x - wg_begin[si]) * kPcgNodes + threadIdx.x / 6;
    const int a = threadIdx.x % 6;
    const bool live = node < uint32_t(d.N);
    const size_t row = size_t(node) * 6 + a;
    float rn = 0.f;
    if (live)
    {
        const double dn = double(dnew[row]);
        d.x[row] = float(fma(alpha, dn, double(d.x[row])));
        rn = float(fma(-alpha, double(d.q[row]), double(d.r[row])));
        d.r[row] = rn;
    }
    rb[threadIdx.x] = rn;
    __syncthreads();
    double prod = 0.0;
    if (live)
    {
        const double *Mi = d.dinv + size_t(node) * 36 + 6 * a;
        const int b0 = threadIdx.x - a;
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            s += Mi[c] * double(rb[b0 + c]);
        const float sf = float(s);
        d.s[row] = sf;
        prod = double(rn) * double(sf);
    }
    const double tot = block_sum_192(prod, sh);
    if (threadIdx.x == 0)
        part_rs[blockIdx.x] = tot;
}


// One wave per solve: the cost after the last update (SH_FLAG_COST_HISTORY).
__global__ void k_cost_final(const SolveDesc *__restrict__ sds, int slot)
{
    reduce_cost(global_desc(sds[blockIdx.x]), slot);
}

// T <- exp(xi) T, stored as a dual quaternion (ApplyWarpFieldUpdateKernel,
// SolverIterationData.cu:19-74)
__global__ __launch_bounds__(256) void k_update(const SolveDesc *__restrict__ sds, const uint32_t *__restrict__ wg_begin,
                                                int n_solve, int parity)
{
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    if (blockIdx.x == wg_begin[si] && threadIdx.x < 64)
        reduce_cost(d, d.gn_iter + parity); // cost of the state this iteration started from
    const uint32_t n = (blockIdx.x - wg_begin[si]) * blockDim.x + threadIdx.x;
    if (n >= uint32_t(d.N))
        return;
    warp_update_node(d, n, parity, d.x + size_t(n) * 6);
}

__global__ void k_identity(const SolveDesc *__restrict__ sds, const uint32_t *__restrict__ wg_begin, int n_solve)
{
    const int si = find_solve(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t n = (blockIdx.x - wg_begin[si]) * blockDim.x + threadIdx.x;
    if (n >= uint32_t(d.N))
        return;
    float4 *o = reinterpret_cast<float4 *>(d.dq[0] + n);
    o[0] = make_float4(0.f, 0.f, 0.f, 1.f);
    o[1] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Identity warp field for one solve (SolverIterationData.cu:78-88).
__global__ void k_identity_dq(float4 *__restrict__ dq, uint32_t n_nodes)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes)
        return;
    dq[2 * n] = make_float4(0.f, 0.f, 0.f, 1.f);
    dq[2 * n + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ============================================================ input producers
// KNN heap: replace the current maximum, then a Lomuto quicksort ascending
// (libs/star/geometry/node_graph/brute_force_knn.cuh:52-135).
template <int K>
struct KnnHeap
{
    float dist[K];
    uint16_t idx[K];
    int max_id;
    float max_d;
    __device__ void init(float d0)
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            dist[k] = d0;
            idx[k] = 0;
        }
        refresh();
    }
    __device__ __forceinline__ void refresh()
    {
        max_id = 0;
        max_d = dist[0];
#pragma unroll
        for (int k = 1; k < K; ++k)
            if (dist[k] > max_d)
            {
                max_d = dist[k];
                max_id = k;
            }
    }
    __device__ __forceinline__ void update(uint16_t id, float d)
    {
        if (d < max_d)
        {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k == max_id)
                {
                    dist[k] = d;
                    idx[k] = id;
                }
            refresh();
        }
    }
    // iterative Lomuto quicksort with the recursion order of the reference
    __device__ void sort()
    {
        float dd[K];
        uint16_t ii[K];
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            dd[k] = dist[k];
            ii[k] = idx[k];
        }
        int stack_lo[2 * K], stack_hi[2 * K];
        int top = 0;
        stack_lo[0] = 0;
        stack_hi[0] = K - 1;
        top = 1;
        while (top > 0)
        {
            --top;
            const int lo = stack_lo[top], hi = stack_hi[top];
            if (lo >= hi)
                continue;
            const float pivot = dd[hi];
            int i = lo - 1;
            for (int j = lo; j <= hi - 1; ++j)
                if (dd[j] < pivot)
                {
                    ++i;
                    const float tf = dd[i];
                    dd[i] = dd[j];
                    dd[j] = tf;
                    const uint16_t ti = ii[i];
                    ii[i] = ii[j];
                    ii[j] = ti;
                }
            {
                const float tf = dd[i + 1];
                dd[i + 1] = dd[hi];
                dd[hi] = tf;
                const uint16_t ti = ii[i + 1];
                ii[i + 1] = ii[hi];
                ii[hi] = ti;
            }
            const int pi = i + 1;
            // recursion order: left part first, then right part
            stack_lo[top] = pi + 1;
            stack_hi[top] = hi;
            ++top;
            stack_lo[top] = lo;
            stack_hi[top] = pi - 1;
            ++top;
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            dist[k] = dd[k];
            idx[k] = ii[k];
        }
    }
};

// fma(dz,dz,fma(dy,dy,dx*dx)) with no further contraction (NodeGraph.cu:32-34)
__device__ __forceinline__ float sqdist_fma(float dx, float dy, float dz)
{
    return __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmul_rn(dx, dx)));
}

constexpr int kSkinTile = 256;

// SkinnerKernel (Skinner.cu:10-94): nodes staged through LDS in tiles.
template <int K>
__global__ __launch_bounds__(256) void k_skin(const float4 *__restrict__ surfels, uint32_t S,
                                              const float4 *__restrict__ nodes, uint32_t N, float r2,
                                              uint16_t *__restrict__ knn, float *__restrict__ spatial,
                                              float *__restrict__ connect)
{
    __shared__ float4 tile[kSkinTile];
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = s < S;
    const float4 v = live ? surfels[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    KnnHeap<K> h;
    h.init(1e6f);
    for (uint32_t base = 0; base < N; base += kSkinTile)
    {
        __syncthreads();
        if (base + threadIdx.x < N)
            tile[threadIdx.x] = nodes[base + threadIdx.x];
        __syncthreads();
        const uint32_t cnt = min(uint32_t(kSkinTile), N - base);
        for (uint32_t i = 0; i < cnt; ++i)
        {
            const float4 nd = tile[i];
            const float dd = sqdist_fma(__fsub_rn(v.x, nd.x), __fsub_rn(v.y, nd.y), __fsub_rn(v.z, nd.z));
            h.update(uint16_t(base + i), dd);
        }
    }
    if (!live)
        return;
    h.sort();
    float w[K];
    float wsum = 0.f;
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        const float4 nd = nodes[h.idx[k]];
        const float dd = sqdist_fma(__fsub_rn(v.x, nd.x), __fsub_rn(v.y, nd.y), __fsub_rn(v.z, nd.z));
        w[k] = __expf(-dd / (2.f * r2));
        wsum += w[k];
    }
    const float inv = 1.f / wsum;
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        knn[size_t(s) * K + k] = h.idx[k];
        spatial[size_t(s) * K + k] = inv * w[k];
        connect[size_t(s) * K + k] = 1.f;
    }
}

// UpdateNodeDistanceKernel + BuildNodeGraphKernel + BuildNodeGraphPairKernel
// + ResetNodeConnectionKernel (NodeGraph.cu:10-145): the FP16 distance of each
// pair is recomputed on the fly instead of being stored as an N x N matrix.
template <int K>
__global__ __launch_bounds__(256) void k_node_graph(const float4 *__restrict__ nodes, uint32_t N, float r2,
                                                    uint16_t *__restrict__ node_knn, float *__restrict__ node_spatial,
                                                    uint16_t *__restrict__ edges, float *__restrict__ node_connect)
{
    __shared__ float4 tile[kSkinTile];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < N;
    const float4 vi = live ? nodes[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    KnnHeap<K> h;
    h.init(1e3f);
    for (uint32_t base = 0; base < N; base += kSkinTile)
    {
        __syncthreads();
        if (base + threadIdx.x < N)
            tile[threadIdx.x] = nodes[base + threadIdx.x];
        __syncthreads();
        const uint32_t cnt = min(uint32_t(kSkinTile), N - base);
        for (uint32_t jj = 0; jj < cnt; ++jj)
        {
            const uint32_t j = base + jj;
            const float4 vj = tile[jj];
            // distance is written for idx >= idy with v1 = node[idx], v2 = node[idy]
            const float4 hi = (i >= j) ? vi : vj;
            const float4 lo = (i >= j) ? vj : vi;
            const float dd = sqdist_fma(__fsub_rn(hi.x, lo.x), __fsub_rn(hi.y, lo.y), __fsub_rn(hi.z, lo.z));
            h.update(uint16_t(j), __half2float(__float2half_rn(dd)));
        }
    }
    if (!live)
        return;
    h.sort();
    float w[K];
    float wsum = 0.f;
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        w[k] = __expf(-h.dist[k] / (2.f * r2));
        wsum += w[k];
    }
    const float inv = 1.f / wsum;
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        node_knn[size_t(i) * K + k] = h.idx[k];
        node_spatial[size_t(i) * K + k] = w[k] * inv;
        node_connect[size_t(i) * K + k] = 1.f;
        const size_t e = size_t(i) * K + k;
        edges[3 * e] = uint16_t(i);
        edges[3 * e + 1] = h.idx[k];
        edges[3 * e + 2] = uint16_t(k);
    }
}

// ============================================================ host launchers
#define SH_GRID(n, b) dim3(unsigned(((n) + (b)-1) / (b)))

uint32_t tile_order_blocks_x(uint32_t width) { return (width + 31u) / 32u; }

uint32_t tile_order_size(uint32_t width, uint32_t height)
{
    return tile_order_blocks_x(width) * ((height + 31u) / 32u) * 1024u;
}

hipError_t launch_pixel_flags(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, bool tiled,
                              const IndexWs &ws, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_pixel_flags, dim3(n_wg), dim3(256), 0, st, g, wmap, tiled, ws);
    return hipGetLastError();
}

hipError_t launch_compact_pixels(int K, const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, bool tiled,
                                 const IndexWs &ws, uint32_t *keys, uint32_t *vals, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    if (K == 8)
        hipLaunchKernelGGL(k_compact_pixels<8>, dim3(n_wg), dim3(256), 0, st, g, wmap, tiled, ws, keys, vals);
    else
        hipLaunchKernelGGL(k_compact_pixels<4>, dim3(n_wg), dim3(256), 0, st, g, wmap, tiled, ws, keys, vals);
    return hipGetLastError();
}

hipError_t launch_n2t_split(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, const uint32_t *skeys,
                            const uint32_t *svals, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_n2t_split, dim3(n_wg), dim3(256), 0, st, g, wmap, skeys, svals);
    return hipGetLastError();
}

// One work-group per job; each thread scans a contiguous run of the array.
__global__ __launch_bounds__(1024) void k_scan_jobs(const ScanJob *__restrict__ jobs)
{
    __shared__ uint32_t wsum[16];
    const ScanJob j = jobs[blockIdx.x];
    const uint32_t n = j.n;
    const uint32_t *__restrict__ in = j.in;
    uint32_t *__restrict__ out = j.out;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = min(n, tid * per), e = min(n, b + per);
    uint32_t own = 0;
    for (uint32_t i = b; i < e; ++i)
        own += in[i];
    uint32_t x = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const uint32_t y = __shfl_up(x, o, 64);
        if (int(lane) >= o)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    uint32_t run = x - own;
    for (uint32_t i = 0; i < w; ++i)
        run += wsum[i];
    for (uint32_t i = b; i < e; ++i)
    {
        const uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
}

hipError_t launch_scan_jobs(const ScanJob *jobs, uint32_t n_jobs, hipStream_t st)
{
    if (n_jobs == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_scan_jobs, dim3(n_jobs), dim3(1024), 0, st, jobs);
    return hipGetLastError();
}

hipError_t launch_index_counts(const IndexGroup &g, int stage, const IndexWs &ws, uint32_t *host, hipStream_t st)
{
    if (g.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_index_counts, dim3((uint32_t(g.n) + 63u) / 64u), dim3(64), 0, st, g, stage, ws, host);
    return hipGetLastError();
}

hipError_t launch_pair_kv(const SolveDesc &d, uint32_t *keys, uint32_t *vals, hipStream_t st)
{
    const uint32_t n = uint32_t(d.P + d.E);
    if (n == 0)
        return hipSuccess;
    if (d.K == 8)
        hipLaunchKernelGGL(k_pair_kv<8>, SH_GRID(n, 256), dim3(256), 0, st, d, keys, vals);
    else
        hipLaunchKernelGGL(k_pair_kv<4>, SH_GRID(n, 256), dim3(256), 0, st, d, keys, vals);
    return hipGetLastError();
}

hipError_t launch_pair_heads(const uint32_t *keys, uint32_t n, uint32_t *head, uint32_t *meta, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_pair_heads, SH_GRID(n, 256), dim3(256), 0, st, keys, n, head, meta);
    return hipGetLastError();
}

hipError_t launch_pair_compact(const uint32_t *keys, const uint32_t *head, const uint32_t *hpos, uint32_t n,
                               const uint32_t *meta, uint32_t *half_keys, uint32_t *half_off, uint32_t *meta_out,
                               hipStream_t st)
{
    hipLaunchKernelGGL(k_pair_compact, SH_GRID(n > 0 ? n : 1, 256), dim3(256), 0, st, keys, head, hpos, n, meta,
                       half_keys, half_off, meta_out);
    return hipGetLastError();
}

hipError_t launch_pair_entries(const SolveDesc &d, const uint32_t *sorted_vals, hipStream_t st)
{
    const uint32_t n = uint32_t(d.n_pair_entries);
    if (n == 0)
        return hipSuccess;
    if (d.K == 8)
        hipLaunchKernelGGL(k_pair_entries<8>, SH_GRID(n, 256), dim3(256), 0, st, d, sorted_vals);
    else
        hipLaunchKernelGGL(k_pair_entries<4>, SH_GRID(n, 256), dim3(256), 0, st, d, sorted_vals);
    return hipGetLastError();
}

hipError_t launch_sym_kv(const SolveDesc &d, uint32_t *keys, uint32_t *vals, hipStream_t st)
{
    if (d.U == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_sym_kv, SH_GRID(d.U, 256), dim3(256), 0, st, d, keys, vals);
    return hipGetLastError();
}

hipError_t launch_sym_scatter(const SolveDesc &d, const uint32_t *skeys, const uint32_t *svals, hipStream_t st)
{
    if (d.U == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_sym_scatter, SH_GRID(2 * d.U, 256), dim3(256), 0, st, d, skeys, svals);
    return hipGetLastError();
}

hipError_t launch_row_ptr(const SolveDesc &d, const uint32_t *skeys, hipStream_t st)
{
    hipLaunchKernelGGL(k_row_ptr, SH_GRID(d.N + 1, 256), dim3(256), 0, st, d, skeys);
    return hipGetLastError();
}

hipError_t launch_identity(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, hipStream_t st)
{
    hipLaunchKernelGGL(k_identity, dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
    return hipGetLastError();
}

hipError_t launch_cost_final(const SolveDesc *sds, int n_solve, int slot, hipStream_t st)
{
    if (n_solve == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cost_final, dim3(n_solve), dim3(64), 0, st, sds, slot);
    return hipGetLastError();
}

hipError_t launch_identity_dq(dq8 *dq, uint32_t n_nodes, hipStream_t st)
{
    if (n_nodes == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_identity_dq, dim3((n_nodes + 255) / 256), dim3(256), 0, st, reinterpret_cast<float4 *>(dq),
                       n_nodes);
    return hipGetLastError();
}

hipError_t launch_terms(int K, bool fp16, const SolveDesc *sds, const uint32_t *pix_map, uint32_t n_pix_wg,
                        const uint32_t *reg_map, uint32_t n_reg_wg, int n_solve, int parity, hipStream_t st, int store)
{
    const dim3 grid(n_pix_wg + n_reg_wg);
    if (grid.x == 0)
        return hipSuccess;
    if (K == 8)
    {
        if (fp16)
            hipLaunchKernelGGL((k_terms<8, __half>), grid, dim3(256), 0, st, sds, pix_map, n_pix_wg, reg_map, n_solve,
                               parity, store);
        else
            hipLaunchKernelGGL((k_terms<8, float>), grid, dim3(256), 0, st, sds, pix_map, n_pix_wg, reg_map, n_solve,
                               parity, store);
    }
    else
    {
        if (fp16)
            hipLaunchKernelGGL((k_terms<4, __half>), grid, dim3(256), 0, st, sds, pix_map, n_pix_wg, reg_map, n_solve,
                               parity, store);
        else
            hipLaunchKernelGGL((k_terms<4, float>), grid, dim3(256), 0, st, sds, pix_map, n_pix_wg, reg_map, n_solve,
                               parity, store);
    }
    return hipGetLastError();
}

hipError_t launch_diag_rhs(int K, bool fp16, const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg,
                           hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    if (K == 8)
    {
        if (fp16)
            hipLaunchKernelGGL((k_diag_rhs<8, __half>), dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
        else
            hipLaunchKernelGGL((k_diag_rhs<8, float>), dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
    }
    else
    {
        if (fp16)
            hipLaunchKernelGGL((k_diag_rhs<4, __half>), dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
        else
            hipLaunchKernelGGL((k_diag_rhs<4, float>), dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
    }
    return hipGetLastError();
}

hipError_t launch_dinv(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, PcgSlot2 *slots2,
                       uint32_t n_slots2, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    if (slots2 && size_t(n_slots2) > size_t(n_wg) * 256)
        return hipErrorInvalidValue; // never: a solve has no more PCG groups than nodes
    hipLaunchKernelGGL(k_dinv, dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve, slots2, n_slots2);
    return hipGetLastError();
}

hipError_t launch_offdiag(int K, bool fp16, const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg,
                          hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    if (K == 8)
    {
        if (fp16)
            hipLaunchKernelGGL((k_offdiag<8, __half>), dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
        else
            hipLaunchKernelGGL((k_offdiag<8, float>), dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
    }
    else
    {
        if (fp16)
            hipLaunchKernelGGL((k_offdiag<4, __half>), dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
        else
            hipLaunchKernelGGL((k_offdiag<4, float>), dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve);
    }
    return hipGetLastError();
}

hipError_t launch_pcg_init(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, double *part_rs,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_pcg_init, dim3(n_wg), dim3(kPcgThreads), 0, st, sds, wmap, n_solve, part_rs);
    return hipGetLastError();
}

hipError_t launch_pcg_spmv(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k,
                           const double *part_rs, double *part_dq, hipStream_t st)
{
    hipLaunchKernelGGL(k_pcg_spmv, dim3(n_wg), dim3(kPcgThreads), 0, st, sds, wmap, n_solve, k, part_rs, part_dq);
    return hipGetLastError();
}

hipError_t launch_pcg_update(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k,
                             const double *part_dq, double *part_rs, hipStream_t st)
{
    hipLaunchKernelGGL(k_pcg_update, dim3(n_wg), dim3(kPcgThreads), 0, st, sds, wmap, n_solve, k, part_dq, part_rs);
    return hipGetLastError();
}

hipError_t launch_pcg_dir(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k,
                          const double *part_rs, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_pcg_dir, dim3(n_wg), dim3(kPcgThreads), 0, st, sds, wmap, n_solve, k, part_rs);
    return hipGetLastError();
}

hipError_t launch_pcg_dot(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k,
                          double *part_dq, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_pcg_dot, dim3(n_wg), dim3(kPcgThreads), 0, st, sds, wmap, n_solve, k, part_dq);
    return hipGetLastError();
}

hipError_t launch_update(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int parity,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_update, dim3(n_wg), dim3(256), 0, st, sds, wmap, n_solve, parity);
    return hipGetLastError();
}

hipError_t launch_skin(uint32_t K, const float4 *surfels, uint32_t S, const float4 *nodes, uint32_t N, float r2,
                       uint16_t *knn, float *spatial, float *connect, hipStream_t st)
{
    if (S == 0)
        return hipSuccess;
    if (K == 8)
        hipLaunchKernelGGL(k_skin<8>, SH_GRID(S, 256), dim3(256), 0, st, surfels, S, nodes, N, r2, knn, spatial,
                           connect);
    else
        hipLaunchKernelGGL(k_skin<4>, SH_GRID(S, 256), dim3(256), 0, st, surfels, S, nodes, N, r2, knn, spatial,
                           connect);
    return hipGetLastError();
}

hipError_t launch_node_graph(uint32_t K, const float4 *nodes, uint32_t N, float r2, uint16_t *node_knn,
                             float *node_spatial, uint16_t *edges, float *node_connect, hipStream_t st)
{
    if (K == 8)
        hipLaunchKernelGGL(k_node_graph<8>, SH_GRID(N, 256), dim3(256), 0, st, nodes, N, r2, node_knn, node_spatial,
                           edges, node_connect);
    else
        hipLaunchKernelGGL(k_node_graph<4>, SH_GRID(N, 256), dim3(256), 0, st, nodes, N, r2, node_knn, node_spatial,
                           edges, node_connect);
    return hipGetLastError();
}

} // namespace sh
