// Pixel-major materialisation of the dense part of J^T J and J^T r:
// per-tile slot partials and a sorted-key segmented reduction into blocks.
//
// The reference gathers each node pair's terms again for every block
// (JtJMaterializer.cu:66-175, PreconditionerRhsBuilder.cu:110-262), so every
// per-pixel record is read once per pair it belongs to.  Here the compacted
// pixels are cut into tiles of TP (64, 128 or 256) consecutive pixels of the GN phase's
// pixel order (tile order: 32x32-pixel blocks, Morton order inside a block),
// so a tile covers a compact image patch and few nodes.
//
// Once per solve (index build), one work-group per tile lists the tile's
// contributions -- (node pair (a <= b) of distinct nodes of a pixel, pixel,
// slot of a, slot of b) -- in (upper block of the pair, pixel) order: by the
// ranks of its node pairs and a pixel mask per pair (or, for a tile with
// too many nodes, an LDS radix sort of the keys).  Equal blocks within a
// group of kSlotPix pixels form a slot; slots are numbered across tiles in
// tile order, and a stable sort of the slots by block gives every block the
// list of its slot partials in tile order.
//
// Per GN iteration, one work-group per tile stages its records, weights and
// contribution list in LDS and sums every slot's contributions in pixel
// order (w_a w_b H, and w_a^2 g on diagonal slots); the combine kernel adds a
// block's slot partials in list order, the row's ARAP and node-translation
// terms and the damping.  Each record is read once per GN iteration, there
// are no atomics, and every sum has a fixed order.
#include <type_traits>

#include <rocprim/block/block_radix_sort.hpp>

#include "sh_common.hpp"
#include "sh_desc.hpp"
#include "sh_kernels.hpp"
#include "sh_terms.hpp"

namespace sh
{
constexpr int kTileThreads = 256;
constexpr int kTileRec = 28; // staged record: 21 H + 6 g + pad = 7 float4
// Tiles hold TP = 64, 128 or 256 pixels (the tuning picks one); a tile's
// contributions are packed as pixel | ka << 8 | kb << 11 and keyed
// block << 8 | pixel.
constexpr int kPixBits = 8;
constexpr uint32_t kPixMask = (1u << kPixBits) - 1u;
static_assert(kTilePixMax <= (1u << kPixBits), "tile pixel bits");
__device__ __forceinline__ uint32_t c_pix(uint32_t c) { return c & kPixMask; }
__device__ __forceinline__ uint32_t c_ka(uint32_t c) { return (c >> kPixBits) & 7u; }
__device__ __forceinline__ uint32_t c_kb(uint32_t c) { return (c >> (kPixBits + 3)) & 7u; }
// A slot is a block's contributions from one group of kSlotPix consecutive
// pixels of a tile, so a slot's list is at most kSlotPix long and k_tiles'
// longest serial sum is bounded.
#ifndef SH_SLOT_SHIFT
#define SH_SLOT_SHIFT 6
#endif
constexpr int kSlotShift = SH_SLOT_SHIFT;
constexpr int kSlotPix = 1 << kSlotShift;
static_assert((1 << kSlotShift) == kSlotPix && kSlotPix >= 32, "slot pixel groups");

// A tile of TP pixels for K nodes per pixel: two threads per pixel.
template <int K, int TP>
struct TileGeom
{
    static constexpr int kPairs = K * (K + 1) / 2; // (ka <= kb) slot pairs per pixel
    static constexpr int kThreads = 2 * TP;
    static constexpr int kItems = kPairs / 2;      // sort items per thread
    static constexpr int kEnt = TP * kPairs;       // contributions a tile can hold
    static_assert(kPairs % 2 == 0 && TP <= int(kTilePixMax), "tile geometry");
};

__device__ __forceinline__ int find_solve_t(const uint32_t *__restrict__ begin, int n, uint32_t wg)
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

// Full-block index of (lo, hi): the diagonal block of row lo for lo == hi,
// else the position of hi among row lo's ascending neighbours, base + 1 + the
// neighbours below hi, counted from row lo's bitmap word that holds hi and
// the neighbour count of the words before it (bitpre: the rows of the
// solve, kBitmapWords words each).  Two independent loads.
__device__ __forceinline__ uint32_t block_of_pair(const uint32_t *__restrict__ row_ptr,
                                                  const uint2 *__restrict__ bitpre, uint32_t lo, uint32_t hi)
{
    const uint32_t base = row_ptr[lo] + lo;
    const uint2 w = bitpre[size_t(lo) * kBitmapWords + (hi >> 5)];
    return lo == hi ? base : base + 1u + w.y + uint32_t(__popc(w.x & ((1u << (hi & 31u)) - 1u)));
}

// (ka, kb) of pair q in 0 .. K(K+1)/2 - 1, ka <= kb, lexicographic; folds to
// constants for a constant q.
template <int K>
__host__ __device__ constexpr int pair_first(int q)
{
    int a = 0, rem = q;
    for (int r = 0; r < K; ++r)
        if (rem >= K - a)
        {
            rem -= K - a;
            ++a;
        }
    return a;
}

template <int K>
__host__ __device__ constexpr int pair_second(int q)
{
    int a = 0, rem = q;
    for (int r = 0; r < K; ++r)
        if (rem >= K - a)
        {
            rem -= K - a;
            ++a;
        }
    return a + rem;
}

// Exclusive scan of one value per thread over the work-group, in thread order;
// returns the thread's prefix, *total gets the sum.
template <int NT>
__device__ __forceinline__ uint32_t group_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t *total)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const uint32_t y = __shfl_up(x, o, 64);
        if (int(lane) >= o)
            x += y;
    }
    if (lane == 63)
        sh[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i)
    {
        before += (uint32_t(i) < w) ? sh[i] : 0u;
        all += sh[i];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Minimum of lo and maximum of hi over the work-group (every thread gets both);
// sh holds 16 values.
template <int NT>
__device__ __forceinline__ void group_min_max(uint32_t &lo, uint32_t &hi, uint32_t *sh)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
        lo = min(lo, uint32_t(__shfl_xor(lo, o, 64)));
        hi = max(hi, uint32_t(__shfl_xor(hi, o, 64)));
    }
    if (lane == 0)
    {
        sh[w] = lo;
        sh[8 + w] = hi;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NT / 64; ++i)
    {
        lo = min(lo, sh[i]);
        hi = max(hi, sh[8 + i]);
    }
    __syncthreads();
}

// ------------------------------------------------------------ index build
// LDS of the tile build's ranking path: the tile's nodes and their ranks,
// its node pairs (keys la * Dn + lb of the ranks la <= lb, Dn distinct nodes)
// and their ranks, and each pair's pixels.
template <int TP>
struct TileRankLds
{
    static constexpr int kMaxNodes = 128;   // distinct nodes of a tile on this path
    static constexpr int kPairWords = kMaxNodes * kMaxNodes / 32;
    static constexpr int kMaxPairs = 1024;  // distinct pairs of a tile on this path
    static constexpr int kMaskWords = TP / 32;
    static constexpr int kGroupWords = (kSlotPix < TP ? kSlotPix : TP) / 32; // mask words of a slot's pixel group
    uint32_t nbits[kBitmapWords];
    uint16_t npre[kBitmapWords];
    uint16_t node_of[kMaxNodes];
    uint32_t pbits[kPairWords];
    uint16_t ppre[kPairWords];
    uint16_t pkey[kMaxPairs];
    uint32_t mask[kMaxPairs][kMaskWords];
    uint16_t coff[kMaxPairs], soff[kMaxPairs];
    uint32_t total[2];
};

// Per tile: the contributions in block order, each block's in pixel order,
// at the tile's fixed stride in the workspace: the packed contributions
// (pixel | ka << 8 | kb << 11), each slot's block and first contribution (a
// slot: the contributions of one block and one group of kSlotPix pixels),
// and the tile's slot and contribution counts.  The order of the blocks of
// pairs lo <= hi is the order of (lo, hi): rows in node order, a row's blocks
// in column order.  So the tile ranks its nodes (an LDS bitmap of node ids
// and its prefix counts), then its pairs by (rank lo, rank hi) the same way,
// and a contribution's place is its pair's first place plus the pair's pixels
// before it (a pixel mask per pair); the block of a pair is looked up once.
// A tile with more nodes or pairs than that LDS holds sorts its
// (block << 8 | pixel) keys instead (a stable LDS radix sort).  Compiled for
// five 128-pixel tiles a CU (at most 102 registers; the kernel waits mostly
// on memory and barriers, so more tiles in flight pay).
template <int K, int TP>
__global__ __launch_bounds__(2 * TP, 5 * 128 / TP) void k_tile_build(IndexGroup g, const uint32_t *__restrict__ wmap,
                                                                      IndexWs ws)
{
    using G = TileGeom<K, TP>;
    using R = TileRankLds<TP>;
    using Sort = rocprim::block_radix_sort<uint32_t, G::kThreads, G::kItems, uint8_t>;
    __shared__ union
    {
        typename Sort::storage_type sort;
        R rank;
    } sm;
    __shared__ uint32_t last_key[G::kThreads];
    __shared__ uint32_t red[16];
    const int it = find_solve_t(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc d = global_desc(g.sds[item.slot]);
    const uint32_t t = blockIdx.x - wmap[it], tid = threadIdx.x;
    const uint32_t pl = tid >> 1, half = tid & 1;
    const uint32_t p = t * TP + pl;
    uint32_t *__restrict__ tile_nslot = ws.tnslot + item.tc0;
    uint32_t *__restrict__ tile_ncontrib = ws.tncontrib + item.tc0;
    // a row structure that overflowed its capacity is rejected by the host
    // after this launch; read none of it
    if (d.row_ptr[d.N] + uint32_t(d.N) > d.blk_cap)
    {
        if (tid == 0)
        {
            tile_nslot[t] = 0;
            tile_ncontrib[t] = 0;
        }
        return; // uniform over the group
    }
    const uint2 *__restrict__ bitpre = ws.bitpre + size_t(item.node0) * kBitmapWords;
    const size_t stride = size_t(G::kEnt);
    const size_t tbase = size_t(item.t0 + t) * stride;
    uint16_t *__restrict__ cout = ws.tcontrib + tbase;

    // this thread's half of the pixel's pairs: q = i or kItems + i, as
    // (lo | hi << 16) and (ka | kb << 3); all ones: no contribution
    uint32_t lohi[G::kItems];
    uint8_t vals[G::kItems];
#pragma unroll
    for (int i = 0; i < G::kItems; ++i)
    {
        lohi[i] = 0xffffffffu;
        vals[i] = 0;
    }
    uint16_t kn[K];
    bool act[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
        act[k] = false, kn[k] = 0;
    if (p < uint32_t(d.P))
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
            kn[k] = d.gpknn[size_t(p) * K + k];
        // slot k is active when no later slot names the same node (last match,
        // dense_image_jacobian.cuh:17-23)
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            act[k] = true;
#pragma unroll
            for (int k2 = k + 1; k2 < K; ++k2)
                if (kn[k2] == kn[k])
                    act[k] = false;
        }
#pragma unroll
        for (int i = 0; i < G::kItems; ++i)
        {
            const int ka0 = pair_first<K>(i), kb0 = pair_second<K>(i);
            const int ka1 = pair_first<K>(G::kItems + i), kb1 = pair_second<K>(G::kItems + i);
            const bool on_a = half ? act[ka1] : act[ka0], on_b = half ? act[kb1] : act[kb0];
            const uint32_t na = half ? kn[ka1] : kn[ka0], nb = half ? kn[kb1] : kn[kb0];
            if (on_a && on_b)
            {
                lohi[i] = min(na, nb) | (max(na, nb) << 16);
                vals[i] = uint8_t(half ? (ka1 | (kb1 << 3)) : (ka0 | (kb0 << 3)));
            }
        }
    }

    // ---- ranking path
    R &r = sm.rank;
    for (uint32_t w = tid; w < kBitmapWords; w += G::kThreads)
        r.nbits[w] = 0u;
    __syncthreads();
    // the pixel's active nodes, half of them per thread
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (act[k] && (k < K / 2) == (half == 0))
            atomicOr(&r.nbits[kn[k] >> 5], 1u << (kn[k] & 31u));
    __syncthreads();
    if (tid < 64)
    {
        // node ranks: exclusive prefix of the word popcounts (two words a lane)
        const uint32_t w0 = 2 * tid, w1 = w0 + 1;
        const uint32_t b0 = r.nbits[w0], b1 = r.nbits[w1];
        const uint32_t c0 = __popc(b0), c1 = __popc(b1);
        uint32_t v = c0 + c1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t u = __shfl_up(v, o, 64);
            if (int(tid) >= o)
                v += u;
        }
        const uint32_t ex = v - c0 - c1;
        r.npre[w0] = uint16_t(ex);
        r.npre[w1] = uint16_t(ex + c0);
        if (tid == 63)
            r.total[0] = v;
        // the node of each rank
        uint32_t at = ex;
        for (uint32_t bb = b0; bb; bb &= bb - 1, ++at)
            if (at < uint32_t(R::kMaxNodes))
                r.node_of[at] = uint16_t(w0 * 32 + __ffs(bb) - 1);
        for (uint32_t bb = b1; bb; bb &= bb - 1, ++at)
            if (at < uint32_t(R::kMaxNodes))
                r.node_of[at] = uint16_t(w1 * 32 + __ffs(bb) - 1);
    }
    __syncthreads();
    const uint32_t dn = r.total[0];
    bool ranked = dn <= uint32_t(R::kMaxNodes);
    uint32_t pk[G::kItems]; // pair keys, then pair ranks
    if (ranked)
    {
        const uint32_t pwords = (dn * dn + 31u) / 32u;
        for (uint32_t w = tid; w < pwords; w += G::kThreads)
            r.pbits[w] = 0u;
        __syncthreads();
        auto nrank = [&](uint32_t n) {
            return uint32_t(r.npre[n >> 5]) + uint32_t(__popc(r.nbits[n >> 5] & ((1u << (n & 31u)) - 1u)));
        };
#pragma unroll
        for (int i = 0; i < G::kItems; ++i)
            if (lohi[i] != 0xffffffffu)
            {
                pk[i] = nrank(lohi[i] & 0xffffu) * dn + nrank(lohi[i] >> 16);
                atomicOr(&r.pbits[pk[i] >> 5], 1u << (pk[i] & 31u));
            }
        __syncthreads();
        // pair ranks: exclusive prefix of the word popcounts, a run of words per thread
        const uint32_t per = (pwords + G::kThreads - 1) / G::kThreads;
        const uint32_t wb = min(pwords, tid * per), we = min(pwords, wb + per);
        uint32_t own = 0;
        for (uint32_t w = wb; w < we; ++w)
            own += __popc(r.pbits[w]);
        uint32_t np = 0;
        uint32_t run = group_exclusive_scan<G::kThreads>(own, red, &np);
        ranked = np <= uint32_t(R::kMaxPairs); // uniform
        if (ranked)
        {
            for (uint32_t w = wb; w < we; ++w)
            {
                const uint32_t bw = r.pbits[w];
                r.ppre[w] = uint16_t(run);
                for (uint32_t bb = bw; bb; bb &= bb - 1, ++run)
                    r.pkey[run] = uint16_t(w * 32 + __ffs(bb) - 1);
            }
            for (uint32_t e = tid; e < np * uint32_t(R::kMaskWords); e += G::kThreads)
                (&r.mask[0][0])[e] = 0u;
            __syncthreads();
#pragma unroll
            for (int i = 0; i < G::kItems; ++i)
                if (lohi[i] != 0xffffffffu)
                {
                    const uint32_t key = pk[i];
                    pk[i] = uint32_t(r.ppre[key >> 5]) + uint32_t(__popc(r.pbits[key >> 5] & ((1u << (key & 31u)) - 1u)));
                    atomicOr(&r.mask[pk[i]][pl >> 5], 1u << (pl & 31u));
                }
            __syncthreads();
            // each pair's contribution count and slot count (non-empty
            // kSlotPix-pixel groups), a run of pairs per thread, and their
            // exclusive scans
            const uint32_t per2 = (np + G::kThreads - 1) / G::kThreads;
            const uint32_t qb = min(np, tid * per2), qe = min(np, qb + per2);
            uint32_t nc = 0, ns = 0;
            for (uint32_t q = qb; q < qe; ++q)
#pragma unroll
                for (int w = 0; w < R::kMaskWords; ++w)
                {
                    const uint32_t m = r.mask[q][w];
                    nc += __popc(m);
                    if (w % R::kGroupWords == 0)
                    {
                        uint32_t grp = 0;
#pragma unroll
                        for (int w2 = 0; w2 < R::kGroupWords; ++w2)
                            grp |= r.mask[q][w + w2];
                        ns += grp ? 1u : 0u;
                    }
                }
            uint32_t total_c = 0, total_s = 0;
            uint32_t c_run = group_exclusive_scan<G::kThreads>(nc, red, &total_c);
            uint32_t s_run = group_exclusive_scan<G::kThreads>(ns, red, &total_s);
            for (uint32_t q = qb; q < qe; ++q)
            {
                r.coff[q] = uint16_t(c_run);
                r.soff[q] = uint16_t(s_run);
                // the pair's slots: block and first contribution
                const uint32_t key = r.pkey[q];
                const uint32_t lo = r.node_of[key / dn], hi = r.node_of[key % dn];
                const uint32_t blk = block_of_pair(d.row_ptr, bitpre, lo, hi);
                uint32_t c = c_run;
#pragma unroll
                for (int w = 0; w < R::kMaskWords; w += R::kGroupWords)
                {
                    uint32_t cnt = 0;
#pragma unroll
                    for (int w2 = 0; w2 < R::kGroupWords; ++w2)
                        cnt += __popc(r.mask[q][w + w2]);
                    if (cnt)
                    {
                        ws.tslot_block[tbase + s_run] = blk;
                        ws.tslot_first[tbase + s_run] = uint16_t(c);
                        ++s_run;
                        c += cnt;
                    }
                }
                c_run = c;
            }
            __syncthreads();
            // each contribution at its pair's first place plus the pair's pixels before it
#pragma unroll
            for (int i = 0; i < G::kItems; ++i)
                if (lohi[i] != 0xffffffffu)
                {
                    const uint32_t q = pk[i], pw = pl >> 5;
                    uint32_t at = r.coff[q] + uint32_t(__popc(r.mask[q][pw] & ((1u << (pl & 31u)) - 1u)));
#pragma unroll
                    for (int w = 0; w < R::kMaskWords; ++w)
                        if (uint32_t(w) < pw)
                            at += __popc(r.mask[q][w]);
                    cout[at] = uint16_t(pl | (uint32_t(vals[i]) << kPixBits));
                }
            if (tid == 0)
            {
                tile_nslot[t] = total_s;
                tile_ncontrib[t] = total_c;
                atomicMax(tile_nslot + uint32_t(d.n_tiles) + 1, total_s); // largest slot count of the solve's tiles
            }
            return; // uniform
        }
    }
    __syncthreads(); // the LDS goes to the sort

    // ---- sorting path: keys (block << 8 | pixel)
    uint32_t keys[G::kItems];
#pragma unroll
    for (int i = 0; i < G::kItems; ++i)
        keys[i] = lohi[i] == 0xffffffffu
                      ? 0xffffffffu
                      : (block_of_pair(d.row_ptr, bitpre, lohi[i] & 0xffffu, lohi[i] >> 16) << kPixBits) | pl;
    // keys are made relative to the tile's smallest block and sorted on their
    // block bits only: the items are generated in pixel order and the radix
    // sort is stable, so each block's items stay in pixel order, as with the
    // whole key, in fewer passes.  An invalid key (all ones) sorts after every
    // valid one: the relative blocks stay below the largest value the sorted
    // bits hold.
    uint32_t bmin = 0xffffffffu, bmax = 0u;
#pragma unroll
    for (int i = 0; i < G::kItems; ++i)
        if (keys[i] != 0xffffffffu)
        {
            bmin = min(bmin, keys[i] >> kPixBits);
            bmax = max(bmax, keys[i] >> kPixBits);
        }
    group_min_max<G::kThreads>(bmin, bmax, red);
    if (bmin == 0xffffffffu)
        bmin = bmax = 0u; // a tile with no valid item
#pragma unroll
    for (int i = 0; i < G::kItems; ++i)
        if (keys[i] != 0xffffffffu)
            keys[i] -= bmin << kPixBits;
    const unsigned block_bits = 32u - unsigned(__clz(bmax - bmin + 1u));
    Sort().sort(keys, vals, sm.sort, unsigned(kPixBits), min(32u, unsigned(kPixBits) + block_bits));

    // slot heads: item gi = tid * kItems + i starts a slot when its block or
    // its pixel group (kSlotPix consecutive pixels of the tile) differs from
    // the previous item's (blocked arrangement after the sort), so no slot
    // holds more than kSlotPix contributions
    last_key[tid] = keys[G::kItems - 1];
    __syncthreads();
    const uint32_t prev0 = tid > 0 ? last_key[tid - 1] : 0xffffffffu;
    uint32_t prev = prev0, nhead = 0, nvalid = 0;
#pragma unroll
    for (int i = 0; i < G::kItems; ++i)
    {
        const bool valid = keys[i] != 0xffffffffu;
        nhead += (valid && (prev == 0xffffffffu || (keys[i] >> kSlotShift) != (prev >> kSlotShift))) ? 1u : 0u;
        nvalid += valid ? 1u : 0u;
        prev = keys[i];
    }
    uint32_t total_heads = 0, total_valid = 0;
    uint32_t slot = group_exclusive_scan<G::kThreads>(nhead, red, &total_heads);
    (void)group_exclusive_scan<G::kThreads>(nvalid, red, &total_valid);
    prev = prev0;
#pragma unroll
    for (int i = 0; i < G::kItems; ++i)
    {
        const uint32_t gi = tid * G::kItems + i;
        const bool valid = keys[i] != 0xffffffffu;
        const bool head = valid && (prev == 0xffffffffu || (keys[i] >> kSlotShift) != (prev >> kSlotShift));
        if (valid)
            cout[gi] = uint16_t((keys[i] & kPixMask) | (uint32_t(vals[i]) << kPixBits));
        if (head)
        {
            ws.tslot_block[tbase + slot] = (keys[i] >> kPixBits) + bmin;
            ws.tslot_first[tbase + slot] = uint16_t(gi);
            ++slot;
        }
        prev = keys[i];
    }
    if (tid == 0)
    {
        tile_nslot[t] = total_heads;
        tile_ncontrib[t] = total_valid; // valid items lead after the sort
        atomicMax(tile_nslot + uint32_t(d.n_tiles) + 1, total_heads); // largest slot count of the solve's tiles
    }
}

// Per tile: move the tile's contributions and slots to their compact places
// (bases from the scans of the counts); slot_first becomes absolute, and
// slot_block / slot id pairs are laid out for the block sort.
template <int K, int TP>
__global__ __launch_bounds__(kTileThreads) void k_tile_compact(IndexGroup g, const uint32_t *__restrict__ wmap,
                                                               IndexWs ws, uint32_t *__restrict__ sort_keys,
                                                               uint32_t *__restrict__ sort_vals)
{
    using G = TileGeom<K, TP>;
    const int it = find_solve_t(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc d = global_desc(g.sds[item.slot]);
    const uint32_t t = blockIdx.x - wmap[it], tid = threadIdx.x;
    const size_t tbase = size_t(item.t0 + t) * size_t(G::kEnt);
    const uint32_t c0 = d.tile_cbase[t], nc = d.tile_cbase[t + 1] - c0;
    const uint32_t s0 = d.tile_sbase[t], ns = d.tile_sbase[t + 1] - s0;
    // the contributions four a 64-bit word: the source is word aligned (the
    // tile stride is a multiple of four), the destination after h < 4 leading
    // entries; destination word j then takes its entries from source words j
    // and j + 1, and the last word it reads holds entry nc - 1, inside the
    // tile's stride
    static_assert(G::kEnt % 4 == 0, "word-aligned tile stride");
    const uint16_t *__restrict__ src = ws.tcontrib + tbase;
    uint16_t *__restrict__ dst = d.tc_contrib + c0;
    const uint32_t h = min(nc, (4u - (c0 & 3u)) & 3u); // entries before the first aligned word
    const uint32_t m = (nc - h) / 4;                  // destination words
    const uint32_t r = nc - h - 4 * m;                // entries after the last
    const uint64_t *__restrict__ sw = reinterpret_cast<const uint64_t *>(src);
    uint64_t *__restrict__ dw = reinterpret_cast<uint64_t *>(dst + h);
    if (h == 0)
        for (uint32_t j = tid; j < m; j += kTileThreads)
            dw[j] = sw[j];
    else
    {
        const uint32_t lo = 16u * h, hi = 64u - lo;
        for (uint32_t j = tid; j < m; j += kTileThreads)
            dw[j] = (sw[j] >> lo) | (sw[j + 1] << hi);
    }
    if (tid < h)
        dst[tid] = src[tid];
    if (tid < r)
        dst[h + 4 * m + tid] = src[h + 4 * m + tid];
    for (uint32_t s = tid; s < ns; s += kTileThreads)
    {
        d.tc_slot_first[s0 + s] = c0 + ws.tslot_first[tbase + s];
        sort_keys[item.sb0 + s0 + s] = item.blk0 + ws.tslot_block[tbase + s];
        sort_vals[item.sb0 + s0 + s] = s0 + s;
    }
    if (t + 1 == uint32_t(d.n_tiles) && tid == 0)
        d.tc_slot_first[d.tile_sbase[t + 1]] = d.tile_cbase[t + 1];
}

// One wave per node, kRowNodes nodes a work-group: for every block (i, c) of
// row i, the index of its transpose (c, i), in row c.
static_assert(kTileThreads == 64 * kRowNodes, "k_block_transpose: one wave per node");
__global__ __launch_bounds__(kTileThreads) void k_block_transpose(IndexGroup g, const uint32_t *__restrict__ wmap,
                                                                  IndexWs ws)
{
    const int it = find_solve_t(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc d = global_desc(g.sds[item.slot]);
    const uint2 *__restrict__ bitpre = ws.bitpre + size_t(item.node0) * kBitmapWords;
    const uint32_t i = (blockIdx.x - wmap[it]) * kRowNodes + (threadIdx.x >> 6);
    if (i >= uint32_t(d.N))
        return;
    const uint32_t base = d.row_ptr[i] + i, L1 = d.row_ptr[i + 1] - d.row_ptr[i] + 1;
    for (uint32_t c = threadIdx.x & 63u; c < L1; c += 64)
    {
        const uint32_t col = d.blk_col[base + c];
        d.blk_transpose[base + c] = col == i ? base + c : block_of_pair(d.row_ptr, bitpre, col, i); // block (col, i) of row col
    }
}

// ------------------------------------------------------------ GN phase
// One work-group per tile.  The records (28 floats: 21 H, 6 g, pad) and the
// folded weights of the tile's pixels are staged in LDS with the tile's
// contribution list.  A thread sums a whole slot, all seven float4 groups, over
// the slot's contributions in list order (increasing pixel), coef * record:
// each contribution's entry and two weights are read once for its 28 products
// (the earlier form, one task per (slot, group), read them seven times and
// spent about five VALU instructions per product on decoding, addressing and
// predication).  The slots are visited longest list first (a counting sort of
// the list lengths in LDS), so the lanes of a wave run lists of about the
// same length; which lane sums a slot does not change its result.  The slot
// lists take LDS for the launch's largest tile only.

template <int K, class JT, int TP, bool FUSED>
__global__ __launch_bounds__(2 * TP) void k_tiles(const SolveDesc *__restrict__ sds,
                                                        const uint32_t *__restrict__ wg_begin, int n_solve,
                                                        uint32_t slot_cap, int parity)
{
    using G = TileGeom<K, TP>;
    constexpr int NT = G::kThreads;
    constexpr int kEnt = G::kEnt;
    constexpr uint32_t kNB = 2 * (TP + 1); // visiting-order buckets
    __shared__ float4 Hs[TP * (kTileRec / 4)];
    __shared__ __align__(16) float wS[TP * K];
    __shared__ __align__(8) uint16_t ent[kEnt];
    __shared__ uint32_t hist[kNB];
    // slot list starts sst[slot_cap + 2], then the visiting order perm[slot_cap]
    extern __shared__ uint16_t slot_lds[];
    uint16_t *__restrict__ sst = slot_lds;
    uint16_t *__restrict__ perm = slot_lds + slot_cap + 2;
    const int si = find_solve_t(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t t = blockIdx.x - wg_begin[si];
    if (t >= uint32_t(d.n_tiles))
        return; // uniform over the group
    const uint32_t tid = threadIdx.x;
    const uint32_t p0 = t * TP;
    const uint32_t np = min(uint32_t(TP), uint32_t(d.P) - p0);
    // the tile's records and weights first: their loads need only the
    // descriptor, so they are in flight with the tile's list bounds
    float hg[kHG];
    float wk[K];
    double cost = 0.0;
    if (tid < np)
    {
        const uint32_t p = p0 + tid;
        if constexpr (FUSED)
            cost = dense_pixel_hg<K>(d, p, parity, hg); // the record, built here instead of read
        else
            load_hg<JT>(d.jr, p, hg);
        // K weights a pixel, 16-byte aligned (the compaction stores them as float4)
        const float4 *__restrict__ w4 = reinterpret_cast<const float4 *>(d.gpw + size_t(p) * K);
#pragma unroll
        for (int q = 0; q < K / 4; ++q)
        {
            const float4 v = w4[q];
            wk[4 * q] = v.x, wk[4 * q + 1] = v.y, wk[4 * q + 2] = v.z, wk[4 * q + 3] = v.w;
        }
    }
    if constexpr (FUSED)
    {
        // the cost partials of the tile's 64-pixel groups, as the term kernel writes them
        if (tid < uint32_t(TP))
        {
            const double wc = wave_sum(cost);
            if ((tid & 63u) == 0 && p0 + tid < uint32_t(d.P))
                d.cost_part[(p0 + tid) / 64u] = wc;
        }
    }
    const uint32_t s0 = d.tile_sbase[t], ns = d.tile_sbase[t + 1] - s0;
    const uint32_t c0 = d.tile_cbase[t], nc = d.tile_cbase[t + 1] - c0;
    for (uint32_t i = tid; i < kNB; i += NT)
        hist[i] = 0u;
    if (tid < np)
    {
#pragma unroll
        for (int v = 0; v < kTileRec / 4; ++v)
            Hs[tid * (kTileRec / 4) + v] = make_float4(hg[4 * v], hg[4 * v + 1], hg[4 * v + 2], hg[4 * v + 3]);
#pragma unroll
        for (int q = 0; q < K / 4; ++q)
            reinterpret_cast<float4 *>(wS)[tid * (K / 4) + q] =
                make_float4(wk[4 * q], wk[4 * q + 1], wk[4 * q + 2], wk[4 * q + 3]);
    }
    if (ns > slot_cap)
        return; // never taken: slot_cap is the largest slot count of the launch's tiles
    // the contribution list (at most kEnt entries) and the slot starts, staged
    // by the upper TP threads (no record of their own) while the lower TP
    // build or read the records: every load issued before the first LDS
    // store, one round trip for all of them
    if (tid >= uint32_t(TP))
    {
        // four entries a 64-bit word: word w of the list takes its entries
        // from the aligned source words w and w + 1 when the list starts off a
        // word boundary (h entries in); the second is read only when it holds
        // an entry of the list, and the words hold no entry past nc - 1 that
        // is ever read back
        static_assert(kEnt % 4 == 0, "whole words");
        constexpr int kPerW = (kEnt / 4 + TP - 1) / TP; // w < nw <= kEnt / 4 bounds each
        const uint32_t u = tid - uint32_t(TP);
        const uint32_t h = c0 & 3u, nw = (nc + 3u) / 4u;
        const uint64_t *__restrict__ sw = reinterpret_cast<const uint64_t *>(d.tc_contrib + (c0 - h));
        uint64_t cv[kPerW];
#pragma unroll
        for (int j = 0; j < kPerW; ++j)
        {
            const uint32_t w = u + uint32_t(j) * TP;
            uint64_t lo = 0, hi = 0;
            if (w < nw)
                lo = sw[w];
            if (w < nw && h != 0 && 4u * w + 4u - h < nc)
                hi = sw[w + 1];
            cv[j] = h == 0 ? lo : (lo >> (16u * h)) | (hi << (64u - 16u * h));
        }
        uint32_t sv = 0;
        if (u <= ns)
            sv = d.tc_slot_first[s0 + u];
        uint64_t *__restrict__ ew = reinterpret_cast<uint64_t *>(ent);
#pragma unroll
        for (int j = 0; j < kPerW; ++j)
        {
            const uint32_t w = u + uint32_t(j) * TP;
            if (w < nw)
                ew[w] = cv[j];
        }
        if (u <= ns)
            sst[u] = uint16_t(sv - c0);
        for (uint32_t s = u + TP; s <= ns; s += TP)
            sst[s] = uint16_t(d.tc_slot_first[s0 + s] - c0);
    }
    __syncthreads();
    // visiting order: the diagonal slots (ka == kb, the only ones whose g
    // part is used) first, then the off-diagonal ones, each longest list
    // first -- a counting sort of (kind, length)
    auto bucket = [&](uint32_t s) {
        const uint32_t c = ent[sst[s]];
        const uint32_t off = c_ka(c) != c_kb(c) ? uint32_t(TP) + 1u : 0u;
        return off + uint32_t(TP) - uint32_t(sst[s + 1] - sst[s]);
    };
    for (uint32_t s = tid; s < ns; s += NT)
        atomicAdd(&hist[bucket(s)], 1u);
    __syncthreads();
    if (tid < 64)
    {
        // exclusive scan of hist[0 .. kNB): kPer consecutive buckets per lane
        constexpr uint32_t kPer = (kNB + 63) / 64;
        const uint32_t b0 = kPer * tid;
        uint32_t h[kPer], v = 0;
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j)
        {
            h[j] = b0 + j < kNB ? hist[b0 + j] : 0u;
            v += h[j];
        }
        const uint32_t own = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t u = __shfl_up(v, o, 64);
            if (int(tid) >= o)
                v += u;
        }
        uint32_t ex = v - own;
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j)
            if (b0 + j < kNB)
            {
                hist[b0 + j] = ex;
                ex += h[j];
            }
    }
    __syncthreads();
    for (uint32_t s = tid; s < ns; s += NT)
        perm[atomicAdd(&hist[bucket(s)], 1u)] = uint16_t(s);
    __syncthreads();

    // Each thread sums whole slots, all of a slot's float4 groups at once, over
    // the slot's contributions in list (pixel) order: the contribution's entry
    // and weights are read once for all groups.  An off-diagonal slot needs
    // H only (groups 0..5: H[20] shares group 5 with g[0..2]); group 6 (g[3..5],
    // pad) is summed and stored for diagonal slots only, and a wave with no
    // diagonal slot skips it altogether.  A slot's partial goes to its list
    // position, so a block's partials are contiguous for the combine.
    float4 *__restrict__ out = reinterpret_cast<float4 *>(d.tpart);
    auto sum_slots = [&](uint32_t k, auto n_groups) {
        constexpr int NG = decltype(n_groups)::value;
        const uint32_t s = perm[k];
        const uint32_t dst = d.slot_pos[s0 + s];
        float4 acc[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g)
            acc[g] = make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t e0 = sst[s], e1 = sst[s + 1];
        // two entries a step: both records are read before the first one's
        // products, so the second read's latency hides under them (the sums
        // stay in list order)
        const uint32_t cfirst = ent[e0];
        const bool diag = c_ka(cfirst) == c_kb(cfirst);
        auto fetch = [&](uint32_t cc, float &coef, float4 (&v)[NG]) {
            const uint32_t pl = c_pix(cc), ka = c_ka(cc), kb = c_kb(cc);
            coef = wS[pl * K + ka] * wS[pl * K + kb];
            const float4 *__restrict__ h = Hs + pl * (kTileRec / 4);
#pragma unroll
            for (int g = 0; g < NG; ++g)
                v[g] = h[g];
        };
        auto madd = [&](float coef, const float4 (&v)[NG]) {
#pragma unroll
            for (int g = 0; g < NG; ++g)
            {
                acc[g].x = fmaf(coef, v[g].x, acc[g].x);
                acc[g].y = fmaf(coef, v[g].y, acc[g].y);
                acc[g].z = fmaf(coef, v[g].z, acc[g].z);
                acc[g].w = fmaf(coef, v[g].w, acc[g].w);
            }
        };
        // the next step's two list entries are read a step ahead
        uint32_t e = e0;
        uint32_t xa = cfirst, xb = ent[min(e0 + 1, e1 - 1)];
        for (; e + 1 < e1; e += 2)
        {
            const uint32_t na = ent[min(e + 2, e1 - 1)], nb = ent[min(e + 3, e1 - 1)];
            float ca, cb;
            float4 va[NG], vb[NG];
            fetch(xa, ca, va);
            fetch(xb, cb, vb);
            madd(ca, va);
            madd(cb, vb);
            xa = na;
            xb = nb;
        }
        if (e < e1)
        {
            float ca;
            float4 va[NG];
            fetch(xa, ca, va);
            madd(ca, va);
        }
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < kTileRec / 4 - 1 || diag)
                out[size_t(dst) * (kPartStride / 4) + g] = acc[g];
    };
    for (uint32_t k = tid; k < ns; k += NT)
    {
        const uint32_t c = ent[sst[perm[k]]];
        const bool diag = c_ka(c) == c_kb(c);
        if (__any(diag))
            sum_slots(k, std::integral_constant<int, kTileRec / 4>{});
        else
            sum_slots(k, std::integral_constant<int, kTileRec / 4 - 1>{});
    }
}

// ARAP entries of the x-side row [-[p]x, I] (see reg_elem in the row kernels).
__device__ __forceinline__ float reg_elem_t(f3 p, int c, int q)
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

__device__ __forceinline__ void sym_pair_t(int e, int &a, int &b)
{
    a = (e >= 6) + (e >= 11) + (e >= 15) + (e >= 18) + (e >= 20);
    b = e - (a * 6 - a * (a - 1) / 2) + a;
}

// One ARAP or node-translation term of a node row, as the combine kernels
// use it: the features every entry of its share of the row is made of.  With
// p this node's warped point, o the other node's and r = p - o, the ARAP
// Jacobian rows of this node are J = [-[p]x, I] (times the side's sign, which
// cancels in every product below; node_graph_reg_jacobian.cuh:250-275), so
//   diagonal block  J^T J:       [(p.p) I - p p^T,  [p]x ; [p]x^T, I]
//   cross block     -J^T J_o:    [-(p.o) I + o p^T, [p]x^T ; [o]x, -I]
//   g               J^T r:       [p x r ; r]
// and each entry is s0 + s1 f[i1] + s2 f[i2] for per-entry constants.  A node
// translation term (node_motion_jacobian.cuh:7-35) is the same with p = o = 0
// and r = t - pred: diagonal I on the translation part, g = [0 ; t - pred].
constexpr int kRegF = 32;
// feature offsets: p.p, p.o, p_a p_b (9), o_a p_b (9), p (3), o (3), p x r (3), r (3)
constexpr int kFpp = 0, kFpo = 1, kFpp9 = 2, kFop9 = 11, kFp = 20, kFo = 23, kFpxr = 26, kFr = 29;
struct RegTerm
{
    float f[kRegF];
    float ws; // lambda^2 cw^2 (0: invalid edge); translation: lambda^2 c^2
    int col;  // block column of the other node, -1 for none (translation term)
};
constexpr uint32_t kTermSub = 32;     // terms staged in LDS at a time (work-group form)
constexpr uint32_t kTermSubWave = 16; // the same in the wave form (less LDS per row; same sums in the same order)

// What a term's features are made of, as loaded from its records.
struct RawTerm
{
    float4 u0, u1;   // ARAP: the edge record (Ti_xj, Tj_xj, weight, valid); translation: t and pred
    uint32_t other;  // ARAP: the edge's other node (the row node itself for none); translation: ~0u
    bool is_x;       // ARAP: the row node is the edge's x side
};

// The loads of ARAP edge or node-translation term t of row i's list
// (node_graph_reg_jacobian.cuh:166-358, node_motion_jacobian.cuh:7-35).
__device__ __forceinline__ RawTerm load_reg_term(const SolveDesc &d, uint32_t i, uint32_t t)
{
    RawTerm w{};
    const uint32_t P = d.P, E = d.E;
    if (t < P + E)
    {
        const uint32_t e = t - P;
        const float4 *rec = reinterpret_cast<const float4 *>(d.reg + size_t(e) * kReg);
        w.u0 = rec[0];
        w.u1 = rec[1];
        const uint32_t ex = d.edges[3 * e], ey = d.edges[3 * e + 1];
        w.is_x = (ex == i);
        w.other = w.is_x ? ey : ex;
    }
    else
    {
        const uint32_t m = t - P - E;
        w.u0 = d.ttrans[m];
        w.u1 = d.node_pred ? d.node_pred[m] : make_float4(0.f, 0.f, 0.f, 0.f);
        w.other = ~0u;
    }
    return w;
}

// The features of a loaded term of row i; col_of(j) gives node j's block
// position in row i, for an edge's cross block.
template <class ColOf>
__device__ __forceinline__ RegTerm make_reg_term(const SolveDesc &d, uint32_t i, const RawTerm &w, ColOf col_of)
{
    RegTerm r{};
    float pm[3] = {0.f, 0.f, 0.f}, po[3] = {0.f, 0.f, 0.f}, res[3];
    if (w.other != ~0u)
    {
        const float pa[3] = {w.u0.x, w.u0.y, w.u0.z}, pb[3] = {w.u0.w, w.u1.x, w.u1.y};
#pragma unroll
        for (int q = 0; q < 3; ++q)
        {
            pm[q] = w.is_x ? pa[q] : pb[q];
            po[q] = w.is_x ? pb[q] : pa[q];
            res[q] = pm[q] - po[q];
        }
        r.ws = (w.u1.w == 0.f) ? 0.f : d.lam2_reg * (w.u1.z * w.u1.z);
        r.col = w.other != i ? col_of(w.other) : -1; // every ARAP edge of the row is a block of it
    }
    else
    {
        const float4 tt = w.u0, pred = w.u1;
        r.ws = d.lam2_nt * (pred.w * pred.w);
        res[0] = tt.x - pred.x;
        res[1] = tt.y - pred.y;
        res[2] = tt.z - pred.z;
        r.col = -1;
    }
    r.f[kFpp] = pm[0] * pm[0] + pm[1] * pm[1] + pm[2] * pm[2];
    r.f[kFpo] = pm[0] * po[0] + pm[1] * po[1] + pm[2] * po[2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
#pragma unroll
        for (int b = 0; b < 3; ++b)
        {
            r.f[kFpp9 + 3 * a + b] = pm[a] * pm[b];
            r.f[kFop9 + 3 * a + b] = po[a] * pm[b];
        }
        r.f[kFp + a] = pm[a];
        r.f[kFo + a] = po[a];
        r.f[kFr + a] = res[a];
    }
    r.f[kFpxr + 0] = pm[1] * res[2] - pm[2] * res[1];
    r.f[kFpxr + 1] = pm[2] * res[0] - pm[0] * res[2];
    r.f[kFpxr + 2] = pm[0] * res[1] - pm[1] * res[0];
    return r;
}

// An entry of a term's share as s0 + s1 f[i1] + s2 f[i2].
struct EntryForm
{
    int i1, i2;
    float s0, s1, s2;
};

// [v]x[rr][cc] (v at feature offset off) as a signed feature: (index, sign)
__device__ __forceinline__ void skew_feature(int off, int rr, int cc, int &idx, float &sgn)
{
    // [v]x = [[0, -v2, v1], [v2, 0, -v0], [-v1, v0, 0]]
    const int k = 3 - rr - cc; // the component off the diagonal
    idx = off + (rr == cc ? 0 : k);
    sgn = rr == cc ? 0.f : (((rr + 1) % 3 == cc) ? -1.f : 1.f);
}

// Row entry ent (0..35 the diagonal block, 36..71 the cross block, 72..77 g)
// of a term's share, in the form above.
__device__ __forceinline__ EntryForm entry_form(int ent)
{
    EntryForm f{0, 0, 0.f, 0.f, 0.f};
    if (ent < 72)
    {
        const bool cross = ent >= 36;
        const int ab = cross ? ent - 36 : ent;
        const int a = ab / 6, b = ab % 6;
        if (a < 3 && b < 3)
        {
            // (p.p) d_ab - p_a p_b, or -(p.o) d_ab + o_a p_b
            f.i1 = cross ? kFpo : kFpp;
            f.s1 = a == b ? (cross ? -1.f : 1.f) : 0.f;
            f.i2 = (cross ? kFop9 : kFpp9) + 3 * a + b;
            f.s2 = cross ? 1.f : -1.f;
        }
        else if (a < 3)
            skew_feature(kFp, cross ? b - 3 : a, cross ? a : b - 3, f.i1, f.s1); // [p]x^T or [p]x
        else if (b < 3)
        {
            if (cross)
                skew_feature(kFo, a - 3, b, f.i1, f.s1); // [o]x
            else
                skew_feature(kFp, b, a - 3, f.i1, f.s1); // [p]x^T
        }
        else
            f.s0 = a == b ? (cross ? -1.f : 1.f) : 0.f;
    }
    else
    {
        const int q = ent - 72;
        f.i1 = q < 3 ? kFpxr + q : kFr + q - 3;
        f.s1 = 1.f;
    }
    return f;
}

// Term r's share of row entry ent (form ef): the diagonal and g shares are
// added to diag and gq, the cross share straight into the edge's block of
// regblk (entry ent - 36 of every block belongs to one lane or thread, so the
// adds do not collide).
__device__ __forceinline__ void reg_term_entry(const RegTerm &r, int ent, const EntryForm &ef, float &diag,
                                               float &gq, float *__restrict__ regblk)
{
    if (r.ws == 0.f)
        return;
    const float v = fmaf(ef.s2, r.f[ef.i2], fmaf(ef.s1, r.f[ef.i1], ef.s0));
    if (ent < 36)
        diag += r.ws * v;
    else if (ent < 72)
    {
        if (r.col >= 0)
            regblk[r.col * 36 + ent - 36] += r.ws * v;
    }
    else
        gq += r.ws * v;
}

// Adds the terms [beg, end) of row i's list into regblk (diagonal block at
// column 0, cross blocks at colmap[other]) and g[0..5]: in rounds of 256
// terms, every entry summed over the round's terms in list order and the
// round's sum then added; the terms are staged kTermSub at a time.  Must be
// reached by all 256 threads.
__device__ void row_reg_terms(const SolveDesc &d, uint32_t i, uint32_t beg, uint32_t end,
                              const uint16_t *__restrict__ colmap, float *__restrict__ regblk, float *__restrict__ g,
                              RegTerm *__restrict__ rt)
{
    const uint32_t tid = threadIdx.x;
    const int ent = int(tid);
    const EntryForm ef = entry_form(ent < 78 ? ent : 0);
    for (uint32_t r0 = beg; r0 < end; r0 += kTileThreads)
    {
        const uint32_t r1 = min(r0 + uint32_t(kTileThreads), end);
        // threads 0..35: diagonal entries, 36..71: cross entries, 72..77: g
        float diag = 0.f, gq = 0.f;
        for (uint32_t s0 = r0; s0 < r1; s0 += kTermSub)
        {
            const uint32_t n = min(kTermSub, r1 - s0);
            if (tid < n)
                rt[tid] = make_reg_term(d, i, load_reg_term(d, i, d.n2t_terms[s0 + tid]),
                                        [&](uint32_t j) { return int(colmap[j]); });
            __syncthreads();
            if (tid < 78)
                for (uint32_t k = 0; k < n; ++k)
                    reg_term_entry(rt[k], ent, ef, diag, gq, regblk);
            __syncthreads();
        }
        if (ent < 36)
            regblk[ent] += diag;
        else if (ent >= 72 && ent < 78)
            g[ent - 72] += gq;
        __syncthreads();
    }
}


// List entries fetched per round of dependent loads (matrix-free row kernel).
constexpr int kBatch = 8;

// The combine cuts every block's slot list into chunks of kChunk entries; a
// chunk's sum (list order) is one task, and a block's sum adds its chunk sums
// in chunk order, so the result depends on the list alone.  Chunk sums are
// kept for kChunkCap chunks at a time (a window), in LDS that the ARAP term
// parameters use before them.
constexpr uint32_t kChunk = 16;
constexpr uint32_t kChunkCap = 128;
constexpr uint32_t kCombUnionF4 = kChunkCap * (kTileRec / 4); // float4s of the shared region
static_assert(sizeof(RegTerm) * kTermSub <= kCombUnionF4 * sizeof(float4), "combine LDS union");


// Sum of float4 group g over list positions [l0, l1), at most kChunk of them,
// in list order; k_tiles stored each slot's partial at its list position, so
// they load in one round.
__device__ __forceinline__ float4 sum_chunk4(const float4 *__restrict__ part, uint32_t l0, uint32_t l1, uint32_t g)
{
    float4 v[kChunk];
#pragma unroll
    for (uint32_t j = 0; j < kChunk; ++j)
        v[j] = (l0 + j < l1) ? part[size_t(l0 + j) * (kPartStride / 4) + g] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (uint32_t j = 0; j < kChunk; ++j)
        if (l0 + j < l1)
        {
            acc.x += v[j].x;
            acc.y += v[j].y;
            acc.z += v[j].z;
            acc.w += v[j].w;
        }
    return acc;
}

// Orders the LDS accesses of one wave's lanes.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// invert6_fp64 (Gauss-Jordan with partial pivoting, FP64) run by one wave on
// an LDS copy of the augmented 6 x 12 matrix: lane l owns entries l and l + 64.
// The pivots and the elimination order are those of invert6_fp64.
__device__ void invert6_fp64_wave(const float *__restrict__ A, double (*gm)[12], double *__restrict__ out)
{
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h)
    {
        const uint32_t e = lane + 64u * h, r = e / 12u, k = e % 12u;
        if (e < 72u)
            gm[r][k] = k < 6u ? double(A[6u * r + k]) : (k - 6u == r ? 1.0 : 0.0);
    }
    wave_lds_sync();
    for (int i = 0; i < 6; ++i)
    {
        int piv = i;
        double pv = fabs(gm[i][i]);
        for (int j = i + 1; j < 6; ++j)
            if (pv < fabs(gm[j][i]))
            {
                pv = fabs(gm[j][i]);
                piv = j;
            }
        if (piv != i) // uniform over the wave
        {
            double a = 0.0, b = 0.0;
            if (lane < 12u)
            {
                a = gm[i][lane];
                b = gm[piv][lane];
            }
            wave_lds_sync();
            if (lane < 12u)
            {
                gm[i][lane] = b;
                gm[piv][lane] = a;
            }
            wave_lds_sync();
        }
        const double inv = 1.0 / gm[i][i];
        wave_lds_sync();
        if (lane < 12u)
            gm[i][lane] *= inv;
        wave_lds_sync();
        double f[2], p[2];
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h)
        {
            const uint32_t e = lane + 64u * h, r = e / 12u, k = e % 12u;
            f[h] = (e < 72u && int(r) != i) ? gm[r][i] : 0.0;
            p[h] = e < 72u ? gm[i][k] : 0.0;
        }
        wave_lds_sync();
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h)
        {
            const uint32_t e = lane + 64u * h, r = e / 12u, k = e % 12u;
            if (e < 72u && int(r) != i)
                gm[r][k] -= f[h] * p[h];
        }
        wave_lds_sync();
    }
    if (lane < 36u)
        out[lane] = gm[lane / 6u][6u + lane % 6u];
}

// One work-group per node row: every block's slot partials summed (chunks in
// list order, then the chunk sums in order), the ARAP terms
// (node_graph_reg_jacobian.cuh:166-358) and the node-translation term
// (node_motion_jacobian.cuh:7-35) of the row added in list order, the damping
// on the diagonal; full 6x6 blocks and rhs_i = -(g of the diagonal slots +
// ARAP/translation g).  The diagonal block and the upper blocks (columns above
// i) are summed; each upper block (i, j) is also written transposed into
// (j, i), so every block's slot partials are read once.
__global__ __launch_bounds__(kTileThreads) void k_tiles_combine_wg(const SolveDesc *__restrict__ sds,
                                                                const uint32_t *__restrict__ wg_begin, int n_solve,
                                                                PcgSlot2 *__restrict__ slots2, uint32_t n_slots2)
{
    extern __shared__ float4 smem4[];
    __shared__ float dg[36];       // the diagonal block, for its inverse
    __shared__ double gm[6][12];   // its augmented matrix during the inversion
    // a pipelined PCG that follows publishes into zeroed slots (the row map
    // has a work-group for every one of its groups)
    if (slots2 && blockIdx.x < n_slots2 && threadIdx.x < 4)
        slots2[blockIdx.x].v[threadIdx.x >> 1][threadIdx.x & 1] = 0.0;
    const int si = find_solve_t(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t i = blockIdx.x - wg_begin[si];
    if (i >= uint32_t(d.N))
        return; // uniform over the group
    const uint32_t tid = threadIdx.x;
    const uint32_t base = d.row_ptr[i] + i;
    const uint32_t L1 = d.row_ptr[i + 1] - d.row_ptr[i] + 1;
    constexpr uint32_t kG = kTileRec / 4;
    // LDS: union{chunk sums csum[kChunkCap][7] | term parameters rt[256]} |
    // run[L1][7] | regblk[L1][36] | red[16] | ls[L1] | le[L1] | cc[L1+1] | colmap u16[N]
    float4 *csum = smem4;
    RegTerm *rt = reinterpret_cast<RegTerm *>(smem4);
    float4 *run = smem4 + kCombUnionF4;
    float *regblk = reinterpret_cast<float *>(run + size_t(L1) * kG);
    float *red = regblk + size_t(L1) * 36;
    uint32_t *ls = reinterpret_cast<uint32_t *>(red + 16);
    uint32_t *le = ls + L1;
    uint32_t *cc = le + L1;
    uint16_t *colmap = reinterpret_cast<uint16_t *>(cc + L1 + 1);

    // columns (and how many lie below i: the upper blocks start after them)
    uint32_t n_below = 0;
    for (uint32_t t0 = 0; t0 < L1; t0 += kTileThreads)
    {
        const uint32_t t = t0 + tid;
        bool below = false;
        if (t < L1)
        {
            const uint32_t col = d.blk_col[base + t];
            colmap[col] = uint16_t(t);
            below = t > 0 && col < i;
        }
        n_below += uint32_t(__syncthreads_count(below));
    }
    const uint32_t c_up = 1 + n_below; // blocks c_up .. L1-1 have columns above i
    const uint32_t nq = 1 + (L1 - c_up); // summed blocks: q = 0 the diagonal, q >= 1 block c_up + q - 1
    for (uint32_t t = tid; t < L1 * 36; t += kTileThreads)
        regblk[t] = 0.f;
    for (uint32_t t = tid; t < nq * kG; t += kTileThreads)
        run[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid < 16)
        red[tid] = 0.f;
    for (uint32_t q = tid; q < nq; q += kTileThreads)
    {
        const uint32_t c = q == 0 ? 0u : c_up + q - 1;
        ls[q] = d.list_off[base + c];
        le[q] = d.list_off[base + c + 1];
    }
    __syncthreads();
    // chunk offsets cc[q] (exclusive prefix of the chunk counts), by wave 0
    if (tid < 64)
    {
        uint32_t carry = 0;
        for (uint32_t q0 = 0; q0 < nq; q0 += 64)
        {
            const uint32_t q = q0 + tid;
            const uint32_t n = q < nq ? (le[q] - ls[q] + kChunk - 1) / kChunk : 0u;
            uint32_t v = n;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1)
            {
                const uint32_t u = __shfl_up(v, o, 64);
                if (int(tid) >= o)
                    v += u;
            }
            if (q < nq)
                cc[q] = carry + v - n;
            carry += __shfl(v, 63, 64);
        }
        if (tid == 0)
            cc[nq] = carry;
    }

    const uint32_t dend = d.n2t_off[i] + d.dense_cnt[i];
    const uint32_t end = d.n2t_off[i + 1];
    // ARAP and node-translation terms of the row: parameters loaded one term
    // per thread, then every entry summed over the terms in list order (ends
    // with a group barrier, after which the shared region holds chunk sums)
    row_reg_terms(d, i, dend, end, colmap, regblk, red + 8, rt);
    __syncthreads();

    const float4 *__restrict__ part = reinterpret_cast<const float4 *>(d.tpart);
    const uint32_t total = cc[nq];
    for (uint32_t w0 = 0; w0 < total; w0 += kChunkCap)
    {
        const uint32_t w1 = min(total, w0 + kChunkCap);
        for (uint32_t task = tid; task < (w1 - w0) * kG; task += kTileThreads)
        {
            const uint32_t k = w0 + task / kG, g = task % kG;
            uint32_t lo = 0, hi = nq; // the block of chunk k: last q with cc[q] <= k
            while (hi - lo > 1)
            {
                const uint32_t m = (lo + hi) >> 1;
                if (cc[m] <= k)
                    lo = m;
                else
                    hi = m;
            }
            const uint32_t l0 = ls[lo] + (k - cc[lo]) * kChunk;
            // group 6 (g[3..5], pad) is written for diagonal slots only
            csum[task] = (lo > 0 && g == kG - 1) ? make_float4(0.f, 0.f, 0.f, 0.f)
                                                 : sum_chunk4(part, l0, min(l0 + kChunk, le[lo]), g);
        }
        __syncthreads();
        for (uint32_t task = tid; task < nq * kG; task += kTileThreads)
        {
            const uint32_t q = task / kG, g = task % kG;
            const uint32_t k0 = max(cc[q], w0), k1 = min(cc[q + 1], w1);
            float4 a = run[task];
            for (uint32_t k = k0; k < k1; ++k)
            {
                const float4 v = csum[(k - w0) * kG + g];
                a.x += v.x;
                a.y += v.y;
                a.z += v.z;
                a.w += v.w;
            }
            run[task] = a;
        }
        __syncthreads();
    }

    const size_t out0 = size_t(base) * 36;
    // stores block q's float4 group g of sums, plus ARAP and damping, and its
    // transpose; the diagonal block also goes to dg and its g part to red
    auto store_task = [&](uint32_t task) {
        const uint32_t q = task / kG, g = task % kG;
        const uint32_t c = q == 0 ? 0u : c_up + q - 1;
        const size_t tr = size_t(d.blk_transpose[base + c]) * 36; // (j, i) for c > 0
        const float4 acc = run[task];
        const float av[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
            const int e = int(g) * 4 + r;
            if (e < 21)
            {
                int a, b;
                sym_pair_t(e, a, b);
                float u = av[r] + regblk[c * 36 + a * 6 + b];
                if (c == 0 && a == b)
                    u += d.damping;
                d.blk[out0 + size_t(c) * 36 + a * 6 + b] = u;
                const float v = av[r] + regblk[c * 36 + b * 6 + a];
                if (a != b)
                    d.blk[out0 + size_t(c) * 36 + b * 6 + a] = v;
                if (c == 0)
                {
                    dg[a * 6 + b] = u;
                    dg[b * 6 + a] = a != b ? v : u;
                }
                if (c > 0)
                {
                    // the transposed block: (j, i)[b][a] = (i, j)[a][b]
                    d.blk[tr + b * 6 + a] = u;
                    if (a != b)
                        d.blk[tr + a * 6 + b] = v;
                }
            }
            else if (c == 0 && e < 27)
                red[e - 21] = av[r]; // g of the diagonal slots
        }
    };
    if (tid < 64)
    {
        // wave 0: the diagonal block, rhs, and the FP64 inverse of the block
        // (the same float values the block store holds), while the other
        // waves store the off-diagonal blocks
        if (tid < kG)
            store_task(tid);
        wave_lds_sync();
        if (tid < 6)
            d.rhs[size_t(i) * 6 + tid] = -(red[tid] + red[8 + tid]);
        invert6_fp64_wave(dg, gm, d.dinv + size_t(i) * 36);
    }
    else
    {
        for (uint32_t task = kG + tid - 64; task < nq * kG; task += kTileThreads - 64)
            store_task(task);
        // the row's zeroed pipelined-PCG vectors
        if (tid >= 128 && tid < 140)
            d.mv[(tid - 128) / 6][size_t(i) * 6 + (tid - 128) % 6] = 0.0;
    }
}


// Block-group tasks a lane of the wave combine holds in registers: rows of up
// to 64 * kWaveTasks / 7 summed blocks (the engine takes the work-group form
// for longer rows).
constexpr uint32_t kWaveTasks = 8;
static_assert(kWaveCombineMaxBlocks * (kTileRec / 4) <= 64 * kWaveTasks, "wave combine tasks");

// One wave per node row: the same sums as k_tiles_combine_wg, in the same
// order and therefore to the same bits -- every block's slot partials summed
// (chunks in list order, then the chunk sums in order), the ARAP terms
// (node_graph_reg_jacobian.cuh:166-358) and the node-translation term
// (node_motion_jacobian.cuh:7-35) of the row added in list order, rounds of
// 256 terms, the damping on the diagonal; full 6x6 blocks, rhs and the FP64
// inverse of the diagonal block.  A row is a few hundred values of work, so
// one wave holds it whole: its steps are ordered by wave-level LDS syncs, with
// no work-group barrier, and many rows share a CU.  For launches with many
// rows (batches, the server's steps); a single solve's few rows finish sooner
// with the work-group form.
__global__ __launch_bounds__(64) void k_tiles_combine_wave(const SolveDesc *__restrict__ sds,
                                                           const uint32_t *__restrict__ wg_begin, int n_solve,
                                                           PcgSlot2 *__restrict__ slots2, uint32_t n_slots2)
{
    extern __shared__ float4 smem4[];
    __shared__ float dg[36];     // the diagonal block, for its inverse
    __shared__ double gm[6][12]; // its augmented matrix during the inversion
    __shared__ float red[16];    // g of the diagonal slots (0..5), of the terms (8..13)
    __shared__ RegTerm rt[kTermSubWave]; // one sub-round of ARAP / node-translation terms
    // a pipelined PCG that follows publishes into zeroed slots (the row map
    // has a work-group for every one of its groups)
    if (slots2 && blockIdx.x < n_slots2 && threadIdx.x < 4)
        slots2[blockIdx.x].v[threadIdx.x >> 1][threadIdx.x & 1] = 0.0;
    const int si = find_solve_t(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t i = blockIdx.x - wg_begin[si];
    if (i >= uint32_t(d.N))
        return; // uniform over the wave
    const uint32_t lane = threadIdx.x;
    const uint32_t base = d.row_ptr[i] + i;
    const uint32_t L1 = d.row_ptr[i + 1] - d.row_ptr[i] + 1;
    constexpr uint32_t kG = kTileRec / 4;
    // LDS: regblk[L1][36] | cols[L1] | ls[L1] | le[L1] | trb[L1]; the block
    // sums stay in registers (a lane sums and stores the same tasks)
    float *regblk = reinterpret_cast<float *>(smem4);
    float4 run[kWaveTasks];
    uint32_t *cols = reinterpret_cast<uint32_t *>(regblk + size_t(L1) * 36);
    uint32_t *ls = cols + L1;
    uint32_t *le = ls + L1;
    uint32_t *trb = le + L1;

    // columns: the diagonal first, then the ascending neighbours; how many lie
    // below i (the upper blocks start after them)
    uint32_t n_below = 0;
    for (uint32_t c0 = 0; c0 < L1; c0 += 64)
    {
        const uint32_t c = c0 + lane;
        bool below = false;
        if (c < L1)
        {
            const uint32_t col = d.blk_col[base + c];
            cols[c] = col;
            below = c > 0 && col < i;
        }
        n_below += uint32_t(__popcll(__ballot(below)));
    }
    const uint32_t c_up = 1 + n_below; // blocks c_up .. L1-1 have columns above i
    const uint32_t nq = 1 + (L1 - c_up); // summed blocks: q = 0 the diagonal, q >= 1 block c_up + q - 1
    for (uint32_t t = lane; t < L1 * 36; t += 64)
        regblk[t] = 0.f;
    if (lane < 16)
        red[lane] = 0.f;
    for (uint32_t q = lane; q < nq; q += 64)
    {
        const uint32_t c = q == 0 ? 0u : c_up + q - 1;
        ls[q] = d.list_off[base + c];
        le[q] = d.list_off[base + c + 1];
        trb[q] = d.blk_transpose[base + c];
    }
    wave_lds_sync();

    // every summed block's float4 group g: its slot partials in chunks of
    // kChunk (list order inside a chunk), the chunk sums added in order -- a
    // task per (block, group), its chunks in sequence (a batch launch has rows
    // enough to hide the latency of the sequence)
    const float4 *__restrict__ part = reinterpret_cast<const float4 *>(d.tpart);
#pragma unroll
    for (uint32_t k = 0; k < kWaveTasks; ++k)
    {
        const uint32_t task = lane + 64u * k;
        run[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (task >= nq * kG)
            continue;
        const uint32_t q = task / kG, g = task % kG;
        // group 6 (g[3..5], pad) is written for diagonal slots only
        const uint32_t l1 = (q > 0 && g == kG - 1) ? ls[q] : le[q];
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        for (uint32_t l0 = ls[q]; l0 < l1; l0 += kChunk)
        {
            const float4 v = sum_chunk4(part, l0, min(l0 + kChunk, l1), g);
            a.x += v.x;
            a.y += v.y;
            a.z += v.z;
            a.w += v.w;
        }
        run[k] = a;
    }

    // ARAP and node-translation terms of the row (after its dense terms in
    // the node's term list), in rounds of 256 as in the work-group form: every
    // entry summed over a round's terms in list order, the round's sum then
    // added to the row.  The terms are staged kTermSubWave at a time (one per
    // lane); a lane's two entry sums carry over the sub-rounds of a round.
    const uint32_t dend = d.n2t_off[i] + d.dense_cnt[i];
    const uint32_t tend = d.n2t_off[i + 1];
    const EntryForm ef[2] = {entry_form(int(lane)), entry_form(int(lane + 64u) < 78 ? int(lane + 64u) : 0)};
    for (uint32_t r0 = dend; r0 < tend; r0 += kTileThreads)
    {
        const uint32_t r1 = min(r0 + uint32_t(kTileThreads), tend);
        float diag[2] = {0.f, 0.f}, gq[2] = {0.f, 0.f};
        for (uint32_t s0 = r0; s0 < r1; s0 += kTermSubWave)
        {
            const uint32_t n = min(kTermSubWave, r1 - s0);
            if (lane < n)
                rt[lane] = make_reg_term(d, i, load_reg_term(d, i, d.n2t_terms[s0 + lane]), [&](uint32_t j) {
                    // the ascending neighbours cols[1..L1)
                    uint32_t lo = 1, hi = L1;
                    while (lo < hi)
                    {
                        const uint32_t m = (lo + hi) >> 1;
                        if (cols[m] < j)
                            lo = m + 1;
                        else
                            hi = m;
                    }
                    return int(lo);
                });
            wave_lds_sync();
            // entries 0..35: diagonal, 36..71: cross, 72..77: g (two per lane)
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h)
            {
                const int ent = int(lane + 64u * h);
                if (ent >= 78)
                    continue;
                for (uint32_t k = 0; k < n; ++k)
                    reg_term_entry(rt[k], ent, ef[h], diag[h], gq[h], regblk);
            }
            wave_lds_sync();
        }
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h)
        {
            const int ent = int(lane + 64u * h);
            if (ent < 36)
                regblk[ent] += diag[h];
            else if (ent >= 72 && ent < 78)
                red[8 + ent - 72] += gq[h];
        }
        wave_lds_sync();
    }

    // block q's float4 group g of sums, plus ARAP and damping, and its
    // transpose; the diagonal block also goes to dg and its g part to red
    const size_t out0 = size_t(base) * 36;
#pragma unroll
    for (uint32_t k = 0; k < kWaveTasks; ++k)
    {
        const uint32_t task = lane + 64u * k;
        if (task >= nq * kG)
            continue;
        const uint32_t q = task / kG, g = task % kG;
        const uint32_t c = q == 0 ? 0u : c_up + q - 1;
        const size_t tr = size_t(trb[q]) * 36; // (j, i) for c > 0
        const float4 acc = run[k];
        const float av[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
            const int e = int(g) * 4 + r;
            if (e < 21)
            {
                int a, b;
                sym_pair_t(e, a, b);
                float u = av[r] + regblk[c * 36 + a * 6 + b];
                if (c == 0 && a == b)
                    u += d.damping;
                d.blk[out0 + size_t(c) * 36 + a * 6 + b] = u;
                const float v = av[r] + regblk[c * 36 + b * 6 + a];
                if (a != b)
                    d.blk[out0 + size_t(c) * 36 + b * 6 + a] = v;
                if (c == 0)
                {
                    dg[a * 6 + b] = u;
                    dg[b * 6 + a] = a != b ? v : u;
                }
                if (c > 0)
                {
                    // the transposed block: (j, i)[b][a] = (i, j)[a][b]
                    d.blk[tr + b * 6 + a] = u;
                    if (a != b)
                        d.blk[tr + a * 6 + b] = v;
                }
            }
            else if (c == 0 && e < 27)
                red[e - 21] = av[r]; // g of the diagonal slots
        }
    }
    // the row's zeroed pipelined-PCG vectors
    if (lane >= 32 && lane < 44)
        d.mv[(lane - 32) / 6][size_t(i) * 6 + (lane - 32) % 6] = 0.0;
    wave_lds_sync();
    if (lane < 6)
        d.rhs[size_t(i) * 6 + lane] = -(red[lane] + red[8 + lane]);
    // the FP64 inverse of the diagonal block (the same float values the block store holds)
    invert6_fp64_wave(dg, gm, d.dinv + size_t(i) * 36);
}

// ------------------------------------------------------------ matrix-free J^T W J x
// ApplyJtJ without the blocks (the reference names it in its sanity checker,
// JtJMaterializerSanityCheck.cu:47-79, but has no implementation): per pixel
// z = sum_k w_k x_k over its distinct nodes and u = H z (H = sum_c lam_c^2
// J_c^T J_c, the same record the materialisation reads); each diagonal slot of
// a tile (node i, its pixels in order) sums w_i u; the row kernel adds node
// i's diagonal-slot partials in tile order, the ARAP terms J_i^T (J_i x_i +
// J_j x_j), the node-translation term and the damping.  x is the solve's
// next PCG direction d_{k+1} (dv[(k+1)&1]) unless x_user is given.
template <int K, class JT, int TP>
__global__ __launch_bounds__(2 * TP) void k_mf_tiles(const SolveDesc *__restrict__ sds,
                                                           const uint32_t *__restrict__ wg_begin, int n_solve, int k_it,
                                                           const float *__restrict__ x_user)
{
    using G = TileGeom<K, TP>;
    constexpr int NT = G::kThreads;
    __shared__ float uS[TP * 6];
    __shared__ float wS[TP * K];
    __shared__ uint16_t ent[G::kEnt];
    __shared__ uint16_t sst[G::kEnt + 1];
    const int si = find_solve_t(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t t = blockIdx.x - wg_begin[si];
    if (t >= uint32_t(d.n_tiles))
        return; // uniform over the group
    const float *__restrict__ x = x_user ? x_user : pick2(d.dv, (k_it + 1) & 1);
    const uint32_t tid = threadIdx.x;
    const uint32_t p0 = t * TP;
    const uint32_t np = min(uint32_t(TP), uint32_t(d.P) - p0);
    if (tid < np)
    {
        const uint32_t p = p0 + tid;
        uint16_t kn[K];
        float w[K];
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            kn[k] = d.gpknn[size_t(p) * K + k];
            w[k] = d.gpw[size_t(p) * K + k];
            wS[tid * K + k] = w[k];
        }
        float z[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            bool act = true; // last matching slot of its node
#pragma unroll
            for (int k2 = k + 1; k2 < K; ++k2)
                act = act && kn[k2] != kn[k];
            if (act)
            {
                const float *xv = x + size_t(kn[k]) * 6;
#pragma unroll
                for (int q = 0; q < 6; ++q)
                    z[q] = fmaf(w[k], xv[q], z[q]);
            }
        }
        float hg[kHG];
        load_hg<JT>(d.jr, p, hg);
#pragma unroll
        for (int a = 0; a < 6; ++a)
        {
            float u = 0.f;
#pragma unroll
            for (int b = 0; b < 6; ++b)
            {
                int e;
                if (a <= b)
                    e = a * 6 - a * (a - 1) / 2 + (b - a);
                else
                    e = b * 6 - b * (b - 1) / 2 + (a - b);
                u = fmaf(hg[e], z[b], u);
            }
            uS[tid * 6 + a] = u;
        }
    }
    const uint32_t c0 = d.tile_cbase[t], nc = d.tile_cbase[t + 1] - c0;
    const uint32_t s0 = d.tile_sbase[t], ns = d.tile_sbase[t + 1] - s0;
    for (uint32_t i = tid; i < nc; i += NT)
        ent[i] = d.tc_contrib[c0 + i];
    for (uint32_t s = tid; s <= ns; s += NT)
        sst[s] = uint16_t(d.tc_slot_first[s0 + s] - c0);
    __syncthreads();
    for (uint32_t task = tid; task < ns * 6; task += NT)
    {
        const uint32_t s = task / 6, comp = task % 6;
        const uint32_t e0 = sst[s], e1 = sst[s + 1];
        const uint32_t first = ent[e0];
        if (c_ka(first) != c_kb(first))
            continue; // off-diagonal slot: not needed
        float acc = 0.f;
        for (uint32_t e = e0; e < e1; ++e)
        {
            const uint32_t c = ent[e];
            const uint32_t pl = c_pix(c), ka = c_ka(c);
            acc = fmaf(wS[pl * K + ka], uS[pl * 6 + comp], acc);
        }
        d.tpart[size_t(s0 + s) * kPartStride + comp] = acc;
    }
}

// One wave per node row: y_i = damping x_i + diagonal-slot partials (tile
// order) + ARAP + node translation.  y is d.q unless y_user is given.
__global__ __launch_bounds__(64) void k_mf_rows(const SolveDesc *__restrict__ sds,
                                                const uint32_t *__restrict__ wg_begin, int n_solve, int k_it,
                                                const float *__restrict__ x_user, float *__restrict__ y_user)
{
    const int si = find_solve_t(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t i = blockIdx.x - wg_begin[si];
    if (i >= uint32_t(d.N))
        return;
    const float *__restrict__ x = x_user ? x_user : pick2(d.dv, (k_it + 1) & 1);
    float *__restrict__ y = y_user ? y_user : d.q;
    const int lane = threadIdx.x;
    const int a = lane < 6 ? lane : 0;
    const uint32_t base = d.row_ptr[i] + i;
    float xi[6];
#pragma unroll
    for (int q = 0; q < 6; ++q)
        xi[q] = x[size_t(i) * 6 + q];
    // dense: diagonal-slot partials of block (i, i)
    float dense = 0.f;
    {
        const uint32_t l0 = d.list_off[base], l1 = d.list_off[base + 1];
        for (uint32_t l = l0; l < l1; l += kBatch)
        {
            uint32_t it[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j)
                it[j] = (l + j < l1) ? d.list_items[l + j] : 0u;
            float v[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j)
                v[j] = (l + j < l1) ? d.tpart[size_t(it[j]) * kPartStride + a] : 0.f;
#pragma unroll
            for (int j = 0; j < kBatch; ++j)
                if (l + j < l1)
                    dense += v[j];
        }
    }
    // ARAP and node translation: each lane takes one term of the list (64 at a
    // time), then lanes 0..5 add the terms' contributions in list order
    __shared__ float contrib[64][6];
    float reg = 0.f;
    const uint32_t dend = d.n2t_off[i] + d.dense_cnt[i], end = d.n2t_off[i + 1];
    const uint32_t P = d.P, E = d.E;
    for (uint32_t c0 = dend; c0 < end; c0 += 64)
    {
        const uint32_t n = min(64u, end - c0);
        float yt[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (uint32_t(lane) < n)
        {
            const uint32_t t = d.n2t_terms[c0 + lane];
            if (t < P + E)
            {
                const uint32_t e = t - P;
                const float4 *rec = reinterpret_cast<const float4 *>(d.reg + size_t(e) * kReg);
                const float4 u0 = rec[0], u1 = rec[1];
                if (u1.w != 0.f)
                {
                    const f3 pa = mk3(u0.x, u0.y, u0.z), pb = mk3(u0.w, u1.x, u1.y);
                    const float cw = u1.z;
                    const uint32_t ex = d.edges[3 * e], ey = d.edges[3 * e + 1];
                    const bool is_x = (ex == i);
                    const uint32_t other = is_x ? ey : ex;
                    const float ws = d.lam2_reg * (cw * cw);
                    const f3 pm = is_x ? pa : pb, po = is_x ? pb : pa;
                    const float sm = is_x ? 1.f : -1.f;
                    float xo[6];
#pragma unroll
                    for (int q = 0; q < 6; ++q)
                        xo[q] = x[size_t(other) * 6 + q];
                    float rc[3];
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch)
                    {
                        float v = 0.f; // (J_m x_i + J_o x_o)[ch]
#pragma unroll
                        for (int b = 0; b < 6; ++b)
                            v += (sm * reg_elem_t(pm, ch, b)) * xi[b] + (-sm * reg_elem_t(po, ch, b)) * xo[b];
                        rc[ch] = v;
                    }
#pragma unroll
                    for (int q = 0; q < 6; ++q)
                    {
                        float v = 0.f;
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch)
                            v += (sm * reg_elem_t(pm, ch, q)) * rc[ch];
                        yt[q] = ws * v;
                    }
                }
            }
            else
            {
                const uint32_t m = t - P - E;
                const float4 pred = d.node_pred ? d.node_pred[m] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int q = 3; q < 6; ++q)
                    yt[q] = d.lam2_nt * (pred.w * pred.w) * xi[q];
            }
        }
#pragma unroll
        for (int q = 0; q < 6; ++q)
            contrib[lane][q] = yt[q];
        __syncthreads();
        if (lane < 6)
            for (uint32_t k = 0; k < n; ++k)
                reg += contrib[k][a];
        __syncthreads();
    }
    if (lane < 6)
        y[size_t(i) * 6 + a] = d.damping * xi[a] + dense + reg;
}

// Materialised y = A x over the blocks of each row (one wave per row), for
// sh_solver_apply_jtj.
__global__ __launch_bounds__(64) void k_spmv_rows(const SolveDesc *__restrict__ sds,
                                                  const uint32_t *__restrict__ wg_begin, int n_solve,
                                                  const float *__restrict__ x, float *__restrict__ y)
{
    const int si = find_solve_t(wg_begin, n_solve, blockIdx.x);
    const SolveDesc d = global_desc(sds[si]);
    const uint32_t i = blockIdx.x - wg_begin[si];
    if (i >= uint32_t(d.N))
        return;
    const int lane = threadIdx.x;
    const int a = lane < 6 ? lane : 0;
    float acc = 0.f;
    const uint32_t b0 = d.row_ptr[i] + i, b1 = d.row_ptr[i + 1] + i + 1;
    for (uint32_t bb = b0; bb < b1; ++bb)
    {
        const uint32_t col = d.blk_col[bb];
        const float *A = d.blk + size_t(bb) * 36 + 6 * a;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            acc = fmaf(A[c], x[size_t(col) * 6 + c], acc);
    }
    if (lane < 6)
        y[size_t(i) * 6 + a] = acc;
}

// ------------------------------------------------------------ launchers
// Calls f(std::integral_constant<int, TP>) for the tile size tp (64, 128 or 256).
template <class F>
static hipError_t with_tile_pix(uint32_t tp, F &&f)
{
    switch (tp)
    {
    case 64:
        f(std::integral_constant<int, 64>{});
        break;
    case 128:
        f(std::integral_constant<int, 128>{});
        break;
    case 256:
        f(std::integral_constant<int, 256>{});
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_tile_build(int K, uint32_t tile_pix, const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg,
                             const IndexWs &ws, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    return with_tile_pix(tile_pix, [&](auto tp) {
        constexpr int TP = decltype(tp)::value;
        if (K == 8)
            hipLaunchKernelGGL((k_tile_build<8, TP>), dim3(n_wg), dim3(2 * TP), 0, st, g, wmap, ws);
        else
            hipLaunchKernelGGL((k_tile_build<4, TP>), dim3(n_wg), dim3(2 * TP), 0, st, g, wmap, ws);
    });
}

hipError_t launch_tile_compact(int K, uint32_t tile_pix, const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg,
                               const IndexWs &ws, uint32_t *sort_keys, uint32_t *sort_vals, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    return with_tile_pix(tile_pix, [&](auto tp) {
        constexpr int TP = decltype(tp)::value;
        if (K == 8)
            hipLaunchKernelGGL((k_tile_compact<8, TP>), dim3(n_wg), dim3(kTileThreads), 0, st, g, wmap, ws,
                               sort_keys, sort_vals);
        else
            hipLaunchKernelGGL((k_tile_compact<4, TP>), dim3(n_wg), dim3(kTileThreads), 0, st, g, wmap, ws,
                               sort_keys, sort_vals);
    });
}

hipError_t launch_block_transpose(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, const IndexWs &ws,
                                  hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_block_transpose, dim3(n_wg), dim3(kTileThreads), 0, st, g, wmap, ws);
    return hipGetLastError();
}

hipError_t launch_tiles(int K, bool fp16, uint32_t tile_pix, bool fused, int parity, const SolveDesc *sds,
                        const uint32_t *wmap, int n_solve, uint32_t n_wg, uint32_t slot_cap, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    if (fused && fp16)
        return hipErrorInvalidValue; // FP16 records are streamed, never built in the tile kernel
    // the slot lists in LDS are sized for the launch's largest tile
    const size_t lds = ((size_t(slot_cap) * 2 + 2) * sizeof(uint16_t) + 3) & ~size_t(3);
    return with_tile_pix(tile_pix, [&](auto tp) {
        constexpr int TP = decltype(tp)::value;
        const dim3 grid(n_wg), block(2 * TP);
        if (K == 8 && fused)
            hipLaunchKernelGGL((k_tiles<8, float, TP, true>), grid, block, lds, st, sds, wmap, n_solve, slot_cap,
                               parity);
        else if (K == 8 && fp16)
            hipLaunchKernelGGL((k_tiles<8, __half, TP, false>), grid, block, lds, st, sds, wmap, n_solve, slot_cap,
                               parity);
        else if (K == 8)
            hipLaunchKernelGGL((k_tiles<8, float, TP, false>), grid, block, lds, st, sds, wmap, n_solve, slot_cap,
                               parity);
        else if (fused)
            hipLaunchKernelGGL((k_tiles<4, float, TP, true>), grid, block, lds, st, sds, wmap, n_solve, slot_cap,
                               parity);
        else if (fp16)
            hipLaunchKernelGGL((k_tiles<4, __half, TP, false>), grid, block, lds, st, sds, wmap, n_solve, slot_cap,
                               parity);
        else
            hipLaunchKernelGGL((k_tiles<4, float, TP, false>), grid, block, lds, st, sds, wmap, n_solve, slot_cap,
                               parity);
    });
}

hipError_t launch_mf_tiles(int K, bool fp16, uint32_t tile_pix, const SolveDesc *sds, const uint32_t *wmap,
                           int n_solve, uint32_t n_wg, int k_it, const float *x_user, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    return with_tile_pix(tile_pix, [&](auto tp) {
        constexpr int TP = decltype(tp)::value;
        const dim3 grid(n_wg), block(2 * TP);
        if (K == 8 && fp16)
            hipLaunchKernelGGL((k_mf_tiles<8, __half, TP>), grid, block, 0, st, sds, wmap, n_solve, k_it, x_user);
        else if (K == 8)
            hipLaunchKernelGGL((k_mf_tiles<8, float, TP>), grid, block, 0, st, sds, wmap, n_solve, k_it, x_user);
        else if (fp16)
            hipLaunchKernelGGL((k_mf_tiles<4, __half, TP>), grid, block, 0, st, sds, wmap, n_solve, k_it, x_user);
        else
            hipLaunchKernelGGL((k_mf_tiles<4, float, TP>), grid, block, 0, st, sds, wmap, n_solve, k_it, x_user);
    });
}

hipError_t launch_mf_rows(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k_it,
                          const float *x_user, float *y_user, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_mf_rows, dim3(n_wg), dim3(64), 0, st, sds, wmap, n_solve, k_it, x_user, y_user);
    return hipGetLastError();
}

hipError_t launch_spmv_rows(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, const float *x,
                            float *y, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_spmv_rows, dim3(n_wg), dim3(64), 0, st, sds, wmap, n_solve, x, y);
    return hipGetLastError();
}

size_t tiles_combine_lds(uint32_t max_nodes, uint32_t max_row_len, bool wave)
{
    const size_t L1 = size_t(max_row_len) + 1;
    if (wave)
        return sizeof(float) * L1 * 36 + sizeof(uint32_t) * 4 * L1;
    return sizeof(float4) * (kCombUnionF4 + L1 * (kTileRec / 4)) + sizeof(float) * (L1 * 36 + 16) +
           sizeof(uint32_t) * (3 * L1 + 1) + sizeof(uint16_t) * ((max_nodes + 1) & ~1u);
}

// After the group's block sort (keys blk0 + block, values the solve's slot
// ids, stable: tile order within a block): list_items[l] and
// slot_pos[list_items[l]] = l for l < S (each slot's position in the block
// lists, where k_tiles puts its partial), and list_off[b] = the first
// position of block b for b <= N + 2U (a lower bound in the sorted blocks):
// sorted position l writes it for the blocks after the key below it up to its
// own, the last position also for the blocks after its key (no searches).
__global__ __launch_bounds__(256) void k_slot_lists(IndexGroup g, const uint32_t *__restrict__ wmap,
                                                    const uint32_t *__restrict__ skeys,
                                                    const uint32_t *__restrict__ svals)
{
    const int it = find_solve_t(wmap, g.n, blockIdx.x);
    const IndexItem item = g.items[it];
    const SolveDesc d = global_desc(g.sds[item.slot]);
    const uint32_t S = d.tile_sbase[d.n_tiles];
    const uint32_t n_blocks = uint32_t(d.N) + 2u * uint32_t(d.U) + 1u;
    const uint32_t l = (blockIdx.x - wmap[it]) * blockDim.x + threadIdx.x;
    if (S == 0)
    {
        if (l < n_blocks)
            d.list_off[l] = 0u;
        return;
    }
    if (l >= S)
        return;
    const uint32_t v = svals[item.sb0 + l];
    d.list_items[l] = v;
    d.slot_pos[v] = l;
    // a slot's block is below N + 2U; the bound keeps every store in the array
    const uint32_t key = min(skeys[item.sb0 + l] - item.blk0, n_blocks - 1u);
    const uint32_t first = l == 0 ? 0u : skeys[item.sb0 + l - 1] - item.blk0 + 1u;
    for (uint32_t b = first; b <= key; ++b)
        d.list_off[b] = l;
    if (l + 1 == S)
        for (uint32_t b = key + 1; b < n_blocks; ++b)
            d.list_off[b] = S;
}

hipError_t launch_slot_lists(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, const uint32_t *skeys,
                             const uint32_t *svals, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_slot_lists, dim3(n_wg), dim3(256), 0, st, g, wmap, skeys, svals);
    return hipGetLastError();
}

hipError_t launch_tiles_combine(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, size_t lds,
                                bool wave, PcgSlot2 *slots2, uint32_t n_slots2, hipStream_t st)
{
    if (n_wg == 0)
        return hipSuccess;
    const void *fn = wave ? reinterpret_cast<const void *>(k_tiles_combine_wave)
                          : reinterpret_cast<const void *>(k_tiles_combine_wg);
    if (lds > 65536)
    {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess)
            return e;
    }
    if (slots2 && n_slots2 > n_wg)
        return hipErrorInvalidValue; // never: a solve has no more PCG groups than rows
    if (wave)
        hipLaunchKernelGGL(k_tiles_combine_wave, dim3(n_wg), dim3(64), lds, st, sds, wmap, n_solve, slots2, n_slots2);
    else
        hipLaunchKernelGGL(k_tiles_combine_wg, dim3(n_wg), dim3(kTileThreads), lds, st, sds, wmap, n_solve, slots2,
                           n_slots2);
    return hipGetLastError();
}
} // namespace sh
