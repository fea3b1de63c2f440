// Per-solve descriptor: everything a GN-phase kernel needs for one solve,
// resident in device memory so that one launch can serve a whole batch of
// uneven solves (each work-group looks up its solve in a work map).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sh
{
struct dq8;

struct SolveDesc
{
    // camera / extrinsic (world -> camera, row-major)
    int W, H;
    float fx, fy, cx, cy;
    float w2c_r[9];
    float w2c_t[3];
    // squared penalties (PenaltyConstants.h:210-234) and damping
    float lam2_depth, lam2_flow, lam2_reg, lam2_nt;
    float damping;
    float node_radius_square;
    int K;
    int N, E;            // nodes, ARAP edge slots
    uint32_t S;          // surfels: index_map values below S name a surfel
    int P;               // valid pixels after compaction
    int U;               // unique upper node pairs
    int n_pair_entries;  // valid (pair, term) entries
    int pcg_iters;
    int gn_iter;         // GN iterations done: node DQs are read from dq[(gn_iter + launch parity) & 1]
    uint32_t blk_cap;    // capacity of blk_col / blocks (N_max + 2 U_max)

    // caller-owned inputs
    const float4 *meas_v, *meas_n, *ref_v, *ref_n;
    const uint32_t *index_map;
    const float2 *flow;
    const uint16_t *surfel_knn;
    const float *surfel_spatial, *surfel_connect;
    const float4 *node_coords;
    const uint16_t *edges; // ushort3
    const float *node_connect;
    const float4 *node_pred;

    // compacted pixel streams in raster order (the reference's order: Node2Term
    // term ids and every index kernel use it)
    ushort4 *pix;     // (x, y, tx, ty)
    uint16_t *pknn;   // P*K
    float *pw;        // P*K, spatial * connect
    // the same streams in the order the GN phase reads them (the per-pixel
    // records follow it): tile order -- 32x32-pixel blocks in raster order,
    // Morton order inside a block -- on the tile path, so that 64 to 256 consecutive
    // pixels cover a compact patch; the raster arrays above on the others
    ushort4 *gpix;
    uint16_t *gpknn;
    float *gpw;
    // node -> term CSR
    uint32_t *n2t_off;   // N+1
    uint32_t *n2t_terms; // sorted term ids
    // node pairs
    uint32_t *half_keys;    // U
    uint32_t *half_off;     // U+1
    uint32_t *pair_entries; // packed, n_pair_entries
    uint32_t *upper_pos;    // U: position of (i,j) among the sorted symmetric keys
    uint32_t *lower_pos;    // U: position of (j,i)
    uint32_t *row_ptr;      // N+1 into the sorted symmetric keys
    uint32_t *blk_col;      // N + 2U block columns, row-major, diagonal block first
    uint32_t *dense_cnt;    // N: number of dense terms leading each node's term list
    uint32_t *seg_ptr;      // N+1: prefix of each node's dense-term segments (kSegTerms terms each)
    int n_seg;              // dense-term segments in all
    uint32_t part_stride;   // floats per segment partial: 8 (rhs) + 21 per block of the longest row
    float *part;            // n_seg * part_stride partial block rows
    // pixel tiles (tile_pix compacted pixels each) and their slot partials
    int n_tiles;
    int tile_pix;            // pixels per tile (64, 128 or 256) of the index build
    uint32_t *tile_cbase;    // n_tiles+1: first contribution of each tile
    uint32_t *tile_sbase;    // n_tiles+1: first slot of each tile
    uint16_t *tc_contrib;    // contributions, per slot in pixel order: pixel | ka << 7 | kb << 10
    uint32_t *tc_slot_first; // S+1: first contribution of each slot
    uint32_t *list_off;      // N+2U+1: first entry of each block in list_items
    uint32_t *list_items;    // S: slot ids sorted by block (stable: tile order)
    uint32_t *slot_pos;      // S: each slot's position in list_items (where k_tiles puts its partial)
    uint32_t *blk_transpose; // N+2U: the index of each block's transpose (block (j, i) for (i, j))
    float *tpart;            // S * 32 slot partials (21 H, 6 g, pad; 128 B each)
    // GN iteration state
    dq8 *dq[2];       // double buffer (SolverIterationData.h)
    void *jr;         // P*24 residual/Jacobian records, FP32 or FP16
    float *reg;       // E*8 ARAP records
    float4 *ttrans;   // N node translations
    float *blk;       // (N + 2U)*36 block values, [a*6+b]
    double *dinv;     // N*36 inverse diagonal blocks
    float *rhs;       // N*6  = -J^T W r
    // PCG vectors (6N each)
    float *x, *r, *s, *q, *dv[2];
    double *mv[2];    // pipelined PCG: the exchanged vector m = M^-1 w, 6N each, by iteration parity
    double *delta;    // pcg_iters + 1 : r.s per iteration
    double *cost;     // weighted cost (dense + ARAP) at the start of each GN iteration, then the final one
    double *cost_part; // partials: ceil(P/64) dense (one per 64-pixel group), then ceil((E+N)/256) ARAP
    uint8_t *rot_cut;  // gn_iterations * N: 1 where the update dropped a node's rotation (|w|_1 < 1e-4)
    float *twist_l1;   // gn_iterations * N: the FP32 |w|_1 that decision was taken on
    // matrix-free scratch
    float *mf_z;      // P*6 per-pixel H_p s_p
    float *mf_reg;    // E*3 per-edge J x
};

// A generic pointer that is known to name device (global) memory.  The pointers
// a kernel reads out of a descriptor array are generic, so their accesses
// compile to flat_* instructions, which count against the LDS counter as well:
// every LDS wait then also waits for the outstanding flat loads and stores,
// and a flat load cannot pass an earlier LDS store it might alias.  Passing
// the address through an integer into the global address space and back lets
// the compiler's address-space inference emit global_* accesses instead.
template <class T>
__device__ __forceinline__ T *as_global(T *p)
{
    typedef __attribute__((address_space(1))) T global_t;
    return (T *)(global_t *)(uintptr_t)p;
}

// Entry i & 1 of a two-entry pointer array of a descriptor (both entries are
// uniform), marked global.  A descriptor copy indexed at run time would be
// kept in scratch memory, and the compiler turns a select between two loads
// of the array into such an indexed load; the empty asm makes the two loaded
// values opaque, so the select stays a select of registers.
template <class T, class I>
__device__ __forceinline__ T *pick2(T *const (&a)[2], I i)
{
    uintptr_t u0 = reinterpret_cast<uintptr_t>(a[0]), u1 = reinterpret_cast<uintptr_t>(a[1]);
    asm("" : "+s"(u0), "+s"(u1));
    return as_global(reinterpret_cast<T *>((i & 1) ? u1 : u0));
}

// A kernel's copy of its solve's descriptor with every pointer marked global
// (only the fields a kernel uses are loaded).
__device__ __forceinline__ SolveDesc global_desc(const SolveDesc &src)
{
    SolveDesc d = src;
#define SH_G(f) d.f = as_global(d.f)
    SH_G(meas_v); SH_G(meas_n); SH_G(ref_v); SH_G(ref_n); SH_G(index_map); SH_G(flow);
    SH_G(surfel_knn); SH_G(surfel_spatial); SH_G(surfel_connect); SH_G(node_coords); SH_G(edges);
    SH_G(node_connect); SH_G(node_pred); SH_G(pix); SH_G(pknn); SH_G(pw); SH_G(gpix); SH_G(gpknn); SH_G(gpw);
    SH_G(n2t_off); SH_G(n2t_terms);
    SH_G(half_keys); SH_G(half_off); SH_G(pair_entries); SH_G(upper_pos); SH_G(lower_pos); SH_G(row_ptr);
    SH_G(blk_col); SH_G(dense_cnt); SH_G(seg_ptr); SH_G(part); SH_G(tile_cbase); SH_G(tile_sbase);
    SH_G(tc_contrib); SH_G(tc_slot_first); SH_G(list_off); SH_G(list_items); SH_G(slot_pos); SH_G(blk_transpose); SH_G(tpart);
    SH_G(dq[0]); SH_G(dq[1]); SH_G(jr); SH_G(reg); SH_G(ttrans); SH_G(blk); SH_G(dinv); SH_G(rhs);
    SH_G(x); SH_G(r); SH_G(s); SH_G(q); SH_G(dv[0]); SH_G(dv[1]); SH_G(mv[0]); SH_G(mv[1]);
    SH_G(delta); SH_G(cost); SH_G(cost_part); SH_G(rot_cut); SH_G(twist_l1); SH_G(mf_z); SH_G(mf_reg);
#undef SH_G
    return d;
}

} // namespace sh
