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

    // compacted pixel streams
    ushort4 *pix;     // (x, y, tx, ty)
    uint16_t *pknn;   // P*K
    float *pw;        // P*K, spatial * connect
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
    // pixel tiles (kTilePix compacted pixels each) and their slot partials
    int n_tiles;
    uint32_t *tile_cbase;    // n_tiles+1: first contribution of each tile
    uint32_t *tile_sbase;    // n_tiles+1: first slot of each tile
    uint16_t *tc_contrib;    // contributions, per slot in pixel order: pixel | ka << 7 | kb << 10
    uint32_t *tc_slot_first; // S+1: first contribution of each slot
    uint32_t *list_off;      // N+2U+1: first entry of each block in list_items
    uint32_t *list_items;    // S: slot ids sorted by block (stable: tile order)
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
    double *cost_part; // work-group partials: ceil(P/256) dense, then ceil((E+N)/256) ARAP
    // matrix-free scratch
    float *mf_z;      // P*6 per-pixel H_p s_p
    float *mf_reg;    // E*3 per-edge J x
};

} // namespace sh
