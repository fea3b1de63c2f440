// Host-side launchers of the solver kernels (defined in sh_kernels.hip) and
// the rocPRIM-backed sort / scan helpers (sh_sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "sh_desc.hpp"

namespace sh
{
// Tile order (the GN phase's pixel order on the tile path): 32x32-pixel
// blocks in raster order, Morton order inside a block; index space size.
uint32_t tile_order_blocks_x(uint32_t width);
uint32_t tile_order_size(uint32_t width, uint32_t height);

// ------------------------------------------------------------ index build
// The index build runs over a group of solves: every stage is one launch
// over all of them.  A solve's arrays in the concatenated workspace start at
// its item's offsets; its own outputs go to the arrays its descriptor names.
struct IndexItem
{
    uint32_t slot;  // the solve's descriptor in the group's descriptor array
    uint32_t m0;    // first raster flag (W * H per solve)
    uint32_t h0;    // first tile-order flag (tile_order_size per solve)
    uint32_t kv0;   // first Node2Term pair (P K + 2 E + N per solve)
    uint32_t node0; // first node key (N + 1 per solve): node sort keys, bitmap rows, row and segment counts
    uint32_t t0;    // first tile of the tile workspace (n_tiles per solve, each at the full tile stride)
    uint32_t tc0;   // first tile count (n_tiles + 2 per solve: counts, scan total, largest count)
    uint32_t sb0;   // first slot of the block sort (the solve's slot count)
    uint32_t blk0;  // first block key (N + 2 U + 1 per solve)
    uint32_t pad[7];
};
static_assert(sizeof(IndexItem) == 64, "index item");
// The group's device tables.
struct IndexGroup
{
    const SolveDesc *sds;
    const IndexItem *items;
    int n;
};
// The group's workspace.  errs and the per-item meta words are cleared by the
// kernels that read them last, so they are zero when the next build starts.
struct IndexWs
{
    uint32_t *flags, *pos;         // raster valid flags and their exclusive scan
    uint32_t *tflags, *tpos;       // the same in tile order
    uint32_t *errs;                // one range-check word per item
    uint32_t *meta;                // 16 words per item: [0] longest row
    uint32_t *bitmap;              // kBitmapWords per node key
    uint2 *bitpre;                 // the same words with the neighbour count of the words before each
    uint32_t *rowlen, *segcnt;     // per node key
    uint16_t *tcontrib;            // per tile, at its fixed stride: sorted contributions
    uint32_t *tslot_block;         // per tile: each slot's block
    uint16_t *tslot_first;         // per tile: each slot's first contribution
    uint32_t *tnslot, *tncontrib;  // per tile count
};
// Exclusive scan of one short array (counts per node or per tile) by one
// work-group: out[i] = in[0] + .. + in[i - 1] for i < n.
struct ScanJob
{
    const uint32_t *in;
    uint32_t *out;
    uint32_t n;
    uint32_t pad;
};
// Every launch takes a work map: wmap[i] is the first work-group of item i,
// wmap[n] = n_wg.
// Pixels (or tile-order indices) a thread of the flag kernel takes
// (work-groups of 256 threads).
constexpr uint32_t kPixPer = 4;
// Valid-pixel flags (tiled: also in tile order) and the range checks.
hipError_t launch_pixel_flags(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, bool tiled,
                              const IndexWs &ws, hipStream_t st);
// Compaction of the valid pixels into each solve's streams (tiled: also the
// tile-order copies), the Node2Term (node key, term) pairs at kv0 (keys
// node0 + node, node0 + N for a self edge), and the clearing of the words
// later stages accumulate into.
hipError_t launch_compact_pixels(int K, const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, bool tiled,
                                 const IndexWs &ws, uint32_t *keys, uint32_t *vals, hipStream_t st);
// After the node sort: each solve's n2t_terms and n2t_off, kSplitPer sorted
// pairs a thread (work-groups of 256 threads).
constexpr uint32_t kSplitPer = 8;
hipError_t launch_n2t_split(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, const uint32_t *skeys,
                            const uint32_t *svals, hipStream_t st);
hipError_t launch_scan_jobs(const ScanJob *jobs, uint32_t n_jobs, hipStream_t st);
// Counts for the host, host[item * 8 + k]: stage 1 -> [0] valid pixels,
// [1] range-check flag (then cleared); stage 2 -> [0] segments, [1] longest
// row, [2] U, [3] slots, [4] contributions, [5] largest tile slot count.
hipError_t launch_index_counts(const IndexGroup &g, int stage, const IndexWs &ws, uint32_t *host, hipStream_t st);
hipError_t launch_pair_kv(const SolveDesc &d, uint32_t *keys, uint32_t *vals, hipStream_t st);
hipError_t launch_pair_heads(const uint32_t *keys, uint32_t n, uint32_t *head, uint32_t *meta, hipStream_t st);
hipError_t launch_pair_compact(const uint32_t *keys, const uint32_t *head, const uint32_t *hpos, uint32_t n,
                               const uint32_t *meta, uint32_t *half_keys, uint32_t *half_off, uint32_t *meta_out,
                               hipStream_t st);
hipError_t launch_pair_entries(const SolveDesc &d, const uint32_t *sorted_vals, hipStream_t st);
hipError_t launch_sym_kv(const SolveDesc &d, uint32_t *keys, uint32_t *vals, hipStream_t st);
hipError_t launch_sym_scatter(const SolveDesc &d, const uint32_t *skeys, const uint32_t *svals, hipStream_t st);
hipError_t launch_row_ptr(const SolveDesc &d, const uint32_t *skeys, hipStream_t st);

hipError_t launch_identity(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, hipStream_t st);
hipError_t launch_identity_dq(dq8 *dq, uint32_t n_nodes, hipStream_t st);
hipError_t launch_cost_final(const SolveDesc *sds, int n_solve, int slot, hipStream_t st);
// The dense (per compacted pixel) and ARAP (per edge, then per node) terms as
// one launch; store = 0: only the cost partials (the records of the last
// linearisation stay)
hipError_t launch_terms(int K, bool fp16, const SolveDesc *sds, const uint32_t *pix_map, uint32_t n_pix_wg,
                        const uint32_t *reg_map, uint32_t n_reg_wg, int n_solve, int parity, hipStream_t st,
                        int store = 1);
hipError_t launch_diag_rhs(int K, bool fp16, const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg,
                           hipStream_t st);
struct PcgSlot2;
// Inverse diagonal blocks; also zeroes slots2[0 .. n_slots2) and every solve's
// published vectors for a pipelined PCG launch that follows (slots2 may be null).
hipError_t launch_dinv(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, PcgSlot2 *slots2,
                       uint32_t n_slots2, hipStream_t st);
hipError_t launch_offdiag(int K, bool fp16, const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg,
                          hipStream_t st);
hipError_t launch_pcg_init(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, double *part_rs,
                           hipStream_t st);
hipError_t launch_pcg_spmv(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k,
                           const double *part_rs, double *part_dq, hipStream_t st);
hipError_t launch_pcg_update(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k,
                             const double *part_dq, double *part_rs, hipStream_t st);
hipError_t launch_pcg_dir(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k,
                          const double *part_rs, hipStream_t st);
hipError_t launch_pcg_dot(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k,
                          double *part_dq, hipStream_t st);
hipError_t launch_update(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int parity,
                         hipStream_t st);
hipError_t launch_skin(uint32_t K, const float4 *surfels, uint32_t S, const float4 *nodes, uint32_t N, float r2,
                       uint16_t *knn, float *spatial, float *connect, hipStream_t st);
hipError_t launch_node_graph(uint32_t K, const float4 *nodes, uint32_t N, float r2, uint16_t *node_knn,
                             float *node_spatial, uint16_t *edges, float *node_connect, hipStream_t st);

// nodes_per_wg (1, or kRowNodes: one wave each) nodes a work-group of 256
// threads: the row bitmap, row length, dense-term count and segment count;
// the longest row into the item's meta[0].
constexpr uint32_t kRowNodes = 4;
hipError_t launch_row_bitmap(int K, uint32_t nodes_per_wg, const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg,
                             const IndexWs &ws, hipStream_t st);
// One wave per node, kRowNodes nodes a work-group: the bitmap words with
// their word prefix counts (bitpre) and, with cols, the block columns from
// them and row_ptr.
hipError_t launch_row_cols(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, const IndexWs &ws, bool cols,
                           hipStream_t st);
// Dense terms per segment of a node row (one work-group of the segment kernel).
constexpr uint32_t kSegTerms = 256;
struct RowsPlan
{
    uint32_t chunk = 0;      // dense terms staged per round of the segment kernel
    size_t lds = 0;          // its dynamic LDS bytes per work-group
    size_t lds_combine = 0;  // the combine kernel's
};
RowsPlan rows_plan(uint32_t max_nodes, uint32_t max_row_len);
// Segment kernel over seg_map (one work-group per segment) and the combine
// kernel over row_map (one work-group per node row), launched in that order.
hipError_t launch_rows_seg(int K, bool fp16, const SolveDesc *sds, const uint32_t *seg_map, int n_solve,
                           uint32_t n_seg_wg, const RowsPlan &plan, hipStream_t st);
hipError_t launch_rows_combine(const SolveDesc *sds, const uint32_t *row_map, int n_solve, uint32_t n_row_wg,
                               const RowsPlan &plan, hipStream_t st);
// Barrier slot of one persistent-PCG work-group.  Every barrier wait of the
// persistent kernels polls at most `wait_limit` times (0: 2^22 polls, about
// 0.1 s) and then sets *err, after which every wait of the launch returns at
// once: the launch always drains, with meaningless sums.
struct PcgSlot
{
    double v[2];
    uint32_t epoch;
    uint32_t pad[3];
};
// Blocks of the normal matrix per persistent-PCG work-group (sets the
// number of groups per solve, at most kPcgMaxGroups).
constexpr uint32_t kPcgBlocksPerGroup = 1024;
constexpr uint32_t kPcgMaxGroups = 256;
hipError_t launch_pcg_coop(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int iters,
                           PcgSlot *slots, uint32_t *epoch_base, uint32_t *err, uint32_t wait_limit, uint32_t tile,
                           hipStream_t st);
int pcg_coop_max_groups(size_t lds);
// Register-resident persistent PCG: groups of at most kPcgResRows rows and
// kPcgResBlocks blocks, split evenly by rows.
constexpr uint32_t kPcgResRows = 85;
constexpr uint32_t kPcgResBlocks = 1536;
hipError_t launch_pcg_resident(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int iters,
                               PcgSlot *slots, uint32_t *epoch_base, uint32_t *err, uint32_t wait_limit,
                               hipStream_t st);
int pcg_resident_max_groups();
// Barrier slot of the pipelined PCG (two sums per barrier).
struct PcgSlot2
{
    double v[2][2];
    uint32_t epoch;
    uint32_t pad[7];
};
// Pipelined resident PCG: one barrier per iteration, FP64 vectors; same
// groups as the resident kernel.
// update_parity >= 0: the kernel also applies the warp update (k_update's
// work for that GN-iteration offset) to each group's own nodes.
hipError_t launch_pcg_pipe(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int iters,
                           PcgSlot2 *slots, uint32_t *epoch_base, uint32_t *err, uint32_t wait_limit,
                           int update_parity, hipStream_t st);
int pcg_pipe_max_groups();
constexpr uint32_t kBitmapWords = 128;
// Floats per tile-slot partial in HBM: the 28 of the record (21 H, 6 g, pad)
// padded to 32, so a partial is one 128-byte line of an aligned buffer.
constexpr int kPartStride = 32;

// work-group geometry of the GN-phase kernels
constexpr uint32_t kThreads = 256;
constexpr uint32_t kWavesPerWg = 4;
constexpr uint32_t kPcgNodesPerWg = 32;  // split PCG kernels

// Forward warp of surfels and nodes (sh_warp.hip)
// optical-flow input and node-motion prior (sh_flow.hip)
// Nodes the allocation-free node-motion entry point can take (its sums live in
// a fixed device array of this many nodes x 32 bytes).
constexpr uint32_t kNoAllocMaxNodes = 4096;
// Device address of that array on the current device (looked up once per device).
hipError_t noalloc_sums(unsigned long long **out);
hipError_t launch_flow_from_of16(const uint16_t *of16, int W, int H, float scale, float *flow2, hipStream_t st);
hipError_t launch_node_motion(const float *rgbd_prev4, const float *rgbd4, const float *flow2,
                              const uint32_t *index_prev, int width, int height, const uint16_t *surfel_knn,
                              const float *surfel_spatial, uint32_t num_surfels, int K, const float cam2world[16],
                              float fx, float fy, float cx, float cy, uint32_t num_nodes, unsigned long long *sums,
                              float *out4, hipStream_t st);
// solver maps by point splatting (sh_render.hip)
hipError_t launch_render_solver_maps(const float *vertex4, const float *normal4, const float *color_time4,
                                     uint32_t num_surfels, const float world2camera[16], int width, int height,
                                     float fx, float fy, float cx, float cy, float max_depth, float current_time,
                                     bool with_recent_observation, unsigned long long *zbuf, float *out_vertex4,
                                     float *out_normal4, uint32_t *out_index, hipStream_t st);
// depth -> measurement maps (sh_measure.hip)
hipError_t launch_filter_depth(const uint16_t *raw, int W, int H, float fx, float fy, float cx, float cy,
                               float clip_near, float clip_far, float *out, hipStream_t st);
hipError_t launch_scaled_vertex(const float *filtered, int W, int H, int w, int h, float scale, float fx, float fy,
                                float cx, float cy, float clip_near, float clip_far, float *vertex4, hipStream_t st);
hipError_t launch_normal_radius(const float *vertex4, int w, int h, float radius_scale, float *normal4,
                                hipStream_t st);
hipError_t launch_vertex_confid(const uint16_t *raw, int W, int H, float fx, float fy, float cx, float cy,
                                float *vertex4, hipStream_t st);
hipError_t launch_forward_warp(uint32_t K, const float4 *warp, const float4 *ref_v, const float4 *ref_n,
                               const uint16_t *s_knn, const float *s_spatial, const float *s_connect,
                               uint32_t num_surfels, const float4 *ref_node, const uint16_t *n_knn,
                               const float *n_spatial, const float *n_connect, uint32_t num_nodes, float4 *live_v,
                               float4 *live_n, float4 *live_node, hipStream_t st);

// Pixel tiles (sh_tiles.hip): 64, 128 (default) or 256 pixels, per the tuning
constexpr uint32_t kTilePixMin = 64, kTilePixDefault = 128, kTilePixMax = 256;
// One work-group per tile: the tile's sorted contributions, slots and counts.
hipError_t launch_tile_build(int K, uint32_t tile_pix, const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg,
                             const IndexWs &ws, hipStream_t st);
// One work-group per tile: contributions and slots to their compact places,
// and the (blk0 + block, slot) pairs of the block sort at sb0.
hipError_t launch_tile_compact(int K, uint32_t tile_pix, const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg,
                               const IndexWs &ws, uint32_t *sort_keys, uint32_t *sort_vals, hipStream_t st);
// One wave per node, kRowNodes nodes a work-group: the transpose map of the
// row's blocks.
hipError_t launch_block_transpose(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, const IndexWs &ws,
                                  hipStream_t st);
// fused: the tile kernel builds the FP32 per-pixel records itself (and the
// dense cost partials) from the warp field of GN iteration offset parity,
// instead of reading them; the term kernel then runs the ARAP part only.
hipError_t launch_tiles(int K, bool fp16, uint32_t tile_pix, bool fused, int parity, const SolveDesc *sds,
                        const uint32_t *wmap, int n_solve, uint32_t n_wg, uint32_t slot_cap, hipStream_t st);
// Longest row (blocks, diagonal included) the one-wave-per-row combine takes:
// it holds a row's block sums in registers.
constexpr uint32_t kWaveCombineMaxBlocks = 73;
// LDS of the tile combine: one work-group per row (wave = false) or one wave per row
size_t tiles_combine_lds(uint32_t max_nodes, uint32_t max_row_len, bool wave);
// After the block sort: each solve's list_items (slot ids by block),
// slot_pos[list_items[l]] = l, and list_off[b] = the first list position of
// block b (b <= N + 2U).
hipError_t launch_slot_lists(const IndexGroup &g, const uint32_t *wmap, uint32_t n_wg, const uint32_t *skeys,
                             const uint32_t *svals, hipStream_t st);
// matrix-free J^T W J x (x = d_{k+1} of each solve unless x_user; y = d.q unless y_user)
hipError_t launch_mf_tiles(int K, bool fp16, uint32_t tile_pix, const SolveDesc *sds, const uint32_t *wmap,
                           int n_solve, uint32_t n_wg, int k_it, const float *x_user, hipStream_t st);
hipError_t launch_mf_rows(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, int k_it,
                          const float *x_user, float *y_user, hipStream_t st);
hipError_t launch_spmv_rows(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, const float *x,
                            float *y, hipStream_t st);
// Also the FP64 inverse diagonal blocks (as k_dinv) and the zeroed state of a
// pipelined PCG launch that follows (slots2[0 .. n_slots2), may be null).
hipError_t launch_tiles_combine(const SolveDesc *sds, const uint32_t *wmap, int n_solve, uint32_t n_wg, size_t lds,
                                bool wave, PcgSlot2 *slots2, uint32_t n_slots2, hipStream_t st);

// rocPRIM wrappers (sh_sort.hip); temp == nullptr queries the temp size
hipError_t sort_pairs_u32_u32(void *temp, size_t &temp_bytes, const uint32_t *kin, uint32_t *kout,
                              const uint32_t *vin, uint32_t *vout, uint32_t n, int end_bit, hipStream_t st);
hipError_t exclusive_scan_u32(void *temp, size_t &temp_bytes, const uint32_t *in, uint32_t *out, uint32_t n,
                              hipStream_t st);
} // namespace sh
