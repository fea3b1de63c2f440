// Host engine: capacity, per-slot solve state, index build and GN loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <initializer_list>
#include <vector>

#include "../../include/sh/warp_solver.h"
#include "sh_desc.hpp"

namespace sh
{
struct dq8;
extern hipError_t last_hip_error;

class Engine
{
public:
    Engine() = default;
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    int init(const sh_solver_config &cfg, uint32_t max_slots);
    int set_inputs(uint32_t slot, const sh_camera *cam, const sh_measure_maps *meas, const sh_render_maps *rend,
                   const sh_geometry *geo, const sh_node_graph *ng, const sh_node_flow *nf, const sh_flow_map *flow,
                   const sh_keypoints *kp, const float cam2world[16]);
    int solve(uint32_t num_solves, hipStream_t st);
    // continuous batching: admit() builds a slot's index and identity start
    // (index workspace); step() runs one GN iteration over a set of admitted
    // slots (launch state).  The two may run on different threads and streams.
    int admit(uint32_t slot, hipStream_t st);
    // admit several slots at once (their index builds share the lanes); a
    // slot whose inputs are rejected gets its status in rcs and the others
    // proceed; the return value is SH_OK or a HIP error that failed them all
    int admit_many(const uint32_t *slots, uint32_t n, hipStream_t st, int *rcs);
    uint32_t num_lanes() const { return uint32_t(lanes_.size()); }
    int step(const uint32_t *slots, uint32_t n, hipStream_t st);
    int gn_iter(uint32_t slot) const { return slots_[slot].d.gn_iter; }
    int cost_history(uint32_t slot, double *out, uint32_t capacity, uint32_t *count) const;
    int apply_jtj(uint32_t slot, const float *x, float *y, bool matrix_free, hipStream_t st);
    int set_tuning(const sh_tuning &t);
    sh_tuning tuning() const;
    int autotune(uint32_t ns, hipStream_t st, uint32_t repeats, sh_tuning *best);
    int time_candidate(const sh_tuning &c, uint32_t ns, hipStream_t st, uint32_t repeats, hipEvent_t e0,
                       hipEvent_t e1, double *ms_out);
    // Called before a slot's grown buffer (tile lists and partials, segment
    // partials) is allocated, with the bytes the engine's holding grows by,
    // and with the negative size of what it releases; returning false refuses
    // the growth, and the index build that needed it fails with SH_E_CAPACITY.
    using GrowthHook = bool (*)(void *ctx, int64_t delta_bytes);
    void set_growth_hook(GrowthHook fn, void *ctx)
    {
        grow_fn_ = fn;
        grow_ctx_ = ctx;
    }
    // polls per barrier wait of the persistent PCG kernels (0: the default)
    void set_wait_limit(uint32_t polls) { wait_limit_ = polls; }
    // solves (or steps) re-run with the split PCG after a barrier wait ran out
    uint32_t pcg_fallbacks() const { return pcg_fallbacks_; }
    // device bytes held: fixed allocations plus the grown per-slot buffers
    size_t device_bytes() const
    {
        size_t b = bytes_;
        for (const Slot &sl : slots_)
            b += sl.part_bytes + sl.tile_bytes;
        return b;
    }
    const dq8 *result(uint32_t slot) const;
    bool solved(uint32_t slot) const { return slot < slots_.size() && slots_[slot].solved; }
    // nodes in the slot's result view: max_nodes identity transforms until the
    // first solve (SolverIterationData.cu:78-88), then the last solve's count
    uint32_t result_nodes(uint32_t slot) const { return slot < slots_.size() ? slots_[slot].result_n : 0u; }
    const SolveDesc &desc(uint32_t slot) const { return slots_[slot].d; }
    uint32_t num_slots() const { return uint32_t(slots_.size()); }
    int debug_view(uint32_t slot, const char *name, const void **ptr, size_t *count) const;
    const sh_solver_config &config() const { return cfg_; }
    // GN-phase kernel time (SH_FLAG_PHASE_TIMING), summed since the last reset
    enum Phase
    {
        kPhDense = 0,
        kPhReg,
        kPhRowsSeg,
        kPhRowsCombine,
        kPhDiagRhs,
        kPhOffdiag,
        kPhDinv,
        kPhPcg,
        kPhUpdate,
        kPhTiles,
        kPhTilesCombine,
        kNumPhases
    };
    int phase_times(double *ms, uint32_t *calls, uint32_t n, bool reset);
    static const char *phase_name(uint32_t p);

private:
    struct Slot
    {
        SolveDesc d;
        bool has_inputs = false;
        bool indexed = false;
        bool solved = false;
        uint32_t result_n = 0;    // nodes in the result view (see result_nodes)
        ushort4 *tpix = nullptr;  // tile-order pixel streams (the GN phase's on the tile path)
        uint16_t *tpknn = nullptr;
        float *tpw = nullptr;
        uint32_t max_row_len = 0;
        uint32_t tile_slot_max = 0; // largest slot count of one tile
        bool records_fused = false;  // the last GN iteration built its records in the tile kernel (not stored)
        void *part_buf = nullptr; // segment partials, grown on demand
        size_t part_bytes = 0;
        void *tile_buf = nullptr; // tile contributions, slots, lists and partials, grown on demand
        size_t tile_bytes = 0;
    };
    enum
    {
        kMapPixel = 0,
        kMapReg,
        kMapNodeWave,
        kMapPairWave,
        kMapPcg,
        kMapNode,
        kMapRow,
        kMapCoop,
        kMapSeg,
        kMapTile,
        kNumMaps
    };

    // Index-build workspace of one lane: the arrays a slot's index build
    // uses between its host readbacks.  Lane 0 is the ws_* set below; a batch
    // builds the indexes of its slots on several lanes at once, each with its
    // own stream, so one slot's host readback overlaps the others' kernels.
    struct IndexLane
    {
        uint32_t *flags = nullptr, *pos = nullptr, *v32a = nullptr, *meta = nullptr;
        uint32_t *tflags = nullptr, *tpos = nullptr; // tile-order flags and their scan
        uint16_t *k16a = nullptr, *k16b = nullptr;
        uint32_t *k32a = nullptr, *k32b = nullptr, *p32a = nullptr;
        uint32_t *bitmap = nullptr, *rowlen = nullptr, *segcnt = nullptr;
        uint16_t *tcontrib = nullptr, *tslot_first = nullptr;
        uint32_t *tslot_block = nullptr, *tnslot = nullptr, *tncontrib = nullptr;
        void *sort_temp = nullptr;
        uint32_t *err = nullptr;    // device: the range-check flag
        uint32_t *h_meta = nullptr; // pinned host readbacks
        hipStream_t stream = nullptr;
        hipEvent_t ev = nullptr;
        int slot = -1; // scheduler state: the slot in progress, its position in the list and next stage
        int item = -1;
        int stage = 0;
    };

    int alloc(void **p, size_t bytes);
    void release();
    int alloc_lane(IndexLane &ln, size_t npix);
    int build_index(uint32_t slot, hipStream_t st);
    int build_indexes(const uint32_t *slots, uint32_t n, hipStream_t st, int *rcs = nullptr);
    int index_pixels(uint32_t slot, IndexLane &ln, hipStream_t st);
    int index_rows(uint32_t slot, IndexLane &ln, hipStream_t st);
    int index_tiles(uint32_t slot, IndexLane &ln, hipStream_t st);
    int size_partials(Slot &sl);
    int build_tiles(Slot &sl, IndexLane &ln, hipStream_t st);
    // the tile build and its count scans; their totals, with up to five more
    // device words (more_src to h_meta[more_dst]), go to the lane's host buffer
    hipError_t launch_tile_build_async(SolveDesc &d, IndexLane &ln, hipStream_t st,
                                       std::initializer_list<const uint32_t *> more_src,
                                       std::initializer_list<uint32_t> more_dst);
    int size_tiles(Slot &sl, uint32_t n_slots, uint32_t n_contrib);
    hipError_t mark(int phase, hipStream_t st);
    void collect_marks();
    void build_maps(uint32_t ns);
    int upload(uint32_t ns, hipStream_t st);
    int enqueue_gn(uint32_t ns, hipStream_t st, uint32_t iters, bool identity);
    int wait_stream(hipStream_t st, hipEvent_t ev);
    int finish(hipStream_t st);
    int fall_back_to_split_pcg(hipStream_t st);
    const uint32_t *map(int m, uint32_t ns) const;

    sh_solver_config cfg_{};
    int K_ = 8;
    bool fp16_ = false;
    uint32_t pcap_ = 0, ncap_ = 0, ecap_ = 0, n2t_cap_ = 0;
    uint32_t tcap_ = 0; // tile-order index space of the largest image
    uint64_t ucap_ = 0, pair_cap_ = 0;
    std::vector<Slot> slots_;
    std::vector<void *> allocs_;
    size_t bytes_ = 0;
    // index workspace
    void *ws_flags_ = nullptr, *ws_pos_ = nullptr, *ws_k16a_ = nullptr, *ws_k16b_ = nullptr, *ws_v32a_ = nullptr;
    void *ws_k32a_ = nullptr, *ws_k32b_ = nullptr, *ws_p32a_ = nullptr, *ws_p32b_ = nullptr;
    void *ws_head_ = nullptr, *ws_hpos_ = nullptr, *ws_meta_ = nullptr;
    void *ws_bitmap_ = nullptr, *ws_rowlen_ = nullptr, *ws_segcnt_ = nullptr;
    void *ws_tcontrib_ = nullptr, *ws_tslot_block_ = nullptr, *ws_tslot_first_ = nullptr;
    void *ws_tnslot_ = nullptr, *ws_tncontrib_ = nullptr;
    void *ws_tflags_ = nullptr, *ws_tpos_ = nullptr;
    bool tile_order() const; // the GN phase reads the tile-order pixel streams
    uint32_t tiles_cap_ = 0;
    std::vector<IndexLane> lanes_;
    int coop_max_groups_ = 0;
    size_t coop_lds_ = 0;
    bool coop_resident_ = false; // every solve of the launch fits the register-resident PCG
    int res_max_groups_ = -1;
    int pipe_max_groups_ = -1;
    uint32_t res_rows_ = 85; // resident-PCG rows per group (tunable, <= kPcgResRows)
    // pixels per tile of the tile path (tunable: SH_FLAG_TILE_PIX_64 / _256, else 128)
    uint32_t tile_pix() const
    {
        return (cfg_.flags & SH_FLAG_TILE_PIX_64) ? 64u : (cfg_.flags & SH_FLAG_TILE_PIX_256) ? 256u : 128u;
    }
    void *d_slots_ = nullptr; // persistent-PCG barrier slots
    void *d_slots2_ = nullptr; // pipelined-PCG barrier slots (two sums each)
    void *d_flags_ = nullptr; // [0] PCG epoch base, [1] PCG barrier error
    uint32_t wait_limit_ = 0;
    GrowthHook grow_fn_ = nullptr;
    void *grow_ctx_ = nullptr;
    int grow_buffer(void **buf, size_t *held, size_t bytes);
    uint32_t pcg_fallbacks_ = 0;
    bool barrier_timeout_ = false; // the last finish() saw the PCG barrier flag
    void *sort_temp_ = nullptr;
    size_t sort_temp_bytes_ = 0;
    uint32_t *h_meta_ = nullptr; // index-build readbacks
    uint32_t *h_flag_ = nullptr; // launch-side readback (PCG barrier flag)
    hipEvent_t sync_ev_ = nullptr;  // polled by finish (the launch half)
    hipEvent_t index_ev_ = nullptr; // polled by build_index (the index half)
    std::vector<uint32_t> batch_; // slots of the current launch, in order
    void *h_stage_ = nullptr;
    size_t stage_bytes_ = 0;
    // launch state
    void *d_descs_ = nullptr, *d_wmaps_ = nullptr, *d_part_rs_ = nullptr, *d_part_dq_ = nullptr;
    std::vector<uint32_t> maps_;
    uint32_t map_total_[kNumMaps] = {0};
    hipGraphExec_t graph_exec_ = nullptr;
    uint64_t graph_key_[5] = {0, 0, 0, 0, 0};
    hipStream_t cap_stream_ = nullptr; // graph capture stream
    hipError_t enqueue_pcg_launches(uint32_t ns, bool matrix_free, hipStream_t st);
    hipError_t pcg_graph(uint32_t ns, bool matrix_free, hipStream_t st);
    struct Mark
    {
        hipEvent_t ev;
        int phase;
    };
    std::vector<Mark> marks_;
    size_t n_marks_ = 0;
    double phase_ms_[kNumPhases] = {0.0};
    uint32_t phase_calls_[kNumPhases] = {0};
    std::mutex phase_mu_; // the sums are read from other threads while a server steps
};
} // namespace sh
