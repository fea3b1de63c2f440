// Host engine: capacity, per-slot solve state, index build and GN loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <initializer_list>
#include <vector>

#include "../../include/sh/warp_solver.h"
#include "sh_desc.hpp"
#include "sh_kernels.hpp"

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
    // admit several slots at once (their index builds run as groups); a
    // slot whose inputs are rejected gets its status in rcs and the others
    // proceed; the return value is SH_OK or a HIP error that failed them all
    int admit_many(const uint32_t *slots, uint32_t n, hipStream_t st, int *rcs);
    // solves of the largest frame one index-build group holds
    uint32_t index_units() const { return group_units_; }
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

    // Index-build workspace of a group of solves: concatenated arrays sized
    // for group_units_ solves of the largest frame (sums of per-solve upper
    // bounds decide how many solves a group takes), the group's device tables
    // (descriptors, items, work maps, scan jobs) and their pinned staging.
    struct GroupWs
    {
        IndexWs ws{};                               // the kernels' view
        uint32_t *ka = nullptr, *kb = nullptr;      // sort keys in / out
        uint32_t *va = nullptr, *vb = nullptr;      // sort values in / out
        uint32_t *head = nullptr, *hpos = nullptr;  // the reference pair index only
        void *sort_temp = nullptr;
        size_t sort_temp_bytes = 0;
        uint64_t cap_m = 0, cap_h = 0, cap_kv = 0, cap_nodes = 0, cap_tcnt = 0, cap_tent = 0;
        uint32_t cap_items = 0;
        char *d_tab = nullptr;                      // device tables
        char *h_tab[3] = {nullptr, nullptr, nullptr}; // pinned staging, one per stage
        size_t off_items = 0, off_maps = 0, off_jobs = 0, tab_bytes = 0;
        uint32_t *h_out = nullptr;                  // pinned counts, 8 words per item
    };
    enum
    {
        kGmFlags = 0,
        kGmCompact,
        kGmSplit,
        kGmNode,
        kGmRowBits,
        kGmTile,
        kGmLists,
        kNumGroupMaps
    };
    int alloc(void **p, size_t bytes);
    void release();
    int alloc_group_ws(uint32_t units, uint32_t max_items);
    int build_index(uint32_t slot, hipStream_t st);
    int build_indexes(const uint32_t *slots, uint32_t n, hipStream_t st, int *rcs = nullptr);
    int build_group(const uint32_t *slots, uint32_t n, hipStream_t st, int *rcs);
    int reference_pair_index(Slot &sl, hipStream_t st);
    int upload_group(int stage, hipStream_t st);
    IndexGroup group_view() const;
    const uint32_t *group_map(int m) const;
    int size_partials(Slot &sl);
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
    // index workspace and the group being built: its slots, items, work maps
    // (kNumGroupMaps of n + 1 entries) and scan jobs (row scans, tile scans)
    GroupWs gws_;
    uint32_t group_units_ = 1;
    std::vector<uint32_t> g_slots_;
    std::vector<IndexItem> g_items_;
    std::vector<uint32_t> g_maps_;
    std::vector<ScanJob> g_jobs_[2];
    bool tile_order() const; // the GN phase reads the tile-order pixel streams
    uint32_t tiles_cap_ = 0;
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
