// Serving loop: tracking-solve requests from many camera streams, continuous
// batching at Gauss-Newton-iteration granularity.
//
// The reference solves one frame of one camera per call
// (OptimizationProcessorWarpSolver::ProcessFrame, app.cu:132-141).  Here:
//  * each registered stream keeps its solver-side state resident in
//    server-owned HBM: surfel KNN ids and weights, the node graph, and the
//    warp field of its last solve;
//  * memory is budgeted: the solve slots (the number of solves in flight) are
//    sized to fit the HBM budget at creation, and a stream is admitted only
//    while its resident state still fits;
//  * an admission thread takes the oldest waiting request whose stream has no
//    solve in flight (frames of one stream stay in order), puts it in a free
//    slot and builds its index on its own HIP stream;
//  * a step thread runs one GN iteration over every admitted slot -- solves at
//    different iterations share each launch -- and retires the solves that
//    completed their last iteration, freeing the slot.  Requests that arrive
//    while others are on the device join at the next iteration boundary.
//  * Every kernel works per solve with fixed-order reductions, so a request's
//    result does not depend on what it was batched with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sh/warp_solver.h"
#include "sh_engine.hpp"
#include "sh_kernels.hpp"

namespace
{
using clk = std::chrono::steady_clock;

struct StreamState
{
    uint32_t id = 0;
    sh_camera cam{};
    float cam2world[16];
    uint32_t num_surfels = 0, num_nodes = 0, num_edges = 0, K = 0;
    float node_radius_square = 0.f;
    // resident device copies
    void *buf = nullptr;
    size_t bytes = 0;
    uint16_t *surfel_knn = nullptr;
    float *surfel_spatial = nullptr, *surfel_connect = nullptr;
    float *node_coords = nullptr;
    uint16_t *edges = nullptr;
    float *node_connect = nullptr, *node_spatial = nullptr;
    uint16_t *node_knn = nullptr;
    sh_dual_quat *warp = nullptr; // warp field of the last solve
    // reference surfels and the live ones (forward-warped after every solve), optional
    float *ref_v = nullptr, *ref_n = nullptr, *live_v = nullptr, *live_n = nullptr, *live_node = nullptr;
    bool busy = false;            // a request of this stream is admitted
    uint64_t solved = 0;
};

struct Request
{
    uint64_t ticket = 0;
    uint32_t stream = 0;
    sh_measure_maps meas{};
    sh_render_maps rend{};
    sh_flow_map flow{};
    sh_node_flow nflow{};
    sh_dual_quat *out = nullptr;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = SH_OK;
    bool done = false;
    uint32_t slot = 0xffffffffu;
};
} // namespace

struct sh_server
{
    std::unique_ptr<sh::Engine> engine;
    sh_server_config cfg{};
    clk::time_point t0;
    hipStream_t s_admit = nullptr, s_step = nullptr;
    hipEvent_t ev_step = nullptr; // polled by the step thread after a step's copies
    std::mutex mu;
    std::condition_variable cv_admit, cv_step, cv_done;
    std::unordered_map<uint32_t, StreamState> streams;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> waiting;  // submitted, not admitted (arrival order)
    std::vector<uint32_t> free_slots;
    std::vector<uint32_t> active;  // admitted slots
    std::vector<uint64_t> slot_req; // ticket per slot
    uint64_t next_ticket = 1;
    uint64_t reserved = 0; // bytes held: engine + resident streams
    sh_server_stats stats{};
    bool stop = false;
    int fatal = SH_OK;
    std::thread th_admit, th_step;

    double now_ms() const { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }
    void admit_loop();
    void step_loop();
};

// The engine's growth hook: a slot's tile and partial buffers grow during index
// builds (to the largest frame it has held), under the same budget as the
// slots and the resident streams.
static bool server_grow(void *ctx, int64_t delta)
{
    sh_server *s = static_cast<sh_server *>(ctx);
    std::lock_guard<std::mutex> lk(s->mu);
    if (delta > 0 && s->cfg.hbm_budget_bytes && s->reserved + uint64_t(delta) > s->cfg.hbm_budget_bytes)
    {
        s->stats.growth_refused += 1;
        return false;
    }
    s->reserved = uint64_t(int64_t(s->reserved) + delta);
    s->stats.hbm_reserved = s->reserved;
    return true;
}

static int hip_status(hipError_t e)
{
    if (e != hipSuccess)
    {
        sh::last_hip_error = e;
        return SH_E_HIP;
    }
    return SH_OK;
}

// Requests one admission round takes per index-build unit (one solve of the
// largest frame in an index group), at most: the round's index builds run as
// one group.  Smaller rounds reach the GN steps sooner, so more admitted
// solves share each step (measured on the 64-stream mix with eight units:
// 2,726 solves/s at one per unit, 2,720 at two, 2,717 at three, 2,681 at
// four, 2,480 at eight); under a light load a round holds what is waiting.
static constexpr uint32_t kAdmitPerUnit = 2;

void sh_server::admit_loop()
{
    std::unique_lock<std::mutex> lk(mu);
    const uint32_t units = std::max(1u, engine->index_units());
    std::vector<uint64_t> picked;
    std::vector<uint32_t> pslots;
    std::vector<int> rcs;
    for (;;)
    {
        // the oldest waiting requests whose streams have nothing in flight, up
        // to kAdmitPerUnit per index unit and one per free slot (frames of one
        // stream stay in order)
        picked.clear();
        for (auto it = waiting.begin(); it != waiting.end() && picked.size() < kAdmitPerUnit * units &&
                                        picked.size() < free_slots.size();)
        {
            StreamState &ss = streams[reqs[*it].stream];
            if (!ss.busy)
            {
                ss.busy = true;
                picked.push_back(*it);
                it = waiting.erase(it);
            }
            else
                ++it;
        }
        if (stop)
            return;
        if (picked.empty())
        {
            cv_admit.wait(lk);
            continue;
        }
        // inputs: resident stream state + each frame's maps
        struct In
        {
            sh_camera cam;
            float pose[16];
            sh_geometry geo;
            sh_node_graph ng;
            sh_measure_maps meas;
            sh_render_maps rend;
            sh_flow_map flow;
            sh_node_flow nflow;
        };
        std::vector<In> in(picked.size());
        pslots.resize(picked.size());
        for (size_t k = 0; k < picked.size(); ++k)
        {
            Request &rq = reqs[picked[k]];
            StreamState &ss = streams[rq.stream];
            const uint32_t slot = free_slots.back();
            free_slots.pop_back();
            rq.slot = slot;
            rq.t_admit = now_ms();
            pslots[k] = slot;
            In &x = in[k];
            x.cam = ss.cam;
            std::memcpy(x.pose, ss.cam2world, sizeof(x.pose));
            x.geo = sh_geometry{ss.surfel_knn, ss.surfel_spatial, ss.surfel_connect, ss.num_surfels};
            x.ng = sh_node_graph{ss.node_coords,  ss.edges,     ss.node_connect, ss.node_knn,
                                 ss.node_spatial, ss.num_nodes, ss.num_edges,    ss.node_radius_square};
            x.meas = rq.meas;
            x.rend = rq.rend;
            x.flow = rq.flow;
            x.nflow = rq.nflow;
        }
        lk.unlock();
        sh_keypoints kp{0};
        rcs.assign(picked.size(), SH_OK);
        std::vector<uint32_t> ok_slots;
        std::vector<size_t> ok_pos;
        for (size_t k = 0; k < picked.size(); ++k)
        {
            In &x = in[k];
            rcs[k] = engine->set_inputs(pslots[k], &x.cam, &x.meas, &x.rend, &x.geo, &x.ng, &x.nflow, &x.flow, &kp,
                                        x.pose);
            if (rcs[k] == SH_OK)
            {
                ok_slots.push_back(pslots[k]);
                ok_pos.push_back(k);
            }
        }
        std::vector<int> rc_admit(ok_slots.size(), SH_OK);
        int fatal_rc = SH_OK;
        if (!ok_slots.empty())
            fatal_rc = engine->admit_many(ok_slots.data(), uint32_t(ok_slots.size()), s_admit, rc_admit.data());
        for (size_t j = 0; j < ok_pos.size(); ++j)
            rcs[ok_pos[j]] = fatal_rc != SH_OK ? fatal_rc : rc_admit[j];
        lk.lock();
        for (size_t k = 0; k < picked.size(); ++k)
        {
            Request &rq = reqs[picked[k]];
            StreamState &ss = streams[rq.stream];
            if (rcs[k] != SH_OK)
            {
                rq.status = rcs[k];
                rq.done = true;
                rq.t_done = now_ms();
                ss.busy = false;
                free_slots.push_back(pslots[k]);
                stats.completed += 1;
                stats.failed += 1;
                cv_done.notify_all();
                continue;
            }
            slot_req[pslots[k]] = picked[k];
            active.push_back(pslots[k]);
        }
        cv_step.notify_one();
    }
}

void sh_server::step_loop()
{
    std::unique_lock<std::mutex> lk(mu);
    std::vector<uint32_t> batch;
    for (;;)
    {
        while (!stop && active.empty())
            cv_step.wait(lk);
        if (stop)
            return;
        batch = active;
        lk.unlock();
        const int rc = engine->step(batch.data(), uint32_t(batch.size()), s_step);
        // retire the finished solves: copy the warp field out, free the slot
        std::vector<uint32_t> finished;
        if (rc == SH_OK)
            for (uint32_t slot : batch)
                if (engine->solved(slot))
                    finished.push_back(slot);
        int copy_rc = SH_OK;
        for (uint32_t slot : finished)
        {
            std::unique_lock<std::mutex> l2(mu);
            const Request &rq = reqs[slot_req[slot]];
            const StreamState &ss = streams[rq.stream];
            sh_dual_quat *out = rq.out;
            sh_dual_quat *warp = ss.warp;
            const size_t nb = size_t(ss.num_nodes) * sizeof(sh_dual_quat);
            l2.unlock();
            const void *res = engine->result(slot);
            if (copy_rc == SH_OK)
                copy_rc = hip_status(hipMemcpyAsync(warp, res, nb, hipMemcpyDeviceToDevice, s_step));
            if (copy_rc == SH_OK && out)
                copy_rc = hip_status(hipMemcpyAsync(out, res, nb, hipMemcpyDeviceToDevice, s_step));
            if (copy_rc == SH_OK && ss.ref_v)
            {
                // roll the resident surfels and nodes forward with the new warp field
                copy_rc = hip_status(sh::launch_forward_warp(
                    ss.K, reinterpret_cast<const float4 *>(warp), reinterpret_cast<const float4 *>(ss.ref_v),
                    reinterpret_cast<const float4 *>(ss.ref_n), ss.surfel_knn, ss.surfel_spatial, ss.surfel_connect,
                    ss.num_surfels, reinterpret_cast<const float4 *>(ss.node_coords), ss.node_knn, ss.node_spatial,
                    ss.node_connect, ss.num_nodes, reinterpret_cast<float4 *>(ss.live_v),
                    reinterpret_cast<float4 *>(ss.live_n), reinterpret_cast<float4 *>(ss.live_node), s_step));
            }
        }
        if (!finished.empty() && copy_rc == SH_OK)
        {
            // a polled event: a blocking synchronisation wakes tens of
            // microseconds after the copies end
            copy_rc = hip_status(hipEventRecord(ev_step, s_step));
            hipError_t q = hipErrorNotReady;
            while (copy_rc == SH_OK && (q = hipEventQuery(ev_step)) == hipErrorNotReady)
                __builtin_ia32_pause();
            if (copy_rc == SH_OK)
                copy_rc = hip_status(q);
        }
        lk.lock();
        stats.steps += 1;
        stats.max_batch = std::max<uint64_t>(stats.max_batch, batch.size());
        stats.slot_iterations += batch.size();
        if (rc != SH_OK || copy_rc != SH_OK)
        {
            // a failed launch fails every solve that was in it
            fatal = rc != SH_OK ? rc : copy_rc;
            finished = batch;
        }
        const double t = now_ms();
        for (uint32_t slot : finished)
        {
            Request &rq = reqs[slot_req[slot]];
            rq.status = (rc != SH_OK) ? rc : copy_rc;
            rq.done = true;
            rq.t_done = t;
            StreamState &ss = streams[rq.stream];
            ss.busy = false;
            ss.solved += rq.status == SH_OK ? 1 : 0;
            stats.completed += 1;
            stats.failed += rq.status == SH_OK ? 0 : 1;
            active.erase(std::find(active.begin(), active.end(), slot));
            free_slots.push_back(slot);
        }
        if (!finished.empty())
        {
            cv_done.notify_all();
            cv_admit.notify_one();
        }
    }
}

extern "C"
{
    int sh_server_create(const sh_server_config *config, sh_server **out)
    {
        if (!config || !out || config->max_active == 0)
            return SH_E_ARG;
        *out = nullptr;
        sh_server *s = new sh_server();
        s->cfg = *config;
        if (config->hbm_budget_bytes == SH_HBM_BUDGET_FREE)
        {
            // admission by free HBM: what the device has free now, less 1/16 headroom
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
            {
                delete s;
                return SH_E_HIP;
            }
            s->cfg.hbm_budget_bytes = uint64_t(free_b) - uint64_t(free_b) / 16;
        }
        // solve slots: as many as the budget holds, at most max_active
        uint32_t n = config->max_active;
        int rc = SH_E_CAPACITY;
        while (n > 0)
        {
            s->engine.reset(new sh::Engine());
            rc = s->engine->init(config->solver, n);
            if (rc == SH_OK &&
                (s->cfg.hbm_budget_bytes == 0 || s->engine->device_bytes() <= s->cfg.hbm_budget_bytes))
                break;
            s->engine.reset();
            if (rc != SH_OK && rc != SH_E_NOMEM)
                break;
            n = n > 1 ? n * 3 / 4 : 0; // does not fit: fewer slots
            rc = SH_E_CAPACITY;
        }
        if (rc != SH_OK)
        {
            delete s;
            return rc;
        }
        // Index builds (the admission thread, on several streams) run while
        // GN steps run.  The persistent PCG kernels need all their groups
        // resident at once, which work on other streams can delay without
        // bound, so the server's PCG is the split one (kernels with no
        // cross-group waits, replayed from a hipGraph).
        {
            sh_tuning t = s->engine->tuning();
            t.flags = (t.flags & ~uint32_t(SH_FLAG_STREAMING_PCG | SH_FLAG_PIPELINED_PCG)) | SH_FLAG_NO_PERSISTENT_PCG;
            if (t.spmv_mode == SH_SPMV_AUTO)
                t.spmv_mode = SH_SPMV_MATERIALIZED;
            if ((rc = s->engine->set_tuning(t)) != SH_OK)
            {
                delete s;
                return rc;
            }
        }
        s->reserved = s->engine->device_bytes();
        s->engine->set_growth_hook(&server_grow, s);
        s->stats.slots = n;
        s->stats.hbm_budget = s->cfg.hbm_budget_bytes;
        s->stats.hbm_reserved = s->reserved;
        s->slot_req.assign(n, 0);
        for (uint32_t i = 0; i < n; ++i)
            s->free_slots.push_back(n - 1 - i);
        if (hipStreamCreateWithFlags(&s->s_admit, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&s->s_step, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_step, hipEventDisableTiming) != hipSuccess)
        {
            delete s;
            return SH_E_HIP;
        }
        s->t0 = clk::now();
        s->th_admit = std::thread(&sh_server::admit_loop, s);
        s->th_step = std::thread(&sh_server::step_loop, s);
        *out = s;
        return SH_OK;
    }

    int sh_server_destroy(sh_server *s)
    {
        if (!s)
            return SH_OK;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_admit.notify_all();
        s->cv_step.notify_all();
        if (s->th_admit.joinable())
            s->th_admit.join();
        if (s->th_step.joinable())
            s->th_step.join();
        if (s->s_admit)
            (void)hipStreamSynchronize(s->s_admit);
        if (s->s_step)
            (void)hipStreamSynchronize(s->s_step);
        for (auto &kv : s->streams)
            if (kv.second.buf)
                (void)hipFree(kv.second.buf);
        if (s->s_admit)
            (void)hipStreamDestroy(s->s_admit);
        if (s->s_step)
            (void)hipStreamDestroy(s->s_step);
        if (s->ev_step)
            (void)hipEventDestroy(s->ev_step);
        delete s;
        return SH_OK;
    }

    int sh_server_add_stream(sh_server *s, uint32_t stream_id, const sh_camera *camera, const sh_geometry *geometry,
                             const sh_node_graph *node_graph, const float cam2world[16],
                             const float *reference_vertex4, const float *reference_normal4)
    {
        if ((reference_vertex4 == nullptr) != (reference_normal4 == nullptr) ||
            (reference_vertex4 && !node_graph->node_knn))
            return SH_E_ARG;
        if (!s || !camera || !geometry || !node_graph || !cam2world || !geometry->surfel_knn ||
            !geometry->surfel_knn_spatial_weight || !geometry->surfel_knn_connect_weight ||
            !node_graph->reference_node_coords || !node_graph->node_graph || !node_graph->node_knn_connect_weight)
            return SH_E_ARG;
        const sh_solver_config &c = s->cfg.solver;
        if (camera->width > c.max_width || camera->height > c.max_height || node_graph->num_nodes > c.max_nodes ||
            node_graph->num_nodes == 0 || node_graph->num_edges > node_graph->num_nodes * c.knn)
            return SH_E_CAPACITY;
        StreamState ss;
        ss.id = stream_id;
        ss.cam = *camera;
        std::memcpy(ss.cam2world, cam2world, sizeof(ss.cam2world));
        ss.K = c.knn;
        ss.num_surfels = geometry->num_surfels;
        ss.num_nodes = node_graph->num_nodes;
        ss.num_edges = node_graph->num_edges;
        ss.node_radius_square = node_graph->node_radius_square;
        const size_t S = ss.num_surfels, N = ss.num_nodes, E = ss.num_edges, K = ss.K;
        auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t b_sk = up(S * K * 2), b_sw = up(S * K * 4), b_nc = up(N * 16), b_e = up(E * 6),
                     b_nw = up(N * K * 4), b_nk = up(N * K * 2), b_w = up(N * sizeof(sh_dual_quat));
        const bool surf = reference_vertex4 != nullptr;
        const size_t b_sv = surf ? up(S * 16) : 0, b_node = surf ? up(N * 16) : 0;
        ss.bytes = b_sk + 2 * b_sw + b_nc + b_e + 2 * b_nw + b_nk + b_w + 4 * b_sv + b_node;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->streams.count(stream_id))
                return SH_E_STATE;
            if (s->cfg.hbm_budget_bytes && s->reserved + ss.bytes > s->cfg.hbm_budget_bytes)
            {
                s->stats.streams_refused += 1;
                return SH_E_CAPACITY; // admission by HBM budget
            }
            s->reserved += ss.bytes;
        }
        if (hipMalloc(&ss.buf, ss.bytes) != hipSuccess)
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->reserved -= ss.bytes;
            return SH_E_NOMEM;
        }
        char *p = static_cast<char *>(ss.buf);
        ss.surfel_knn = reinterpret_cast<uint16_t *>(p);
        ss.surfel_spatial = reinterpret_cast<float *>(p += b_sk);
        ss.surfel_connect = reinterpret_cast<float *>(p += b_sw);
        ss.node_coords = reinterpret_cast<float *>(p += b_sw);
        ss.edges = reinterpret_cast<uint16_t *>(p += b_nc);
        ss.node_connect = reinterpret_cast<float *>(p += b_e);
        ss.node_spatial = reinterpret_cast<float *>(p += b_nw);
        ss.node_knn = reinterpret_cast<uint16_t *>(p += b_nw);
        ss.warp = reinterpret_cast<sh_dual_quat *>(p += b_nk);
        if (surf)
        {
            ss.ref_v = reinterpret_cast<float *>(p += b_w);
            ss.ref_n = reinterpret_cast<float *>(p += b_sv);
            ss.live_v = reinterpret_cast<float *>(p += b_sv);
            ss.live_n = reinterpret_cast<float *>(p += b_sv);
            ss.live_node = reinterpret_cast<float *>(p += b_sv);
        }
        hipStream_t st = s->s_admit;
        int rc = SH_OK;
        auto cp = [&](void *dst, const void *src, size_t bytes) {
            if (rc == SH_OK && src)
                rc = hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
        };
        cp(ss.surfel_knn, geometry->surfel_knn, S * K * 2);
        cp(ss.surfel_spatial, geometry->surfel_knn_spatial_weight, S * K * 4);
        cp(ss.surfel_connect, geometry->surfel_knn_connect_weight, S * K * 4);
        cp(ss.node_coords, node_graph->reference_node_coords, N * 16);
        cp(ss.edges, node_graph->node_graph, E * 6);
        cp(ss.node_connect, node_graph->node_knn_connect_weight, N * K * 4);
        cp(ss.node_spatial, node_graph->node_knn_spatial_weight, N * K * 4);
        cp(ss.node_knn, node_graph->node_knn, N * K * 2);
        if (surf)
        {
            cp(ss.ref_v, reference_vertex4, S * 16);
            cp(ss.ref_n, reference_normal4, S * 16);
            cp(ss.live_v, reference_vertex4, S * 16); // live = reference until the first solve
            cp(ss.live_n, reference_normal4, S * 16);
            cp(ss.live_node, node_graph->reference_node_coords, N * 16);
        }
        if (rc == SH_OK)
            rc = hip_status(hipMemsetAsync(ss.warp, 0, N * sizeof(sh_dual_quat), st));
        if (rc == SH_OK)
            rc = hip_status(hipStreamSynchronize(st));
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc != SH_OK)
        {
            (void)hipFree(ss.buf);
            s->reserved -= ss.bytes;
            return rc;
        }
        s->streams.emplace(stream_id, ss);
        s->stats.streams += 1;
        s->stats.hbm_reserved = s->reserved;
        return SH_OK;
    }

    int sh_server_submit(sh_server *s, uint32_t stream_id, const sh_measure_maps *measure,
                         const sh_render_maps *render, const sh_flow_map *flow, const sh_node_flow *node_flow,
                         sh_dual_quat *out_device, uint64_t *ticket)
    {
        if (!s || !measure || !render || !ticket)
            return SH_E_ARG;
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->streams.count(stream_id))
            return SH_E_STATE;
        if (s->fatal != SH_OK)
            return s->fatal;
        Request rq;
        rq.ticket = s->next_ticket++;
        rq.stream = stream_id;
        rq.meas = *measure;
        rq.rend = *render;
        if (flow)
            rq.flow = *flow;
        if (node_flow)
            rq.nflow = *node_flow;
        rq.out = out_device;
        rq.t_submit = s->now_ms();
        s->reqs.emplace(rq.ticket, rq);
        s->waiting.push_back(rq.ticket);
        s->stats.submitted += 1;
        *ticket = rq.ticket;
        s->cv_admit.notify_one();
        return SH_OK;
    }

    int sh_server_wait(sh_server *s, uint64_t ticket, sh_request_result *result)
    {
        if (!s)
            return SH_E_ARG;
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end())
            return SH_E_ARG;
        s->cv_done.wait(lk, [&] { return it->second.done; });
        const Request &rq = it->second;
        if (result)
        {
            result->ticket = rq.ticket;
            result->stream_id = rq.stream;
            result->status = rq.status;
            result->submit_ms = rq.t_submit;
            result->admit_ms = rq.t_admit;
            result->done_ms = rq.t_done;
        }
        const int st = rq.status;
        s->reqs.erase(it);
        return st;
    }

    int sh_server_warp_field(sh_server *s, uint32_t stream_id, const sh_dual_quat **device_view, uint32_t *num_nodes)
    {
        if (!s || !device_view || !num_nodes)
            return SH_E_ARG;
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->streams.find(stream_id);
        if (it == s->streams.end())
            return SH_E_STATE;
        *device_view = it->second.warp;
        *num_nodes = it->second.num_nodes;
        return SH_OK;
    }

    int sh_server_live_surfels(sh_server *s, uint32_t stream_id, const float **live_vertex4,
                               const float **live_normal4, const float **live_node4, uint32_t *num_surfels)
    {
        if (!s || !live_vertex4 || !live_normal4 || !live_node4 || !num_surfels)
            return SH_E_ARG;
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->streams.find(stream_id);
        if (it == s->streams.end() || !it->second.ref_v)
            return SH_E_STATE;
        *live_vertex4 = it->second.live_v;
        *live_normal4 = it->second.live_n;
        *live_node4 = it->second.live_node;
        *num_surfels = it->second.num_surfels;
        return SH_OK;
    }

    int sh_server_phase_times(sh_server *s, double *ms, uint32_t *calls, uint32_t num_phases, uint32_t reset)
    {
        if (!s)
            return SH_E_ARG;
        return s->engine->phase_times(ms, calls, num_phases, reset != 0);
    }

    int sh_server_get_stats(sh_server *s, sh_server_stats *out)
    {
        if (!s || !out)
            return SH_E_ARG;
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->stats;
        out->hbm_reserved = s->reserved;
        out->now_ms = s->now_ms();
        return SH_OK;
    }
}
