// C++ drop-in for the reference's warp-solver call, on top of the C-ABI of
// warp_solver.h.  Mirrors
//   star::OptimizationProcessorWarpSolver::ProcessFrame / SolvedSE3
//     (examples/mono_star/opt/OptimizationProcessorWarpSolver.h:18-30)
//   star::WarpSolver::SetSolverInputs / SolveStreamed / SolvedNodeSE3
//     (examples/mono_star/opt/WarpSolver.h:34-52)
// with the same argument meaning and order.  The reference reports failures by
// exit()/throw (e.g. OptimizationProcessorWarpSolver.cu:51, GBufferArray.h:93-100);
// here every failing C call throws sh::Error carrying the status, so a caller
// that relied on the reference aborting sees an exception instead.
//
// Header-only; link libstarhub_amd.so.  No HIP headers are needed: streams are
// passed as opaque handles (sh_stream_t).
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

#include "warp_solver.h"

namespace sh
{
class Error : public std::runtime_error
{
public:
    Error(int status, const char *what)
        : std::runtime_error(std::string(what) + ": " + sh_status_string(status)), status_(status)
    {
    }
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status, const char *what)
{
    if (status != SH_OK)
        throw Error(status, what);
}

// GArrayView<DualQuaternion> of the reference: device pointer + size.
struct NodeSE3View
{
    const sh_dual_quat *ptr = nullptr;
    uint32_t size = 0;
};

// Solver input bundles (star/opt/solver_types.h:21-120) as the C structs.
using Measure4Solver = sh_measure_maps;
using Render4Solver = sh_render_maps;
using Geometry4Solver = sh_geometry;
using NodeGraph4Solver = sh_node_graph;
using NodeFlow4Solver = sh_node_flow;
using OpticalFlow4Solver = sh_flow_map;
using KeyPoint4Solver = sh_keypoints;

// star::WarpSolver: one solve handle.  Capacity comes from the config at
// construction (the reference's AllocateBuffer, WarpSolver.cu:82-103).
class WarpSolver
{
public:
    using Ptr = std::shared_ptr<WarpSolver>;
    WarpSolver(const sh_solver_config &config, const sh_camera &camera) : camera_(camera)
    {
        check(sh_solver_create(&config, &h_), "sh_solver_create");
    }
    ~WarpSolver() { sh_solver_destroy(h_); }
    WarpSolver(const WarpSolver &) = delete;
    WarpSolver &operator=(const WarpSolver &) = delete;

    void SetSolverInputs(const Measure4Solver &measure4solver, const Render4Solver &render4solver,
                         const Geometry4Solver &geometry4solver, const NodeGraph4Solver &node_graph4solver,
                         const NodeFlow4Solver &nodeflow4solver, const OpticalFlow4Solver &opticalflow4solver,
                         const KeyPoint4Solver &keypoint4solver, const float camera2world[16])
    {
        check(sh_solver_set_inputs(h_, &camera_, &measure4solver, &render4solver, &geometry4solver,
                                   &node_graph4solver, &nodeflow4solver, &opticalflow4solver, &keypoint4solver,
                                   camera2world),
              "sh_solver_set_inputs");
    }
    // Enqueues the whole solve on `stream` and returns after synchronising it.
    void SolveStreamed(sh_stream_t stream = nullptr) { check(sh_solver_solve(h_, stream), "sh_solver_solve"); }
    NodeSE3View SolvedNodeSE3() const
    {
        NodeSE3View v;
        check(sh_solver_node_se3(h_, &v.ptr, &v.size), "sh_solver_node_se3");
        return v;
    }
    sh_solver *handle() const { return h_; }

private:
    sh_solver *h_ = nullptr;
    sh_camera camera_;
};

// star::OptimizationProcessorWarpSolver: keeps the camera-to-world extrinsic
// and forwards one frame to the solver.
class OptimizationProcessorWarpSolver
{
public:
    using Ptr = std::shared_ptr<OptimizationProcessorWarpSolver>;
    OptimizationProcessorWarpSolver(const sh_solver_config &config, const sh_camera &camera,
                                    const float camera2world[16])
        : solver_(std::make_shared<WarpSolver>(config, camera))
    {
        for (int i = 0; i < 16; ++i)
            cam2world_[i] = camera2world[i];
    }

    void ProcessFrame(Measure4Solver &measure4solver, Render4Solver &render4solver, Geometry4Solver &geometry4solver,
                      NodeGraph4Solver &node_graph4solver, NodeFlow4Solver &nodeflow4solver,
                      OpticalFlow4Solver &opticalflow4solver, KeyPoint4Solver &keypoint4solver,
                      const unsigned frame_idx, sh_stream_t stream)
    {
        (void)frame_idx; // the reference only logs it
        solver_->SetSolverInputs(measure4solver, render4solver, geometry4solver, node_graph4solver, nodeflow4solver,
                                 opticalflow4solver, keypoint4solver, cam2world_);
        solver_->SolveStreamed(stream);
    }
    NodeSE3View SolvedSE3() const { return solver_->SolvedNodeSE3(); }
    const WarpSolver::Ptr &solver() const { return solver_; }

private:
    WarpSolver::Ptr solver_;
    float cam2world_[16];
};
} // namespace sh
