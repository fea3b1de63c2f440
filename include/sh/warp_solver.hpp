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
#include <utility>

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

// star::ApplySpMVBase<6> (libs/star/pcg_solver/ApplySpMVBase.h:9-47): the
// operator a BlockPCG applies.  Derive and override ApplySpMV (and InitResidual
// with HasInitResidual() returning true, or leave both and the solver forms
// b - A x through ApplySpMV).  Vectors are device arrays of MatrixSize() floats;
// work is enqueued on `stream`.  An exception thrown by an override is caught
// at the C boundary and the solve returns SH_E_ARG (sh::Error in BlockPCG).
class ApplySpMVBase
{
public:
    using Ptr = std::shared_ptr<ApplySpMVBase>;
    ApplySpMVBase() = default;
    virtual ~ApplySpMVBase() = default;
    ApplySpMVBase(const ApplySpMVBase &) = delete;
    ApplySpMVBase &operator=(const ApplySpMVBase &) = delete;

    virtual size_t MatrixSize() const = 0;
    virtual void ApplySpMV(const float *x, float *spmv_x, sh_stream_t stream) = 0;
    virtual bool HasInitResidual() const { return false; }
    // residual <- b - A x_init
    virtual void InitResidual(const float *x_init, const float *b, float *residual, sh_stream_t stream)
    {
        (void)x_init, (void)b, (void)residual, (void)stream;
        throw Error(SH_E_STATE, "InitResidual is not implemented");
    }

    // The C callback table that forwards to the virtuals above.
    sh_spmv_ops Ops()
    {
        sh_spmv_ops o{};
        o.user = this;
        o.matrix_size = uint32_t(MatrixSize());
        o.apply_spmv = &ApplySpMVBase::apply_cb;
        o.init_residual = HasInitResidual() ? &ApplySpMVBase::init_cb : nullptr;
        return o;
    }

private:
    static int apply_cb(void *user, const float *x, float *y, sh_stream_t stream)
    {
        try
        {
            static_cast<ApplySpMVBase *>(user)->ApplySpMV(x, y, stream);
            return SH_OK;
        }
        catch (const Error &e)
        {
            return e.status();
        }
        catch (...)
        {
            return SH_E_ARG;
        }
    }
    static int init_cb(void *user, const float *x, const float *b, float *r, sh_stream_t stream)
    {
        try
        {
            static_cast<ApplySpMVBase *>(user)->InitResidual(x, b, r, stream);
            return SH_OK;
        }
        catch (const Error &e)
        {
            return e.status();
        }
        catch (...)
        {
            return SH_E_ARG;
        }
    }
};

// star::ApplySpMVBinBlockCSR<6> (libs/star/pcg_solver/ApplySpMVBinBlockCSR.h:8-52)
// over a caller-owned 6x6 block-CSR in this library's layout (sh_spmv_create_block_csr).
class ApplySpMVBlockCSR : public ApplySpMVBase
{
public:
    ApplySpMVBlockCSR(uint32_t num_block_rows, uint32_t num_blocks, const uint32_t *row_ptr, const uint32_t *block_col,
                      const float *blocks)
    {
        check(sh_spmv_create_block_csr(num_block_rows, num_blocks, row_ptr, block_col, blocks, &h_),
              "sh_spmv_create_block_csr");
    }
    ~ApplySpMVBlockCSR() override { sh_spmv_destroy(h_); }
    size_t MatrixSize() const override
    {
        uint32_t n = 0;
        check(sh_spmv_matrix_size(h_, &n), "sh_spmv_matrix_size");
        return n;
    }
    void ApplySpMV(const float *x, float *spmv_x, sh_stream_t stream) override
    {
        check(sh_spmv_apply(h_, x, spmv_x, stream), "sh_spmv_apply");
    }
    bool HasInitResidual() const override { return true; }
    void InitResidual(const float *x_init, const float *b, float *residual, sh_stream_t stream) override
    {
        check(sh_spmv_init_residual(h_, x_init, b, residual, stream), "sh_spmv_init_residual");
    }
    sh_spmv *handle() const { return h_; }

private:
    sh_spmv *h_ = nullptr;
};

// star::BlockPCG<6> (libs/star/pcg_solver/BlockPCG.h:14-115): the same methods
// and argument meaning.  Solve returns the solution view (the caller's x).
class BlockPCG
{
public:
    using Ptr = std::shared_ptr<BlockPCG>;
    explicit BlockPCG(size_t max_matrix_size, sh_stream_t stream = nullptr) : stream_(stream)
    {
        check(sh_block_pcg_create(uint32_t(max_matrix_size), &h_), "sh_block_pcg_create");
    }
    ~BlockPCG() { sh_block_pcg_destroy(h_); }
    BlockPCG(const BlockPCG &) = delete;
    BlockPCG &operator=(const BlockPCG &) = delete;

    // inv_diag_blks: MatrixSize()/6 row-major FP32 6x6 blocks; x_init: the
    // solution vector (read as the start unless zero_init)
    bool SetSolverInput(const float *inv_diag_blks, ApplySpMVBase::Ptr spmv_handler, const float *b, float *x_init,
                        size_t actual_size = 0)
    {
        spmv_ = std::move(spmv_handler);
        ops_ = spmv_->Ops();
        x_ = x_init;
        check(sh_block_pcg_set_input(h_, inv_diag_blks, &ops_, b, x_init, uint32_t(actual_size)),
              "sh_block_pcg_set_input");
        return true;
    }
    void UpdateCudaStream(sh_stream_t stream) { stream_ = stream; }
    const float *Solve(size_t max_iterations = 10, bool zero_init = true)
    {
        return SolveNoConvergeCheck(max_iterations, zero_init);
    }
    const float *SolveNoConvergeCheck(size_t max_iterations, bool zero_init = true)
    {
        check(sh_block_pcg_solve(h_, uint32_t(max_iterations), zero_init ? 1 : 0, stream_), "sh_block_pcg_solve");
        return x_;
    }
    const float *SolveConvergeChecked(size_t max_iterations, bool zero_init = true, float epsilon = 2e-3f)
    {
        uint32_t done = 0;
        check(sh_block_pcg_solve_converge_checked(h_, uint32_t(max_iterations), zero_init ? 1 : 0, epsilon, stream_,
                                                  &done),
              "sh_block_pcg_solve_converge_checked");
        return x_;
    }

private:
    sh_block_pcg *h_ = nullptr;
    sh_stream_t stream_ = nullptr;
    ApplySpMVBase::Ptr spmv_;
    sh_spmv_ops ops_{};
    float *x_ = nullptr;
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
        // the reference starts operating at frame 1: frame 0 returns without
        // solving (OptimizationProcessorWarpSolver.cu:35-36), and SolvedSE3()
        // keeps the solver's field (identity before any solve)
        if (frame_idx == 0)
            return;
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
