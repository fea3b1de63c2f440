// A caller of the C++ drop-in (include/sh/warp_solver.hpp) written the way the
// reference's app calls its optimizer (examples/mono_star/app.cu:132-150) and
// its linear solver (WarpSolverPCG.cu:7-34): test infrastructure for
// tests/test_gpu_dropin.py, which writes the input file and reads the output.
//
//   dropin_check <inputs.bin> <outputs.bin>
//
// 1. ProcessFrame at frame 0 returns without solving: SolvedSE3() is then the
//    solver's initial field, max_nodes identity transforms.
// 2. ProcessFrame at frame 1 solves; SolvedSE3() is written out.
// 3. BlockPCG<6> with an ApplySpMVBlockCSR operator, 10 iterations from zero;
//    x is written out.
#include <hip/hip_runtime_api.h>
#include <sh/warp_solver.hpp>

#include <cstdio>
#include <cstring>
#include <vector>

namespace
{
struct Reader
{
    FILE *f;
    template <class T>
    T get()
    {
        T v{};
        if (std::fread(&v, sizeof(T), 1, f) != 1)
            throw std::runtime_error("truncated input");
        return v;
    }
    // one length-prefixed array, copied to a new device allocation; an empty
    // one is an absent input (NULL)
    void *array(size_t *bytes_out = nullptr)
    {
        const uint64_t n = get<uint64_t>();
        if (bytes_out)
            *bytes_out = n;
        if (n == 0)
            return nullptr;
        std::vector<char> h(n);
        if (std::fread(h.data(), 1, n, f) != n)
            throw std::runtime_error("truncated array");
        void *d = nullptr;
        if (hipMalloc(&d, n) != hipSuccess || hipMemcpy(d, h.data(), n, hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("device copy failed");
        return d;
    }
};
} // namespace

int main(int argc, char **argv)
{
    if (argc != 3)
        return 64;
    try
    {
        FILE *fi = std::fopen(argv[1], "rb");
        if (!fi)
            return 65;
        Reader rd{fi};
        if (rd.get<uint32_t>() != 0x49444853u) // "SHDI"
            return 66;
        sh_solver_config cfg{};
        sh_camera cam{};
        cam.width = rd.get<uint32_t>();
        cam.height = rd.get<uint32_t>();
        cam.fx = rd.get<float>();
        cam.fy = rd.get<float>();
        cam.cx = rd.get<float>();
        cam.cy = rd.get<float>();
        float pose[16];
        for (float &v : pose)
            v = rd.get<float>();
        const uint32_t num_surfels = rd.get<uint32_t>(), num_nodes = rd.get<uint32_t>(),
                       num_edges = rd.get<uint32_t>();
        cfg.knn = rd.get<uint32_t>();
        const float r2 = rd.get<float>();
        cfg.gn_iterations = rd.get<uint32_t>();
        cfg.pcg_iterations = rd.get<uint32_t>();
        cfg.lambda_depth = rd.get<float>();
        cfg.lambda_flow = rd.get<float>();
        cfg.lambda_reg = rd.get<float>();
        cfg.lambda_node_translation = rd.get<float>();
        cfg.lambda_feature = 1.f;
        cfg.damping = rd.get<float>();
        cfg.stream_precision = rd.get<uint32_t>();
        cfg.max_nodes = rd.get<uint32_t>();
        cfg.max_width = cam.width;
        cfg.max_height = cam.height;

        sh::Measure4Solver measure{};
        sh::Render4Solver render{};
        sh::Geometry4Solver geometry{};
        sh::NodeGraph4Solver graph{};
        sh::NodeFlow4Solver nodeflow{};
        sh::OpticalFlow4Solver flow{};
        sh::KeyPoint4Solver keypoints{};
        sh::OptimizationProcessorWarpSolver proc(cfg, cam, pose);

        // frame 0: the reference returns before touching its inputs
        proc.ProcessFrame(measure, render, geometry, graph, nodeflow, flow, keypoints, 0, nullptr);
        const sh::NodeSE3View v0 = proc.SolvedSE3();
        std::vector<float> h0(size_t(v0.size) * 8);
        if (hipMemcpy(h0.data(), v0.ptr, h0.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return 67;
        bool identity = v0.size == cfg.max_nodes;
        for (uint32_t n = 0; n < v0.size; ++n)
            for (int c = 0; c < 8; ++c)
                identity = identity && h0[size_t(n) * 8 + c] == (c == 3 ? 1.f : 0.f);
        std::printf("frame 0: %u identity transforms: %s\n", v0.size, identity ? "yes" : "no");

        measure.vertex_confid = static_cast<const float *>(rd.array());
        measure.normal_radius = static_cast<const float *>(rd.array());
        render.reference_vertex = static_cast<const float *>(rd.array());
        render.reference_normal = static_cast<const float *>(rd.array());
        render.index_map = static_cast<const uint32_t *>(rd.array());
        flow.flow = static_cast<const float *>(rd.array());
        geometry.surfel_knn = static_cast<const uint16_t *>(rd.array());
        geometry.surfel_knn_spatial_weight = static_cast<const float *>(rd.array());
        geometry.surfel_knn_connect_weight = static_cast<const float *>(rd.array());
        geometry.num_surfels = num_surfels;
        graph.reference_node_coords = static_cast<const float *>(rd.array());
        graph.node_graph = static_cast<const uint16_t *>(rd.array());
        graph.node_knn_connect_weight = static_cast<const float *>(rd.array());
        graph.node_knn = static_cast<const uint16_t *>(rd.array());
        graph.node_knn_spatial_weight = static_cast<const float *>(rd.array());
        graph.num_nodes = num_nodes;
        graph.num_edges = num_edges;
        graph.node_radius_square = r2;
        nodeflow.num_nodes = num_nodes;

        proc.ProcessFrame(measure, render, geometry, graph, nodeflow, flow, keypoints, 1, nullptr);
        const sh::NodeSE3View v1 = proc.SolvedSE3();
        std::vector<float> h1(size_t(v1.size) * 8);
        if (hipMemcpy(h1.data(), v1.ptr, h1.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return 68;

        // the linear-solver plug-in on a block-CSR system
        size_t rp_bytes = 0, col_bytes = 0;
        const uint32_t *row_ptr = static_cast<const uint32_t *>(rd.array(&rp_bytes));
        const uint32_t *col = static_cast<const uint32_t *>(rd.array(&col_bytes));
        const float *blocks = static_cast<const float *>(rd.array());
        const float *dinv = static_cast<const float *>(rd.array());
        size_t b_bytes = 0;
        const float *b = static_cast<const float *>(rd.array(&b_bytes));
        std::fclose(fi);
        const uint32_t rows = uint32_t(rp_bytes / 4 - 1), nnz = uint32_t(col_bytes / 4);
        auto op = std::make_shared<sh::ApplySpMVBlockCSR>(rows, nnz, row_ptr, col, blocks);
        float *x = nullptr;
        if (hipMalloc(&x, b_bytes) != hipSuccess)
            return 69;
        sh::BlockPCG pcg(op->MatrixSize());
        pcg.SetSolverInput(dinv, op, b, x);
        const float *xs = pcg.Solve(10, true);
        std::vector<float> hx(b_bytes / 4);
        if (hipMemcpy(hx.data(), xs, b_bytes, hipMemcpyDeviceToHost) != hipSuccess)
            return 70;

        FILE *fo = std::fopen(argv[2], "wb");
        if (!fo)
            return 71;
        const uint32_t head[3] = {identity ? 1u : 0u, v1.size, uint32_t(hx.size())};
        std::fwrite(head, 4, 3, fo);
        std::fwrite(h1.data(), 4, h1.size(), fo);
        std::fwrite(hx.data(), 4, hx.size(), fo);
        std::fclose(fo);
        std::printf("frame 1: %u nodes solved; block PCG: %zu unknowns\n", v1.size, hx.size());
    }
    catch (const sh::Error &e)
    {
        std::printf("sh::Error %d: %s\n", e.status(), e.what());
        return 2;
    }
    catch (const std::exception &e)
    {
        std::printf("error: %s\n", e.what());
        return 3;
    }
    return 0;
}
