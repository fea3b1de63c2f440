"""The C++ drop-in (include/sh/warp_solver.hpp) on the GPU: a C++ caller
(tests/cpp/dropin_check.cpp, built by the package build) drives
OptimizationProcessorWarpSolver::ProcessFrame at frame 0 and frame 1 and
BlockPCG<6> over ApplySpMVBlockCSR, the way the reference's app and warp
solver call theirs.  Frame 0 returns without solving (the reference starts at
frame 1, OptimizationProcessorWarpSolver.cu:35-36) and leaves the identity
field; frame 1 gives the bits of the C-ABI solve; the PCG gives the bits of the
C-ABI plug-in."""
import subprocess
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "_build" / "dropin_check"


def _write_inputs(path, frame, cfg, params, max_nodes, system):
    f32, u32 = np.float32, np.uint32
    with open(path, "wb") as f:
        w = lambda a, dt: f.write(np.asarray(a, dt).tobytes())
        w([0x49444853], u32)
        w([frame.width, frame.height], u32)
        w([frame.fx, frame.fy, frame.cx, frame.cy], f32)
        w(np.asarray(frame.cam2world, f32).reshape(16), f32)
        w([frame.num_surfels, frame.num_nodes, frame.num_edges, cfg.K], u32)
        w([frame.node_radius_square], f32)
        w([params.gn_iterations, params.pcg_iterations], u32)
        w([params.lambda_depth, params.lambda_flow, params.lambda_reg, params.lambda_node_translation,
           params.damping], f32)
        w([1 if params.fp16 else 0, max_nodes], u32)
        arrays = [frame.measure_vertex, frame.measure_normal, frame.ref_vertex, frame.ref_normal, frame.index_map,
                  frame.flow, frame.surfel_knn, frame.surfel_spatial, frame.surfel_connect, frame.node_coords,
                  frame.edges, frame.node_connect, frame.node_knn, frame.node_spatial]
        for t in arrays:
            b = b"" if t is None else t.contiguous().cpu().numpy().tobytes()
            w([len(b)], np.uint64)
            f.write(b)
        for a in system:
            b = np.ascontiguousarray(a).tobytes()
            w([len(b)], np.uint64)
            f.write(b)


def test_cpp_dropin_frames_and_block_pcg(tmp_path):
    import torch
    import sho
    from block_system import block_csr_of
    from starhub_amd import native, scene, workload
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    assert EXE.exists(), f"{EXE} is missing: run the package build"
    cfg = scene.config_c1()
    sc = scene.generate(cfg)
    frame, prod = workload.device_frame(sc)
    N = cfg.num_nodes
    res = sho.solve(workload.host_inputs(sc, prod), replace(cfg.solver, gn_iterations=1), 64, sho.max_threads())
    rp, col, blk, dinv, b = block_csr_of(res, N)
    system = [rp, col, blk.astype(np.float32), dinv.astype(np.float32), b.astype(np.float32)]
    max_nodes = N + 64
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_inputs(inp, frame, cfg, cfg.solver, max_nodes, system)
    print("running", EXE.name, flush=True)
    r = subprocess.run([str(EXE), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(out, np.uint8)
    ident, n1, nx = raw[:12].view(np.uint32)
    body = raw[12:].view(np.float32)
    field, x = body[:n1 * 8].reshape(n1, 8), body[n1 * 8:]
    assert ident == 1 and n1 == N and nx == 6 * N
    # the same frame through the C-ABI (same capacities)
    s = native.WarpSolver(native.make_config(cfg.solver, frame.width, frame.height, max_nodes, cfg.K))
    s.process_frame(frame)
    assert np.array_equal(field, s.node_se3().cpu().numpy())
    s.close()
    # the same system through the C-ABI plug-in
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32) if a.dtype == np.uint32 else a).cuda()
    A = native.BlockCsr(dev(rp), dev(col), dev(system[2]))
    pcg = native.BlockPCG(6 * N)
    xd = torch.zeros(6 * N, device="cuda")
    assert pcg.set_input(dev(system[3]), A.ops(), dev(system[4]), xd) == 0
    pcg.solve(10)
    assert np.array_equal(x, xd.cpu().numpy())
