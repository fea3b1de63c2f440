"""The allocation-free entry point (sh_node_motion_from_flow_noalloc) against
the workspace entry point it mirrors (sh_node_motion_from_flow), through the
C-ABI: bit for bit on two calls in a row, the second with a smaller input
(fewer nodes and a smaller image), so the shared device array is re-cleared
and reused; and a node count above the array's maximum is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(rng, H, W, S, N):
    f32 = np.float32
    depth_p = rng.uniform(0.5, 1.5, (H, W)).astype(f32)
    P = np.zeros((H, W, 4), f32)
    P[..., 3] = f32(1) / depth_p
    C = np.zeros((H, W, 4), f32)
    C[..., 3] = f32(1) / (depth_p + rng.normal(0, 0.02, (H, W))).astype(f32)
    C[rng.random((H, W)) < 0.05, 3] = 0
    F = rng.normal(0, 3, (H, W, 2)).astype(f32)
    I = rng.integers(0, S, (H, W)).astype(np.uint32)
    I[rng.random((H, W)) < 0.3] = 0xFFFFFFFF
    knn = rng.integers(0, N, (S, 8)).astype(np.uint16)
    sw = rng.uniform(0.05, 1, (S, 8)).astype(f32)
    M = np.eye(4, dtype=f32)
    M[:3, :3] = np.linalg.qr(rng.standard_normal((3, 3)))[0].astype(f32)
    return P, C, F, I, knn, sw, M


def test_noalloc_node_motion_equals_workspace_entry_point():
    import torch
    from starhub_amd import native
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    rng = np.random.default_rng(17)
    t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else a.view(dt))).cuda()
    for (H, W, S, N) in ((240, 320, 60000, 4096), (120, 160, 9000, 300)):
        print("noalloc node motion", H, W, S, N, flush=True)
        P, C, F, I, knn, sw, M = _case(rng, H, W, S, N)
        args = (t(P), t(C), t(F), t(I, np.int32), t(knn, np.int16), t(sw), M, 280.0, 280.0, W / 2 - 0.5,
                H / 2 - 0.5, N)
        want = native.node_motion_from_flow(*args).cpu().numpy()
        got = native.node_motion_from_flow_noalloc(*args).cpu().numpy()
        assert int((want[:, 3] != 0).sum()) > N // 2  # most nodes received motion
        assert np.array_equal(got, want)
    with pytest.raises(native.SolverError, match="capacity"):
        native.node_motion_from_flow_noalloc(*args[:-1], 4097)
