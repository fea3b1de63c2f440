"""GPU parity of the flow input and the node-motion prior (through the C-ABI)
against the oracle: the decoded, sampled flow map bit for bit; the node-motion
prior bit for bit (its sums are exact integers, so the order the GPU's atomics
land in cannot change them), and the same result on a second run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    return torch


@pytest.mark.parametrize("scale", [1.0, 0.5, 1 / 3])
def test_flow_map_from_of16(torch_dev, scale):
    import sho
    from starhub_amd import native
    torch = torch_dev
    rng = np.random.default_rng(1)
    flow = rng.uniform(-60, 60, (480, 640, 2))
    img = np.concatenate([flow, np.zeros((480, 640, 1))], 2)
    enc = ((img + 1000.0) * 10.0).astype(np.uint16)
    got = native.flow_map_from_of16(torch.from_numpy(enc.view(np.int16)).cuda(), scale).cpu().numpy()
    assert np.array_equal(got, sho.flow_map_from_of16(enc, scale))


def test_node_motion_prior(torch_dev):
    import sho
    from starhub_amd import native
    torch = torch_dev
    rng = np.random.default_rng(5)
    H, W, S, N = 240, 320, 60000, 900
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
    t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else a.view(dt))).cuda()
    args = (t(P), t(C), t(F), t(I, np.int32), t(knn, np.int16), t(sw), M, 280.0, 280.0, 159.5, 119.5, N)
    a = native.node_motion_from_flow(*args).cpu().numpy()
    b = native.node_motion_from_flow(*args).cpu().numpy()
    ref = sho.node_motion_from_flow(P, C, F, I, knn, sw, M, 280.0, 280.0, 159.5, 119.5, N)
    print("nodes with motion", int((ref[:, 3] != 0).sum()), flush=True)
    assert np.array_equal(a, b)
    assert np.array_equal(a, ref)
