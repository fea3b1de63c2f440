"""GPU parity of the solver-map rasteriser (sh_render_solver_maps) against the
oracle: index maps bit-exact (integer work), the copied reference vertices and
normals bit-exact, on the synthetic scenes' surfels seen from their cameras
and on random clouds with ties, culled points and the confidence/time test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    return torch


def _compare(torch, V, N, T, M, w, h, fx, fy, cx, cy, maxd, now, recent):
    import sho
    from starhub_amd import native
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda() if a is not None else None
    v, n, idx = native.render_solver_maps(t(V), t(N), t(T), M, w, h, fx, fy, cx, cy, maxd, now, recent)
    rv, rn, ri = sho.render_solver_maps(V, N, T, M, w, h, fx, fy, cx, cy, maxd, now, recent)
    got = idx.cpu().numpy().view(np.uint32)
    print("render", w, h, "points", V.shape[0], "covered", int((ri != 0xFFFFFFFF).sum()), flush=True)
    assert np.array_equal(got, ri)
    assert np.array_equal(v.cpu().numpy(), rv) and np.array_equal(n.cpu().numpy(), rn)


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_render_scene_surfels(torch_dev, cfg):
    from starhub_amd import scene
    sc = scene.generate({"c1": scene.config_c1, "c2": scene.config_c2}[cfg]())
    c = sc.cfg
    M = np.linalg.inv(sc.cam2world.astype(np.float64)).astype(np.float32)
    V = np.concatenate([sc.surfel_pos[:, :3], np.ones((sc.surfel_pos.shape[0], 1))], 1).astype(np.float32)
    N = sc.surfel_normal.astype(np.float32)
    _compare(torch_dev, V, N, None, M, c.width, c.height, c.fx, c.fy, (c.width - 1) / 2.0, (c.height - 1) / 2.0,
             5.0, 0.0, True)


@pytest.mark.parametrize("recent", [True, False])
def test_render_random_cloud(torch_dev, recent):
    rng = np.random.default_rng(7)
    n = 200000
    V = np.zeros((n, 4), np.float32)
    V[:, 0] = rng.uniform(-1.2, 1.2, n)
    V[:, 1] = rng.uniform(-0.9, 0.9, n)
    V[:, 2] = rng.uniform(-0.5, 3.0, n)
    V[:, 3] = rng.choice([0.5, 3.0], n)
    V[1000:1100] = V[17]  # exact duplicates: the lowest index wins
    N = rng.standard_normal((n, 4)).astype(np.float32)
    T = np.zeros((n, 4), np.float32)
    T[:, 2] = rng.integers(0, 12, n)
    M = np.eye(4, dtype=np.float32)
    M[:3, 3] = [0.05, -0.03, 0.4]
    _compare(torch_dev, V, N, T, M, 320, 240, 280.0, 280.0, 159.5, 119.5, 2.5, 8.0, recent)


def test_render_empty(torch_dev):
    from starhub_amd import native
    torch = torch_dev
    z = torch.zeros((0, 4), dtype=torch.float32, device="cuda")
    v, n, idx = native.render_solver_maps(z, z, None, np.eye(4), 16, 8, 10.0, 10.0, 8.0, 4.0, 2.0, 0.0)
    assert bool((idx == -1).all()) and float(v.abs().max()) == 0 and float(n.abs().max()) == 0
