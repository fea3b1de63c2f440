"""The synthetic workload generator: determinism and the workload shapes."""
import numpy as np

from starhub_amd import scene


def test_deterministic_for_a_seed():
    cfg = scene.SceneConfig(3000, 40, 64, 48, 60.0, 60.0, with_flow=True, seed=7)
    a, b = scene.generate(cfg), scene.generate(cfg)
    for f in ("surfel_pos", "nodes", "gt_dq", "ref_vertex", "index_map", "measure_vertex", "measure_normal", "flow"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    c = scene.generate(scene.SceneConfig(3000, 40, 64, 48, 60.0, 60.0, with_flow=True, seed=8))
    assert not np.array_equal(a.surfel_pos, c.surfel_pos)


def test_c1_shape():
    sc = scene.generate(scene.config_c1())
    H, W = sc.index_map.shape
    assert (W, H) == (320, 240)
    valid = sc.index_map != 0xFFFFFFFF
    # silhouette radius ~0.4 H
    assert 0.25 < valid.mean() < 0.5
    n = np.linalg.norm(sc.surfel_normal[:, :3], axis=1)
    assert np.allclose(n, 1, atol=1e-5)
    # ground truth: rotations <= 5 deg, translations (about node centres) <= 1 cm
    q0 = sc.gt_dq[:, :4].astype(np.float64)
    ang = 2 * np.degrees(np.arccos(np.clip(np.abs(q0[:, 0]), 0, 1)))
    assert ang.max() <= 5.0 + 1e-3
    # every surfel seen in the reference map has a matching index
    idx = sc.index_map[valid]
    assert np.allclose(sc.ref_vertex[valid][:, :3], sc.surfel_pos[idx, :3])


def test_flow_mask_fraction():
    cfg = scene.SceneConfig(20000, 64, 160, 120, 150.0, 150.0, with_flow=True, seed=3)
    sc = scene.generate(cfg)
    valid = sc.index_map != 0xFFFFFFFF
    zero = np.all(sc.flow[valid] == 0, axis=1)
    assert 0.25 < zero.mean() < 0.35


def test_batch_workloads():
    c4 = scene.config_c4()
    assert len(c4) == 32
    ns = np.array([c.num_nodes for c in c4])
    assert ns.min() >= 128 and ns.max() <= 4096
    assert all(c.num_surfels == 300 * c.num_nodes for c in c4)
    c5 = scene.config_c5_streams()
    assert len(c5) == 64
    assert min(c.num_nodes for c in c5) >= 256 and max(c.num_nodes for c in c5) <= 4096
