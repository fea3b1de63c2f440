"""Pins the oracle's forward warp of surfels and nodes
(SurfelNodeDeformer::ForwardWarpSurfelsAndNodes, SurfelNodeDeformer.cu:13-80)
with closed forms and invariants, not with a restatement of its formula."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import sho


def _dq(R, t):
    """Unit dual quaternion (q0 wxyz, q1 wxyz) of x -> R x + t."""
    x, y, z, w = Rotation.from_matrix(R).as_quat()
    q0 = np.array([w, x, y, z])
    tq = np.array([0.0, *t])
    # q1 = 0.5 * t * q0 (quaternion product)
    a, b = tq, q0
    q1 = 0.5 * np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                         a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                         a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                         a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]])
    return np.concatenate([q0, q1])


@pytest.fixture(scope="module")
def cloud():
    rng = np.random.default_rng(4)
    S, N, K = 300, 24, 8
    nodes = np.c_[rng.uniform(-1, 1, (N, 3)), np.ones(N)].astype(np.float32)
    surf = np.c_[rng.uniform(-1, 1, (S, 3)), np.ones(S)].astype(np.float32)
    normals = rng.standard_normal((S, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    nrm = np.c_[normals, rng.uniform(0.001, 0.01, S)].astype(np.float32)
    knn, w, c = sho.skin(surf, nodes, K, 0.25)
    nk, nw, _, nc = sho.node_graph(nodes, K, 0.25)
    return dict(S=S, N=N, K=K, nodes=nodes, surf=surf, nrm=nrm, knn=knn, w=w, c=c, nk=nk, nw=nw, nc=nc)


def _warp(cl, field, precision=64):
    return sho.forward_warp(field, cl["K"], cl["surf"], cl["nrm"], cl["knn"], cl["w"], cl["c"],
                            cl["nodes"], cl["nk"], cl["nw"], cl["nc"], precision)


def test_identity_field_keeps_everything(cloud):
    field = np.tile([1.0, 0, 0, 0, 0, 0, 0, 0], (cloud["N"], 1))
    lv, ln, lnode = _warp(cloud, field)
    assert np.allclose(lv, cloud["surf"], atol=1e-7)
    assert np.allclose(ln, cloud["nrm"], atol=1e-7)
    assert np.allclose(lnode, cloud["nodes"], atol=1e-7)


def test_one_rigid_motion_moves_everything_rigidly(cloud):
    R = Rotation.from_rotvec([0.3, -0.2, 0.5]).as_matrix()
    t = np.array([0.05, -0.02, 0.1])
    field = np.tile(_dq(R, t), (cloud["N"], 1))
    lv, ln, lnode = _warp(cloud, field)
    P = cloud["surf"][:, :3].astype(np.float64)
    assert np.allclose(lv[:, :3], P @ R.T + t, atol=1e-6)
    assert np.allclose(ln[:, :3], cloud["nrm"][:, :3].astype(np.float64) @ R.T, atol=1e-6)
    assert np.allclose(lnode[:, :3], cloud["nodes"][:, :3].astype(np.float64) @ R.T + t, atol=1e-6)
    assert np.array_equal(lv[:, 3], cloud["surf"][:, 3]) and np.array_equal(ln[:, 3], cloud["nrm"][:, 3])


def test_each_node_dq_is_normalised_before_the_blend(cloud):
    """Scaling node DQs by different positive factors changes nothing; a blend
    without the per-node normalisation would move toward the larger ones."""
    rng = np.random.default_rng(5)
    field = np.stack([_dq(Rotation.from_rotvec(rng.normal(0, 0.3, 3)).as_matrix(), rng.normal(0, 0.05, 3))
                      for _ in range(cloud["N"])])
    base = _warp(cloud, field)
    scaled = field * rng.uniform(0.3, 3.0, (cloud["N"], 1))
    again = _warp(cloud, scaled)
    for a, b in zip(base, again):
        assert np.allclose(a, b, atol=1e-10)


def test_symmetric_rotations_cancel():
    """A surfel weighted equally between +theta and -theta about z stays put."""
    nodes = np.array([[1, 0, 0, 1], [-1, 0, 0, 1]], np.float32)
    surf = np.array([[0.3, 0.4, 0.5, 1.0]], np.float32)
    nrm = np.array([[0.0, 0.6, 0.8, 0.01]], np.float32)
    knn = np.array([[0, 1, 0, 1]], np.uint16)
    w = np.full((1, 4), 0.25, np.float32)
    c = np.ones((1, 4), np.float32)
    th = 0.4
    field = np.stack([_dq(Rotation.from_rotvec([0, 0, th]).as_matrix(), np.zeros(3)),
                      _dq(Rotation.from_rotvec([0, 0, -th]).as_matrix(), np.zeros(3))])
    lv, ln, _ = sho.forward_warp(field, 4, surf, nrm, knn, w, c)
    assert np.allclose(lv, surf, atol=1e-12) and np.allclose(ln, nrm, atol=1e-12)


def test_fp32_restatement_close_to_fp64(cloud):
    rng = np.random.default_rng(6)
    field = np.stack([_dq(Rotation.from_rotvec(rng.normal(0, 0.2, 3)).as_matrix(), rng.normal(0, 0.03, 3))
                      for _ in range(cloud["N"])])
    for a, b in zip(_warp(cloud, field, 64), _warp(cloud, field, 32)):
        assert np.abs(a - b).max() < 2e-6
