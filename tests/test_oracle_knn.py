"""Pins the input producers (surfel skinning, node graph) of the oracle.

Known answers derived by hand from the reference's replace-max heap + Lomuto
quicksort (libs/star/geometry/node_graph/brute_force_knn.cuh:52-135) pin the
tie behaviour; tie-free random inputs are pinned by brute force.
"""
import numpy as np

import sho


def _f4(pts):
    p = np.zeros((len(pts), 4), np.float32)
    p[:, :3] = pts
    p[:, 3] = 1
    return p


def test_skin_tie_known_answer():
    # node 0 far away, nodes 1..8 on the unit cube corners: all at squared
    # distance 0.75 from the surfel.  Node 8 replaces the far node (slot 0), all
    # eight distances tie, and Lomuto's partition rotates the last slot forward.
    corners = [[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)]
    nodes = _f4([[5, 5, 5]] + corners)
    surfel = _f4([[0.5, 0.5, 0.5]])
    knn, w, c = sho.skin(surfel, nodes, 8, 1.0, threads=1)
    assert knn[0].tolist() == [7, 8, 1, 2, 3, 4, 5, 6]
    assert np.allclose(w[0], 1 / 8)
    assert np.all(c == 1)


def test_node_graph_tie_known_answer():
    # node 0 at the origin, five unit neighbours; K = 4: heap keeps nodes 1..3
    # (node 4, 5 tie with the max and are not inserted), Lomuto gives [0, 3, 1, 2].
    nodes = _f4([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0]])
    knn, w, edges, c = sho.node_graph(nodes, 4, 1.0, threads=1)
    assert knn[0].tolist() == [0, 3, 1, 2]
    assert edges[:4].tolist() == [[0, 0, 0], [0, 3, 1], [0, 1, 2], [0, 2, 3]]
    e = np.exp(-np.array([0, 1, 1, 1]) / 2.0)
    assert np.allclose(w[0], e / e.sum(), rtol=1e-6)


def test_skin_matches_brute_force():
    rng = np.random.default_rng(4)
    nodes = _f4(rng.normal(size=(200, 3)))
    surf = _f4(rng.normal(size=(3000, 3)))
    r2 = 0.3
    for K in (4, 8):
        knn, w, _ = sho.skin(surf, nodes, K, r2)
        d = ((surf[:, None, :3].astype(np.float64) - nodes[None, :, :3]) ** 2).sum(-1)
        order = np.argsort(d, axis=1, kind="stable")
        ds = np.take_along_axis(d, order, 1)
        # skip near-ties at the K-th boundary, where FP32 rounding may decide
        gap = (ds[:, K] - ds[:, K - 1]) / ds[:, K]
        ok = gap > 1e-5
        assert ok.mean() > 0.99
        assert np.array_equal(knn[ok], order[ok, :K])
        ww = np.exp(-ds[:, :K] / (2 * r2))
        ww /= ww.sum(1, keepdims=True)
        assert np.allclose(w[ok], ww[ok], rtol=2e-5, atol=1e-7)


def test_node_graph_self_first_and_fp16_distances():
    rng = np.random.default_rng(5)
    nodes = _f4(rng.uniform(-1, 1, size=(300, 3)))
    knn, w, edges, c = sho.node_graph(nodes, 8, 0.2)
    assert np.array_equal(knn[:, 0], np.arange(300))
    d = ((nodes[:, None, :3].astype(np.float64) - nodes[None, :, :3]) ** 2).sum(-1)
    d16 = d.astype(np.float32).astype(np.float16).astype(np.float64)
    for i in range(300):
        got = d16[i, knn[i]]
        assert np.all(np.diff(got) >= 0)  # ascending FP16 distances
        kth = got[-1]
        # every node strictly closer (in FP16) than the K-th is in the list
        assert set(np.nonzero(d16[i] < kth)[0]) <= set(knn[i].tolist())
    assert np.array_equal(edges[:, 0], np.repeat(np.arange(300), 8))
    assert np.array_equal(edges[:, 1], knn.reshape(-1))
    assert np.array_equal(edges[:, 2], np.tile(np.arange(8), 300))
