"""Pins the solver-map rasteriser oracle (oracle/sho.py render_solver_maps)
by brute force on scenes small enough to enumerate: for every pixel the
winner is found by scanning all surfels with the rules the reference's
shader and depth test fix (culling, pixel rounding, nearest depth, lower id
on a tie, clear values)."""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import sho  # noqa: E402

f32 = np.float32


def brute(V, N, T, M, w, h, fx, fy, cx, cy, maxd, now, recent):
    vert = np.zeros((h, w, 4), f32)
    nrm = np.zeros((h, w, 4), f32)
    idx = np.full((h, w), 0xFFFFFFFF, np.uint32)
    best = {}
    for i in range(V.shape[0]):
        x, y, z = (f32(c) for c in V[i, :3])
        p = [((M[r, 0] * x + M[r, 1] * y) + M[r, 2] * z) + M[r, 3] for r in range(3)]
        if p[2] > f32(maxd) or p[2] <= 0:
            continue
        if T is not None and V[i, 3] < 1 and abs(T[i, 2] - f32(now)) >= (f32(3) if recent else f32(-1)):
            continue
        u = int(np.floor(((f32(fx) * p[0]) / p[2] + f32(cx)) + f32(0.5)))
        v = int(np.floor(((f32(fy) * p[1]) / p[2] + f32(cy)) + f32(0.5)))
        if not (0 <= u < w and 0 <= v < h):
            continue
        d = (p[2] / f32(maxd)) * f32(0.5) + f32(0.5)
        if not d < 1:
            continue
        if (v, u) not in best or (d, i) < best[(v, u)]:
            best[(v, u)] = (d, i)
    for (v, u), (d, i) in best.items():
        vert[v, u] = [*V[i, :3], 1]
        nrm[v, u] = [*N[i, :3], 0]
        idx[v, u] = i
    return vert, nrm, idx


def scene(n, seed, w=24, h=18):
    rng = np.random.default_rng(seed)
    V = np.zeros((n, 4), f32)
    V[:, 0] = rng.uniform(-0.6, 0.6, n)
    V[:, 1] = rng.uniform(-0.45, 0.45, n)
    V[:, 2] = rng.uniform(-0.5, 2.5, n)  # some behind, some beyond max depth
    V[:, 3] = rng.choice([0.5, 2.0], n)
    N = rng.standard_normal((n, 4)).astype(f32)
    T = np.zeros((n, 4), f32)
    T[:, 2] = rng.integers(0, 10, n)
    # duplicates at identical positions: the tie goes to the lower index
    V[n // 2:n // 2 + 5] = V[3]
    return V, N, T


def test_render_matches_brute_force():
    M = np.eye(4, dtype=f32)
    M[:3, 3] = [0.01, -0.02, 0.3]
    for seed in range(4):
        V, N, T = scene(400, seed)
        for recent in (True, False):
            args = (M, 24, 18, f32(20.0), f32(20.0), f32(11.5), f32(8.5), f32(2.0), 5.0, recent)
            got = sho.render_solver_maps(V, N, T, *args)
            want = brute(V, N, T, *args)
            for g, e in zip(got, want):
                assert np.array_equal(g, e)


def test_render_rules():
    M = np.eye(4, dtype=f32)
    V = np.array([[0, 0, 1.0, 2], [0, 0, 0.5, 2], [0, 0, 0.5, 2], [0, 0, -1, 2], [0, 0, 3.0, 2],
                  [0.001, 0, 0.8, 0.5]], f32)
    N = np.arange(24, dtype=f32).reshape(6, 4)
    T = np.zeros((6, 4), f32)
    T[5, 2] = 10.0
    v, n, idx = sho.render_solver_maps(V, N, T, M, 8, 8, 10.0, 10.0, 3.5, 3.5, 2.0, 10.0, True)
    # pixel (4, 4): points 0, 1, 2, 5 land there; 1 and 2 are nearest and tie: 1 wins
    assert idx[4, 4] == 1 and np.array_equal(v[4, 4], [0, 0, 0.5, 1]) and np.array_equal(n[4, 4], [4, 5, 6, 0])
    assert (idx != 0xFFFFFFFF).sum() == 1  # behind the camera and beyond max depth are culled
    # low confidence, observed 0 frames ago: kept with recent observations (and nearer than 0 only when 1, 2 go)
    V2 = V[[0, 5]]
    v2, _, idx2 = sho.render_solver_maps(V2, N[[0, 5]], T[[0, 5]], M, 8, 8, 10.0, 10.0, 3.5, 3.5, 2.0, 10.0, True)
    assert idx2[4, 4] == 1
    _, _, idx3 = sho.render_solver_maps(V2, N[[0, 5]], T[[0, 5]], M, 8, 8, 10.0, 10.0, 3.5, 3.5, 2.0, 10.0, False)
    assert idx3[4, 4] == 0
    # a point at exactly max depth fails the depth test against the clear value
    _, _, idx4 = sho.render_solver_maps(np.array([[0, 0, 2.0, 2]], f32), N[:1], None, M, 8, 8, 10.0, 10.0, 3.5,
                                        3.5, 2.0, 0.0, True)
    assert np.all(idx4 == 0xFFFFFFFF)
