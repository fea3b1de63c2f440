"""Pins the flow-input and node-motion oracles (oracle/sho.py):

* the .of.png round trip: the reference's encoder (tools/transfer_hdf5.py:
  (flow + 1000) * 10 truncated to u16) followed by the decode recovers the
  flow to its 0.1-pixel step, and the sampling / scaling follows the closed
  form;
* the node-motion prior against a per-pixel Python loop, and against a rigid
  translation, where every node's averaged motion is the translation itself;
* the quantisation of the exact sums (round half to even, symmetric).
Order independence is the GPU test's: atomics in any order against the
oracle's sequential sums, bit for bit.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import sho  # noqa: E402

f32 = np.float32


def encode_of16(flow):
    """tools/transfer_hdf5.py:42-52: pad a zero channel, (x + 1000) * 10, astype(uint16)."""
    img = np.concatenate([flow, np.zeros(flow.shape[:2] + (1,))], 2)
    return ((img + 1000.0) * 10.0).astype(np.uint16)


def test_of16_round_trip_and_sampling():
    rng = np.random.default_rng(0)
    flow = rng.uniform(-40, 40, (48, 64, 2))
    enc = encode_of16(flow)
    dec = sho.flow_map_from_of16(enc, 1.0)
    assert dec.shape == (48, 64, 2)
    assert np.all(np.abs(dec - flow) <= 0.1 + 1e-4)  # truncation step of the encoder
    half = sho.flow_map_from_of16(enc, 0.5)
    assert half.shape == (24, 32, 2)
    # scaled pixel (x, y) samples raw (2x, 2y) and scales the flow by 0.5
    raw = (enc[::2, ::2, :2].astype(np.float64) * 0.1 - 1000.0).astype(f32)
    assert np.array_equal(half, (raw * f32(0.5)).astype(f32))
    # a 1/3 scale rounds x / scale half to even: x = 1 -> 3, x = 2 -> 6
    third = sho.flow_map_from_of16(enc, 1 / 3)
    assert third.shape == (16, 21, 2)
    expect = (enc[6, 3, :2].astype(np.float64) * 0.1 - 1000.0).astype(f32) * f32(1 / 3)
    assert np.array_equal(third[2, 1], expect.astype(f32))


def _scene(rng, H=30, W=40, S=500, N=20):
    depth_p = rng.uniform(0.8, 1.2, (H, W)).astype(f32)
    depth = (depth_p + rng.normal(0, 0.01, (H, W))).astype(f32)
    P = np.zeros((H, W, 4), f32)
    P[..., 3] = f32(1) / depth_p
    C = np.zeros((H, W, 4), f32)
    C[..., 3] = f32(1) / depth
    C[rng.random((H, W)) < 0.1, 3] = 0  # holes in the current frame
    F = rng.normal(0, 1.5, (H, W, 2)).astype(f32)
    I = rng.integers(0, S, (H, W)).astype(np.uint32)
    I[rng.random((H, W)) < 0.2] = 0xFFFFFFFF
    knn = rng.integers(0, N, (S, 8)).astype(np.uint16)
    sw = rng.uniform(0.1, 1, (S, 8)).astype(f32)
    th = 0.3
    M = np.eye(4, dtype=f32)
    M[:3, :3] = [[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]]
    return P, C, F, I, knn, sw, M


def test_node_motion_matches_loop():
    rng = np.random.default_rng(4)
    P, C, F, I, knn, sw, M = _scene(rng)
    fx = fy = f32(50)
    cx, cy = f32(19.5), f32(14.5)
    got = sho.node_motion_from_flow(P, C, F, I, knn, sw, M, fx, fy, cx, cy, 20)
    sums = np.zeros((20, 4))
    H, W = I.shape
    for y in range(H):
        for x in range(W):
            xf, yf = f32(x) + F[y, x, 0], f32(y) + F[y, x, 1]
            tx, ty = int(np.floor(xf)), int(np.floor(yf))
            cw = C[ty, tx, 3] if 0 <= tx < W and 0 <= ty < H else f32(0)
            s = I[y, x]
            if s == 0xFFFFFFFF or cw == 0:
                continue
            d, dp = f32(1) / cw, f32(1) / P[y, x, 3]
            m = np.array([(xf - cx) / fx * d - (f32(x) - cx) / fx * dp,
                          (yf - cy) / fy * d - (f32(y) - cy) / fy * dp, d - dp], np.float64)
            if abs(m[2]) > 0.05:
                continue
            w = float(sw[s, 0])
            sums[knn[s, 0], :3] += M[:3, :3].astype(np.float64) @ (m * w)
            sums[knn[s, 0], 3] += w
    want = np.where(sums[:, 3:] != 0, sums / np.where(sums[:, 3:] != 0, sums[:, 3:], 1), 0)
    assert np.allclose(got, want, atol=1e-5)


def test_node_motion_rigid_translation():
    """A fronto-parallel plane moving by t in the camera: every node's average is R t."""
    H, W, f, t = 30, 40, 50.0, np.array([0.02, -0.01, 0.03])
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    z0 = 1.0
    X, Y = (xs - 19.5) / f * z0, (ys - 14.5) / f * z0
    u1 = (X + t[0]) / (z0 + t[2]) * f + 19.5
    v1 = (Y + t[1]) / (z0 + t[2]) * f + 14.5
    F = np.stack([u1 - xs, v1 - ys], -1).astype(f32)
    P = np.zeros((H, W, 4), f32)
    P[..., 3] = f32(1 / z0)
    C = np.zeros((H, W, 4), f32)
    C[..., 3] = f32(1 / (z0 + t[2]))
    I = np.arange(H * W, dtype=np.uint32).reshape(H, W)
    knn = (np.arange(H * W) % 7).astype(np.uint16)[:, None].repeat(4, 1)
    sw = np.full((H * W, 4), 0.5, f32)
    M = np.eye(4, dtype=f32)
    got = sho.node_motion_from_flow(P, C, F, I, knn, sw, M, f, f, 19.5, 14.5, 7)
    # the flowed pixel is not on the integer grid: the back-projection uses the flowed
    # float pixel with the current depth, which is the moved point itself
    assert np.allclose(got[:, :3], t, atol=2e-6) and np.allclose(got[:, 3], 1, atol=1e-6)


def test_node_motion_quantisation_is_symmetric():
    """Contributions are quantised with round-half-even, so x and -x cancel
    exactly; a mirrored pair of pixels feeding one node averages to zero
    motion."""
    P = np.zeros((1, 2, 4), f32)
    P[..., 3] = 1
    C = P.copy()
    F = np.array([[[0.25, 0.0], [-0.25, 0.0]]], f32)
    I = np.array([[0, 1]], np.uint32)
    knn = np.zeros((2, 4), np.uint16)
    sw = np.full((2, 4), 0.7, f32)
    got = sho.node_motion_from_flow(P, C, F, I, knn, sw, np.eye(4), 10.0, 10.0, 0.5, 0.0, 1)
    assert got[0, 0] == 0 and got[0, 1] == 0 and got[0, 2] == 0 and got[0, 3] == 1
