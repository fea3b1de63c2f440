"""Pins the oracle of the depth -> measurement-map path (oracle/sh_measure.hpp)
against what the mathematics fixes and against independent numpy
computations:

* the closed-form eigen solver against numpy.linalg.eigh, including the
  repeated-eigenvalue branches;
* window normals of exact planes, the surfel-radius formula in closed form,
  and the 6x6 covariance window (a planted point that only a 7x7 window
  would see);
* the reliability filter's decisions (clip thresholds in float, grazing and
  facing planes, missing neighbours, background value);
* the nearest downsample and the raw vertex formula;
* the fixture's PNG reader on images it encodes itself.
"""
import sys
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests" / "golden"))
import sho  # noqa: E402

IU = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def pack(C):
    return np.stack([C[..., i, j] for i, j in IU], -1)


def cam(w=64, h=48, f=60.0, near=0.1, far=1.1, scale=0.5, rs=0.8):
    return sho.DepthCamera(w, h, f, f, (w - 1) / 2.0, (h - 1) / 2.0, near, far, scale, rs)


@pytest.mark.parametrize("precision,tol", [(64, 1e-7), (32, 1e-6)])
def test_eigen33_matches_eigh(precision, tol):
    """1 - |cos| to the eigh eigenvector of the smallest eigenvalue.  The
    closed form loses accuracy as the two smallest eigenvalues approach each
    other (measured: up to 1.2e-8 in FP64 and 2.2e-7 in FP32 for relative
    gaps above 1e-3); a wrong vector is off by O(1)."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3000, 3, 3))
    C = A @ A.transpose(0, 2, 1)
    w, V = np.linalg.eigh(C)
    gap = (w[:, 1] - w[:, 0]) / w[:, 2]
    keep = gap > 1e-3
    v = sho.eigen33(pack(C), precision)
    cos = np.abs(np.einsum("ni,ni->n", v, V[:, :, 0]))
    assert np.all(1 - cos[keep] <= tol), (1 - cos[keep]).max()
    assert np.allclose(np.linalg.norm(v, axis=1), 1, atol=10 * tol)


def test_eigen33_repeated_eigenvalues():
    R = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))[0]
    # all equal: the reference returns the x axis
    assert np.allclose(sho.eigen33(pack(np.eye(3) * 2.5))[0], [1, 0, 0])
    # two smallest equal: any unit vector of that plane
    C = R @ np.diag([1.0, 1.0, 3.0]) @ R.T
    v = sho.eigen33(pack(C))[0]
    assert abs(np.linalg.norm(v) - 1) < 1e-9 and np.linalg.norm(C @ v - v) < 1e-9
    # two largest equal: the single smallest direction
    C = R @ np.diag([0.5, 2.0, 2.0]) @ R.T
    v = sho.eigen33(pack(C))[0]
    assert abs(abs(v @ R[:, 0]) - 1) < 1e-9


def plane_vertex_map(w, h, n, d, f=60.0):
    """Exact vertices of the plane n.X = d seen by a pinhole (principal point at the centre)."""
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    rays = np.stack([(xs - (w - 1) / 2) / f, (ys - (h - 1) / 2) / f, np.ones_like(xs)], -1)
    t = d / (rays @ n)
    v = np.zeros((h, w, 4), np.float32)
    v[..., :3] = rays * t[..., None]
    v[..., 3] = 1
    return v


def test_plane_normals_and_radius_closed_form():
    n = np.array([0.3, -0.2, -1.0])
    n /= np.linalg.norm(n)
    V = plane_vertex_map(40, 30, n, -0.8)
    out = sho.normal_radius_map(V, 0.8, precision=64)
    inner = out[3:-3, 3:-3]
    assert np.allclose(inner[..., :3], n, atol=1e-5)  # oriented towards the camera: n . v < 0
    z = V[3:-3, 3:-3, 2].astype(np.float64)
    r = z * np.float32(1000 * np.float32(1.414) / np.float32(570.0)) * np.float32(0.8) / abs(n[2])
    assert np.allclose(inner[..., 3], r, rtol=1e-6)
    # a steep plane (|n_z| <= 0.5): twice the base radius
    n2 = np.array([0.95, 0.0, -0.3])
    n2 /= np.linalg.norm(n2)
    V2 = plane_vertex_map(40, 30, n2, -0.5)
    o2 = sho.normal_radius_map(V2, 0.8, precision=64)[10, 10]
    base = np.float32(V2[10, 10, 2]) * np.float32(1000 * np.float32(1.414) / np.float32(570.0)) * np.float32(0.8)
    assert np.isclose(o2[3], 2 * base, rtol=1e-6)
    # the corners see fewer than 25 valid points: no normal
    assert np.all(out[0, 0] == 0) and np.all(out[-1, -1] == 0)


def test_covariance_window_is_six_by_six():
    """The covariance loops stop before +3: a point planted on the window's
    last row changes the centroid (7x7) but enters no covariance product.
    numpy computes that definition directly and with a symmetric 7x7
    covariance; the oracle matches the former and not the latter."""
    rng = np.random.default_rng(11)
    V = plane_vertex_map(21, 21, np.array([0.0, 0.0, -1.0]), -0.7)
    V[..., 2] += rng.normal(0, 2e-3, V.shape[:2]).astype(np.float32)
    x = y = 10
    V[y + 3, x] = [0.05, 0.05, 0.9, 1]  # last row of the 7x7 window
    win7 = V[y - 3:y + 4, x - 3:x + 4, :3].reshape(-1, 3).astype(np.float64)
    c = win7.mean(0)
    six = V[y - 3:y + 3, x - 3:x + 3, :3].reshape(-1, 3).astype(np.float64) - c
    seven = win7 - c

    def normal(D):
        w, E = np.linalg.eigh(D.T @ D)
        v = E[:, 0]
        return -v if v @ V[y, x, :3] >= 0 else v

    o = sho.normal_radius_map(V, 0.8, precision=64)[y, x, :3]
    assert np.allclose(o, normal(six), atol=1e-9)
    assert np.linalg.norm(o - normal(seven)) > 1e-3


def test_filter_depth_decisions():
    c = cam(64, 48)
    ys, xs = np.mgrid[0:48, 0:64]
    # facing plane at 0.5 m: interior kept (exact mm), corners lack neighbours
    raw = np.full((48, 64), 500, np.uint16)
    out, dots = sho.filter_depth(raw, c, 64)
    assert np.all(out[4:-4, 4:-4] == np.float32(0.5))
    assert out[0, 0] == 0
    assert np.all(dots[4:-4, 4:-4] < -0.1)
    # grazing plane: depth x 1.25 per pixel against a lateral step of z / f
    # (slope 0.25 f = 15, |cos| to the rays below 0.1) is rejected; a gentle
    # one (x 1.01 per pixel, slope 0.6) is kept
    xs20 = np.mgrid[0:30, 0:20][1]
    graze = np.round(400 * 1.25 ** xs20).astype(np.uint16)
    out_g, dots_g = sho.filter_depth(graze, cam(20, 30, far=100.0), 64)
    assert np.all(out_g[6:-6, 6:-6] == 0) and np.all(dots_g[6:-6, 6:-6] > -0.1)
    gentle = np.round(400 * 1.01 ** xs20).astype(np.uint16)
    out_k, _ = sho.filter_depth(gentle, cam(20, 30, far=100.0), 64)
    assert np.all(out_k[6:-6, 6:-6] == gentle[6:-6, 6:-6].astype(np.float32) / np.float32(1000))
    # beyond clip_far: the background value; missing depth: 0
    far = np.full((48, 64), 1500, np.uint16)
    assert np.all(sho.filter_depth(far, c, 64)[0] == 10.0)
    assert np.all(sho.filter_depth(np.zeros((48, 64), np.uint16), c, 64)[0] == 0)
    # the near clip compares float(raw) / 1000.f with the float clip: 100 mm at
    # clip_near 0.1 is not "too near" (in double it would be), so it is kept
    near = np.full((48, 64), 100, np.uint16)
    kept = sho.filter_depth(near, cam(64, 48, near=0.1), 64)[0]
    assert np.all(kept[4:-4, 4:-4] == np.float32(100) / np.float32(1000))
    assert np.all(sho.filter_depth(np.full((48, 64), 99, np.uint16), cam(64, 48, near=0.1), 64)[0] == 0)


def test_scaled_vertex_map_formula():
    c = cam(64, 48)
    rng = np.random.default_rng(2)
    filt = rng.uniform(0.2, 1.0, (48, 64)).astype(np.float32)
    filt[5, 6] = 0.05   # nearer than clip_near
    filt[8, 8] = 10.0   # background
    v, _ = sho.measure_maps(filt, c, precision=32)
    assert v.shape == (24, 32, 4)
    f32 = np.float32
    for (y, x) in [(0, 0), (3, 7), (23, 31), (11, 19)]:
        d = filt[2 * y, 2 * x]
        ex = (f32(2 * x) - f32(c.cx)) / f32(c.fx) * d
        ey = (f32(2 * y) - f32(c.cy)) / f32(c.fy) * d
        assert v[y, x, 0] == ex and v[y, x, 1] == ey and v[y, x, 2] == d and v[y, x, 3] == 1
    # raw texel (6, 5) is not sampled (odd); (8, 8) is: scaled (4, 4) is cleared
    assert np.all(v[4, 4] == 0)


def test_vertex_confid_map_formula():
    raw = np.arange(12, dtype=np.uint16).reshape(3, 4) * 100
    v = sho.vertex_confid_map(raw, 500.0, 400.0, 1.5, 1.0)
    f32 = np.float32
    ifx = f32(1.0 / np.float64(f32(500.0)))
    z = f32(raw[2, 3]) / f32(1000)
    assert v[2, 3, 0] == (f32(3) - f32(1.5)) * ifx * z and v[2, 3, 2] == z and v[2, 3, 3] == 1


def _png16(img, filt):
    """A 16-bit greyscale PNG with every scanline under filter type `filt`."""
    h, w = img.shape
    rows = img.astype(">u2").view(np.uint8).reshape(h, 2 * w).astype(np.int32)
    out = bytearray()
    prev = np.zeros(2 * w, np.int32)
    for y in range(h):
        cur = rows[y]
        left = np.concatenate([[0, 0], cur[:-2]])
        ul = np.concatenate([[0, 0], prev[:-2]])
        if filt == 0:
            enc = cur
        elif filt == 1:
            enc = cur - left
        elif filt == 2:
            enc = cur - prev
        elif filt == 3:
            enc = cur - ((left + prev) >> 1)
        else:
            pa, pb, pc = np.abs(prev - ul), np.abs(left - ul), np.abs(left + prev - 2 * ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, ul))
            enc = cur - pred
        out += bytes([filt]) + bytes((enc & 255).astype(np.uint8))
        prev = cur

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body))
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 16, 0, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b""))


@pytest.mark.parametrize("filt", [0, 1, 2, 3, 4])
def test_fixture_png_reader(tmp_path, filt):
    import make_depth_fixture
    img = np.random.default_rng(filt).integers(0, 65535, (7, 9)).astype(np.uint16)
    p = tmp_path / "x.png"
    p.write_bytes(_png16(img, filt))
    assert np.array_equal(make_depth_fixture.read_png_u16(p), img)


def test_depth_fixture_is_the_reference_frame():
    z = np.load(ROOT / "tests" / "golden" / "move_dragon_depth.npz")
    d = z["depth_mm"]
    assert d.shape == (2, 480, 640) and d.dtype == np.uint16
    assert np.allclose(z["intrinsic"], [616.992, 616.955, 314.104, 236.412])
    valid = d[0][d[0] > 0]
    assert 0.5 < valid.size / d[0].size < 1 and 300 < np.median(valid) < 3000
