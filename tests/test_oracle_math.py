"""Pins the oracle's scalar / SE(3) algebra against closed forms and library routines."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import sho


def test_round_to_half_matches_ieee_binary16():
    # numpy's float16 conversion is IEEE round-to-nearest-even
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.uniform(0, 1, 4000), 10 ** rng.uniform(-9, 4.8, 4000), rng.uniform(0, 6.2e-5, 2000),
        -(10 ** rng.uniform(-6, 4, 1000)), [0.0, 65504.0, 65519.99, 65520.0, 1e9, 6.103515625e-05, 5.96e-8, 2.98e-8],
    ]).astype(np.float32)
    # exact halfway points between consecutive halves
    h = np.arange(1, 2000, dtype=np.float16).astype(np.float32)
    vals = np.concatenate([vals, (h + np.float32(0.5 ** 11) * h).astype(np.float32)])
    got = np.array([sho.round_to_half(float(v)) for v in vals], np.float32)
    want = vals.astype(np.float16).astype(np.float32)
    assert np.array_equal(got, want)


def test_float2uint_rn_saturating():
    cases = {0.4: 0, 0.5: 0, 1.5: 2, 2.5: 2, 3.5: 4, -0.3: 0, -7.0: 0, 65535.6: 65536, 4.3e9: 0xFFFFFFFF}
    for v, want in cases.items():
        assert sho.float2uint_rn(v) == want, v
    assert sho.float2uint_rn(float("nan")) == 0


def _quat_to_dq(q_wxyz, t):
    w, x, y, z = q_wxyz
    tw, tx, ty, tz = 0.0, 0.5 * t[0], 0.5 * t[1], 0.5 * t[2]
    # (0, t/2) * q
    d = [tw * w - tx * x - ty * y - tz * z, tw * x + tx * w + ty * z - tz * y,
         tw * y - tx * z + ty * w + tz * x, tw * z + tx * y - ty * x + tz * w]
    return np.array([w, x, y, z, *d])


def test_dq_identity_translation_rotation_closed_forms():
    m = sho.dq_se3(np.array([1, 0, 0, 0, 0, 0, 0, 0.0]))
    assert np.allclose(m, np.hstack([np.eye(3), np.zeros((3, 1))]))
    t = np.array([0.3, -0.2, 0.7])
    m = sho.dq_se3(_quat_to_dq([1, 0, 0, 0], t))
    assert np.allclose(m[:, :3], np.eye(3)) and np.allclose(m[:, 3], t)
    rng = np.random.default_rng(1)
    for _ in range(50):
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        theta = rng.uniform(-3, 3)
        q = np.concatenate([[math.cos(theta / 2)], math.sin(theta / 2) * axis])
        t = rng.normal(size=3)
        scale = rng.uniform(0.2, 5.0)  # se3_matrix normalises
        m = sho.dq_se3(scale * _quat_to_dq(q, t))
        R = Rotation.from_rotvec(theta * axis).as_matrix()
        assert np.allclose(m[:, :3], R, atol=1e-12)
        assert np.allclose(m[:, 3], t, atol=1e-12)
        # round trip through the mat34 constructor (dq_from_mat34), up to sign
        d = sho.dq_from_mat34(m)
        m2 = sho.dq_se3(d)
        assert np.allclose(m2, m, atol=1e-12)


def test_dq_from_mat34_all_trace_branches():
    for rv in ([0.1, 0.2, 0.3], [3.0, 0.1, 0.1], [0.1, 3.0, 0.2], [0.2, 0.1, 3.0]):
        R = Rotation.from_rotvec(rv).as_matrix()
        m = np.hstack([R, np.array([[1.0], [2.0], [3.0]])])
        assert np.allclose(sho.dq_se3(sho.dq_from_mat34(m)), m, atol=1e-12)


def test_twist_exp_rodrigues_and_small_angle_cutoff():
    rng = np.random.default_rng(2)
    for _ in range(50):
        w = rng.normal(size=3) * rng.uniform(0.001, 1.0)
        t = rng.normal(size=3)
        m = sho.twist_exp(np.concatenate([w, t]))
        assert np.allclose(m[:, :3], Rotation.from_rotvec(w).as_matrix(), atol=1e-12)
        assert np.allclose(m[:, 3], t)
    # |w|_1 < 1e-4 -> identity rotation (SolverIterationData.cu:42)
    m = sho.twist_exp(np.array([3e-5, -3e-5, 3e-5, 1, 2, 3]))
    assert np.array_equal(m[:, :3], np.eye(3))
    m = sho.twist_exp(np.array([4e-5, -4e-5, 4e-5, 0, 0, 0]))
    assert not np.array_equal(m[:, :3], np.eye(3))


def test_invert6_spd_and_pivoting():
    rng = np.random.default_rng(3)
    for _ in range(20):
        A = rng.normal(size=(6, 6))
        A = A @ A.T + 1e-3 * np.eye(6)
        assert np.allclose(sho.invert6(A) @ A, np.eye(6), atol=1e-8)
    # zero diagonal: needs row exchanges
    P = np.eye(6)[[3, 5, 0, 1, 4, 2]]
    A = P * rng.uniform(1, 2, size=(6, 1))
    assert A[0, 0] == 0.0
    assert np.allclose(sho.invert6(A), np.linalg.inv(A), atol=1e-12)
