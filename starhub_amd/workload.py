"""Moves synthetic scenes onto the device and runs the input producers.

``device_frame`` uploads a :class:`starhub_amd.scene.Scene`, runs the surfel
skinning and node-graph producers on the GPU (``sh_skin_surfels`` /
``sh_build_node_graph``) and returns a :class:`starhub_amd.native.DeviceFrame`
ready for ``WarpSolver.process_frame``.  ``host_inputs`` returns the same
solver inputs as numpy arrays (the layout the CPU restatement consumes).
"""
from __future__ import annotations

import numpy as np

from . import native
from .scene import Scene


def _t(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.view(dtype)
    return t.cuda()


def device_frame(scene: Scene, producers: dict | None = None) -> tuple[native.DeviceFrame, dict]:
    """Upload `scene`.  Producer outputs come from the GPU kernels unless given
    (as numpy arrays: surfel_knn, surfel_spatial, surfel_connect, node_knn,
    node_spatial, edges, node_connect)."""
    import torch
    cfg = scene.cfg
    K = cfg.K
    surfel_pos = _t(scene.surfel_pos)
    nodes = _t(scene.nodes)
    if producers is None:
        nknn, nw, edges, nconn = native.build_node_graph(nodes, K, scene.node_radius_square)
        sknn, sw, sconn = native.skin_surfels(surfel_pos, nodes, K, scene.node_radius_square)
        torch.cuda.synchronize()
    else:
        nknn = _t(producers["node_knn"].astype(np.uint16), torch.int16)
        nw = _t(producers["node_spatial"].astype(np.float32))
        edges = _t(producers["edges"].astype(np.uint16), torch.int16)
        nconn = _t(producers["node_connect"].astype(np.float32))
        sknn = _t(producers["surfel_knn"].astype(np.uint16), torch.int16)
        sw = _t(producers["surfel_spatial"].astype(np.float32))
        sconn = _t(producers["surfel_connect"].astype(np.float32))
    frame = native.DeviceFrame(
        width=cfg.width, height=cfg.height, fx=cfg.fx, fy=cfg.fy, cx=cfg.cx, cy=cfg.cy,
        cam2world=scene.cam2world,
        measure_vertex=_t(scene.measure_vertex), measure_normal=_t(scene.measure_normal),
        ref_vertex=_t(scene.ref_vertex), ref_normal=_t(scene.ref_normal),
        index_map=_t(scene.index_map.view(np.int32)),
        flow=_t(scene.flow) if scene.flow is not None else None,
        surfel_knn=sknn, surfel_spatial=sw, surfel_connect=sconn,
        node_coords=nodes, edges=edges, node_connect=nconn, node_knn=nknn, node_spatial=nw,
        node_radius_square=scene.node_radius_square,
        num_surfels=scene.surfel_pos.shape[0], num_nodes=scene.nodes.shape[0], num_edges=int(edges.shape[0]),
    )
    prod = {
        "surfel_knn": sknn.cpu().numpy().view(np.uint16), "surfel_spatial": sw.cpu().numpy(),
        "surfel_connect": sconn.cpu().numpy(), "node_knn": nknn.cpu().numpy().view(np.uint16),
        "node_spatial": nw.cpu().numpy(), "edges": edges.cpu().numpy().view(np.uint16),
        "node_connect": nconn.cpu().numpy(),
    }
    return frame, prod


def host_inputs(scene: Scene, prod: dict, node_motion_pred: np.ndarray | None = None) -> dict:
    """Solver inputs as numpy arrays (the CPU restatement's layout)."""
    cfg = scene.cfg
    return dict(node_motion_pred=node_motion_pred,fx=cfg.fx, fy=cfg.fy, cx=cfg.cx, cy=cfg.cy, cam2world=scene.cam2world,
                measure_vertex=scene.measure_vertex, measure_normal=scene.measure_normal,
                ref_vertex=scene.ref_vertex, ref_normal=scene.ref_normal, index_map=scene.index_map,
                flow=scene.flow, K=cfg.K, surfel_knn=prod["surfel_knn"], surfel_spatial=prod["surfel_spatial"],
                surfel_connect=prod["surfel_connect"], node_coords=scene.nodes, edges=prod["edges"],
                node_connect=prod["node_connect"], node_radius_square=scene.node_radius_square)


def previous_frame_maps(scene: Scene) -> tuple[np.ndarray, np.ndarray]:
    """The rgbd maps (float4, w = 1 / depth, 0 where there is none) of the
    synthetic pair: the previous frame from the reference vertex map (world,
    brought into the camera), the current one from the measured vertex map
    (camera frame).  The node-motion prior's inputs besides the flow."""
    c2w = np.asarray(scene.cam2world, np.float64)
    R, t = c2w[:3, :3], c2w[:3, 3]
    ref = scene.ref_vertex[..., :3].astype(np.float64)
    z_prev = ((ref - t) @ R)[..., 2]
    valid = scene.index_map != 0xFFFFFFFF
    prev = np.zeros(scene.ref_vertex.shape, np.float32)
    prev[..., 3] = np.where(valid & (z_prev > 0), 1.0 / np.where(z_prev > 0, z_prev, 1.0), 0.0)
    z = scene.measure_vertex[..., 2].astype(np.float64)
    cur = np.zeros(scene.measure_vertex.shape, np.float32)
    cur[..., 3] = np.where(z > 0, 1.0 / np.where(z > 0, z, 1.0), 0.0)
    return prev, cur


def solver_for(scene: Scene, spmv_mode: int = native.SH_SPMV_MATERIALIZED, flags: int = 0) -> native.WarpSolver:
    cfg = scene.cfg
    c = native.make_config(cfg.solver, cfg.width, cfg.height, cfg.num_nodes, cfg.K, spmv_mode, flags)
    return native.WarpSolver(c)


def dq_to_mat34(dq: np.ndarray) -> np.ndarray:
    """N x 8 dual quaternions (q0 wxyz, q1 wxyz) -> N x 3 x 4 rigid transforms (float64)."""
    dq = np.asarray(dq, np.float64)
    q0, q1 = dq[:, :4], dq[:, 4:]
    n = np.linalg.norm(q0, axis=1, keepdims=True)
    q0 = q0 / n
    q1 = q1 / n
    q1 = q1 - np.sum(q0 * q1, axis=1, keepdims=True) * q0
    w, x, y, z = q0.T
    R = np.stack([
        np.stack([1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * z * w, 2 * x * z + 2 * y * w], -1),
        np.stack([2 * x * y + 2 * z * w, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * x * w], -1),
        np.stack([2 * x * z - 2 * y * w, 2 * y * z + 2 * x * w, 1 - 2 * x * x - 2 * y * y], -1)], 1)
    # t = 2 vec(q1 * conj(q0))
    bw, bx, by, bz = q1.T
    cw, cx, cy, cz = w, -x, -y, -z
    tx = bw * cx + bx * cw + by * cz - bz * cy
    ty = bw * cy - bx * cz + by * cw + bz * cx
    tz = bw * cz + bx * cy - by * cx + bz * cw
    t = 2 * np.stack([tx, ty, tz], -1)
    return np.concatenate([R, t[:, :, None]], axis=2)


def device_dq_to_wxyz(dq_dev) -> np.ndarray:
    """Device DQ rows (q0 xyzw, q1 xyzw) -> numpy N x 8 (q0 wxyz, q1 wxyz)."""
    a = dq_dev.cpu().numpy().astype(np.float64)
    return np.concatenate([a[:, 3:4], a[:, 0:3], a[:, 7:8], a[:, 4:7]], axis=1)


def warp_error(dq_a: np.ndarray, dq_b: np.ndarray) -> float:
    """max over nodes of ||T_a - T_b||_F / ||T_b||_F on the 3x4 transforms."""
    Ta, Tb = dq_to_mat34(dq_a), dq_to_mat34(dq_b)
    num = np.linalg.norm((Ta - Tb).reshape(len(Ta), -1), axis=1)
    den = np.linalg.norm(Tb.reshape(len(Tb), -1), axis=1)
    return float(np.max(num / den))
