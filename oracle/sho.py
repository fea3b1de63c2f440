"""ctypes front end of the CPU restatement (``oracle/sh_oracle.hpp``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker; the product path never
imports it.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

_FP = ctypes.POINTER(ctypes.c_float)
_U16P = ctypes.POINTER(ctypes.c_uint16)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_DP = ctypes.POINTER(ctypes.c_double)


class ShoInputs(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
        ("cam2world", ctypes.c_float * 16),
        ("measure_vertex", _FP), ("measure_normal", _FP), ("ref_vertex", _FP), ("ref_normal", _FP),
        ("index_map", _U32P), ("flow", _FP),
        ("num_surfels", ctypes.c_int), ("K", ctypes.c_int),
        ("surfel_knn", _U16P), ("surfel_spatial", _FP), ("surfel_connect", _FP),
        ("num_nodes", ctypes.c_int), ("node_coords", _FP),
        ("num_edges", ctypes.c_int), ("edges", _U16P), ("node_connect", _FP), ("node_motion_pred", _FP),
        ("node_radius_square", ctypes.c_float),
    ]


class ShoParams(ctypes.Structure):
    _fields_ = [
        ("gn_iters", ctypes.c_int), ("pcg_iters", ctypes.c_int),
        ("lambda_depth", ctypes.c_float), ("lambda_flow", ctypes.c_float), ("lambda_reg", ctypes.c_float),
        ("lambda_node_translation", ctypes.c_float), ("damping", ctypes.c_float),
        ("rot_cut", ctypes.POINTER(ctypes.c_uint8)),
    ]


_VARIANTS = {}


def _host_has_avx512() -> bool:
    try:
        flags = next(line for line in open("/proc/cpuinfo") if line.startswith("flags")).split()
    except (OSError, StopIteration):
        return False
    return all(f in flags for f in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"))


def variant_name(variant: str = "default") -> str:
    """Which build a variant loads: "default" is the portable x86-64-v3 build
    (the checks); "native" is the x86-64-v4 build on hosts with AVX-512, else
    the portable one (the CPU baseline)."""
    if variant == "native" and _host_has_avx512() and (_HERE / "_build" / "libsh_oracle_v4.so").exists():
        return "x86-64-v4 (AVX-512)"
    return "x86-64-v3"


def lib(variant: str = "default"):
    global _LIB
    if variant != "default":
        name = variant_name(variant)
        if name == "x86-64-v3":
            return lib()
        if name not in _VARIANTS:
            _VARIANTS[name] = _load(_HERE / "_build" / "libsh_oracle_v4.so")
        return _VARIANTS[name]
    if _LIB is None:
        _LIB = _load(_HERE / "_build" / "libsh_oracle.so")
    return _LIB


def _load(path):
    if not path.exists():
        raise RuntimeError(f"{path} is missing: run `make -C oracle` or `python -m starhub_amd.build`")
    L = ctypes.CDLL(str(path))
    L.sho_skin.argtypes = [_FP, ctypes.c_int, _FP, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                           _U16P, _FP, _FP, ctypes.c_int]
    L.sho_node_graph.argtypes = [_FP, ctypes.c_int, ctypes.c_int, ctypes.c_float, _U16P, _FP, _U16P, _FP,
                                 ctypes.c_int]
    L.sho_solve.argtypes = [ctypes.POINTER(ShoInputs), ctypes.POINTER(ShoParams), ctypes.c_int, ctypes.c_int]
    L.sho_solve.restype = ctypes.c_void_p
    L.sho_result_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                                 ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)]
    L.sho_result_free.argtypes = [ctypes.c_void_p]
    L.sho_round_to_half.argtypes = [ctypes.c_float]
    L.sho_round_to_half.restype = ctypes.c_float
    L.sho_float2uint_rn.argtypes = [ctypes.c_float]
    L.sho_float2uint_rn.restype = ctypes.c_uint32
    L.sho_dq_se3.argtypes = [_DP, _DP]
    L.sho_dq_from_mat34.argtypes = [_DP, _DP]
    L.sho_twist_exp.argtypes = [_DP, _DP]
    L.sho_invert6.argtypes = [_DP, _DP]
    L.sho_max_threads.restype = ctypes.c_int
    return L


def max_threads() -> int:
    return int(lib().sho_max_threads())


def skin(surfel_pos: np.ndarray, nodes: np.ndarray, K: int, r2: float, threads: int = 0):
    S, N = surfel_pos.shape[0], nodes.shape[0]
    sp = np.ascontiguousarray(surfel_pos, np.float32)
    nd = np.ascontiguousarray(nodes, np.float32)
    knn = np.zeros((S, K), np.uint16)
    w = np.zeros((S, K), np.float32)
    c = np.zeros((S, K), np.float32)
    rc = lib().sho_skin(sp.ctypes.data_as(_FP), S, nd.ctypes.data_as(_FP), N, K, r2, knn.ctypes.data_as(_U16P),
                        w.ctypes.data_as(_FP), c.ctypes.data_as(_FP), threads or max_threads())
    if rc != 0:
        raise ValueError("sho_skin rejected its arguments")
    return knn, w, c


def node_graph(nodes: np.ndarray, K: int, r2: float, threads: int = 0):
    N = nodes.shape[0]
    nd = np.ascontiguousarray(nodes, np.float32)
    knn = np.zeros((N, K), np.uint16)
    w = np.zeros((N, K), np.float32)
    edges = np.zeros((N * K, 3), np.uint16)
    c = np.zeros((N, K), np.float32)
    rc = lib().sho_node_graph(nd.ctypes.data_as(_FP), N, K, r2, knn.ctypes.data_as(_U16P), w.ctypes.data_as(_FP),
                              edges.ctypes.data_as(_U16P), c.ctypes.data_as(_FP), threads or max_threads())
    if rc != 0:
        raise ValueError("sho_node_graph rejected its arguments")
    return knn, w, edges, c


_DTYPES = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.float64}

RESULT_FIELDS = ["pixel_pair", "pixel_knn", "n2t_offset", "n2t_terms", "half_keys", "half_offset", "pair_terms",
                 "sym_keys", "sym_range", "row_offset", "residual", "jacobian", "reg_Ti_xj", "reg_Tj_xj",
                 "reg_valid", "diag", "rhs", "offdiag", "diag_inv", "pcg_x", "cost", "dq_iter", "twist_l1", "rot_cut", "dq"]


def _inputs_struct(inputs: dict):
    keep = []

    def fp(a):
        if a is None:
            return _FP()
        a = np.ascontiguousarray(a, np.float32)
        keep.append(a)
        return a.ctypes.data_as(_FP)

    def up16(a):
        a = np.ascontiguousarray(a, np.uint16)
        keep.append(a)
        return a.ctypes.data_as(_U16P)

    def up32(a):
        a = np.ascontiguousarray(a, np.uint32)
        keep.append(a)
        return a.ctypes.data_as(_U32P)

    H, W = inputs["index_map"].shape
    c = ShoInputs()
    c.width, c.height = W, H
    c.fx, c.fy, c.cx, c.cy = inputs["fx"], inputs["fy"], inputs["cx"], inputs["cy"]
    c.cam2world[:] = [float(v) for v in np.asarray(inputs["cam2world"], np.float32).reshape(16)]
    c.measure_vertex = fp(inputs["measure_vertex"])
    c.measure_normal = fp(inputs["measure_normal"])
    c.ref_vertex = fp(inputs["ref_vertex"])
    c.ref_normal = fp(inputs["ref_normal"])
    c.index_map = up32(inputs["index_map"])
    c.flow = fp(inputs.get("flow"))
    c.K = int(inputs["K"])
    c.num_surfels = int(inputs["surfel_knn"].shape[0])
    c.surfel_knn = up16(inputs["surfel_knn"])
    c.surfel_spatial = fp(inputs["surfel_spatial"])
    c.surfel_connect = fp(inputs["surfel_connect"])
    c.num_nodes = int(inputs["node_coords"].shape[0])
    c.node_coords = fp(inputs["node_coords"])
    c.num_edges = int(inputs["edges"].shape[0])
    c.edges = up16(inputs["edges"])
    c.node_connect = fp(inputs["node_connect"])
    c.node_motion_pred = fp(inputs.get("node_motion_pred"))
    c.node_radius_square = float(inputs["node_radius_square"])
    return c, keep


def _params(params, rot_cut=None):
    p = ShoParams(params.gn_iterations, params.pcg_iterations, params.lambda_depth, params.lambda_flow,
                  params.lambda_reg, params.lambda_node_translation, params.damping)
    if rot_cut is not None:
        p.rot_cut = rot_cut.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    return p


def solve(inputs: dict, params, precision: int = 64, threads: int = 1, rot_cut=None, variant: str = "default") -> dict:
    """Run the restated GN solve.  `inputs` holds numpy arrays (see starhub_amd.workload.host_inputs).

    rot_cut: optional gn_iterations x num_nodes array of 0/1 decisions replayed in
    place of the oracle's own rotation cutoff (|w|_1 < 1e-4, SolverIterationData.cu:42)."""
    c, keep = _inputs_struct(inputs)
    if rot_cut is not None:
        rot_cut = np.ascontiguousarray(rot_cut, np.uint8).reshape(-1)
        N = int(inputs["node_coords"].shape[0])
        if rot_cut.size != params.gn_iterations * N:
            raise ValueError(f"rot_cut holds {rot_cut.size} decisions, expected {params.gn_iterations} x {N}")
        keep.append(rot_cut)
    p = _params(params, rot_cut)
    L = lib(variant)
    h = L.sho_solve(ctypes.byref(c), ctypes.byref(p), precision, threads)
    try:
        out = {}
        for name in RESULT_FIELDS + ["num_pixels", "solve_ms"]:
            ptr = ctypes.c_void_p()
            cnt = ctypes.c_size_t()
            dt = ctypes.c_int()
            rc = L.sho_result_get(h, name.encode(), ctypes.byref(ptr), ctypes.byref(cnt), ctypes.byref(dt))
            if rc != 0:
                continue
            typ = _DTYPES[dt.value]
            if cnt.value == 0:
                out[name] = np.zeros(0, typ)
                continue
            buf = (ctypes.c_char * (cnt.value * np.dtype(typ).itemsize)).from_address(ptr.value)
            out[name] = np.frombuffer(buf, typ).copy()
        out["num_pixels"] = int(out["num_pixels"][0])
        out["solve_ms"] = float(out["solve_ms"][0])
        return out
    finally:
        L.sho_result_free(h)


def dq_se3(dq8) -> np.ndarray:
    a = np.ascontiguousarray(dq8, np.float64)
    o = np.zeros(12, np.float64)
    lib().sho_dq_se3(a.ctypes.data_as(_DP), o.ctypes.data_as(_DP))
    return o.reshape(3, 4)


def dq_from_mat34(m34) -> np.ndarray:
    a = np.ascontiguousarray(m34, np.float64).reshape(12)
    o = np.zeros(8, np.float64)
    lib().sho_dq_from_mat34(a.ctypes.data_as(_DP), o.ctypes.data_as(_DP))
    return o


def twist_exp(xi6) -> np.ndarray:
    a = np.ascontiguousarray(xi6, np.float64)
    o = np.zeros(12, np.float64)
    lib().sho_twist_exp(a.ctypes.data_as(_DP), o.ctypes.data_as(_DP))
    return o.reshape(3, 4)


def invert6(A) -> np.ndarray:
    a = np.ascontiguousarray(A, np.float64).reshape(36)
    o = np.zeros(36, np.float64)
    lib().sho_invert6(a.ctypes.data_as(_DP), o.ctypes.data_as(_DP))
    return o.reshape(6, 6)


def block_pcg(row_ptr, col, blocks, inv_diag, b, x0=None, iters: int = 10, zero_init: bool = True,
              precision: int = 64):
    """BlockPCG<6>::Solve over a 6x6 block-CSR (row-major blocks), restated (sho::block_pcg).
    Returns (x, delta history of iters + 1 values)."""
    L = lib()
    if not hasattr(L, "_bpcg_set"):
        L.sho_block_pcg.argtypes = [ctypes.c_int, _U32P, _U32P, _DP, _DP, _DP, _DP, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _DP]
        L._bpcg_set = True
    row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    blocks = np.ascontiguousarray(blocks, np.float64).reshape(-1)
    inv_diag = np.ascontiguousarray(inv_diag, np.float64).reshape(-1)
    b = np.ascontiguousarray(b, np.float64).reshape(-1)
    nblk = len(row_ptr) - 1
    assert b.size == 6 * nblk and inv_diag.size == 36 * nblk and blocks.size == 36 * int(row_ptr[-1])
    assert col.size == int(row_ptr[-1]) and (col.size == 0 or int(col.max()) < nblk)
    x = np.zeros(6 * nblk) if x0 is None else np.array(x0, np.float64).reshape(-1)
    hist = np.zeros(iters + 1)
    L.sho_block_pcg(nblk, row_ptr.ctypes.data_as(_U32P), col.ctypes.data_as(_U32P), blocks.ctypes.data_as(_DP),
                    inv_diag.ctypes.data_as(_DP), b.ctypes.data_as(_DP), x.ctypes.data_as(_DP), iters,
                    int(zero_init), precision, hist.ctypes.data_as(_DP))
    return x, hist


def round_to_half(f: float) -> float:
    return float(lib().sho_round_to_half(f))


def float2uint_rn(f: float) -> int:
    return int(lib().sho_float2uint_rn(f))


def eval_terms(inputs: dict, params, dq_wxyz: np.ndarray):
    """Residual/Jacobian per pixel and ARAP records at the warp field `dq_wxyz` (N x 8, FP64)."""
    L = lib()
    if not hasattr(L, "_eval_set"):
        L.sho_eval_terms.argtypes = [ctypes.POINTER(ShoInputs), ctypes.POINTER(ShoParams), _DP, _DP, ctypes.c_size_t,
                                     _DP, _DP, _DP, ctypes.POINTER(ctypes.c_uint8)]
        L.sho_eval_terms.restype = ctypes.c_int
        L._eval_set = True
    c, keep = _inputs_struct(inputs)
    p = _params(params)
    H, W = inputs["index_map"].shape
    cap = H * W
    E = inputs["edges"].shape[0]
    r = np.zeros(cap * 3)
    J = np.zeros(cap * 18)
    ti = np.zeros(E * 3)
    tj = np.zeros(E * 3)
    va = np.zeros(E, np.uint8)
    dq = np.ascontiguousarray(dq_wxyz, np.float64)
    P = L.sho_eval_terms(ctypes.byref(c), ctypes.byref(p), dq.ctypes.data_as(_DP), r.ctypes.data_as(_DP), r.size,
                         J.ctypes.data_as(_DP), ti.ctypes.data_as(_DP), tj.ctypes.data_as(_DP),
                         va.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    assert P >= 0
    return r[:3 * P].reshape(P, 3), J[:18 * P].reshape(P, 18), ti.reshape(E, 3), tj.reshape(E, 3), va


def forward_warp(warp_wxyz, K, ref_v, ref_n, s_knn, s_spatial, s_connect, ref_node=None, n_knn=None,
                 n_spatial=None, n_connect=None, precision: int = 64):
    """SurfelNodeDeformer::ForwardWarpSurfelsAndNodes restated: live surfel vertex/normal (S x 4) and live
    node coordinates (N x 4) under the warp field warp_wxyz (N x 8, q0 wxyz, q1 wxyz)."""
    L = lib()
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    u16 = lambda a: np.ascontiguousarray(a, np.uint16)
    warp = np.ascontiguousarray(warp_wxyz, np.float64)
    ref_v, ref_n, s_knn, s_spatial, s_connect = f32(ref_v), f32(ref_n), u16(s_knn), f32(s_spatial), f32(s_connect)
    S = ref_v.shape[0]
    Nn = 0 if ref_node is None else ref_node.shape[0]
    if Nn:
        ref_node, n_knn, n_spatial, n_connect = f32(ref_node), u16(n_knn), f32(n_spatial), f32(n_connect)
    lv = np.zeros((S, 4)); ln = np.zeros((S, 4)); lnode = np.zeros((max(Nn, 1), 4))
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None
    L.sho_forward_warp.restype = ctypes.c_int
    rc = L.sho_forward_warp(P(warp, ctypes.c_double), K, P(ref_v, ctypes.c_float), P(ref_n, ctypes.c_float),
                            P(s_knn, ctypes.c_uint16), P(s_spatial, ctypes.c_float), P(s_connect, ctypes.c_float),
                            ctypes.c_uint64(S), P(ref_node, ctypes.c_float) if Nn else None,
                            P(n_knn, ctypes.c_uint16) if Nn else None, P(n_spatial, ctypes.c_float) if Nn else None,
                            P(n_connect, ctypes.c_float) if Nn else None, ctypes.c_uint64(Nn),
                            P(lv, ctypes.c_double), P(ln, ctypes.c_double), P(lnode, ctypes.c_double), precision)
    if rc != 0:
        raise ValueError("sho_forward_warp rejected its arguments")
    return lv, ln, lnode[:Nn]


# ------------------------------------------------------------ measurement maps
class DepthCamera(ctypes.Structure):
    """Raw depth camera: image size, intrinsics, clip range [m], downsample scale, surfel radius scale."""
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("fx", ctypes.c_float), ("fy", ctypes.c_float),
                ("cx", ctypes.c_float), ("cy", ctypes.c_float), ("clip_near", ctypes.c_float),
                ("clip_far", ctypes.c_float), ("scale", ctypes.c_float), ("radius_scale", ctypes.c_float)]


def scaled_size(cam: DepthCamera) -> tuple[int, int]:
    """(w, h) of the downsampled maps: truncation of the float product, as filterAndScaleVertex."""
    f = np.float32
    return int(f(cam.width) * f(cam.scale)), int(f(cam.height) * f(cam.scale))


def eigen33(cov6, precision: int = 64):
    """eigen33::compute restated: the smallest-eigenvalue eigenvector of packed PSD 3x3 matrices (n x 6)."""
    cov = np.ascontiguousarray(np.asarray(cov6, np.float64).reshape(-1, 6))
    out = np.zeros((cov.shape[0], 3))
    lib().sho_eigen33(cov.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cov.shape[0],
                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), precision)
    return out


def filter_depth(raw_mm, cam: DepthCamera, precision: int = 32):
    """filterUnreliableDepth restated: (filtered depth [m], camera_dot statistic) at the raw size."""
    raw = np.ascontiguousarray(raw_mm, np.uint16)
    out = np.zeros(raw.shape, np.float32)
    dots = np.zeros(raw.shape, np.float32)
    lib().sho_filter_depth(raw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), ctypes.byref(cam),
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                           dots.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), precision)
    return out, dots


def measure_maps(filtered_m, cam: DepthCamera, precision: int = 32):
    """Filtered depth -> (vertex_confid, normal_radius), each h x w x 4 at the downsampled size."""
    f = np.ascontiguousarray(filtered_m, np.float32)
    w, h = scaled_size(cam)
    v = np.zeros((h, w, 4))
    n = np.zeros((h, w, 4))
    lib().sho_measure_maps(f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(cam),
                           v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           n.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), precision)
    return v, n


def normal_radius_map(vertex4, radius_scale: float, precision: int = 32):
    """createNormalRadiusMap restated on an h x w x 4 float vertex map."""
    v = np.ascontiguousarray(vertex4, np.float32)
    h, w = v.shape[:2]
    n = np.zeros((h, w, 4))
    lib().sho_normal_radius_map(v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w, h,
                                ctypes.c_float(radius_scale), n.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                precision)
    return n


def vertex_confid_map(raw_mm, fx, fy, cx, cy):
    """createVertexConfidMap restated: h x w x 4 float32."""
    raw = np.ascontiguousarray(raw_mm, np.uint16)
    h, w = raw.shape
    v = np.zeros((h, w, 4), np.float32)
    lib().sho_vertex_confid_map(raw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), w, h, ctypes.c_float(fx),
                                ctypes.c_float(fy), ctypes.c_float(cx), ctypes.c_float(cy),
                                v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return v


# ------------------------------------------------------------ solver-map rasteriser
def render_solver_maps(ref_vertex4, ref_normal4, color_time4, world2camera, width, height, fx, fy, cx, cy,
                       max_depth, current_time, with_recent_observation=True):
    """Renderer::drawSolverMaps restated (libs/star/geometry/render/RendererDraw.cu:104-155 with
    shaders/solver_map.vert:21-53 and solver_map.frag:1-17): every model surfel is a one-pixel point.

    Per surfel, in float32 and in the order written:
      camera point p = R v + t of the row-major world2camera (rows summed left to right);
      culled if p.z > max_depth, p.z <= 0, or (confidence < 1 and |time - current_time| >= 3;
      the time bound is -1 without recent observations), with color_time4 None skipping this test;
      pixel (floor(fx p.x / p.z + cx + 0.5), floor(fy p.y / p.z + cy + 0.5)), dropped outside the image;
      depth (p.z / max_depth) * 0.5 + 0.5 (the [0, 1] depth range of a 32-bit float depth buffer
      cleared to 1, GL_LESS): the nearest point wins, equal depths go to the lower vertex id (the
      in-order depth test).
    Returns (reference_vertex h x w x 4 with w = 1, reference_normal h x w x 4 with w = 0,
    index_map h x w uint32, 0xffffffff where empty)."""
    f = np.float32
    V = np.ascontiguousarray(ref_vertex4, f).reshape(-1, 4)
    Nn = np.ascontiguousarray(ref_normal4, f).reshape(-1, 4)
    M = np.asarray(world2camera, f).reshape(4, 4)
    x, y, z = V[:, 0], V[:, 1], V[:, 2]
    px = ((M[0, 0] * x + M[0, 1] * y) + M[0, 2] * z) + M[0, 3]
    py = ((M[1, 0] * x + M[1, 1] * y) + M[1, 2] * z) + M[1, 3]
    pz = ((M[2, 0] * x + M[2, 1] * y) + M[2, 2] * z) + M[2, 3]
    keep = ~((pz > f(max_depth)) | (pz <= f(0)))
    if color_time4 is not None:
        T = np.ascontiguousarray(color_time4, f).reshape(-1, 4)
        tthr = f(3) if with_recent_observation else f(-1)
        keep &= ~((V[:, 3] < f(1)) & (np.abs(T[:, 2] - f(current_time)) >= tthr))
    with np.errstate(divide="ignore", invalid="ignore"):
        wx = ((f(fx) * px) / pz + f(cx)) + f(0.5)
        wy = ((f(fy) * py) / pz + f(cy)) + f(0.5)
        u = np.floor(np.where(keep, wx, f(-1)))
        v = np.floor(np.where(keep, wy, f(-1)))
    keep &= (u >= 0) & (u < width) & (v >= 0) & (v < height)
    ids = np.nonzero(keep)[0]
    depth = (pz[ids] / f(max_depth)) * f(0.5) + f(0.5)
    ids, depth = ids[depth < f(1)], depth[depth < f(1)]  # depth 1 does not pass GL_LESS against the clear value
    key = (depth.astype(f).view(np.uint32).astype(np.uint64) << np.uint64(32)) | ids.astype(np.uint64)
    zbuf = np.full(width * height, np.iinfo(np.uint64).max, np.uint64)
    np.minimum.at(zbuf, v[ids].astype(np.int64) * width + u[ids].astype(np.int64), key)
    hit = zbuf != np.iinfo(np.uint64).max
    win = (zbuf[hit] & np.uint64(0xFFFFFFFF)).astype(np.int64)
    vert = np.zeros((width * height, 4), f)
    nrm = np.zeros((width * height, 4), f)
    index = np.full(width * height, 0xFFFFFFFF, np.uint32)
    vert[hit, :3] = V[win, :3]
    vert[hit, 3] = 1
    nrm[hit, :3] = Nn[win, :3]
    index[hit] = win.astype(np.uint32)
    return vert.reshape(height, width, 4), nrm.reshape(height, width, 4), index.reshape(height, width)


# ------------------------------------------------------------ optical flow and node motion
def flow_map_from_of16(of16, scale: float):
    """FetchOFImage's decode (u16 / 10 - 1000, the double expression rounded once to float) of the first two
    channels, sampled at (round-half-even(x / scale), ...) and times scale (createScaledOpticalFlowMap)."""
    f = np.float32
    img = np.asarray(of16, np.uint16)
    H, W = img.shape[:2]
    w, h = int(np.floor(f(W) * f(scale))), int(np.floor(f(H) * f(scale)))
    win = f(1) / f(scale)
    rx = np.rint(np.arange(w, dtype=f) * win).astype(np.int64)  # numpy rint: round half to even
    ry = np.rint(np.arange(h, dtype=f) * win).astype(np.int64)
    ok = (rx[None, :] < W) & (ry[:, None] < H)
    sub = img[np.minimum(ry, H - 1)][:, np.minimum(rx, W - 1), :2].astype(np.float64)
    flow = (sub * 0.1 - 1000.0).astype(f)
    flow[~ok] = 0
    return (flow * f(scale)).astype(f)


FIX_SCALE = 2.0 ** 31


def node_motion_from_flow(rgbd_prev4, rgbd4, flow2, index_prev, surfel_knn, surfel_spatial, cam2world, fx, fy, cx,
                          cy, num_nodes):
    """AtomicSumIdealNodeMotionFromOpticalFlowKernel + AverageNodeMotionKernel restated with the exact
    fixed-point sums (integers of 2^-31) the HIP kernels use; float32 arithmetic in the order written."""
    f = np.float32
    P = np.asarray(rgbd_prev4, f)
    C = np.asarray(rgbd4, f)
    F = np.asarray(flow2, f)
    I = np.asarray(index_prev, np.uint32)
    knn = np.asarray(surfel_knn, np.uint16)
    sw = np.asarray(surfel_spatial, f)
    R = np.asarray(cam2world, f).reshape(4, 4)[:3, :3]
    H, W = I.shape
    S = knn.shape[0]
    fx, fy, cx, cy = f(fx), f(fy), f(cx), f(cy)
    ys, xs = np.mgrid[0:H, 0:W]
    xf = xs.astype(f) + F[..., 0]
    yf = ys.astype(f) + F[..., 1]
    tx, ty = np.floor(xf).astype(np.int64), np.floor(yf).astype(np.int64)
    inside = (tx >= 0) & (ty >= 0) & (tx < W) & (ty < H)
    cw = np.zeros((H, W), f)
    cw[inside] = C[ty[inside], tx[inside], 3]
    keep = (I != 0xFFFFFFFF) & (cw != 0) & (I < S)
    s = I[keep].astype(np.int64)
    node = knn[s, 0].astype(np.int64)
    ok = node < num_nodes
    s, node = s[ok], node[ok]
    sel = np.nonzero(keep)
    sel = (sel[0][ok], sel[1][ok])
    wgt = sw[s, 0]
    with np.errstate(divide="ignore", invalid="ignore"):
        depth = f(1) / cw[sel]
        depth_p = f(1) / P[sel][:, 3]
        vx = (xf[sel] - cx) / fx * depth
        vy = (yf[sel] - cy) / fy * depth
        ux = (xs[sel].astype(f) - cx) / fx * depth_p
        uy = (ys[sel].astype(f) - cy) / fy * depth_p
        mx, my, mz = vx - ux, vy - uy, depth - depth_p
        good = ~(np.abs(mz) > f(0.05))
    mx, my, mz, wgt, node = mx[good] * wgt[good], my[good] * wgt[good], mz[good] * wgt[good], wgt[good], node[good]
    wx = (R[0, 0] * mx + R[0, 1] * my) + R[0, 2] * mz
    wy = (R[1, 0] * mx + R[1, 1] * my) + R[1, 2] * mz
    wz = (R[2, 0] * mx + R[2, 1] * my) + R[2, 2] * mz
    sums = np.zeros((num_nodes, 4), np.int64)
    for c, v in enumerate((wx, wy, wz, wgt)):
        np.add.at(sums[:, c], node, np.rint(v.astype(np.float64) * FIX_SCALE).astype(np.int64))
    m = (sums.astype(np.float64) / FIX_SCALE).astype(f)
    nz = m[:, 3] != 0
    k = f(1) / m[nz, 3]
    m[nz] = m[nz] * k[:, None]
    return m
