"""ctypes binding of the HIP solver library (``include/sh/warp_solver.h``).

Python-side mirror of the reference's processor API
(``OptimizationProcessorWarpSolver::ProcessFrame`` / ``SolvedSE3``,
examples/mono_star/opt/OptimizationProcessorWarpSolver.h:18-30).  Device
memory and streams come from PyTorch (ROCm); every solve runs in the native
library, and loading fails loudly when that library is missing.
"""
from __future__ import annotations

import atexit
import ctypes
import sys
import weakref
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_LIB = None

# Every open handle, closed by an atexit hook: interpreter exit then destroys
# the library's HIP objects (frees, graph executables, events, streams, the
# server's worker threads) while the HIP runtime -- and any profiler attached
# to it -- is still up, rather than leaving them to the runtime's own teardown
# inside exit().  The hook is registered when the library is loaded, after
# torch, so it runs before any exit hook torch registered (last in, first out).  A __del__ that runs during interpreter finalisation does
# nothing: the atexit hook has already closed the handle.
_OPEN = weakref.WeakSet()


def close_all() -> None:
    """Closes every handle still open (solvers, batches, servers, block PCGs)."""
    for h in list(_OPEN):
        try:
            h.close()
        except Exception:
            pass


class _Handle:
    """Owner of one native handle: close() destroys it once."""

    _destroy = None  # name of the sh_*_destroy entry point

    def _register(self):
        _OPEN.add(self)

    def close(self):
        if self._h:
            getattr(lib(), self._destroy)(self._h)
            self._h = ctypes.c_void_p()
        _OPEN.discard(self)

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

SH_OK = 0
SH_STREAM_FP32, SH_STREAM_FP16 = 0, 1
SH_SPMV_MATERIALIZED, SH_SPMV_MATRIX_FREE, SH_SPMV_AUTO = 0, 1, 2
SH_FLAG_REFERENCE_INDEX, SH_FLAG_PAIR_MAJOR, SH_FLAG_NO_PERSISTENT_PCG, SH_FLAG_STREAMING_PCG = 1, 2, 4, 8
SH_FLAG_PHASE_TIMING, SH_FLAG_ROW_SEGMENTS, SH_FLAG_COST_HISTORY, SH_FLAG_NO_GRAPH = 16, 32, 64, 128
SH_FLAG_PIPELINED_PCG = 256
SH_FLAG_TILE_PIX_64 = 512
SH_FLAG_TILE_PIX_256 = 1024

_VP = ctypes.c_void_p


class SolverConfig(ctypes.Structure):
    _fields_ = [
        ("max_width", ctypes.c_uint32), ("max_height", ctypes.c_uint32), ("max_nodes", ctypes.c_uint32),
        ("knn", ctypes.c_uint32), ("gn_iterations", ctypes.c_uint32), ("pcg_iterations", ctypes.c_uint32),
        ("lambda_depth", ctypes.c_float), ("lambda_flow", ctypes.c_float), ("lambda_reg", ctypes.c_float),
        ("lambda_node_translation", ctypes.c_float), ("lambda_feature", ctypes.c_float),
        ("damping", ctypes.c_float), ("stream_precision", ctypes.c_uint32), ("spmv_mode", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
    ]


class Camera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float)]


class MeasureMaps(ctypes.Structure):
    _fields_ = [("vertex_confid", _VP), ("normal_radius", _VP)]


class RenderMaps(ctypes.Structure):
    _fields_ = [("reference_vertex", _VP), ("reference_normal", _VP), ("index_map", _VP)]


class Geometry(ctypes.Structure):
    _fields_ = [("surfel_knn", _VP), ("surfel_knn_spatial_weight", _VP), ("surfel_knn_connect_weight", _VP),
                ("num_surfels", ctypes.c_uint32)]


class NodeGraph(ctypes.Structure):
    _fields_ = [("reference_node_coords", _VP), ("node_graph", _VP), ("node_knn_connect_weight", _VP),
                ("node_knn", _VP), ("node_knn_spatial_weight", _VP), ("num_nodes", ctypes.c_uint32),
                ("num_edges", ctypes.c_uint32), ("node_radius_square", ctypes.c_float)]


class NodeFlow(ctypes.Structure):
    _fields_ = [("node_motion_pred", _VP), ("num_nodes", ctypes.c_uint32)]


class FlowMap(ctypes.Structure):
    _fields_ = [("flow", _VP)]


class KeyPoints(ctypes.Structure):
    _fields_ = [("num_matches", ctypes.c_uint32)]


def lib_path() -> Path:
    return Path(__file__).resolve().parent / "_build" / "libstarhub_amd.so"


def lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not path.exists():
            raise RuntimeError(f"native solver library {path} is missing: run `python -m starhub_amd.build`")
        import torch  # noqa: F401  -- load the process's HIP runtime before ours resolves against it
        L = ctypes.CDLL(str(path))
        atexit.register(close_all)
        P = ctypes.POINTER
        L.sh_version.restype = ctypes.c_char_p
        L.sh_status_string.restype = ctypes.c_char_p
        L.sh_status_string.argtypes = [ctypes.c_int]
        L.sh_solver_create.argtypes = [P(SolverConfig), P(_VP)]
        L.sh_solver_destroy.argtypes = [_VP]
        L.sh_solver_set_inputs.argtypes = [_VP, P(Camera), P(MeasureMaps), P(RenderMaps), P(Geometry),
                                           P(NodeGraph), P(NodeFlow), P(FlowMap), P(KeyPoints),
                                           P(ctypes.c_float)]
        L.sh_solver_solve.argtypes = [_VP, _VP]
        L.sh_solver_node_se3.argtypes = [_VP, P(_VP), P(ctypes.c_uint32)]
        L.sh_solver_sizes.argtypes = [_VP, P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        L.sh_solver_debug_view.argtypes = [_VP, ctypes.c_char_p, P(_VP), P(ctypes.c_size_t)]
        L.sh_solver_cost_history.argtypes = [_VP, P(ctypes.c_double), ctypes.c_uint32, P(ctypes.c_uint32)]
        L.sh_solver_apply_jtj.argtypes = [_VP, _VP, _VP, ctypes.c_uint32, _VP]
        L.sh_skin_surfels.argtypes = [_VP, ctypes.c_uint32, _VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                      _VP, _VP, _VP, _VP]
        L.sh_build_node_graph.argtypes = [_VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _VP, _VP, _VP,
                                          _VP, _VP]
        L.sh_batch_create.argtypes = [P(SolverConfig), ctypes.c_uint32, P(_VP)]
        L.sh_batch_destroy.argtypes = [_VP]
        L.sh_batch_set_inputs.argtypes = [_VP, ctypes.c_uint32, P(Camera), P(MeasureMaps), P(RenderMaps),
                                          P(Geometry), P(NodeGraph), P(NodeFlow), P(FlowMap), P(KeyPoints),
                                          P(ctypes.c_float)]
        L.sh_batch_solve.argtypes = [_VP, ctypes.c_uint32, _VP]
        L.sh_batch_node_se3.argtypes = [_VP, ctypes.c_uint32, P(_VP), P(ctypes.c_uint32)]
        _LIB = L
    return _LIB


def exported_symbols() -> list[str]:
    """Names declared in include/sh/warp_solver.h that the library must export."""
    hdr = Path(__file__).resolve().parent.parent / "include" / "sh" / "warp_solver.h"
    import re
    names = []
    for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sh_\w+)\s*\(", hdr.read_text(), re.M):
        names.append(m.group(1))
    return names


class SolverError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != SH_OK:
        msg = lib().sh_status_string(rc).decode()
        if rc == 3:  # SH_E_HIP
            L = lib()
            L.sh_last_hip_error.restype = ctypes.c_char_p
            msg += ": " + L.sh_last_hip_error().decode()
        raise SolverError(f"{what} failed: {msg} ({rc})")


def make_config(params, max_width: int, max_height: int, max_nodes: int, K: int = 8,
                spmv_mode: int = SH_SPMV_MATERIALIZED, flags: int = 0) -> SolverConfig:
    return SolverConfig(max_width, max_height, max_nodes, K, params.gn_iterations, params.pcg_iterations,
                        params.lambda_depth, params.lambda_flow, params.lambda_reg, params.lambda_node_translation,
                        params.lambda_feature, params.damping,
                        SH_STREAM_FP16 if params.fp16 else SH_STREAM_FP32, spmv_mode, flags)


@dataclass
class DeviceFrame:
    """Device-resident inputs of one solve (torch tensors on the GPU)."""
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    cam2world: np.ndarray
    measure_vertex: object
    measure_normal: object
    ref_vertex: object
    ref_normal: object
    index_map: object
    flow: object | None
    surfel_knn: object
    surfel_spatial: object
    surfel_connect: object
    node_coords: object
    edges: object
    node_connect: object
    node_knn: object
    node_spatial: object
    node_radius_square: float
    num_surfels: int
    num_nodes: int
    num_edges: int
    node_motion_pred: object | None = None  # N x 4 float32 (NodeFlow4Solver), optional

    def structs(self):
        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()
        cam = Camera(self.width, self.height, self.fx, self.fy, self.cx, self.cy)
        meas = MeasureMaps(ptr(self.measure_vertex), ptr(self.measure_normal))
        rend = RenderMaps(ptr(self.ref_vertex), ptr(self.ref_normal), ptr(self.index_map))
        geo = Geometry(ptr(self.surfel_knn), ptr(self.surfel_spatial), ptr(self.surfel_connect), self.num_surfels)
        ng = NodeGraph(ptr(self.node_coords), ptr(self.edges), ptr(self.node_connect), ptr(self.node_knn),
                       ptr(self.node_spatial), self.num_nodes, self.num_edges, self.node_radius_square)
        nf = NodeFlow(ptr(self.node_motion_pred), self.num_nodes)
        fl = FlowMap(ptr(self.flow))
        kp = KeyPoints(0)
        c2w = (ctypes.c_float * 16)(*[float(v) for v in np.asarray(self.cam2world, np.float32).reshape(16)])
        return cam, meas, rend, geo, ng, nf, fl, kp, c2w


def _stream_ptr(stream) -> ctypes.c_void_p:
    if stream is None:
        return ctypes.c_void_p()
    return ctypes.c_void_p(int(stream.cuda_stream))


class WarpSolver(_Handle):
    """One solve handle (sh_solver)."""

    _destroy = "sh_solver_destroy"

    def __init__(self, config: SolverConfig):
        self._h = ctypes.c_void_p()
        check(lib().sh_solver_create(ctypes.byref(config), ctypes.byref(self._h)), "sh_solver_create")
        self._register()
        self.config = config
        self._frame = None

    def set_inputs(self, frame: DeviceFrame) -> None:
        s = frame.structs()
        self._frame = frame
        check(lib().sh_solver_set_inputs(self._h, *[ctypes.byref(x) for x in s[:-1]], s[-1]), "sh_solver_set_inputs")

    def solve(self, stream=None) -> None:
        check(lib().sh_solver_solve(self._h, _stream_ptr(stream)), "sh_solver_solve")

    def process_frame(self, frame: DeviceFrame, stream=None) -> None:
        """ProcessFrame: set the inputs and solve (OptimizationProcessorWarpSolver.cu:22-52)."""
        self.set_inputs(frame)
        self.solve(stream)

    def node_se3(self):
        """SolvedSE3(): N x 8 float32 (q0 xyzw, q1 xyzw) copy as a torch tensor on the device."""
        import torch
        view = ctypes.c_void_p()
        n = ctypes.c_uint32()
        check(lib().sh_solver_node_se3(self._h, ctypes.byref(view), ctypes.byref(n)), "sh_solver_node_se3")
        return _device_copy(view.value, n.value * 8, torch.float32).view(n.value, 8)

    def sizes(self) -> tuple[int, int, int]:
        a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().sh_solver_sizes(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "sh_solver_sizes")
        return a.value, b.value, c.value

    def debug(self, name: str, dtype):
        """Copy of an internal device array (see sh_solver_debug_view)."""
        view = ctypes.c_void_p()
        cnt = ctypes.c_size_t()
        check(lib().sh_solver_debug_view(self._h, name.encode(), ctypes.byref(view), ctypes.byref(cnt)),
              f"debug_view({name})")
        return _device_copy(view.value, cnt.value, dtype)

    def apply_jtj(self, x, matrix_free: bool):
        """y = J^T W J x (+ damping) with the last GN iteration's system; x: 6N float32 device tensor."""
        import torch
        y = torch.empty_like(x)
        check(lib().sh_solver_apply_jtj(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                        int(matrix_free), ctypes.c_void_p()), "sh_solver_apply_jtj")
        return y

    def cost_history(self) -> np.ndarray:
        """FP64 weighted cost at the start of every GN iteration (+ the final one with SH_FLAG_COST_HISTORY)."""
        n = ctypes.c_uint32()
        check(lib().sh_solver_cost_history(self._h, None, 0, ctypes.byref(n)), "sh_solver_cost_history")
        out = np.zeros(n.value, np.float64)
        check(lib().sh_solver_cost_history(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n.value,
                                           ctypes.byref(n)), "sh_solver_cost_history")
        return out

    def phase_times(self, reset: bool = False) -> dict:
        """{phase: (total ms, launches)} since the last reset (needs SH_FLAG_PHASE_TIMING)."""
        return _phase_times(lib().sh_solver_phase_times, self._h, reset)


NUM_PHASES = 11


def _phase_times(fn, handle, reset: bool) -> dict:
    ms = (ctypes.c_double * NUM_PHASES)()
    calls = (ctypes.c_uint32 * NUM_PHASES)()
    check(fn(handle, ms, calls, NUM_PHASES, int(reset)), "phase_times")
    L = lib()
    L.sh_phase_name.restype = ctypes.c_char_p
    return {L.sh_phase_name(p).decode(): (ms[p], calls[p]) for p in range(NUM_PHASES)}


class BatchSolver(_Handle):
    """Batched uneven solves (sh_batch): one launch per GN phase covers every slot."""

    _destroy = "sh_batch_destroy"

    def __init__(self, config: SolverConfig, max_solves: int):
        self._h = ctypes.c_void_p()
        check(lib().sh_batch_create(ctypes.byref(config), max_solves, ctypes.byref(self._h)), "sh_batch_create")
        self._register()
        self.config = config
        self.max_solves = max_solves

    def set_inputs(self, slot: int, frame: DeviceFrame) -> None:
        s = frame.structs()
        check(lib().sh_batch_set_inputs(self._h, slot, *[ctypes.byref(x) for x in s[:-1]], s[-1]),
              "sh_batch_set_inputs")

    def solve(self, num: int, stream=None) -> None:
        check(lib().sh_batch_solve(self._h, num, _stream_ptr(stream)), "sh_batch_solve")

    def node_se3(self, slot: int):
        import torch
        view = ctypes.c_void_p()
        n = ctypes.c_uint32()
        check(lib().sh_batch_node_se3(self._h, slot, ctypes.byref(view), ctypes.byref(n)), "sh_batch_node_se3")
        return _device_copy(view.value, n.value * 8, torch.float32).view(n.value, 8)

    def phase_times(self, reset: bool = False) -> dict:
        return _phase_times(lib().sh_batch_phase_times, self._h, reset)


class Tuning(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint32), ("spmv_mode", ctypes.c_uint32), ("pcg_rows_per_group", ctypes.c_uint32),
                ("ms", ctypes.c_double)]


def autotune(handle, batch_solves: int | None = None, repeats: int = 3) -> dict:
    """Runs the autotuner on a WarpSolver (or a BatchSolver with batch_solves) and returns the chosen tuning."""
    t = Tuning()
    if batch_solves is None:
        check(lib().sh_solver_autotune(handle._h, ctypes.c_void_p(), repeats, ctypes.byref(t)), "sh_solver_autotune")
    else:
        check(lib().sh_batch_autotune(handle._h, batch_solves, ctypes.c_void_p(), repeats, ctypes.byref(t)),
              "sh_batch_autotune")
    return {"flags": t.flags, "spmv_mode": t.spmv_mode, "pcg_rows_per_group": t.pcg_rows_per_group, "ms": t.ms}


def set_tuning(solver, flags: int, spmv_mode: int, pcg_rows_per_group: int) -> None:
    """Applies a tuning to a WarpSolver without timing anything (sh_solver_set_tuning)."""
    t = Tuning(flags, spmv_mode, pcg_rows_per_group, 0.0)
    check(lib().sh_solver_set_tuning(solver._h, ctypes.byref(t)), "sh_solver_set_tuning")


def set_batch_tuning(batch, flags: int, spmv_mode: int, pcg_rows_per_group: int) -> None:
    """Applies a tuning to a BatchSolver without timing anything (sh_batch_set_tuning)."""
    t = Tuning(flags, spmv_mode, pcg_rows_per_group, 0.0)
    check(lib().sh_batch_set_tuning(batch._h, ctypes.byref(t)), "sh_batch_set_tuning")


def get_tuning(solver) -> dict:
    """The tuning a WarpSolver uses now (sh_solver_get_tuning)."""
    t = Tuning()
    check(lib().sh_solver_get_tuning(solver._h, ctypes.byref(t)), "sh_solver_get_tuning")
    return {"flags": t.flags, "spmv_mode": t.spmv_mode, "pcg_rows_per_group": t.pcg_rows_per_group}


def set_wait_limit(solver, polls: int) -> None:
    """Polls per persistent-PCG barrier wait before the solve falls back to the
    split PCG (0 restores the default, 2^22)."""
    check(lib().sh_solver_set_wait_limit(solver._h, ctypes.c_uint32(polls)), "sh_solver_set_wait_limit")


def pcg_fallbacks(solver) -> int:
    """Solves re-run with the split PCG after a persistent-PCG wait ran out."""
    n = ctypes.c_uint32()
    check(lib().sh_solver_pcg_fallbacks(solver._h, ctypes.byref(n)), "sh_solver_pcg_fallbacks")
    return n.value


HBM_BUDGET_FREE = (1 << 64) - 1  # SH_HBM_BUDGET_FREE: the free HBM at creation


class ServerConfig(ctypes.Structure):
    _fields_ = [("solver", SolverConfig), ("max_active", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_uint64)]


class RequestResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("stream_id", ctypes.c_uint32), ("status", ctypes.c_int),
                ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("slots", "streams", "streams_refused", "submitted", "completed",
                                                "failed", "steps", "max_batch", "slot_iterations", "hbm_reserved",
                                                "hbm_budget", "growth_refused")]
    _fields_ += [("now_ms", ctypes.c_double)]


class Server(_Handle):
    """Serving loop (sh_server): continuous batching of tracking solves from many streams."""

    _destroy = "sh_server_destroy"

    def __init__(self, solver_config: SolverConfig, max_active: int, hbm_budget_bytes: int = 0):
        L = lib()
        L.sh_server_wait.argtypes = [_VP, ctypes.c_uint64, ctypes.POINTER(RequestResult)]
        L.sh_server_submit.argtypes = [_VP, ctypes.c_uint32, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(ctypes.c_uint64)]
        self._h = ctypes.c_void_p()
        self.cfg = ServerConfig(solver_config, max_active, hbm_budget_bytes)
        check(L.sh_server_create(ctypes.byref(self.cfg), ctypes.byref(self._h)), "sh_server_create")
        self._register()
        self._frames = {}

    def add_stream(self, stream_id: int, frame: "DeviceFrame", ref_vertex=None, ref_normal=None) -> int:
        """Registers a stream; returns the status (SH_E_CAPACITY when the HBM budget refuses it).
        With reference surfels (S x 4 float32 device tensors) they are kept resident and forward-warped."""
        cam, meas, rend, geo, ng, nf, fl, kp, c2w = frame.structs()
        pv = ctypes.c_void_p(ref_vertex.data_ptr()) if ref_vertex is not None else ctypes.c_void_p()
        pn = ctypes.c_void_p(ref_normal.data_ptr()) if ref_normal is not None else ctypes.c_void_p()
        return lib().sh_server_add_stream(self._h, stream_id, ctypes.byref(cam), ctypes.byref(geo),
                                          ctypes.byref(ng), c2w, pv, pn)

    def live_surfels(self, stream_id: int):
        """(live vertex S x 4, live normal S x 4, live nodes N x 4) copies as device tensors."""
        import torch
        v, n, nd = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        cnt = ctypes.c_uint32()
        check(lib().sh_server_live_surfels(self._h, stream_id, ctypes.byref(v), ctypes.byref(n), ctypes.byref(nd),
                                           ctypes.byref(cnt)), "sh_server_live_surfels")
        S = cnt.value
        N = self.warp_field(stream_id).shape[0]
        return (_device_copy(v.value, S * 4, torch.float32).view(S, 4),
                _device_copy(n.value, S * 4, torch.float32).view(S, 4),
                _device_copy(nd.value, N * 4, torch.float32).view(N, 4))

    def submit(self, stream_id: int, frame: "DeviceFrame", out=None) -> int:
        cam, meas, rend, geo, ng, nf, fl, kp, c2w = frame.structs()
        self._frames[stream_id] = frame  # keep the maps alive
        t = ctypes.c_uint64()
        out_p = ctypes.c_void_p(out.data_ptr()) if out is not None else ctypes.c_void_p()
        check(lib().sh_server_submit(self._h, stream_id, ctypes.byref(meas), ctypes.byref(rend), ctypes.byref(fl),
                                     ctypes.byref(nf), out_p, ctypes.byref(t)), "sh_server_submit")
        return t.value

    def wait(self, ticket: int) -> dict:
        r = RequestResult()
        rc = lib().sh_server_wait(self._h, ticket, ctypes.byref(r))
        return {"ticket": r.ticket, "stream": r.stream_id, "status": rc, "submit_ms": r.submit_ms,
                "admit_ms": r.admit_ms, "done_ms": r.done_ms}

    def warp_field(self, stream_id: int):
        import torch
        view = ctypes.c_void_p()
        n = ctypes.c_uint32()
        check(lib().sh_server_warp_field(self._h, stream_id, ctypes.byref(view), ctypes.byref(n)),
              "sh_server_warp_field")
        return _device_copy(view.value, n.value * 8, torch.float32).view(n.value, 8)

    def phase_times(self, reset: bool = False) -> dict:
        """{phase: (total ms, launches)} of the server's GN steps (needs SH_FLAG_PHASE_TIMING)."""
        return _phase_times(lib().sh_server_phase_times, self._h, reset)

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().sh_server_get_stats(self._h, ctypes.byref(st)), "sh_server_get_stats")
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}


def _device_copy(addr: int, count: int, dtype):
    """Copy `count` elements at device address `addr` into a new torch tensor."""
    import torch
    out = torch.empty(count, dtype=dtype, device="cuda")
    if count:
        nbytes = count * out.element_size()
        hip = _hip()
        rc = hip.hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(addr), ctypes.c_size_t(nbytes), 3)
        if rc != 0:
            raise SolverError(f"hipMemcpy failed ({rc})")
    return out


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        import torch  # noqa: F401  -- the HIP runtime torch loaded (same soname as ours)
        _HIP = ctypes.CDLL("libamdhip64.so.7")
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _HIP


def skin_surfels(surfel_pos, nodes, K: int, r2: float, stream=None):
    """SkinnerKernel on the device; returns (knn u16, spatial f32, connect f32) torch tensors."""
    import torch
    S, N = surfel_pos.shape[0], nodes.shape[0]
    knn = torch.empty((S, K), dtype=torch.int16, device="cuda")
    w = torch.empty((S, K), dtype=torch.float32, device="cuda")
    c = torch.empty((S, K), dtype=torch.float32, device="cuda")
    check(lib().sh_skin_surfels(ctypes.c_void_p(surfel_pos.data_ptr()), S, ctypes.c_void_p(nodes.data_ptr()), N, K,
                                r2, ctypes.c_void_p(knn.data_ptr()), ctypes.c_void_p(w.data_ptr()),
                                ctypes.c_void_p(c.data_ptr()), _stream_ptr(stream)), "sh_skin_surfels")
    return knn, w, c


def build_node_graph(nodes, K: int, r2: float, stream=None):
    """Node-graph KNN + edges on the device; returns (knn, spatial, edges (N*K x 3), connect)."""
    import torch
    N = nodes.shape[0]
    knn = torch.empty((N, K), dtype=torch.int16, device="cuda")
    w = torch.empty((N, K), dtype=torch.float32, device="cuda")
    edges = torch.empty((N * K, 3), dtype=torch.int16, device="cuda")
    c = torch.empty((N, K), dtype=torch.float32, device="cuda")
    check(lib().sh_build_node_graph(ctypes.c_void_p(nodes.data_ptr()), N, K, r2, ctypes.c_void_p(knn.data_ptr()),
                                    ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(edges.data_ptr()),
                                    ctypes.c_void_p(c.data_ptr()), _stream_ptr(stream)), "sh_build_node_graph")
    return knn, w, edges, c


# ------------------------------------------------------------ depth -> measurement maps
class DepthCamera(ctypes.Structure):
    """sh_depth_camera: raw depth image size, raw intrinsics, clip range [m], downsample scale, surfel radius scale."""
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("clip_near", ctypes.c_float), ("clip_far", ctypes.c_float), ("downsample_scale", ctypes.c_float),
                ("surfel_radius_scale", ctypes.c_float)]

    def scaled_size(self) -> tuple[int, int]:
        f = np.float32
        return int(f(self.width) * f(self.downsample_scale)), int(f(self.height) * f(self.downsample_scale))


def filter_depth(depth_mm, cam: DepthCamera, stream=None):
    """sh_filter_depth: H x W int16 (uint16 bits) device tensor of depth [mm] -> H x W float32 filtered depth [m]."""
    import torch
    out = torch.empty((cam.height, cam.width), dtype=torch.float32, device="cuda")
    check(lib().sh_filter_depth(ctypes.c_void_p(depth_mm.data_ptr()), ctypes.byref(cam),
                                ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)), "sh_filter_depth")
    return out


def measure_maps(filtered_m, cam: DepthCamera, stream=None):
    """sh_measure_maps_from_depth: filtered depth -> (vertex_confid, normal_radius), each h x w x 4 float32."""
    import torch
    w, h = cam.scaled_size()
    v = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    n = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    check(lib().sh_measure_maps_from_depth(ctypes.c_void_p(filtered_m.data_ptr()), ctypes.byref(cam),
                                           ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(n.data_ptr()),
                                           _stream_ptr(stream)), "sh_measure_maps_from_depth")
    return v, n


def normal_radius_map(vertex4, radius_scale: float, stream=None):
    """sh_normal_radius_map on an h x w x 4 float32 device vertex map."""
    import torch
    h, w = vertex4.shape[:2]
    n = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    check(lib().sh_normal_radius_map(ctypes.c_void_p(vertex4.data_ptr()), w, h, ctypes.c_float(radius_scale),
                                     ctypes.c_void_p(n.data_ptr()), _stream_ptr(stream)), "sh_normal_radius_map")
    return n


def vertex_confid_map(depth_mm, fx: float, fy: float, cx: float, cy: float, stream=None):
    """sh_vertex_confid_map: H x W depth [mm] -> H x W x 4 float32 vertex map."""
    import torch
    h, w = depth_mm.shape
    v = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    check(lib().sh_vertex_confid_map(ctypes.c_void_p(depth_mm.data_ptr()), w, h, ctypes.c_float(fx),
                                     ctypes.c_float(fy), ctypes.c_float(cx), ctypes.c_float(cy),
                                     ctypes.c_void_p(v.data_ptr()), _stream_ptr(stream)), "sh_vertex_confid_map")
    return v


# ------------------------------------------------------------ solver maps (point splatting)
def render_solver_maps(ref_vertex4, ref_normal4, color_time4, world2camera, width: int, height: int, fx: float,
                       fy: float, cx: float, cy: float, max_depth: float, current_time: float,
                       with_recent_observation: bool = True, stream=None):
    """sh_render_solver_maps: (reference_vertex h x w x 4, reference_normal h x w x 4, index_map h x w int32
    holding uint32 bits) as device tensors."""
    import torch
    n = ref_vertex4.shape[0]
    zbuf = torch.empty(width * height, dtype=torch.int64, device="cuda")
    v = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
    nr = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
    idx = torch.empty((height, width), dtype=torch.int32, device="cuda")
    m = (ctypes.c_float * 16)(*np.asarray(world2camera, np.float32).reshape(-1).tolist())
    cam = Camera(width, height, fx, fy, cx, cy)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()
    check(lib().sh_render_solver_maps(ptr(ref_vertex4), ptr(ref_normal4), ptr(color_time4), ctypes.c_uint32(n), m,
                                      ctypes.byref(cam), ctypes.c_float(max_depth), ctypes.c_float(current_time),
                                      int(with_recent_observation), ptr(zbuf), ptr(v), ptr(nr), ptr(idx),
                                      _stream_ptr(stream)), "sh_render_solver_maps")
    return v, nr, idx


# ------------------------------------------------------------ optical flow and node motion
def flow_map_from_of16(of16, scale: float, stream=None):
    """sh_flow_map_from_of16: H x W x 3 int16 (uint16 bits) device tensor -> h x w x 2 float32 flow map."""
    import torch
    H, W = of16.shape[:2]
    f = np.float32
    w, h = int(np.floor(f(W) * f(scale))), int(np.floor(f(H) * f(scale)))
    out = torch.empty((h, w, 2), dtype=torch.float32, device="cuda")
    check(lib().sh_flow_map_from_of16(ctypes.c_void_p(of16.data_ptr()), W, H, ctypes.c_float(scale),
                                      ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)), "sh_flow_map_from_of16")
    return out


def node_motion_from_flow(rgbd_prev4, rgbd4, flow2, index_prev, surfel_knn, surfel_spatial, cam2world, fx, fy, cx,
                          cy, num_nodes: int, stream=None):
    """sh_node_motion_from_flow: device tensors in, num_nodes x 4 float32 node-motion prior out."""
    import torch
    H, W = index_prev.shape
    S, K = surfel_knn.shape
    ws = torch.empty(max(1, num_nodes) * 4, dtype=torch.int64, device="cuda")
    out = torch.empty((max(1, num_nodes), 4), dtype=torch.float32, device="cuda")
    m = (ctypes.c_float * 16)(*np.asarray(cam2world, np.float32).reshape(-1).tolist())
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    check(lib().sh_node_motion_from_flow(p(rgbd_prev4), p(rgbd4), p(flow2), p(index_prev), W, H, p(surfel_knn),
                                         p(surfel_spatial), S, K, m, ctypes.c_float(fx), ctypes.c_float(fy),
                                         ctypes.c_float(cx), ctypes.c_float(cy), num_nodes, p(ws), p(out),
                                         _stream_ptr(stream)), "sh_node_motion_from_flow")
    return out[:num_nodes]


def node_motion_from_flow_noalloc(rgbd_prev4, rgbd4, flow2, index_prev, surfel_knn, surfel_spatial, cam2world, fx,
                                  fy, cx, cy, num_nodes: int, stream=None):
    """sh_node_motion_from_flow_noalloc: as node_motion_from_flow, with the node sums in the library's
    fixed device array instead of a caller workspace (num_nodes <= 4096)."""
    import torch
    H, W = index_prev.shape
    S, K = surfel_knn.shape
    out = torch.empty((max(1, num_nodes), 4), dtype=torch.float32, device="cuda")
    m = (ctypes.c_float * 16)(*np.asarray(cam2world, np.float32).reshape(-1).tolist())
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    check(lib().sh_node_motion_from_flow_noalloc(p(rgbd_prev4), p(rgbd4), p(flow2), p(index_prev), W, H,
                                                 p(surfel_knn), p(surfel_spatial), S, K, m, ctypes.c_float(fx),
                                                 ctypes.c_float(fy), ctypes.c_float(cx), ctypes.c_float(cy),
                                                 num_nodes, p(out), _stream_ptr(stream)),
          "sh_node_motion_from_flow_noalloc")
    return out[:num_nodes]


# ------------------------------------------------------------ linear-solver plug-in
SPMV_APPLY_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
SPMV_INIT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p)


class SpmvOps(ctypes.Structure):
    """sh_spmv_ops: the ApplySpMVBase<6> callback table."""
    _fields_ = [("user", ctypes.c_void_p), ("matrix_size", ctypes.c_uint32), ("apply_spmv", SPMV_APPLY_FN),
                ("init_residual", SPMV_INIT_FN)]


class BlockCsr(_Handle):
    """sh_spmv over a caller-owned 6x6 block-CSR (torch device tensors, kept alive here)."""

    _destroy = "sh_spmv_destroy"

    def __init__(self, row_ptr, block_col, blocks):
        self._h = ctypes.c_void_p()
        self._keep = (row_ptr, block_col, blocks)
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        check(lib().sh_spmv_create_block_csr(row_ptr.numel() - 1, block_col.numel(), p(row_ptr), p(block_col),
                                             p(blocks), ctypes.byref(self._h)), "sh_spmv_create_block_csr")
        self._register()

    def check(self) -> int:
        return lib().sh_spmv_check(self._h, ctypes.c_void_p())

    def ops(self) -> SpmvOps:
        o = SpmvOps()
        check(lib().sh_spmv_ops_of(self._h, ctypes.byref(o)), "sh_spmv_ops_of")
        return o

    def apply(self, x, y, stream=None) -> int:
        return lib().sh_spmv_apply(self._h, ctypes.c_void_p(x), ctypes.c_void_p(y), _stream_ptr(stream))


def solver_spmv_ops(solver, matrix_free: bool = False) -> SpmvOps:
    """sh_solver_spmv_ops: the solver's own last-iteration system as an operator."""
    o = SpmvOps()
    check(lib().sh_solver_spmv_ops(solver._h, int(matrix_free), ctypes.byref(o)), "sh_solver_spmv_ops")
    return o


class BlockPCG(_Handle):
    """sh_block_pcg: BlockPCG<6> (SetSolverInput / Solve / SolveConvergeChecked)."""

    _destroy = "sh_block_pcg_destroy"

    def __init__(self, max_matrix_size: int):
        self._h = ctypes.c_void_p()
        check(lib().sh_block_pcg_create(max_matrix_size, ctypes.byref(self._h)), "sh_block_pcg_create")
        self._register()
        self._keep = None

    def set_input(self, inv_diag, ops: SpmvOps, b, x, actual_size: int = 0) -> int:
        self._keep = (inv_diag, ops, b, x)
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        return lib().sh_block_pcg_set_input(self._h, p(inv_diag), ctypes.byref(ops), p(b), p(x), actual_size)

    def solve(self, iterations: int = 10, zero_init: bool = True, stream=None) -> None:
        check(lib().sh_block_pcg_solve(self._h, iterations, int(zero_init), _stream_ptr(stream)), "sh_block_pcg_solve")

    def solve_converge_checked(self, iterations: int, zero_init: bool, epsilon: float, stream=None) -> int:
        n = ctypes.c_uint32()
        check(lib().sh_block_pcg_solve_converge_checked(self._h, iterations, int(zero_init), ctypes.c_float(epsilon),
                                                        _stream_ptr(stream), ctypes.byref(n)),
              "sh_block_pcg_solve_converge_checked")
        return n.value

    def delta_history(self) -> np.ndarray:
        n = ctypes.c_uint32()
        check(lib().sh_block_pcg_delta_history(self._h, None, 0, ctypes.byref(n)), "sh_block_pcg_delta_history")
        out = np.zeros(n.value, np.float64)
        check(lib().sh_block_pcg_delta_history(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n.value,
                                               ctypes.byref(n)), "sh_block_pcg_delta_history")
        return out
