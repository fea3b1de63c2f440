"""Benchmark: warp-field Gauss-Newton tracking solves per second on one MI355X.

Workload (default): the full-frame plain solve -- 300k surfels, 1024 nodes,
640x480 depth + optical flow, ARAP weight 2.3, 10 GN x 20 PCG, K = 8.  One
"step" is one complete solve (index build + GN loop) of that frame with every
input already resident in HBM.  Prints one JSON line with:

* value / ms_per_step / latency_ms: solves per second over the timed steps and
  the p50/p99 of the per-solve wall time (each solve ends in a stream sync);
* roofline: the GN-phase kernel with the largest share of device time, its
  average launch time from HIP events recorded on the solve's stream around
  every launch in the timed region, and its algorithmic bytes per launch
  (each necessary input read once, each output written once, at the storage
  precision used) -> achieved GB/s against the 8 TB/s HBM peak.  `traffic`
  is the measured HBM bytes per launch from the rocprofv3 FETCH_SIZE /
  WRITE_SIZE summary committed under profiles/ for this kernel and workload,
  when one is there (null otherwise);
* cpu_baseline: the FP32 CPU restatement of the same solve (OpenMP over the
  host cores), timed on a bounded sample of the same frame;
* warp_error_vs_fp64: max over nodes of ||T - T_fp64||_F / ||T_fp64||_F of the
  last timed solve against the FP64 CPU oracle.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "GN tracking solves/s + p50/p99 latency at ≤1e-4 err vs FP64; served vs plain"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64-oracle check of the warp field")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU-baseline sample budget")
    ap.add_argument("--no-served", action="store_true", help="skip the serving (64 streams) and batch (32 solves) legs")
    ap.add_argument("--served-requests", type=int, default=10000)
    ap.add_argument("--tuning", default=None, metavar="FLAGS,SPMV,ROWS",
                    help="use this tuning instead of running the autotuner (profiled runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 plain device copies of 256 MiB (the reference reading for HBM counters)")
    return ap.parse_args()


def copy_calibration():
    import torch
    x = torch.ones(64 << 20, device="cuda")  # 256 MiB of float32
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": x.numel() * 4, "launches": 5}), flush=True)


def phase_bytes(name: str, cfg, P: int, U: int) -> float:
    """Algorithmic bytes of one launch of a GN-phase kernel.

    Per-unit figures (K slots, record = the per-pixel H/g record: 28 values,
    4 B each in FP32 streams, 2 B in FP16 streams):
      dense_terms  P * (8 pixel pair + 2K ids + 4K folded weights + 64 texels + record) + N * 32 (node DQs)
                   + E * 40 + N * 48 (the ARAP terms, same launch: k_terms)
      reg_terms    E * 40 + N * 48 (its own launch, other paths)
      tiles        P * (record + 2K + 4K)  (each record and its slots read once)
      tiles_combine E * 40 + (N + 2U) * 144 (blocks out) + N * 24 (rhs out) + N * 288 (FP64 inverse out)
      rows_seg, rows_combine: as tiles, tiles_combine (the row-segment variant)
      dinv         N * (144 in + 288 FP64 out)
      pcg          iterations * ((N + 2U) * 148 + N * (10 * 24 + 144))
      update       N * 88
    """
    K, N = cfg.K, cfg.num_nodes
    E = N * K
    rec = 28 * (2 if cfg.solver.fp16 else 4)
    it = cfg.solver.pcg_iterations
    return {
        "dense_terms": P * (8 + 2 * K + 4 * K + 64 + rec) + N * 32 + E * 40 + N * 48,
        "reg_terms": E * 40 + N * 48,
        "rows_seg": P * (rec + 6 * K),
        "rows_combine": E * 40 + (N + 2 * U) * 144 + N * 24,
        "tiles": P * (rec + 6 * K),
        "tiles_combine": E * 40 + (N + 2 * U) * 144 + N * 24 + N * 288,
        "dinv": N * (144 + 288),
        "pcg": it * ((N + 2 * U) * 148 + N * (10 * 24 + 144)),
        "update": N * 88,
    }.get(name, 0.0)


KERNEL_OF_PHASE = {"dense_terms": "k_terms", "reg_terms": "k_terms", "rows_seg": "k_rows_seg",
                   "rows_combine": "k_rows_combine", "dinv": "k_dinv", "pcg": "k_pcg_resident",
                   "update": "k_update", "tiles": "k_tiles", "tiles_combine": "k_tiles_combine"}


def pcg_kernel(tuning: dict) -> str:
    if tuning["spmv_mode"] == 1 or tuning["flags"] & 4:
        return "k_pcg_spmv/k_pcg_update"
    if tuning["flags"] & 8:
        return "k_pcg_coop"
    return "k_pcg_pipe" if tuning["flags"] & 256 else "k_pcg_resident"


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch from a committed rocprofv3 counter summary, if any."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_traffic.json"
    if not f.exists():
        return None, None
    try:
        d = json.loads(f.read_text())
        return float(d["hbm_bytes_per_launch"]), f"profiles/{f.name}"
    except (ValueError, KeyError):
        return None, None


def cpu_baseline(sc, prod, cfg, seconds: float):
    sys.path.insert(0, str(ROOT / "oracle"))
    import sho  # the CPU restatement: the baseline, not the thing measured
    from starhub_amd import workload
    threads = sho.max_threads()
    host = workload.host_inputs(sc, prod)
    sho.solve(host, cfg.solver, 32, threads)  # warm-up
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < seconds and len(times) < 50:
        r = sho.solve(host, cfg.solver, 32, threads)
        times.append(r["solve_ms"])
    times.sort()
    med = times[len(times) // 2]
    return {"value": 1000.0 / med, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} full FP32 solves of the same frame (median solve time), OpenMP over "
                      f"{threads} host threads"}


def plain_rate(frames, cfgs, native, reps: int = 1):
    """Plain path on a mix: one solver, the frames solved one after another."""
    import time as _t
    W = max(c.width for c in cfgs)
    H = max(c.height for c in cfgs)
    N = max(c.num_nodes for c in cfgs)
    solver = native.WarpSolver(native.make_config(cfgs[0].solver, W, H, N, cfgs[0].K))
    for f in frames[:2]:
        solver.process_frame(f)
    t0 = _t.perf_counter()
    for _ in range(reps):
        for f in frames:
            solver.process_frame(f)
    dt = _t.perf_counter() - t0
    solver.close()
    return reps * len(frames) / dt


def batched_leg(native, scene, workload):
    """32 uneven solves (N log-uniform in [128, 4096], 300 surfels/node) in one
    batch: one launch per GN phase for all of them."""
    import time as _t
    cfgs = scene.config_c4()
    frames = [workload.device_frame(scene.generate(c))[0] for c in cfgs]
    W = max(c.width for c in cfgs)
    H = max(c.height for c in cfgs)
    N = max(c.num_nodes for c in cfgs)
    b = native.BatchSolver(native.make_config(cfgs[0].solver, W, H, N, cfgs[0].K), len(cfgs))
    for i, f in enumerate(frames):
        b.set_inputs(i, f)
    tuning = native.autotune(b, batch_solves=len(cfgs), repeats=2)
    b.solve(len(cfgs))
    reps = 5
    t0 = _t.perf_counter()
    for _ in range(reps):
        b.solve(len(cfgs))
    dt = _t.perf_counter() - t0
    b.close()
    batched = reps * len(cfgs) / dt
    plain = plain_rate(frames, cfgs, native)
    return {"workload": "32 solves, N log-uniform [128, 4096], 300 surfels/node, 320x240..1280x720, "
                        f"{cfgs[0].solver.gn_iterations} GN x {cfgs[0].solver.pcg_iterations} PCG",
            "solves_per_s": batched, "plain_solves_per_s": plain, "vs_plain": batched / plain, "tuning": tuning}


def served_leg(native, scene, workload, n_requests: int):
    """64 camera streams, Poisson arrivals at 30 Hz each, continuous batching."""
    import ctypes
    import time as _t
    import numpy as np
    import torch
    cfgs = scene.config_c5_streams()
    scs = [scene.generate(c) for c in cfgs]
    frames = [workload.device_frame(s)[0] for s in scs]
    refs = [(torch.from_numpy(s.surfel_pos).cuda(), torch.from_numpy(s.surfel_normal).cuda()) for s in scs]
    W = max(c.width for c in cfgs)
    H = max(c.height for c in cfgs)
    N = max(c.num_nodes for c in cfgs)
    scfg = native.make_config(cfgs[0].solver, W, H, N, cfgs[0].K)
    plain = plain_rate(frames, cfgs, native)
    srv = native.Server(scfg, max_active=64)
    for i, f in enumerate(frames):
        native.check(srv.add_stream(i, f, *refs[i]), "add_stream")  # surfels resident, warped after each solve
    L = native.lib()
    args = []
    for i, f in enumerate(frames):
        cam, meas, rend, geo, ng, nf, fl, kp, c2w = f.structs()
        args.append((meas, rend, fl, nf))
    rng = np.random.default_rng(5)
    arrivals = []
    for i in range(len(frames)):
        t = np.cumsum(rng.exponential(1.0 / 30.0, size=n_requests // len(frames) + 50))
        arrivals += [(float(x), i) for x in t]
    arrivals.sort()
    arrivals = arrivals[:n_requests]
    tickets = []
    tk = ctypes.c_uint64()
    t0 = _t.perf_counter()
    for when, i in arrivals:
        d = when - (_t.perf_counter() - t0)
        if d > 0:
            _t.sleep(d)
        meas, rend, fl, nf = args[i]
        native.check(L.sh_server_submit(srv._h, i, ctypes.byref(meas), ctypes.byref(rend), ctypes.byref(fl),
                                        ctypes.byref(nf), None, ctypes.byref(tk)), "submit")
        tickets.append(tk.value)
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    ok = [r for r in res if r["status"] == 0]
    lat = np.array([r["done_ms"] - r["submit_ms"] for r in ok])
    span_s = (max(r["done_ms"] for r in ok) - min(r["submit_ms"] for r in ok)) / 1e3
    served = len(ok) / span_s
    return {"workload": f"64 streams, N log-uniform [256, 4096], 300 surfels/node, Poisson 30 Hz/stream, "
                        f"resident surfels forward-warped after each solve, "
                        f"{len(res)} requests, {cfgs[0].solver.gn_iterations} GN x {cfgs[0].solver.pcg_iterations} PCG",
            "solves_per_s": served, "offered_per_s": len(res) / arrivals[-1][0],
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "failed": len(res) - len(ok), "slots": st["slots"], "max_batch": st["max_batch"],
            "mean_batch": st["slot_iterations"] / max(1, st["steps"]),
            "plain_solves_per_s": plain, "vs_plain": served / plain}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): solves are independent and shard by "
              f"request across GPUs with no exchange, so the one-GPU figure is the per-GPU figure.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    import numpy as np
    import torch
    from starhub_amd import native, scene, workload

    cfg = {"c1": scene.config_c1, "c2": scene.config_c2, "c3": scene.config_c3}[args.config]()
    sc = scene.generate(cfg)
    frame, prod = workload.device_frame(sc)
    solver = workload.solver_for(sc, flags=native.SH_FLAG_PHASE_TIMING)
    solver.set_inputs(frame)
    if args.tuning:
        f, m, r = (int(v) for v in args.tuning.split(","))
        native.set_tuning(solver, f, m, r)
        tuning = {"flags": f, "spmv_mode": m, "pcg_rows_per_group": r, "ms": None}
    else:
        tuning = native.autotune(solver, repeats=9)  # fastest variant on this frame (median of 9), kept for the timed steps
    for _ in range(args.warmup):
        solver.solve()
    torch.cuda.synchronize()
    solver.phase_times(reset=True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        solver.solve()  # ends in a stream synchronisation
        lat.append(time.perf_counter() - s0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    phases = solver.phase_times(reset=True)
    P, U, _ = solver.sizes()

    lat_ms = np.sort(np.array(lat) * 1e3)
    out = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 streams, fp32 accumulate, fp64 PCG scalars + 6x6 inverse" if cfg.solver.fp16
        else "fp32 (fp64 PCG scalars + 6x6 inverse)",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: plain single solve, {cfg.num_surfels} surfels, {cfg.num_nodes} "
                               f"nodes, {cfg.width}x{cfg.height}, K={cfg.K}, {cfg.solver.gn_iterations} GN x "
                               f"{cfg.solver.pcg_iterations} PCG",
                   "valid_pixels": P, "node_pairs": U},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "tuning": tuning,
    }

    # roofline of the kernel with the largest device-time share
    total = {k: v[0] for k, v in phases.items() if v[1] > 0}
    dom = max(total, key=total.get)
    avg_s = phases[dom][0] / phases[dom][1] / 1e3
    nbytes = phase_bytes(dom, cfg, P, U)
    achieved = nbytes / avg_s / 1e9
    kernel = pcg_kernel(tuning) if dom == "pcg" else KERNEL_OF_PHASE.get(dom, dom)
    traffic, tsrc = measured_traffic(kernel, args.config)
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                       "avg_launch_us": avg_s * 1e6, "bytes_per_launch": nbytes,
                       "launches": phases[dom][1]}
    if tsrc:
        out["roofline"]["traffic_source"] = tsrc
    out["phase_share"] = {k: round(v / sum(total.values()), 4) for k, v in total.items()}

    if not args.no_check:
        sys.path.insert(0, str(ROOT / "oracle"))
        import sho  # the checker
        gpu = workload.device_dq_to_wxyz(solver.node_se3())
        ref = sho.solve(workload.host_inputs(sc, prod), cfg.solver, 64, sho.max_threads())
        out["warp_error_vs_fp64"] = workload.warp_error(gpu, ref["dq"].reshape(-1, 8))

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, prod, cfg, args.cpu_seconds)
    if not args.no_served:
        del solver
        out["batched"] = batched_leg(native, scene, workload)
        out["served"] = served_leg(native, scene, workload, args.served_requests)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
