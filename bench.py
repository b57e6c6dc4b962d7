#!/usr/bin/env python
"""Headline benchmark: Mcell-updates/s of the fused HIP hydro step on one MI355X.

Workload (2-D Kelvin-Helmholtz, plain method at scale):
2048x2048 periodic on [0,1]^2, 3 ghost cells, gamma 1.4, WENO5-Z + HLLC +
SSPRK3, CFL 0.4, seeded shear-band initial state. A step is one full SSPRK3
step: 3 fused stages, 3 ghost fills, the CFL dt (fused into the last stage)
and the step clock, replayed from a HIP graph with the state resident in HBM.

`value` is the compensated FP32 form's throughput; the FP64 drop-in's is
reported beside it, with the FP32C-vs-FP64 rel L1 per conserved variable after
200 steps (the FP64 GPU path equals the oracle bitwise, checked in the tests).
The roofline is that of the dominant kernel (the fused stage): algorithmic
HBM bytes per launch (read Q^(s), read Q^n for stages >= 1, write Q^(s+1):
256 B per cell-step over 3 launches) over its average launch time from HIP
events on the library's stream. The CPU baseline is the oracle (the
reference's unfused sequence restated, OpenMP, -O3) on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s fp32c vs fp64 @ rel L1<1e-5, 2048² WENO5Z/HLLC; adjoint/fwd"


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
            return
    except Exception:  # torch is only plumbing here
        pass
    hip = C.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()


def _rel_l1(a, b, g):
    a = a[:, 0, g:-g, g:-g].reshape(a.shape[0], -1)
    b = b[:, 0, g:-g, g:-g].reshape(b.shape[0], -1)
    return (np.abs(a - b).sum(1) / np.abs(b).sum(1)).tolist()


def time_form(M, cfg, q0, form, steps, warmup, use_graph=True, span_rows=0, autotune=True):
    """Steps replayed from a graph: wall time of exactly `steps` steps after `warmup`
    (the span height is picked by the library's autotuner unless given)."""
    ctx = M.HydroContext(cfg, precision_form=form, span_rows=span_rows)
    ctx.upload(q0)
    ctx.fill_bcs()
    tune = None
    if autotune and not span_rows:
        best, table = ctx.autotune(steps=3)
        tune = {"span_rows": best, "ms_per_step": {k: round(v, 4) for k, v in table.items()}}
    t, p = ctx.run(warmup, use_graph=use_graph)
    _sync()
    t0 = time.perf_counter()
    t, p = ctx.run(steps, time=t, prev_output_time=p, use_graph=use_graph)
    _sync()
    elapsed = time.perf_counter() - t0
    # per-kernel HIP-event timing on the library's stream (separate pass)
    t, p, tm = ctx.run_timed(steps, time=t, prev_output_time=p)
    ctx.close()
    return elapsed, tm, tune


def parity(M, cfg, q0, nsteps):
    out = {}
    for form in (M.FP64, M.FP32C):
        with M.HydroContext(cfg, precision_form=form) as ctx:
            ctx.upload(q0)
            ctx.fill_bcs()
            ctx.run(nsteps, use_graph=True)
            out[form] = ctx.download()
    return _rel_l1(out[M.FP32C], out[M.FP64], cfg.ng)


def adjoint_ratio(M, problems, solver, n=512, nsteps=200, every=20):
    """KH n^2 (WENO5-Z + `solver` + SSPRK3, FP64), nsteps frozen-dt steps, checkpoint every `every`:
    wall time of the reverse sweep (mh_adjoint_backward: segment recomputation, reverse stages and
    the cotangents' two host transfers) over the wall time of the same nsteps forward steps replayed
    from a graph; the reverse sweep's device time (HIP events) is reported beside it."""
    from mosscap_amd.adjoint import KHAdjointProblem
    prob = KHAdjointProblem(n=n, nsteps=nsteps, checkpoint_every=every, riemann_solver=solver, autotune=True)
    control = 0.01 * (2 * problems.uniform01(17, np.arange(n * n)).reshape(n, n) - 1)
    target = prob.forward(0.01 * (2 * problems.uniform01(18, np.arange(n * n)).reshape(n, n) - 1))
    rho = prob.forward(control)
    diff = rho - target
    lam = diff / np.linalg.norm(diff)
    prob.backward(lam)  # warm-up (allocations of the reverse sweep's scratch)
    _sync()
    t0 = time.perf_counter()
    prob.backward(lam)
    _sync()
    bwd_ms = (time.perf_counter() - t0) * 1e3
    bwd_dev_ms = prob.ctx.adjoint_backward_ms()
    q0 = prob.initial_state(control)
    with M.HydroContext(prob.cfg, precision_form=M.FP64, span_rows=prob.span_rows) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        ctx.run(3, use_graph=True)
        ctx.upload(q0)
        ctx.fill_bcs()
        _sync()
        t0 = time.perf_counter()
        ctx.run(nsteps, use_graph=True)
        _sync()
        fwd_ms = (time.perf_counter() - t0) * 1e3
    prob.ctx.close()
    return {"solver": solver, "grid": n, "steps": nsteps, "checkpoint_every": every, "span_rows": prob.span_rows,
            "backward_ms": round(bwd_ms, 3), "forward_ms": round(fwd_ms, 3), "ratio": round(bwd_ms / fwd_ms, 3),
            "backward_device_ms": round(bwd_dev_ms, 3), "ratio_device_over_wall": round(bwd_dev_ms / fwd_ms, 3)}


def cpu_baseline(cfg, q0, seconds):
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    import oracle
    sim = oracle.OracleSim(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    t, p, steps = 0.0, 0.0, 0
    oracle.profile(reset=True)
    t0 = time.perf_counter()
    while steps < 1 or time.perf_counter() - t0 < seconds:
        t, p, _ = sim.run(time=t, prev_output_time=p, max_steps=1)
        steps += 1
    el = time.perf_counter() - t0
    breakdown = {k: round(v / steps, 2) for k, v in oracle.profile(reset=True).items()}
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": cfg.nx * cfg.ny * steps / el / 1e6, "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} full SSPRK3 steps of the 2048^2 KH workload (oracle: the reference's unfused "
                      f"Q->W/Rx/Fx/Ry/Fy/combine/BC/dt sequence restated, g++ -O3 -fopenmp -ffp-contract=off), "
                      f"{el:.1f} s on {threads} threads of {cpu}",
            "ms_per_step_by_kernel": breakdown}


def copy_probe(mib: int) -> int:
    """A plain device-to-device copy kernel of known size, 5 times (torch's
    elementwise copy), for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))
    return 0


TRAFFIC_JSON = ROOT / "profiles" / "stage2d_kh2048_traffic.json"


def pmc_summary(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, copy_bytes, out_json):
    """Per-dispatch HBM traffic of each kernel from two rocprofv3 --pmc passes
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), scaled by what the same
    counters read for a plain copy of `copy_bytes` bytes."""
    import csv
    from collections import defaultdict

    def load(path):
        d = defaultdict(list)
        for r in csv.DictReader(open(path)):
            d[r["Kernel_Name"]].append(float(r["Counter_Value"]) * 1024.0)
        return d

    cf, cw = load(copy_fetch_csv), load(copy_write_csv)
    big_f = max(cf.items(), key=lambda kv: max(kv[1]))[1]
    big_w = max(cw.items(), key=lambda kv: max(kv[1]))[1]
    kf = copy_bytes / (sum(big_f) / len(big_f))
    kw = copy_bytes / (sum(big_w) / len(big_w))
    f, w = load(fetch_csv), load(write_csv)
    rows = {}
    for name in f:
        if name in w:
            rf = sum(f[name]) / len(f[name]) * kf
            rw = sum(w[name]) / len(w[name]) * kw
            rows[name] = {"dispatches": len(f[name]), "read_bytes": rf, "write_bytes": rw, "traffic": rf + rw}
    out = {"fetch_scale": kf, "write_scale": kw, "kernels": rows,
           "sources": [str(fetch_csv), str(write_csv), str(copy_fetch_csv), str(copy_write_csv)]}
    Path(out_json).write_text(json.dumps(out, indent=1))
    print(json.dumps({"fetch_scale": kf, "write_scale": kw}))
    for n, r in rows.items():
        print(f"{r['dispatches']:4d} read {r['read_bytes'] / 1e6:10.2f} MB  write {r['write_bytes'] / 1e6:10.2f} MB  {n[:110]}")
    return 0


def trace_summary(trace_csv):
    """Per-kernel launch statistics from a rocprofv3 kernel trace (CSV), grouped by kernel and grid
    size (launch shapes, e.g. the autotuner's trial spans, are kept apart), and the gaps between
    consecutive launches inside graph-replayed runs (launches less than 50 us apart)."""
    import csv
    from collections import defaultdict
    rows = list(csv.DictReader(open(trace_csv)))
    if not rows:
        print("empty trace")
        return 1
    grid_key = next((k for k in ("Grid_Size_X", "Grid_Size", "grid_size_x") if k in rows[0]), None)
    recs = []
    for r in rows:
        t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        recs.append((t0, t1, r["Kernel_Name"], r.get(grid_key, "?") if grid_key else "?"))
    recs.sort()
    by = defaultdict(list)
    for t0, t1, name, grid in recs:
        by[(name, grid)].append((t1 - t0) / 1e3)
    print("| kernel | grid | launches | avg (us) | min (us) | max (us) |")
    print("|---|---|---|---|---|---|")
    for (name, grid), d in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        short = name.replace("mh::", "").replace("(mh::Reconstruction)", "R").replace("(mh::SlopeLimiter)", "SL")
        short = short.replace("(mh::RiemannSolver)", "RS").split("(mh::StageGrid")[0][:90]
        print(f"| `{short}` | {grid} | {len(d)} | {sum(d) / len(d):.1f} | {min(d):.1f} | {max(d):.1f} |")
    gaps = defaultdict(list)
    for (a0, a1, an, _), (b0, b1, bn, _) in zip(recs, recs[1:]):
        g = (b0 - a1) / 1e3
        if 0 <= g < 50:
            form = "FormF64" if "FormF64" in an else "FormF32c" if "FormF32c" in an else None
            if form:
                gaps[form].append(g)
    for form, g in gaps.items():
        print(f"{form}: {len(g)} inter-launch gaps under 50 us, mean {sum(g) / len(g):.2f} us, "
              f"max {max(g):.2f} us")
    return 0


def measured_traffic(kernel_prefix, suffixes=("",)):
    """Average HBM bytes per launch of the matching kernel (name = prefix + one of the suffixes, then
    the argument list) from the committed PMC summary."""
    try:
        data = json.loads(TRAFFIC_JSON.read_text())
    except (OSError, ValueError):
        return None
    vals = [r for n, r in data["kernels"].items() if any(n.startswith(kernel_prefix + s) for s in suffixes)]
    if not vals:
        return None
    return sum(v["traffic"] for v in vals) / len(vals)


def bench_batch(a):
    """Ensemble of 4096 uneven 1-D Riemann problems (WENO5-Z + HLLC + SSPRK3, CFL 0.4,
    FP32C): one persistent launch per stage (autotuned grid) against one launch per problem."""
    import mosscap_amd as M
    from mosscap_amd import problems
    lengths, q0, _ = problems.riemann_batch(nprob=4096)
    cells = int(lengths.sum())
    out = {"metric": "Mcell-updates/s (batched 1-D ensemble)", "unit": "Mcell-updates/s", "n_gpus": 1,
           "dtype": "f32c", "data": "synthetic (seeded ensemble)",
           "config": {"workload": f"4096 1-D Riemann problems, lengths log-uniform in [128, 4096] "
                                  f"({cells} cells), WENO5-Z + HLLC + SSPRK3, CFL 0.4, per-problem dt"}}
    # positivity floors on: the plain method goes non-finite on near-vacuum members
    with M.BatchContext(lengths, precision_form=M.FP32C, rho_floor=1e-10, p_floor=1e-12) as b:
        b.upload(q0)
        b.fill_bcs()
        best, table = b.autotune(steps=5)
        out["autotune_waves"] = {"best": best, "ms_per_step_by_waves_per_cu": table}
        b.upload(q0)
        b.fill_bcs()
        b.run(a.warmup)
        _sync()
        t0 = time.perf_counter()
        b.run(a.batch_steps)
        _sync()
        el = time.perf_counter() - t0
        out["steps"] = a.batch_steps
        out["value"] = round(cells * a.batch_steps / el / 1e6, 2)
        out["ms_per_step"] = round(el * 1e3 / a.batch_steps, 4)
        out["floor_hits"] = int(b.stats().floor_hits)
        # one launch per problem and stage (graph-captured too), fewer steps
        n2 = max(2, a.batch_steps // 25)
        b.upload(q0)
        b.fill_bcs()
        b.run(1, persistent=False, use_graph=True)
        _sync()
        t0 = time.perf_counter()
        b.run(n2, persistent=False, use_graph=True)
        _sync()
        el2 = time.perf_counter() - t0
        out["per_problem_launches"] = {"steps": n2, "value": round(cells * n2 / el2 / 1e6, 2),
                                       "ms_per_step": round(el2 * 1e3 / n2, 3)}
        out["persistent_over_per_problem"] = round(out["value"] / out["per_problem_launches"]["value"], 2)
    # 1-D SSPRK3, 3 variables: 192 algorithmic bytes per cell-step (stage 0 reads Q^(s) and writes
    # Q^(s+1), stages 1-2 also read Q^n; 8 bytes per value in both forms)
    gbs = out["value"] * 1e6 * 192.0 / 1e9
    out["algorithmic_GBs"] = round(gbs, 1)
    out["roofline_frac"] = round(gbs / HBM_PEAK_GBS, 4)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(lengths, q0, a.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline_batch(lengths, q0, seconds):
    """The oracle (FP64, no floors) on the ensemble's members one after another, as one reference
    Simulation per problem would run them: 20 steps of each member in seeded order until about
    `seconds` of work, in cell-updates per second."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle
    import mosscap_amd as M
    offs = np.concatenate([[0], np.cumsum(3 * (lengths + 6))])
    cells, members, t0 = 0, 0, time.perf_counter()
    for p, n in enumerate(lengths):
        cfg = M.Config(num_dim=1, nx=int(n), ng=3, bc=("zerograd", "zerograd", "wall", "wall"), max_cfl=0.4,
                       max_time=1e30, output_delta=1e30, reconstruction="weno5z", riemann_solver="hllc",
                       time_stepper="ssprk3")
        sim = oracle.OracleSim(cfg)
        sim.set_q(q0[offs[p]:offs[p + 1]].reshape(cfg.shape))
        sim.fill_bcs()
        sim.run(max_steps=20)
        cells += int(n) * 20
        members += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    return {"value": cells / el / 1e6, "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"20 SSPRK3 steps of each of the first {members} members ({cells // 20} cells), FP64 oracle, "
                      f"{el:.1f} s on {threads} threads"}


def bench_sedov(a):
    """Sedov blast (PPM-MC + HLL + SSPRK3, CFL 0.3, outflow) to t = 0.05: FP32C with FP64 pressure
    recovery, compensated state and positivity floors against all-FP64."""
    import mosscap_amd as M
    from mosscap_amd import problems
    cfg = problems.sedov_config(a.n)
    q0 = problems.sedov(cfg)
    res = {}
    for form in (M.FP64, M.FP32C):
        kw = dict(rho_floor=1e-10, p_floor=1e-12) if form == M.FP32C else {}
        with M.HydroContext(cfg, precision_form=form, **kw) as ctx:
            ctx.upload(q0)
            ctx.fill_bcs()
            _sync()
            t0 = time.perf_counter()
            t, p, steps = 0.0, 0.0, 0
            while t < cfg.max_time:
                t, p = ctx.run(50, time=t, prev_output_time=p)
                steps += 50
            _sync()
            el = time.perf_counter() - t0
            st = ctx.stats()
            res[form] = (ctx.download(), el, int(st.steps), t, int(st.floor_hits))
    q64, el64, s64, t64, _ = res[M.FP64]
    q32, el32, s32, t32, hits = res[M.FP32C]
    cells = cfg.nx * cfg.ny
    out = {"metric": "Sedov FP32C vs FP64", "unit": "Mcell-updates/s", "n_gpus": 1,
           "config": {"workload": f"Sedov {cfg.nx}x{cfg.ny} on [-1,1]^2, PPM(MC) + HLL + SSPRK3, CFL 0.3, "
                                  "zerograd, t_end 0.05 (steps past t_end have dt = 0)"},
           "fp64": {"steps": s64, "t": t64, "value": round(cells * s64 / el64 / 1e6, 1)},
           "fp32c": {"steps": s32, "t": t32, "value": round(cells * s32 / el32 / 1e6, 1), "floor_hits": hits},
           # 256 algorithmic bytes per cell-step (2-D SSPRK3, 4 variables, 8 bytes per value)
           "algorithmic_GBs": {"fp64": round(cells * s64 / el64 * 256 / 1e9, 1),
                               "fp32c": round(cells * s32 / el32 * 256 / 1e9, 1)},
           "roofline_frac": {"fp64": round(cells * s64 / el64 * 256 / 1e9 / HBM_PEAK_GBS, 4),
                             "fp32c": round(cells * s32 / el32 * 256 / 1e9 / HBM_PEAK_GBS, 4)},
           "rel_l1_fp32c_vs_fp64": [float(f"{r:.3e}") for r in _rel_l1(q32, q64, cfg.ng)]}
    print(json.dumps(out), flush=True)
    return 0


def bench_blast3d(a):
    """3-D blast (WENO5-Z + HLLC + SSPRK3, periodic, CFL 0.3) on an n^3 grid: FP64 and FP32C
    throughput of graph-replayed steps, FP32C-vs-FP64 rel L1 after the timed steps, and the
    algorithmic bandwidth (5 variables x 8 B x (2 x 3 + 2) = 320 B per cell-step)."""
    import mosscap_amd as M
    from mosscap_amd import Config, problems
    n = a.n3
    cfg = Config(num_dim=3, nx=n, ny=n, nz=n, ng=3, x_dim=1.0, bc=("periodic",) * 6, reconstruction="weno5z",
                 riemann_solver="hllc", time_stepper="ssprk3", max_cfl=0.3, output_delta=1e30, max_time=1e30)
    q0 = problems.blast_3d(cfg, seed=3)
    res, states = {}, {}
    for form in (M.FP64, M.FP32C):
        el, tm, tune = time_form(M, cfg, q0, form, a.steps, a.warmup)
        res[form] = (el, tm, tune)
        with M.HydroContext(cfg, precision_form=form) as ctx:
            ctx.upload(q0)
            ctx.fill_bcs()
            ctx.run(a.steps)
            states[form] = ctx.download()
    cells = n ** 3
    out = {"metric": "3-D blast Mcell-updates/s", "unit": "Mcell-updates/s", "n_gpus": 1, "steps": a.steps,
           "config": {"workload": f"3-D blast {n}^3 periodic, ng=3, WENO5-Z + HLLC + SSPRK3, CFL 0.3, "
                                  "hipGraph-replayed steps (z sweep + per-plane x-y march)"}}
    for form, name in ((M.FP64, "fp64"), (M.FP32C, "fp32c")):
        el, tm, tune = res[form]
        v = cells * a.steps / el / 1e6
        out[name] = {"value": round(v, 1), "ms_per_step": round(el / a.steps * 1e3, 4),
                     "algorithmic_GBs": round(v * 1e6 * 320 / 1e9, 1),
                     "roofline_frac": round(v * 1e6 * 320 / 1e9 / HBM_PEAK_GBS, 4),
                     "span_rows": tune["span_rows"] if tune else None,
                     "kernel_ms_per_step": {k: round(getattr(tm, k) / a.steps, 4) for k in ("stage_ms", "ghost_ms", "clock_ms")}}
    g = cfg.ng
    a32, a64 = states[M.FP32C][:, g:-g, g:-g, g:-g], states[M.FP64][:, g:-g, g:-g, g:-g]
    out["rel_l1_fp32c_vs_fp64"] = [float(f"{r:.3e}") for r in
                                   (np.abs(a32 - a64).reshape(5, -1).sum(1) / np.abs(a64).reshape(5, -1).sum(1))]
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=2048, help="grid size (the workload is 2048)")
    ap.add_argument("--parity-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch kernels one by one (profiler passes)")
    ap.add_argument("--span-rows", type=int, default=0, help="rows per wavefront (0: autotuned)")
    ap.add_argument("--no-autotune", action="store_true", help="use the built-in span heuristic")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration) and exit")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="summarise rocprofv3 --pmc CSVs into the traffic JSON (CPU only) and exit")
    ap.add_argument("--copy-bytes", type=int, default=512 << 20)
    ap.add_argument("--trace-summary", metavar="KERNEL_TRACE_CSV",
                    help="summarise a rocprofv3 --kernel-trace CSV (CPU only) and exit")
    ap.add_argument("--workload", choices=["kh", "batch", "sedov", "blast3d"], default="kh",
                    help="kh: the headline line; batch / sedov: the other workloads; blast3d: the 3-D path")
    ap.add_argument("--n3", type=int, default=256, help="cells per side of the 3-D blast")
    ap.add_argument("--recon", default="weno5z", choices=["fog", "muscl", "ppm", "weno5z"],
                    help="KH reconstruction (counter passes; the headline is weno5z)")
    ap.add_argument("--solver", default="hllc", choices=["rusanov", "hll", "hllc"],
                    help="KH Riemann solver (counter passes; the headline is hllc)")
    ap.add_argument("--batch-steps", type=int, default=500)
    a = ap.parse_args(argv)
    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "blast3d":
        return bench_blast3d(a)
    if a.workload == "sedov":
        if a.n == 2048:
            a.n = 4096
        return bench_sedov(a)
    if a.pmc_summary:
        return pmc_summary(*a.pmc_summary, a.copy_bytes, TRAFFIC_JSON)
    if a.trace_summary:
        return trace_summary(a.trace_summary)
    if a.copy_probe_mb:
        return copy_probe(a.copy_probe_mb)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU; multi-GPU runs are not supported (a multi-GPU run would shard "
              "the grid in slabs with a halo exchange per stage)", file=sys.stderr)
        return 2

    import mosscap_amd as M
    from mosscap_amd import problems

    cfg = problems.kh_config(a.n)
    if (a.recon, a.solver) != ("weno5z", "hllc"):  # counter passes over other schemes; not the headline
        import dataclasses
        cfg = dataclasses.replace(cfg, reconstruction=a.recon, riemann_solver=a.solver)
    scheme = ("WENO5-Z + HLLC" if (a.recon, a.solver) == ("weno5z", "hllc")
              else f"{cfg.reconstruction} + {cfg.riemann_solver}")
    q0 = problems.kh_band(cfg)
    cells = cfg.nx * cfg.ny

    el64, tm64, tune64 = time_form(M, cfg, q0, M.FP64, a.steps, a.warmup, not a.no_graph, a.span_rows,
                                   not a.no_autotune)
    el32, tm32, tune32 = time_form(M, cfg, q0, M.FP32C, a.steps, a.warmup, not a.no_graph, a.span_rows,
                                   not a.no_autotune)
    v64 = cells * a.steps / el64 / 1e6
    v32 = cells * a.steps / el32 / 1e6
    rel = parity(M, cfg, q0, a.parity_steps) if a.parity_steps > 0 else None
    adj = [adjoint_ratio(M, problems, sv) for sv in ("rusanov", "hllc")] if not a.no_adjoint else None

    # dominant kernel: the fused stage; 256 algorithmic bytes per cell-step over 3 launches
    bytes_per_launch = 256.0 * cells / 3.0

    def roof(tm):
        avg_ms = tm.stage_ms / max(tm.stage_launches, 1)
        gbs = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        return avg_ms, gbs

    avg32, gbs32 = roof(tm32)
    avg64, gbs64 = roof(tm64)

    res = {
        "metric": METRIC,
        "value": round(v32, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(el32 * 1e3 / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32c",
        "data": "synthetic (seeded KH shear band, generated on the host)",
        "config": {"workload": f"KH {cfg.nx}x{cfg.ny} periodic, ng=3, {scheme} + SSPRK3, CFL 0.4, "
                               "hipGraph-replayed steps",
                   "form": "fp32c: FP32 recon/Riemann, hi+lo FP32 state, FP64 pressure recovery and dt"},
        "fp64": {"value": round(v64, 2), "unit": "Mcell-updates/s", "ms_per_step": round(el64 * 1e3 / a.steps, 4),
                 "stage_kernel_ms": round(avg64, 4), "stage_kernel_GBs": round(gbs64, 1)},
        "fp32c_over_fp64": round(v32 / v64, 3),
        "rel_l1_fp32c_vs_fp64": None if rel is None else {"steps": a.parity_steps,
                                                            "rho_mx_my_E": [float(f"{r:.3e}") for r in rel]},
        "adjoint_over_forward": None if adj is None else max(a["ratio"] for a in adj),
        "adjoint": adj,
        "roofline": {"bound": "hbm", "achieved": round(gbs32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs32 / HBM_PEAK_GBS, 4),
                     # the plain instance (StageInst 0; named "false" before there were three)
                     "traffic": measured_traffic("void mh::stage2d_kernel<mh::FormF32c, (mh::Reconstruction)3, "
                                                 "(mh::SlopeLimiter)1, (mh::RiemannSolver)2, ",
                                                 ("0>(", "false>(")),
                     "kernel": "stage2d_kernel<FormF32c, Weno5Z, MC, HLLC, 0 (plain instance)>",
                     "avg_launch_ms": round(avg32, 4), "algorithmic_bytes_per_launch": int(bytes_per_launch)},
        "autotune": {"fp32c": tune32, "fp64": tune64},
        "kernel_split_ms_per_step": {
            "stage": round(tm32.stage_ms / a.steps, 4), "ghost": round(tm32.ghost_ms / a.steps, 4),
            "clock": round(tm32.clock_ms / a.steps, 4)},
    }
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(cfg, q0, a.cpu_seconds)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
