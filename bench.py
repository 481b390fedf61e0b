#!/usr/bin/env python3
"""Chemistry-step benchmark on one MI355X.

Default workload (--workload 3): the seeded 53-species / 325-reaction mechanism on
1048576 cells with an engineered stiffness spread (T 300-2500 K, radicals
log-uniform 1e-12 - 1e-3), the largest single-GPU configuration.  Each cell is
integrated as a constant-pressure adiabatic reactor over dt = 1e-5 s at rtol 1e-7,
atol 1e-12.  A "step" is one chemistry step over all cells, i.e. the batched forAll
loop of chemistry_DI.H.  Every step integrates the same input field, which stays
resident in HBM.  Other workloads:
  --workload 1: the 14-species / 34-reaction mechanism on its 4096 cells (T 900-2400 K,
                Y Dirichlet), dt = 1e-5 s, rtol 1e-7;
  --workload 2: the 53-species mechanism on the 512x512 coflow-like field (262144 cells;
                60% cold 300-600 K, 40% flame 1200-2200 K);
  --workload 4: the 160-species / 1000-reaction mechanism (wide kernel, FP32
                Newton LU with FP64 refinement) on a subsample of its
                131072-cell field, at rtol 1e-6.
  --workload ch4: the reference's CH4 skeletal mechanism (POLIMI_CH4_SKELETAL_1412,
                24 species / 155 reactions, CHEMKIN input from tests/golden) with
                the OdeHomogeneous settings of its MethaneCoflowFlame tutorial
                (rtol 1e-6, atol 1e-12, fullPivoting true: team kernel with row
                pivoting) on a 512x512 coflow field of premixed fuel/air cells
                (40 % flame band 1200-2200 K with radicals, the rest 300-600 K),
                dt = 1e-5 s.

Before the timed steps:
  * the launch shape (cells per wavefront, waves per workgroup, workgroups per CU) is
    autotuned on a 16384-cell sample;
  * one warm-up step runs, and its per-cell step counts give the visiting order
    of the timed steps (most expensive cells first), as the previous CFD step's
    counters would.
With --graph, the step (queue reset + integrator) is captured once in a HIP
graph and replayed.  With --step-cap K every cell stops after K BDF steps, so
that kernels can be compared at equal work where full integrations of the slow
plain kernel would take hours (160-species mechanism).

Prints one JSON line with:
  * cell-integrations/s and the p50/p99 batch latency;
  * the roofline of the integrator kernel: algorithmic FP64 operations from the
    per-cell counters, divided by the kernel time from HIP events on the launch
    stream, against the FP64 vector peak (the kernel issues no MFMA); its HBM
    traffic is the copy-calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE of the same
    configuration, read from the committed profile summary named in traffic_source;
  * the CPU baseline: the oracle on a bounded random sample of the same cells, on
    this host, one warm-up run and the median of five (1 thread; all 16 threads of
    the box's share in extra), and the GPU's deviation from the oracle on that
    sample against the oracle's own spread.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-ODE integrations/s per chemistry step at stated rtol; p50/p99 batch latency"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def progress(msg):
    """Progress of the phases on stderr (the JSON line alone goes to stdout)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def cpu_baseline(mech, y, n_sample, threads, dt, rtol, atol, full_pivoting=False, y_gpu=None, st_gpu=None,
                 jacobian="numerical"):
    """Oracle timing on a seeded sample of the field (one warm-up run, then the median of five);
    with the GPU's outputs of the same cells, also their deviation from the oracle in units of
    rtol |y| + atol, against the oracle's own spread on those cells."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    idx = np.sort(np.random.default_rng(2024).choice(y.shape[1], n_sample, replace=False))
    ys = np.ascontiguousarray(y[:, idx])
    run = lambda: pyoracle.chemistry_step(mech, ys, dt, rtol=rtol, atol=atol, threads=threads,
                                          full_pivoting=full_pivoting)
    progress(f"CPU baseline: {n_sample} cells, {threads} thread(s)")
    yo, st, _, cnt = run()                      # warm-up (its result is the reference below)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    el = float(np.median(times))
    out = dict(value=n_sample / el, unit="cell-integrations/s", cores=threads, kind="port",
               sample=f"{n_sample} cells drawn at random (seed 2024) from the same field; one warm-up run, "
                      f"median of 5 runs = {el:.2f} s (min {min(times):.2f}, max {max(times):.2f}); oracle C "
                      f"restatement, {threads} thread(s), host {platform.processor() or platform.machine()} "
                      f"with {os.cpu_count()} logical CPUs")
    err = None
    if y_gpu is not None:
        dev = lambda a: (np.abs(a - yo) / (rtol * np.abs(yo) + atol)).max(axis=0)
        r = dev(y_gpu[:, idx])
        # the oracle's own spread: 5 seeded draws of 1e-15 relative input noise and, for a Newton
        # matrix that differs on purpose (analytic Jacobian), FD increments scaled by 0.5 and 2
        par = min(16, os.cpu_count() or 1)
        rng = np.random.default_rng(12345)
        env = np.zeros(n_sample)
        for _ in range(5):
            yp = ys * (1.0 + 1e-15 * rng.standard_normal(ys.shape))
            y2, *_ = pyoracle.chemistry_step(mech, yp, dt, rtol=rtol, atol=atol, threads=par,
                                             full_pivoting=full_pivoting)
            env = np.maximum(env, dev(y2))
        if jacobian == "analytic":
            for sc in (0.5, 2.0):
                y2, *_ = pyoracle.chemistry_step(mech, ys, dt, rtol=rtol, atol=atol, threads=par,
                                                 full_pivoting=full_pivoting, jac_increment_scale=sc)
                env = np.maximum(env, dev(y2))
        bound = np.maximum(10.0, 3.0 * env)
        # cells beyond the bound: 256 more noise draws each, as in the parity tests (some cells are
        # bimodal, a borderline step decision going one way or the other with the rounding; on one
        # cell of the 1M-cell field's sample the oracle's minority mode, 858 units away, comes up in
        # about 1 % of the draws, so 64 draws miss it about half the time)
        beyond = np.where(r > bound)[0]
        if beyond.size:
            nd = 256
            yp = np.repeat(ys[:, beyond], nd, axis=1)
            yp = np.ascontiguousarray(yp * (1.0 + 1e-15 * np.random.default_rng(54321).standard_normal(yp.shape)))
            y2, *_ = pyoracle.chemistry_step(mech, yp, dt, rtol=rtol, atol=atol, threads=par,
                                             full_pivoting=full_pivoting)
            d2 = (np.abs(y2 - np.repeat(yo[:, beyond], nd, axis=1)) /
                  (rtol * np.abs(np.repeat(yo[:, beyond], nd, axis=1)) + atol)).max(axis=0)
            env[beyond] = np.maximum(env[beyond], d2.reshape(beyond.size, nd).max(axis=1))
            bound = np.maximum(10.0, 3.0 * env)
        over = float((r / bound).max())
        err = {"cells": int(n_sample), "max": float(r.max()), "p50": float(np.median(r)),
               "within_10x": float(np.mean(r <= 10.0)),
               "oracle_envelope_max": float(env.max()),
               "within_max_10_3env": float(np.mean(r <= bound)),
               "max_over_bound": over, "parity_ok": bool(over <= 1.0 and np.array_equal(st_gpu[idx], st)),
               "statuses_equal": bool(np.array_equal(st_gpu[idx], st)),
               "unit": "max_k |y_gpu - y_oracle| / (rtol |y_oracle| + atol) per cell; bound max(10, 3 x envelope), "
                       "envelope: the oracle against itself under 1e-15 input noise (5 draws, 256 more on cells "
                       "beyond the bound)"
                       + (" and with its FD-Jacobian increments scaled by 0.5 and 2" if jacobian == "analytic" else "")}
        if not err["parity_ok"]:
            print(f"bench.py: GPU result outside the parity bound on the CPU sample (max_over_bound {over:.3g}, "
                  f"statuses equal {err['statuses_equal']})", file=sys.stderr, flush=True)
    return out, err


TRAFFIC_PROFILES = {
    ("cfg3-53sp-spread-1M", "team", "analytic"): "team_kernel_cfg3_traffic.md",
}


def measured_traffic(workload_name, method, jacobian):
    """HBM bytes per launch measured with rocprofv3 (FETCH_SIZE + WRITE_SIZE, copy-calibrated) for this
    configuration, from its committed profile summary: (bytes, file) or (None, None)."""
    name = TRAFFIC_PROFILES.get((workload_name, method, jacobian))
    if name is None:
        return None, None
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1]), "profiles/" + name
    except OSError:
        pass
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="3", choices=["1", "2", "3", "4", "ch4"])
    ap.add_argument("--cells", type=int, default=None, help="subsample of the field (default: all of it; "
                                                               "workload 4: 4096)")
    ap.add_argument("--method", default="auto", choices=["team", "plain", "wide", "auto"])
    ap.add_argument("--jacobian", default="analytic", choices=["numerical", "analytic"])
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cells-per-wave", type=int, default=0,
                    help="team kernel: 1, 2 (ns+1 <= 32) or 4 (ns+1 <= 16) cells per wavefront (default: the "
                         "autotuner's choice, 1 without it)")
    ap.add_argument("--no-order", action="store_true")
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--step-cap", type=int, default=0,
                    help="@MaximumNumberOfSteps for every cell (equal-work kernel comparisons; cells stop with "
                         "status -1 at the cap)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="run each step as consecutive launches over blocks of this many cells, with a progress "
                         "line after each (steps of minutes; implies --no-order)")
    ap.add_argument("--cpu-sample", type=int, default=None, help="cells of the 1-thread CPU sample (the 16-thread "
                                                                 "sample is 16 times larger)")
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(cells shard with no exchange; see the design notes on sharding)", file=sys.stderr)
        sys.exit(2)

    import torch
    from laminarsmoke_amd import api, workmodel
    from laminarsmoke_amd.fields import workload_cells, WORKLOADS

    full_pivoting = False
    if args.workload == "ch4":
        from laminarsmoke_amd.chemkin import load_chemkin
        from laminarsmoke_amd.fields import coflow_premixed_field
        golden = os.path.join(ROOT, "tests", "golden")
        mech = load_chemkin(os.path.join(golden, "POLIMI_CH4_SKELETAL_1412.CKI"),
                            os.path.join(golden, "POLIMI_CH4_SKELETAL_1412_thermo.CKT"))
        w = dict(name="ch4-skeletal-coflow-512x512", rtol=1e-6, atol=1e-12, dt=1e-5)
        rtol, atol, dt = w["rtol"], w["atol"], w["dt"]
        full_pivoting = True
        cells = args.cells
        y, _ = coflow_premixed_field(mech, 512, 512, 77)
        if cells is not None and cells < y.shape[1]:
            y = np.ascontiguousarray(y[:, np.linspace(0, y.shape[1] - 1, cells).astype(np.int64)])
    else:
        args.workload = int(args.workload)
        w = WORKLOADS[args.workload]
        rtol, atol, dt = w["rtol"], w["atol"], w["dt"]
        cells = args.cells if args.cells is not None else (4096 if args.workload == 4 else None)
        mech, y = workload_cells(args.workload, n=cells)
    n = y.shape[1]
    dev = torch.device("cuda:0")
    y_in = torch.from_numpy(y).to(dev)
    dts = torch.full((n,), dt, dtype=torch.float64, device=dev)
    y_out = torch.empty_like(y_in)
    st = torch.zeros(n, dtype=torch.int32, device=dev)
    q = torch.zeros(n, dtype=torch.float64, device=dev)
    cnt = torch.zeros((n, 12), dtype=torch.int32, device=dev)
    order = torch.arange(n, dtype=torch.int32, device=dev)

    ode = api.ODE_Parameters()
    ode.SetType("OpenSMOKE_GPU")
    ode.SetRelativeTolerance(rtol)
    ode.SetAbsoluteTolerance(atol)
    if full_pivoting:
        ode.SetFullPivoting(True)
    if args.step_cap > 0:
        ode.SetMaximumNumberOfSteps(args.step_cap)
    step = api.ChemistryStep(mech, ode, method=args.method, jacobian=args.jacobian)
    stream = torch.cuda.current_stream(dev)

    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = step.autotune(n, dts.data_ptr(), y_in.data_ptr(), 16384, stream.cuda_stream)
        tune["seconds"] = time.perf_counter() - t0
        progress(f"autotuned launch shape {tune}")
    if args.cells_per_wave > 0:
        step.set_cells_per_wave(args.cells_per_wave)

    use_order = False

    blocks = [(a, min(a + args.chunk, n)) for a in range(0, n, args.chunk)] if args.chunk > 0 else [(0, n)]
    if len(blocks) > 1:
        if args.graph:
            print("bench.py: --graph and --chunk cannot be combined", file=sys.stderr)
            sys.exit(2)
        args.no_order = True
        # each block is its own SoA batch [ns+1][cells of the block]
        parts = [dict(y_in=y_in[:, a:b].contiguous(), y_out=torch.empty((y.shape[0], b - a), dtype=torch.float64,
                                                                         device=dev),
                      dt=dts[a:b].contiguous(), st=st[a:b].clone(), q=q[a:b].clone(), cnt=cnt[a:b].clone())
                 for a, b in blocks]

    def run():
        if len(blocks) == 1:
            step.integrate_device(n, dts.data_ptr(), y_in.data_ptr(), y_out.data_ptr(), st.data_ptr(), q.data_ptr(),
                                  cnt.data_ptr(), stream.cuda_stream, order.data_ptr() if use_order else None)
            return
        for i, ((a, b), p) in enumerate(zip(blocks, parts)):
            step.integrate_device(b - a, p["dt"].data_ptr(), p["y_in"].data_ptr(), p["y_out"].data_ptr(),
                                  p["st"].data_ptr(), p["q"].data_ptr(), p["cnt"].data_ptr(), stream.cuda_stream)
            torch.cuda.synchronize()
            progress(f"block {i + 1}/{len(blocks)} done")

    progress(f"{n} cells of workload {w['name']}; warm-up steps: {args.warmup}")
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    if not args.no_order and args.warmup > 0:
        step.cost_order(n, cnt.data_ptr(), order.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        use_order = True

    graph = None
    if args.graph:
        side = torch.cuda.Stream(dev)
        side.wait_stream(stream)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                step.integrate_device(n, dts.data_ptr(), y_in.data_ptr(), y_out.data_ptr(), st.data_ptr(),
                                      q.data_ptr(), cnt.data_ptr(), side.cuda_stream,
                                      order.data_ptr() if use_order else None)
        stream.wait_stream(side)
        torch.cuda.synchronize()

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        if graph is not None:
            graph.replay()
        else:
            run()
        b.record(stream)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    progress(f"{args.steps} timed steps: {total:.2f} s")
    lat_ms = np.array([a.elapsed_time(b) for a, b in evs])

    if len(blocks) > 1:                          # gather the blocks' outputs
        for (a, b), p in zip(blocks, parts):
            y_out[:, a:b] = p["y_out"]
            st[a:b], q[a:b], cnt[a:b] = p["st"], p["q"], p["cnt"]
    status = st.cpu().numpy()
    counters = cnt.cpu().numpy()
    work = workmodel.flops_from_counters(mech, counters, args.jacobian)
    kernel_s = float(lat_ms.mean()) / 1e3
    achieved_tflops = work["flops"] / kernel_s / 1e12
    value = n * args.steps / total
    launch = step.launch()

    name = w["name"] if cells is None else f"{w['name']}-sub{n}"
    traffic, traffic_src = measured_traffic(name, launch["method"], args.jacobian)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-integrations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded mechanism and cell field)",
        "config": {"workload": name, "mechanism": f"{mech.ns} species / {mech.nr} reactions", "cells": n,
                   "dt_s": dt, "rtol": rtol, "atol": atol, "method": launch["method"], "jacobian": args.jacobian,
                   "newton_lu": "fp32+refinement" if launch["method"] == "wide" else "fp64",
                   "pivoting": ("full (plain kernel)" if launch["method"] == "plain" else
                                "rows (fullPivoting requested)") if full_pivoting else "rows",
                   "waves_per_workgroup": launch["waves_per_block"], "grid": launch["grid"],
                   "cells_per_wave": launch["cells_per_wave"],
                   "visit_order": "cost (previous step's counters)" if use_order else "cell index",
                   "hip_graph": graph is not None, "step_cap": args.step_cap or None,
                   "launch_blocks": len(blocks)},
        "p50_batch_ms": float(np.percentile(lat_ms, 50)), "p99_batch_ms": float(np.percentile(lat_ms, 99)),
        "roofline": {"bound": "valu-fp64", "pipe": "fp64 vector (no MFMA in this kernel)", "achieved": achieved_tflops,
                     "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP64_TFLOPS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_flops_per_launch": work["flops"],
                     "special_ops_per_launch": work["special_ops"],
                     "kernel_ms": kernel_s * 1e3},
        "extra": {"status_counts": {str(k): int(v) for k, v in zip(*np.unique(status, return_counts=True))},
                  "steps_per_cell_p50": float(np.median(counters[:, 0])),
                  "steps_per_cell_max": int(counters[:, 0].max()),
                  "rhs_per_cell_mean": float(counters[:, 9].mean()),
                  "min_hbm_GBps": workmodel.hbm_bytes_min(mech, n) / kernel_s / 1e9,
                  "autotune": tune},
    }
    if not args.no_cpu:
        # bounded samples: about 20-30 s of CPU work per timing (warm-up + 5 runs)
        ns = args.cpu_sample if args.cpu_sample is not None else {2: 48, 3: 512, 4: 4, "ch4": 1024}.get(args.workload, 96)
        y_gpu = y_out.cpu().numpy()
        out["cpu_baseline"], err = cpu_baseline(mech, y, min(ns, n), args.cpu_threads, dt, rtol, atol, full_pivoting,
                                                y_gpu=y_gpu, st_gpu=status, jacobian=args.jacobian)
        # the same cells through the oracle: per-cell deviation of the GPU result
        out["extra"]["error_vs_oracle"] = err
        # the oracle on all of the box's CPU share (16 threads, dynamic chunks), on a 16x larger sample
        par, _ = cpu_baseline(mech, y, min(16 * ns, n), 16, dt, rtol, atol, full_pivoting)
        out["extra"]["cpu_baseline_16_threads"] = par
    step.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
