"""Serving benchmark: chemistry-step requests from 8 concurrent synthetic CFD cases.

``python -m laminarsmoke_amd.serve_bench`` prints one JSON line.

Cases:
  * mechanisms: three 14-species, three 53-species and two 160-species cases;
  * mesh sizes: log-uniform in [1e3, 5e5] cells (seed 5), capped at 1e5 cells for
    the 53-species and 1024 cells for the 160-species mechanism (the 14-species
    cases are not capped), so that a request takes seconds: a full step over 5e5
    cells of the 160-species mechanism would take about 40 minutes at the wide
    kernel's rate on these hot cells (about 220 cells/s);
  * initial fields: the workloads' generators at the case's size (14 species:
    T 900-2400 K, Dirichlet composition; 53: a coflow field; 160: hot cells
    800-2400 K);
  * dt of each request: log-uniform in [1e-6, 1e-4] s;
  * every request ships its case's initial field as the new state (what a CFD
    code sends after its transport step), so that each step integrates the stiff
    initial chemistry rather than a field already relaxed by earlier steps.

Procedure:
  1. Capacity: one request per case, all queued at once, drained.  The rate is
     requests/s; the throughput is cells/s.
  2. For offered loads of 50%, 80% and 100% of that capacity: Poisson arrivals at
     that request rate, each for a case drawn at random.  The server admits each
     request by its HBM budget (the cases' handles plus their resident states),
     overlaps cases on the device and keeps every state resident.  For each load
     (100 requests by default, so that p99 is not the largest latency): p50, p90
     and p99 latency (submission to completion), achieved cells/s, and the device
     concurrency from the steps' device timestamps: the sum of the steps' device
     times over the length of the union of their intervals (1: no overlap).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import api
from .fields import coflow_field, dirichlet_cells, hot_cells
from .mechanism import workload_mechanism

RTOL = {14: 1e-7, 53: 1e-7, 160: 1e-6}
CAP = {14: 500000, 53: 100000, 160: 1024}


def make_cases(seed=5):
    rng = np.random.default_rng(seed)
    kinds = [14, 14, 14, 53, 53, 53, 160, 160]
    cases = []
    for k, ns in enumerate(kinds):
        n = min(int(np.exp(rng.uniform(np.log(1e3), np.log(5e5)))), CAP[ns])
        mech = workload_mechanism(ns)
        if ns == 14:
            y = dirichlet_cells(mech, n, 900.0, 2400.0, 100 + k)
        elif ns == 53:
            nx = int(np.sqrt(n))
            y, _ = coflow_field(mech, nx, max(1, n // nx), 100 + k)
        else:
            y = hot_cells(mech, n, 100 + k)
        cases.append(dict(ns=ns, n=y.shape[1], y=np.ascontiguousarray(y), mech=mech))
    return cases


def concurrency(out):
    """Sum of the steps' device times over the length of the union of their device intervals."""
    iv = sorted((o[3], o[4]) for o in out)
    union, cur0, cur1 = 0.0, None, None
    for a, b in iv:
        if cur1 is None or a > cur1:
            if cur1 is not None:
                union += cur1 - cur0
            cur0, cur1 = a, b
        else:
            cur1 = max(cur1, b)
    if cur1 is not None:
        union += cur1 - cur0
    return sum(b - a for a, b in iv) / union if union > 0 else 1.0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=100, help="requests per load level")
    ap.add_argument("--budget-gb", type=float, default=16.0)
    ap.add_argument("--inflight", type=int, default=4)
    ap.add_argument("--loads", default="0.5,0.8,1.0")
    args = ap.parse_args(argv)

    rng = np.random.default_rng(7)
    cases = make_cases()
    srv = api.Server(int(args.budget_gb * (1 << 30)), args.inflight)
    steps, ids = [], []
    for c in cases:
        ode = api.ODE_Parameters()
        ode.SetRelativeTolerance(RTOL[c["ns"]])
        ode.SetAbsoluteTolerance(1e-12)
        st = api.ChemistryStep(c["mech"], ode, method="auto", jacobian="analytic")
        steps.append(st)
        ids.append(srv.add_case(st, c["y"]))
    cells = {cid: c["n"] for cid, c in zip(ids, cases)}
    fields = {cid: c["y"] for cid, c in zip(ids, cases)}
    draw_dt = lambda: float(np.exp(rng.uniform(np.log(1e-6), np.log(1e-4))))

    def run(arrivals):
        """arrivals: list of (time offset s, case id, dt); returns per-request (latency, cells)."""
        t0 = time.perf_counter()
        owner, out, k = {}, [], 0
        while k < len(arrivals) or len(out) < len(arrivals):
            now = time.perf_counter() - t0
            while k < len(arrivals) and arrivals[k][0] <= now:
                _, cid, dt = arrivals[k]
                owner[srv.submit(cid, dt, fields[cid])] = cid
                k += 1
            for ticket, lat, dev, ts, te in srv.poll(timeline=True):
                out.append((lat, cells[owner[ticket]], dev, ts, te, owner[ticket]))
                print(f"  request {len(out)}/{len(arrivals)}: case {owner[ticket]} ({cells[owner[ticket]]} cells) "
                      f"latency {lat:.2f} s", file=sys.stderr, flush=True)
            if k < len(arrivals) or len(out) < len(arrivals):
                time.sleep(0.0005)
        return out, time.perf_counter() - t0

    # warm-up: one step per case (first-call costs, cost orders)
    run([(0.0, cid, 1e-6) for cid in ids])
    # capacity: one request per case, all at once
    cap_req, cap_wall = run([(0.0, cid, draw_dt()) for cid in ids])
    cap_cells = sum(o[1] for o in cap_req) / cap_wall
    cap_rate = len(ids) / cap_wall
    result = dict(metric="serving: p50/p99 chemistry-step latency and cells/s under Poisson load",
                  cases=[dict(species=c["ns"], cells=c["n"]) for c in cases],
                  capacity=dict(requests_per_s=cap_rate, cells_per_s=cap_cells, wall_s=cap_wall), loads=[])
    for load in [float(x) for x in args.loads.split(",")]:
        rate = load * cap_rate
        t, arrivals = 0.0, []
        for _ in range(args.requests):
            t += float(rng.exponential(1.0 / rate))
            arrivals.append((t, ids[int(rng.integers(len(ids)))], draw_dt()))
        out, wall = run(arrivals)
        lat = np.array([o[0] for o in out])
        result["loads"].append(dict(offered=load, requests=len(out), p50_s=float(np.percentile(lat, 50)),
                                    p90_s=float(np.percentile(lat, 90)), p99_s=float(np.percentile(lat, 99)),
                                    max_s=float(lat.max()), cells_per_s=sum(o[1] for o in out) / wall,
                                    wall_s=wall, device_concurrency=concurrency(out)))
        print(f"load {load}: {result['loads'][-1]}", file=sys.stderr, flush=True)
    result["server"] = srv.stats()
    srv.close()
    for st in steps:
        st.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
