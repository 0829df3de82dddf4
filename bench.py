"""Benchmark of the MI355X conic IPM solve path.

A step is one full IPM solve to OPTIMAL at gap_tol = feas_tol = 1e-8 of one seeded instance of
the 2000-variable sparse QP workload (n=2000, m=1000, nonnegative orthant, Q = PᵀP + 1e-3 I with
P at 0.5 % density, B at 1 % density). Every timed instance is distinct (its own seed) and
resident in HBM — problem data, symbolic factor and buffers set up before the timed region — and
is solved from the cold starting point. `--slots S` instances are solved at a time, each from its
own host thread on its own HIP stream, so their kernels overlap on the GPU; the persistent
factorization of each takes 1/S of the CUs. value = solves/s over the K timed solves (whole job);
p50/p99 latency over the same solves; solo_ms = one instance alone on the whole GPU.

roofline: the dominant kernel, the persistent FP64 MFMA Cholesky of the 2000×2000 front,
timed live with HIP events on each solver's stream over the timed region; achieved = its
algorithmic flops r³/3 per launch ÷ its measured average launch time (under the S-way sharing).

cpu_baseline: the oracle (numpy + LAPACK restatement of the reference path) on the host cores,
one instance.

`--serve`: a served stream instead — requests of mixed kind and size (QP and SOCP, n log-uniform
in [50, 4000]) arrive as a Poisson process and go through the continuous-batching loop
(IPMServer); latency = finish − arrival, including host analysis and device setup.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "mumblebee.jl_amd"))

import numpy as np  # noqa: E402

PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X dense FP64 matrix
PEAK_HBM = 8000.0      # GB/s
METRIC = "solves/s + p99 latency at 1e-8 gap (mixed vs plain FP64); factor TFLOP/s"
# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes
# scaled by a 256 MiB device copy (profiles/k_dag_factor_qp2000_fp64.md); keyed by (n, m)
TRAFFIC_BYTES = {(2000, 1000): 144.0e6}


def workload(seed, n, m):
    from mumblebee_hip import generators as G
    return G.qp(n=n, m=m, seed=seed)


def cpu_baseline(n, m, seed):
    from oracle import ipm as O
    from oracle.problem import CSC, make_problem
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        cores = os.cpu_count() or 1
    d = workload(seed, n, m)
    t0 = time.perf_counter()
    prob = make_problem(CSC.from_scipy(d["Q"]), CSC.from_scipy(d["B"]), d["f"], d["g"], 0.0, d["K"], d["s"])
    solver = O.Solver(prob, order="identity")
    t1 = time.perf_counter()
    res = solver.solve()
    t2 = time.perf_counter()
    return dict(value=1.0 / (t2 - t1), unit="solves/s", cores=int(cores), kind="port",
                sample=f"1 solve of the QP n={n} m={m} seed={seed} by the numpy/LAPACK oracle "
                       f"(solve {t2 - t1:.2f} s after {t1 - t0:.2f} s setup; status "
                       f"{O.STATUS_NAMES[res.status]}, {res.niter} iterations)")


def run_concurrent(solvers, slots):
    """Solve every solver from `slots` host threads; returns (elapsed s, latencies s, results)."""
    lat = [0.0] * len(solvers)
    res = [None] * len(solvers)
    nxt = [0]
    lock = threading.Lock()

    def worker():
        while True:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            if i >= len(solvers):
                return
            a = time.perf_counter()
            res[i] = solvers[i].solve()
            lat[i] = time.perf_counter() - a

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker) for _ in range(max(1, slots))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    return time.perf_counter() - t0, lat, res


def bench_resident(args, mh):
    K, W, S = args.steps, args.warmup, max(1, args.slots)
    ncu = mh.device_info()["cus"]
    settings = mh.IPMSettings(dag_grid=max(16, ncu // S))
    kinds = ["factor", "update", "panel", "fsolve", "bsolve", "extend_add", "dag"]
    t_setup = time.perf_counter()
    solvers = []
    for k in range(W + K):
        d = workload(1000 + k, args.n, args.m)
        prob = mh.IPMProblem(d["Q"], d["B"], d["f"], d["g"], 0.0, [mh.PositiveCone()], d["s"])
        solvers.append(mh.init(prob, settings))
    t_setup = (time.perf_counter() - t_setup) / (W + K)
    # one instance alone on the whole GPU (latency reference; not part of the timed job)
    d = workload(999, args.n, args.m)
    solo = mh.init(mh.IPMProblem(d["Q"], d["B"], d["f"], d["g"], 0.0, [mh.PositiveCone()], d["s"]))
    solo.solve()
    solo.reinit()
    a = time.perf_counter()
    r = solo.solve()
    solo_ms = (time.perf_counter() - a) * 1e3
    assert r.status == mh.IPMStatus.OPTIMAL, r.status
    solo.close()
    if W:
        _, _, wres = run_concurrent(solvers[:W], S)
        assert all(x.status == mh.IPMStatus.OPTIMAL for x in wres), [x.status for x in wres]
    timed = solvers[W:]
    for s in timed:
        s.set_profile(kinds)
    elapsed, lat, res = run_concurrent(timed, S)
    assert all(x.status == mh.IPMStatus.OPTIMAL for x in res), [x.status for x in res]
    prof = {}
    for s in timed:
        for k, v in s.profile().items():
            acc = prof.setdefault(k, dict(ms=0.0, launches=0, flops=0.0, bytes=0.0))
            for key in acc:
                acc[key] += v[key]
    st0 = timed[0].stats()
    fac = prof["factor"]
    dom_name = "dag" if prof["dag"]["ms"] >= prof["update"]["ms"] else "update"
    dom = prof[dom_name]
    achieved = dom["flops"] / (dom["ms"] * 1e-3) / 1e12 if dom["ms"] > 0 else 0.0
    lat_ms = np.array(lat) * 1e3
    line = {
        "metric": METRIC,
        "value": K / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"sparse convex QP n={args.n} m={args.m} on the nonnegative orthant, "
                        f"Q=PᵀP+1e-3I (P 0.5% dense), B 1% dense, N(0,1); {K} distinct seeded instances, "
                        f"resident, cold start, gap_tol=feas_tol=1e-8, {S} solved at a time",
            "precision": "fp64",
            "slots": S,
            "dag_grid": settings.dag_grid,
            "p50_ms": float(np.percentile(lat_ms, 50)),
            "p99_ms": float(np.percentile(lat_ms, 99)),
            "solo_ms": solo_ms,
            "iterations_mean": float(np.mean([x.niter for x in res])),
            "factor_tflops_aggregate": fac["flops"] / elapsed / 1e12,
            "factor_ms_per_call": fac["ms"] / max(1, fac["launches"]),
            "factor_exact_flops": st0["exact_flops"],
            "factor_front_flops": st0["front_flops"],
            "nfront": st0["nfront"],
            "setup_s_per_instance": t_setup,
            "kernel_ms": {k: v["ms"] for k, v in prof.items()},
            "kernel_launches": {k: v["launches"] for k, v in prof.items()},
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP64_MFMA,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_MFMA,
            "traffic": TRAFFIC_BYTES.get((args.n, args.m)) if dom_name == "dag" else None,
            "kernel": ("k_dag_factor (persistent left-looking tile FP64 MFMA Cholesky of the large fronts)"
                       if dom_name == "dag" else "k_update (FP64 MFMA trailing update of the frontal Cholesky)"),
            "avg_launch_us": dom["ms"] / max(1, dom["launches"]) * 1e3,
            "flops_per_launch": dom["flops"] / max(1, dom["launches"]),
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.n, args.m, 1000 + W)
    return line


def bench_serve(args, mh):
    """A Poisson stream of mixed requests through IPMServer."""
    from mumblebee_hip import generators as G
    rng = np.random.default_rng(7)
    K = args.steps
    reqs = []
    for k in range(K):
        n = int(G.log_uniform_n(rng, 50, args.serve_nmax))
        if rng.random() < 0.7:
            d = G.qp(n=n, m=max(1, n // 2), seed=5000 + k)
        else:
            d = G.socp(n=n, m=max(1, n // 2), ncone=max(1, n // 40), seed=5000 + k, adens=min(0.2, 20.0 / n))
        cones = [{"pos": mh.PositiveCone, "free": mh.CofreeCone, "soc": mh.SecondOrderCone}[c[0]]() for c in d["K"]]
        reqs.append(mh.IPMProblem(d["Q"], d["B"], d["f"], d["g"], 0.0, cones, d["s"]))
    gaps = rng.exponential(1.0 / args.rate, K)
    server = mh.IPMServer(nslots=args.slots)
    t0 = mh.clock()
    arrivals, tickets = [], []
    for k in range(K):
        target = t0 + float(np.sum(gaps[:k + 1]))
        while mh.clock() < target:
            time.sleep(min(0.002, max(0.0, target - mh.clock())))
        arrivals.append(mh.clock())
        tickets.append(server.submit(reqs[k]))
    out = [server.wait(t) for t in tickets]
    st = server.stats()
    server.close()
    ok = sum(r.status == mh.IPMStatus.OPTIMAL for r, _ in out)
    fin = np.array([t["finished"] for _, t in out])
    lat_ms = (fin - np.array([t["submitted"] for _, t in out])) * 1e3
    elapsed = float(fin.max() - t0)
    return {
        "metric": METRIC, "value": K / elapsed, "unit": "solves/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": elapsed / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"served stream: {K} requests, 70% QP / 30% SOCP, n log-uniform in "
                               f"[50, {args.serve_nmax}], Poisson arrivals at {args.rate}/s, {args.slots} slots",
                   "optimal": ok, "p50_ms": float(np.percentile(lat_ms, 50)),
                   "p99_ms": float(np.percentile(lat_ms, 99)), "peak_reserved_gb": st["peak_reserved_bytes"] / 2 ** 30},
    }


def copy_calibration(mib):
    """A plain device copy of known size (mib MiB read, mib MiB written, 3 launches): the reference
    the HBM traffic counters of the solve kernels are scaled by."""
    import torch
    n = mib * 2 ** 20 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--n", type=int, default=2000)
    ap.add_argument("--m", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", action="store_true")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--serve-nmax", type=int, default=4000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    # HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default); slots beyond that
    # would share a queue and run their kernels one after another, so give each slot its own
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(min(32, max(4, args.slots))))
    import mumblebee_hip as mh
    line = bench_serve(args, mh) if args.serve else bench_resident(args, mh)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
