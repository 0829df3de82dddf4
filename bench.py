"""Benchmark of the MI355X conic IPM solve path.

A step is one solve(prob, settings) — the reference's own unit of measurement
(examples/utils.jl:72-85 times `solve(prob, settings)`): init (symbolic analysis, equilibration,
BᵀB, device buffers and uploads, starting point), the IPM loop to OPTIMAL at gap_tol = feas_tol =
1e-8, and the result, for one seeded instance of the 2000-variable sparse QP workload (n=2000,
m=1000, nonnegative orthant, Q = PᵀP + 1e-3 I with P at 0.5 % density, B at 1 % density). Every
timed instance is distinct (its own seed); the problems are constructed (IPMProblem) before the
timed region, as the reference's prob is. `--threads T` host threads run the solves, so one
instance's host init overlaps the others' device work; each solver has its own HIP stream and the
persistent factorization of each takes 1/`--slots` of the CUs. value = solves/s over the K timed
solves (whole job); p50/p99 latency over the same solves. The IPM loop alone on instances
already resident in HBM is reported beside it (config.resident_loop).

roofline: the dominant kernel, the persistent FP64 MFMA Cholesky of the 2000×2000 front,
timed live with HIP events on each solver's stream over the timed region; achieved = its
algorithmic flops r³/3 per launch ÷ its measured average launch time (under the sharing).

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
ARGS = None
# HBM bytes per large front of the dominant kernel's launches, from rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes scaled by a 256 MiB device copy (FETCH_SIZE ×2, WRITE_SIZE ×1); keyed by (n, m,
# gang): the single-front kernel (profiles/k_dag_factor_qp2000_fp64.md) and the gang's combined
# launch, 16 fronts of which read 3.27 GB and write 0.32 GB (profiles/k_dag_factor_multi_qp2000_gang16.md)
TRAFFIC_PER_FRONT = {(2000, 1000, False): 141.5e6, (2000, 1000, True): 224.4e6}


WORKLOADS = {
    "qp": "sparse convex QP n={n} m={m} on the nonnegative orthant, Q=PᵀP+1e-3I (P 0.5% dense), B 1% dense, N(0,1)",
    "socp": "SOCP n=10000 m=5000: 500 second-order cones (sizes U{{3..50}} scaled to 9000 columns) + nonneg, "
            "B 0.2% dense, N(0,1), strictly feasible g, dual-feasible f",
    "maxcut": "max-cut SDP relaxation, 2000 vertices, band 20 + uniformly random extra edges with probability "
              "{extra} per vertex pair, weights U(0,1), C = -L/4, decomposed over the maximal cliques of a chordal "
              "extension (minimum-degree elimination, clique tree, cliques merged up to 48 vertices), one PSD cone "
              "per clique, consensus rows on clique-tree intersections, diag(X) = 1",
    "exppow": "2000 exponential cones (entropy on the simplex) + 500 power cones (alpha U(0.2,0.8)) with budgets "
              "+ one {psd}x{psd} PSD block (dense symmetrised N(0,1) C, tr X = 1), 3000 sparse equalities at 0.5% "
              "density over all columns, strictly feasible g",
}


def workload(kind, seed, n, m):
    from mumblebee_hip import generators as G
    if kind == "qp":
        return G.qp(n=n, m=m, seed=seed)
    if kind == "socp":
        return G.socp(seed=seed)
    if kind == "maxcut":
        return G.maxcut_chordal(nv=2000, band=20, extra=ARGS.maxcut_extra, seed=seed, merge=48)
    if kind == "exppow":
        return G.exppow(nexp=2000, npow=500, m=3000, psd=ARGS.psd, rows="all", seed=seed)
    raise ValueError(kind)


def hip_problem(mh, d):
    kinds = {"pos": mh.PositiveCone, "free": mh.CofreeCone, "soc": mh.SecondOrderCone,
             "psd": mh.SemidefiniteCone, "exp": mh.ExponentialCone}
    cones = [mh.PowerCone(c[1]) if c[0] == "pow" else kinds[c[0]]() for c in d["K"]]
    return mh.IPMProblem(d["Q"], d["B"], d["f"], d["g"], 0.0, cones, d["s"])


def cpu_baseline(kind, n, m, seed, gpu_iters, budget_s=30.0):
    """The oracle (numpy + LAPACK) on one instance: a whole solve when it ends within the budget,
    otherwise the IPM iterations done within it, converted to solves/s at the GPU's iterations
    per solve."""
    from oracle import ipm as O
    from oracle.problem import CSC, make_problem
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        cores = os.cpu_count() or 1
    d = workload(kind, seed, n, m)
    t0 = time.perf_counter()
    prob = make_problem(CSC.from_scipy(d["Q"]), CSC.from_scipy(d["B"]), d["f"], d["g"], 0.0, d["K"], d["s"])
    solver = O.Solver(prob, order="identity" if kind == "qp" else "mindegree")
    t1 = time.perf_counter()
    it, st = 0, 0
    while st == 0 and time.perf_counter() - t1 < budget_s:
        st = solver.step()
        it += 1
    t2 = time.perf_counter()
    if st != 0:
        return dict(value=1.0 / (t2 - t1), unit="solves/s", cores=int(cores), kind="port",
                    sample=f"1 solve of the {kind} instance seed={seed} by the numpy/LAPACK oracle "
                           f"(solve {t2 - t1:.2f} s after {t1 - t0:.2f} s setup; status {O.STATUS_NAMES[st]}, "
                           f"{it} iterations)")
    per_it = (t2 - t1) / it
    return dict(value=1.0 / (per_it * gpu_iters), unit="solves/s", cores=int(cores), kind="port",
                sample=f"{it} IPM iterations of the {kind} instance seed={seed} by the numpy/LAPACK oracle in "
                       f"{t2 - t1:.1f} s (after {t1 - t0:.1f} s setup), scaled to {gpu_iters:.1f} iterations "
                       f"per solve")


def note(msg):
    print(msg, file=sys.stderr, flush=True)


def run_concurrent(solvers, slots, tag=None):
    """Solve every solver from `slots` host threads; returns (elapsed s, latencies s, results).
    With a tag, each finished solve is noted on stderr."""
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
            if tag:
                note(f"{tag}: solve {i} {res[i].status.name} in {lat[i]:.2f} s ({res[i].niter} iterations)")

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker) for _ in range(max(1, slots))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    return time.perf_counter() - t0, lat, res


def pct(x, q):
    """Percentile of the latencies, None where they were not measured per solve."""
    x = np.asarray(x, dtype=float)
    return None if x.size == 0 or np.isnan(x).any() else float(np.percentile(x, q))


def _gang_stats(g):
    """The combined factorizations' rate from a gang's own events (profile=True)."""
    st = g.stats()
    st["dag_tflops"] = st["dag_flops"] / (st["dag_ms"] * 1e-3) / 1e12 if st["dag_ms"] > 0 else 0.0
    st["dag_us_per_launch"] = st["dag_ms"] / max(1.0, st["dag_launches"]) * 1e3
    return st


def _resident_run(args, mh, precision):
    """Set up W + K resident instances at the given precision, solve the W untimed and then the
    K timed ones (`--gang G`: through one gang of G members; else `--slots` at a time, each on its
    own stream); returns the measurements and the timed results."""
    K, W, S = args.steps, args.warmup, max(1, args.slots)
    ncu = mh.device_info()["cus"]
    grid = args.dag_grid if args.dag_grid > 0 else (max(16, ncu // S) if S > 1 and not args.gang else 0)
    settings = mh.IPMSettings(dag_grid=grid, precision=precision)
    kinds = ["factor", "update", "panel", "fsolve", "bsolve", "extend_add", "dag"]
    # exponential cones can end NEAR_OPTIMAL: the reference's line search stalls within a few
    # noise bands of a cone boundary (a documented property of the path, not of this build)
    good = (mh.IPMStatus.OPTIMAL, mh.IPMStatus.NEAR_OPTIMAL) if args.workload == "exppow" else (mh.IPMStatus.OPTIMAL,)
    t_setup = time.perf_counter()
    solvers = []
    for k in range(W + K):
        solvers.append(mh.init(hip_problem(mh, workload(args.workload, 1000 + k, args.n, args.m)), settings))
        if args.workload != "qp":
            note(f"{args.workload} precision {precision}: instance {k} set up")
    t_setup = (time.perf_counter() - t_setup) / (W + K)
    tag = None if args.workload == "qp" else f"{args.workload} precision {precision}"
    note(f"{args.workload} precision {precision}: {W + K} instances set up ({t_setup:.2f} s each)")
    # one instance alone on the whole GPU (latency reference; not part of the timed job)
    solo = mh.init(hip_problem(mh, workload(args.workload, 999, args.n, args.m)), mh.IPMSettings(precision=precision))
    solo.solve()
    solo.reinit()
    a = time.perf_counter()
    r = solo.solve()
    solo_ms = (time.perf_counter() - a) * 1e3
    assert r.status in good, r.status
    solo.close()
    note(f"{args.workload} precision {precision}: solo solve {solo_ms:.1f} ms")
    gst = None
    if args.gang:
        g = mh.IPMGang(capacity=args.gang, profile=True)
        if W:
            wres = g.solve(solvers[:W])
            assert all(x.status in good for x in wres), [x.status for x in wres]
            for s in solvers[:W]:
                s.close()
        g.close()
        g = mh.IPMGang(capacity=args.gang, profile=True)
        timed = solvers[W:]
        t0 = time.perf_counter()
        res = g.solve(timed)
        elapsed = time.perf_counter() - t0
        lat = [float("nan")] * len(timed)  # one batch: no per-solve latency
        gst = _gang_stats(g)
        g.close()
        note(f"{args.workload} precision {precision}: {K} timed solves in {elapsed:.2f} s (gang of {args.gang})")
        assert all(x.status in good for x in res), [x.status for x in res]
        stats = [s.stats() for s in timed]
        for s in timed:
            s.close()
        return dict(settings=settings, elapsed=elapsed, lat_ms=np.array(lat) * 1e3, res=res, prof={}, stats=stats,
                    solo_ms=solo_ms, t_setup=t_setup, gang=gst)
    if W:
        _, _, wres = run_concurrent(solvers[:W], S, tag)
        assert all(x.status in good for x in wres), [x.status for x in wres]
        for s in solvers[:W]:
            s.close()
    timed = solvers[W:]
    for s in timed:
        s.set_profile(kinds)
    elapsed, lat, res = run_concurrent(timed, S, tag)
    note(f"{args.workload} precision {precision}: {K} timed solves in {elapsed:.2f} s")
    assert all(x.status in good for x in res), [x.status for x in res]
    prof = {}
    for s in timed:
        for k, v in s.profile().items():
            acc = prof.setdefault(k, dict(ms=0.0, launches=0, flops=0.0, bytes=0.0))
            for key in acc:
                acc[key] += v[key]
    stats = [s.stats() for s in timed]
    for s in timed:
        s.close()
    lat_ms = np.array(lat) * 1e3
    return dict(settings=settings, elapsed=elapsed, lat_ms=lat_ms, res=res, prof=prof, stats=stats,
                solo_ms=solo_ms, t_setup=t_setup, gang=None)


def _e2e_run(args, mh, precision):
    """The reference's unit of work, solve(prob, settings) (examples/utils.jl:72-85 times exactly
    this): init (symbolic analysis, equilibration, BᵀB, device buffers and uploads, starting
    point) plus the IPM loop plus the result, from `--threads` host threads. The problems are
    constructed (IPMProblem) before the timed region, as the reference's prob is."""
    K, W, T, S = args.steps, args.warmup, max(1, args.threads), max(1, args.slots)
    ncu = mh.device_info()["cus"]
    grid = args.dag_grid if args.dag_grid > 0 else (max(16, ncu // S) if S > 1 and not args.gang else 0)
    settings = mh.IPMSettings(dag_grid=grid, precision=precision)
    kinds = ["factor", "update", "panel", "fsolve", "bsolve", "extend_add", "dag"]
    good = (mh.IPMStatus.OPTIMAL, mh.IPMStatus.NEAR_OPTIMAL) if args.workload == "exppow" else (mh.IPMStatus.OPTIMAL,)
    gang = [mh.IPMGang(capacity=args.gang, profile=True)] if args.gang else None
    probs = [hip_problem(mh, workload(args.workload, 1000 + k, args.n, args.m)) for k in range(W + K)]
    note(f"{args.workload} precision {precision}: {W + K} problems constructed")
    lat = [0.0] * (W + K)
    res = [None] * (W + K)
    stats = [None] * (W + K)
    prof = {}
    lock = threading.Lock()

    def one(i, profile):
        a = time.perf_counter()
        s = mh.init(probs[i], settings)
        if gang:
            gang[0].solve_in(s)
            r = s.result()
        else:
            if profile:
                s.set_profile(kinds)
            r = s.solve()
        lat[i] = time.perf_counter() - a
        res[i] = r
        stats[i] = s.stats()
        if args.workload != "qp":
            note(f"{args.workload} precision {precision}: solve {i} {r.status.name} in {lat[i]:.2f} s "
                 f"({r.niter} iterations)")
        if profile:
            with lock:
                for k, v in s.profile().items():
                    acc = prof.setdefault(k, dict(ms=0.0, launches=0, flops=0.0, bytes=0.0))
                    for key in acc:
                        acc[key] += v[key]
        s.close()

    def run(lo, hi, profile):
        nxt = [lo]

        def worker():
            while True:
                with lock:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= hi:
                    return
                one(i, profile)

        t0 = time.perf_counter()
        threads = [threading.Thread(target=worker) for _ in range(T)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        return time.perf_counter() - t0

    if W:
        run(0, W, False)
    if gang:  # the timed solves' statistics only
        gang[0].close()
        gang[0] = mh.IPMGang(capacity=args.gang, profile=True)
    elapsed = run(W, W + K, True)
    gst = None
    if gang:
        gst = _gang_stats(gang[0])
        gang[0].close()
    note(f"{args.workload} precision {precision}: {K} timed solve(prob, settings) in {elapsed:.2f} s "
         f"on {T} host threads")
    r = res[W:]
    assert all(x.status in good for x in r), [x.status for x in r]
    return dict(settings=settings, elapsed=elapsed, lat_ms=np.array(lat[W:]) * 1e3, res=r, prof=prof,
                stats=stats[W:], gang=gst)


def bench_resident(args, mh):
    K, W, S = args.steps, args.warmup, max(1, args.slots)
    if args.loop_only:
        # the IPM loop on instances set up before timing: for the large SDP workloads, whose host
        # init (the reference's dense zero-filled blocks between PSD vertices) takes seconds
        plain = _resident_run(args, mh, 0)
        mixed = _resident_run(args, mh, 1) if not args.no_mixed else None
        loop = None
    else:
        plain = _e2e_run(args, mh, 0)
        mixed = _e2e_run(args, mh, 1) if not args.no_mixed else None
        loop = _resident_run(args, mh, 0) if not args.no_loop else None
    settings, elapsed, lat_ms, res, prof = (plain[k] for k in ("settings", "elapsed", "lat_ms", "res", "prof"))
    st0 = plain["stats"][0]
    gst = plain["gang"]
    if gst is not None:
        # the gang's combined factorizations, timed with events on its stream
        fac = dict(ms=gst["dag_ms"], launches=gst["dag_launches"], flops=gst["dag_flops"], bytes=0.0)
        prof = dict(factor=fac, dag=fac)
        dom_name = "dag"
    else:
        fac = prof["factor"]
        dom_name = "dag" if prof["dag"]["ms"] >= prof["update"]["ms"] else "update"
    dom = prof[dom_name]
    achieved = dom["flops"] / (dom["ms"] * 1e-3) / 1e12 if dom["ms"] > 0 else 0.0
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
            "workload": WORKLOADS[args.workload].format(n=args.n, m=args.m, extra=args.maxcut_extra, psd=args.psd) +
                        (f"; {K} distinct seeded instances set up in HBM before timing (init excluded, "
                         f"setup_s_per_instance), each solved from a cold start to gap_tol=feas_tol=1e-8, "
                         f"{S} at a time" if args.loop_only else
                         f"; {K} distinct seeded instances, each a whole solve(prob, settings) (init + IPM "
                         f"loop + result) from a cold start to gap_tol=feas_tol=1e-8, on {args.threads} host "
                         f"threads"),
            "statuses": {mh.IPMStatus(int(x.status)).name: int(sum(int(y.status) == int(x.status) for y in res))
                         for x in res},
            "precision": "fp64 (plain); the mixed form is reported under mixed",
            "threads": args.threads,
            "dag_grid": settings.dag_grid,
            "p50_ms": pct(lat_ms, 50),
            "p99_ms": pct(lat_ms, 99),
            "init_host_ms_mean": float(np.mean([x["init_host_ms"] for x in plain["stats"]])),
            "init_device_ms_mean": float(np.mean([x["init_device_ms"] for x in plain["stats"]])),
            "init_device_parts_ms_mean": {k: float(np.mean([x["init_device_parts_ms"][k] for x in plain["stats"]]))
                                          for k in ("stream", "buffers", "schedule", "start")},
            "init_buffer_marks_ms_mean": {k: float(np.mean([x["build_marks_ms"][k] for x in plain["stats"]]))
                                          for k in ("fronts", "union", "cones", "allocs", "sparse")},
            "iterations_mean": float(np.mean([x.niter for x in res])),
            "gang_members": args.gang,
            "factor_tflops_aggregate": fac["flops"] / elapsed / 1e12,
            "factor_ms_per_call": fac["ms"] / max(1, fac["launches"]),
            "factor_exact_flops": st0["exact_flops"],
            "factor_front_flops": st0["front_flops"],
            "nfront": st0["nfront"],
            "kernel_ms": {k: v["ms"] for k, v in prof.items()},
            "kernel_launches": {k: v["launches"] for k, v in prof.items()},
            "gang": gst,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP64_MFMA,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_MFMA,
            "traffic": (TRAFFIC_PER_FRONT[(args.n, args.m, gst is not None)] *
                        (gst["factor_fronts"] / max(1.0, gst["factor_launches"]) if gst is not None else 1.0)
                        if dom_name == "dag" and args.workload == "qp" and (args.n, args.m, gst is not None)
                        in TRAFFIC_PER_FRONT else None),
            "kernel": (("k_dag_factor_multi (one persistent left-looking tile FP64 MFMA Cholesky launch over "
                        "the large fronts of every gang member)" if gst is not None else
                        "k_dag_factor (persistent left-looking tile FP64 MFMA Cholesky of the large fronts)")
                       if dom_name == "dag" else "k_update (FP64 MFMA trailing update of the frontal Cholesky)"),
            "avg_launch_us": dom["ms"] / max(1, dom["launches"]) * 1e3,
            "flops_per_launch": dom["flops"] / max(1, dom["launches"]),
        },
    }
    if args.loop_only:
        line["config"]["setup_s_per_instance"] = plain["t_setup"]
        line["config"]["solo_ms"] = plain["solo_ms"]
        line["config"]["slots"] = S
    if loop is not None:
        line["config"]["resident_loop"] = {
            "what": f"the IPM loop alone (solve! after init) on the same kind of instances, resident in HBM, "
                    + (f"through one gang of {args.gang} members" if args.gang else f"{S} at a time"),
            "solves_per_s": K / loop["elapsed"],
            "p50_ms": pct(loop["lat_ms"], 50),
            "p99_ms": pct(loop["lat_ms"], 99),
            "solo_ms": loop["solo_ms"],
            "gang": loop["gang"],
        }
    if mixed is not None:
        # the same seeded instances: per-result error of the mixed form against the FP64 results
        err = [max(abs(a.pobj - b.pobj) / max(1.0, abs(b.pobj)), abs(a.dobj - b.dobj) / max(1.0, abs(b.dobj)))
               for a, b in zip(mixed["res"], res)]
        perr = [float(np.linalg.norm(a.p - b.p) / max(1e-300, np.linalg.norm(b.p))) for a, b in zip(mixed["res"], res)]
        mp = mixed["prof"]
        if mixed["gang"] is not None:
            mg = mixed["gang"]
            mp = {"factor": dict(ms=mg["dag_ms"], launches=mg["dag_launches"], flops=mg["dag_flops"], bytes=0.0)}
        line["config"]["mixed"] = {
            "solves_per_s": K / mixed["elapsed"],
            "vs_plain": (K / mixed["elapsed"]) / (K / elapsed),
            "p50_ms": pct(mixed["lat_ms"], 50),
            "p99_ms": pct(mixed["lat_ms"], 99),
            "iterations_mean": float(np.mean([x.niter for x in mixed["res"]])),
            "fp32_factor_share": float(np.sum([s["factor_calls_fp32"] for s in mixed["stats"]]) /
                                       max(1, np.sum([s["factor_calls"] for s in mixed["stats"]]))),
            "factor_ms_per_call": mp["factor"]["ms"] / max(1, mp["factor"]["launches"]),
            "max_rel_obj_err_vs_fp64": float(np.max(err)),
            "max_rel_p_err_vs_fp64": float(np.max(perr)),
            "kernel_ms": {k: v["ms"] for k, v in mp.items()},
        }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.n, args.m, 1000 + W,
                                            float(np.mean([x.niter for x in res])))
    return line


def bench_serve(args, mh):
    """A Poisson stream of mixed requests through IPMServer."""
    from mumblebee_hip import generators as G
    rng = np.random.default_rng(7)
    K = args.steps
    reqs = []
    for k in range(K):
        reqs.append(hip_problem(mh, G.served_request(rng, k, args.serve_nmax)))
    gaps = rng.exponential(1.0 / args.rate, K)
    nslots = args.gang if args.gang else args.slots
    server = mh.IPMServer(nslots=nslots, gang=bool(args.gang))
    settings = mh.IPMSettings(precision=args.precision)
    t0 = mh.clock()
    arrivals, tickets = [], []
    for k in range(K):
        target = t0 + float(np.sum(gaps[:k + 1]))
        while mh.clock() < target:
            time.sleep(min(0.002, max(0.0, target - mh.clock())))
        arrivals.append(mh.clock())
        tickets.append(server.submit(reqs[k], settings))
    note(f"served stream: {K} requests submitted")
    out = [server.wait(t) for t in tickets]
    st = server.stats()
    server.close()
    ok = sum(r.status == mh.IPMStatus.OPTIMAL for r, _ in out)
    statuses = {}
    for r, _ in out:
        statuses[r.status.name] = statuses.get(r.status.name, 0) + 1
    fin = np.array([t["finished"] for _, t in out])
    lat_ms = (fin - np.array([t["submitted"] for _, t in out])) * 1e3
    elapsed = float(fin.max() - t0)
    return {
        "metric": METRIC, "value": K / elapsed, "unit": "solves/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": elapsed / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"served stream: {K} requests, 60% QP / 30% SOCP / 10% chordal max-cut SDP, n "
                               f"log-uniform in [50, {args.serve_nmax}], Poisson arrivals at {args.rate}/s, "
                               + (f"a gang of {nslots} slots per size class (lockstep, shared launches)"
                                  if args.gang else f"{nslots} slots, each on its own stream"),
                   "precision": "mixed" if args.precision else "fp64",
                   "optimal": ok, "statuses": statuses,
                   "not_optimal": [[k, r.status.name, int(r.niter)] for k, (r, _) in enumerate(out)
                                   if r.status != mh.IPMStatus.OPTIMAL],
                   "p50_ms": float(np.percentile(lat_ms, 50)),
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
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--slots", type=int, default=4,
                    help="the CUs of a large-front factorization are 1/slots of the chip; the resident-loop "
                         "measurement solves this many at a time")
    ap.add_argument("--threads", type=int, default=32, help="host threads running solve(prob, settings)")
    ap.add_argument("--gang", type=int, default=16,
                    help="members of the lockstep gang the solves run in (0: each solver on its own stream)")
    ap.add_argument("--no-loop", action="store_true", help="skip the resident-loop measurement")
    ap.add_argument("--loop-only", action="store_true",
                    help="time the IPM loop on instances set up before timing (init excluded and reported)")
    ap.add_argument("--dag-grid", type=int, default=0,
                    help="workgroups of each large-front factorization (0: the CUs shared among the slots)")
    ap.add_argument("--n", type=int, default=2000)
    ap.add_argument("--m", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="qp")
    ap.add_argument("--precision", type=int, default=0, choices=[0, 1], help="--serve: 0 FP64 factors, 1 mixed")
    ap.add_argument("--serve", action="store_true")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--serve-nmax", type=int, default=4000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB")
    ap.add_argument("--maxcut-extra", type=float, default=3e-5,
                    help="max-cut: probability of an extra edge per vertex pair")
    ap.add_argument("--psd", type=int, default=128, help="exp/pow: order of the PSD block")
    args = ap.parse_args()
    global ARGS
    ARGS = args
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    # HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default); slots beyond that
    # would share a queue and run their kernels one after another, so give each slot its own
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(min(32, max(4, args.slots, 0 if args.gang else args.threads))))
    import mumblebee_hip as mh
    line = bench_serve(args, mh) if args.serve else bench_resident(args, mh)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
