"""Benchmark of the MI355X AVBD step: body-steps/s at 10 iterations, batched vs plain.

Workload (default): 4,096 independent worlds in one batch, body counts log-uniform in
[8, 2048] (about 1.55 M bodies), each a seeded random box pile (half extents U[0.25, 1] m,
log-normal density, random orientation, dropped from a jittered grid onto a ground box),
mu 0.5, dt 1/60, 10 iterations, FP32.  One step = one captured-graph step of every world.

Printed (one JSON line): `value` = batched body-steps/s (movable bodies x steps / seconds,
inputs resident in HBM); p50/p99 step latency; the plain single-world path on the same
worlds (each world stepped through avbd_world_step, one after another); the roofline of the
dominant kernel; and the CPU baseline (the oracle, a serial restatement of the reference's
CPU step, timed on this host on a bounded sample of the same worlds).

`--gpus N` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
pkg = importlib.import_module("godot-avbd_amd")

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes (compulsory FP32 SoA traffic, each field touched once per pass):
#   per step      424 N_b + 256 N_m + 124 N_c
#   per iteration 100 N_b + 320 N_m + 144 N_c + 408 N_j  (primal: 100 N_b + 160 N_m + 60 N_c,
#                                                         dual:   160 N_m + 84 N_c + 408 N_j)
def step_bytes(nb, nm, nc, nj, iters):
    return 424 * nb + 256 * nm + 124 * nc + iters * (100 * nb + 320 * nm + 144 * nc + 408 * nj)


def primal_bytes(nb_overflow, nm, nc):
    return 100 * nb_overflow + 160 * nm + 60 * nc


def dual_bytes(nm, nc, nj):
    return 160 * nm + 84 * nc + 408 * nj


def rocprof_window(rows, steps):
    """Rows (dicts in dispatch order) of the last `steps` batch steps of a bench run.  Batch
    launches are told apart from the plain single-world ones that follow by their grid: the
    batch's narrow phase and rebase have the largest grids of those kernels in the run.  The
    window runs from the dispatch after the batch rebase that precedes the first of those
    steps' narrow phases to the rebase that ends the last one."""
    grid = [int(r["Grid_Size"]) if "Grid_Size" in r else
            int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"]) for r in rows]
    name = [r["Kernel_Name"] for r in rows]
    gn = max(g for g, n in zip(grid, name) if n.startswith("k_manifold_init"))
    gr = max(g for g, n in zip(grid, name) if n.startswith("k_rebase"))
    narrow = [i for i, n in enumerate(name) if n.startswith("k_manifold_init") and grid[i] == gn]
    rebase = [i for i, n in enumerate(name) if n.startswith("k_rebase") and grid[i] == gr]
    if len(narrow) < steps:
        raise SystemExit(f"only {len(narrow)} batch steps in the trace")
    first, last = narrow[-steps], narrow[-1]
    before = [i for i in rebase if i < first]
    after = [i for i in rebase if i > last]
    start = before[-1] + 1 if before else 0
    end = after[0] if after else len(rows) - 1
    return rows[start:end + 1]


def plain_window(rows, steps):
    """Rows of the last `steps` steps of the last world stepped in the run (a plain world run
    after the batch, e.g. C3): from the dispatch after the rebase that ends the step before them
    to the last rebase."""
    rebase = [i for i, r in enumerate(rows) if r["Kernel_Name"].startswith("k_rebase")]
    if len(rebase) < steps + 1:
        raise SystemExit(f"only {len(rebase)} steps in the trace")
    return rows[rebase[-steps - 1] + 1:rebase[-1] + 1]


def summarize_rocprof(path, steps, plain=False):
    """Markdown summary of a rocprofv3 CSV (kernel trace or counter collection) over the last
    `steps` batch steps (or, with `plain`, the last steps of the run's last plain world), per
    kernel: launches per step, average duration, time per step, and the average of each counter
    per launch."""
    import csv
    import collections
    with open(path) as f:
        rows = list(csv.DictReader(f))
    for r in rows:   # "avbd_dev::k_dual<true>(avbd_dev::State, float)" -> "k_dual<true>"
        r["Kernel_Name"] = r["Kernel_Name"].split("(")[0].replace("avbd_dev::", "")
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))   # the CSVs are not written in dispatch order
    counters = "Counter_Name" in rows[0]
    if counters:   # one row per (dispatch, counter)
        disp = collections.OrderedDict()
        for r in rows:
            d = disp.setdefault(r["Dispatch_Id"], dict(r, counters={}))
            d["counters"][r["Counter_Name"]] = float(r["Counter_Value"])
        rows = list(disp.values())
    win = plain_window(rows, steps) if plain else rocprof_window(rows, steps)
    agg = collections.OrderedDict()
    for r in win:
        a = agg.setdefault(r["Kernel_Name"], {"n": 0, "ns": 0.0, "c": collections.Counter()})
        a["n"] += 1
        a["ns"] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        for k, v in r.get("counters", {}).items():
            a["c"][k] += v
    total = sum(a["ns"] for a in agg.values())
    names = sorted(agg, key=lambda k: -agg[k]["ns"])
    cnames = sorted({c for a in agg.values() for c in a["c"]})
    out = [f"window: last {steps} {'plain-world' if plain else 'batch'} steps, {len(win)} dispatches, "
           f"{total / steps / 1e6:.3f} ms of kernel time per step", "",
           "| kernel | launches/step | avg us | ms/step | share |" + "".join(f" {c}/launch |" for c in cnames),
           "|---|---|---|---|---|" + "---|" * len(cnames)]
    for k in names:
        a = agg[k]
        out.append(f"| {k} | {a['n'] / steps:g} | {a['ns'] / a['n'] / 1e3:.1f} | {a['ns'] / steps / 1e6:.3f} | "
                   f"{100 * a['ns'] / total:.1f}% |" + "".join(f" {a['c'][c] / a['n']:.1f} |" for c in cnames))
    print("\n".join(out))


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def sample_worlds(sizes, k, seed=0):
    """k world ids stratified over the size distribution (sorted by size)."""
    order = np.argsort(sizes, kind="stable")
    idx = np.linspace(0, len(order) - 1, k).round().astype(int)
    return sorted(set(int(order[i]) for i in idx))


def bench_chains(args, log):
    """C3: one plain world of 4096 chains x 32 unit boxes (ball joints, anchors, ground),
    beta_lin 1e5, k cap 1e9, 10 iterations: colour Gauss-Seidel for the jointed bodies."""
    params = pkg.default_params(iterations=args.iterations)
    params.beta_lin = 1.0e5
    params.penalty_max = 1.0e9
    t0 = time.time()
    bodies, joints = pkg.scenes.chains(n_chains=4096, links=32)
    w = pkg.World(params, len(bodies["mass"]), len(joints["body_b"]))
    w.upload(bodies, joints)
    if args.no_graph:
        w.use_graph(False)
    for _ in range(args.c3_warmup):
        w.step()
    tuning = w.autotune(args.tune_steps) if args.tune_steps > 0 else None
    a = time.perf_counter()
    for _ in range(args.c3_steps):
        w.step()
    dt = (time.perf_counter() - a) / args.c3_steps
    movable = int((bodies["mass"] > 0).sum())
    st = w.stats()
    lists = w.list_stats()
    # roofline of the colour sweep (the dominant kernel here): per iteration it reads each
    # movable body once and each manifold, contact and joint as the primal does
    # (100 N_b + 160 N_m + 60 N_c + 256 N_j); a launch is one of the 8 colours
    prof = w.profile(args.profile_steps)
    cp_ms = prof["avg_ms"]["colour_pass"]
    nj = len(joints["body_b"])
    sweep = 100 * movable + 160 * st["manifolds"] + 60 * st["contacts"] + 256 * nj
    per_launch = sweep / 8
    achieved = per_launch / (cp_ms / 1e3) / 1e9 if cp_ms > 0 else None
    w.close()
    log(f"C3 chains: {dt * 1e3:.3f} ms/step, {movable / dt / 1e6:.2f} M body-steps/s ({time.time() - t0:.1f}s)")
    return {"workload": "C3 plain world: 4096 chains x 32 unit boxes, ball joints, static anchors, ground; "
                        "beta_lin 1e5, k cap 1e9", "bodies": len(bodies["mass"]), "movable_bodies": movable,
            "joints": len(joints["body_b"]), "manifolds": st["manifolds"], "contacts": st["contacts"],
            "warmup": args.c3_warmup, "steps": args.c3_steps, "ms_per_step": dt * 1e3, "tuning": tuning,
            "broadphase_lists": lists,
            "roofline": {"bound": "hbm", "kernel": "colour_pass", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
                         "traffic": None, "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": cp_ms,
                         "launches_per_step": prof["launches"]["colour_pass"] / max(prof["steps"], 1)},
            "kernel_ms": prof["avg_ms"], "kernel_launches_per_step":
                {k: v / max(prof["steps"], 1) for k, v in prof["launches"].items()},
            "value": movable / dt, "unit": "body-steps/s"}


def bench_plain_configs(args, log):
    """C1 (the reference's CPU-runnable case: 406-cube pyramid, 300 steps) and C2 (one world of
    8192 boxes, seeded pile) on the plain single-world path, each beside the oracle on this
    host's CPU; C1 also gives the long-run statistics of the GPU and the oracle (penetration
    depth, the last dual's residuals, energy drift), which must agree within 1e-6 relative."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    params = pkg.default_params(iterations=args.iterations)
    out = {}
    # ---- C1: 300 steps, timed whole (the reference's bench shape), statistics at the end
    b, j = pkg.scenes.pyramid()
    w = pkg.World(params, len(b["mass"]), 0)
    w.upload(b, j)
    w.record_residuals(True)
    w.step()                      # one-time setup (graph capture)
    w2 = pkg.World(params, len(b["mass"]), 0)
    w2.upload(b, j)
    w2.record_residuals(True)
    w2.step()
    a = time.perf_counter()
    w2.step(299)
    gpu_s = time.perf_counter() - a
    o = oracle.OracleWorld(params)
    o.upload(b, j)
    a = time.perf_counter()
    o.step(300)
    cpu_s = time.perf_counter() - a
    w.step(299)
    movable = int((b["mass"] > 0).sum())
    gst = pkg.scenes.world_statistics({**b, **w.download()}, w.manifolds(), w.residuals(), b)
    ost = pkg.scenes.world_statistics({**b, **o.download()}, o.manifolds(), o.residuals(), b)
    agree = all(abs(gst[k] - v) <= 1e-6 * max(1.0, abs(v)) for k, v in ost.items() if isinstance(v, float))
    same = pkg.state_digest(w.download()) == o.digest()
    out["c1"] = {"workload": "C1 plain world: pyramid of 406 unit cubes (28 rows) on a static ground box, "
                             "dt 1/60, 300 steps", "movable_bodies": movable, "steps": 299,
                 "ms_per_step": gpu_s * 1e3 / 299, "value": movable * 299 / gpu_s, "unit": "body-steps/s",
                 "cpu_oracle": {"value": movable * 300 / cpu_s, "unit": "body-steps/s", "cores": 1, "kind": "port",
                                "ms_per_step": cpu_s * 1e3 / 300},
                 "statistics_gpu": gst, "statistics_oracle": ost, "statistics_agree_1e-6": agree,
                 "state_bit_identical": same}
    w.close()
    w2.close()
    log(f"C1 pyramid: {gpu_s * 1e3 / 299:.3f} ms/step on the GPU, {cpu_s * 1e3 / 300:.3f} ms/step oracle; "
        f"statistics agree {agree}, bit-identical {same}")
    # ---- C2: 8192 boxes, 60 settle steps then 60 timed
    b, j = pkg.scenes.pile(8192, seed=2)
    w = pkg.World(params, len(b["mass"]), 0)
    w.upload(b, j)
    for _ in range(60):
        w.step()
    a = time.perf_counter()
    w.step(60)
    gpu_s = time.perf_counter() - a
    movable = int((b["mass"] > 0).sum())
    st, m = {**b, **w.download()}, w.manifolds()
    o = oracle.OracleWorld(params)
    o.upload(st, j)
    o.set_manifolds(m)
    a = time.perf_counter()
    n_cpu = 0
    while time.perf_counter() - a < 5.0 or n_cpu < 1:
        o.step()
        n_cpu += 1
    cpu_s = time.perf_counter() - a
    out["c2"] = {"workload": "C2 plain world: 8192 boxes, half extents U[0.25, 1] m, log-normal density, random "
                             "orientations, dropped on a ground box; mu 0.5", "movable_bodies": movable,
                 "settle_steps": 60, "steps": 60, "ms_per_step": gpu_s * 1e3 / 60, "value": movable * 60 / gpu_s,
                 "unit": "body-steps/s", "manifolds": w.stats()["manifolds"],
                 "cpu_oracle": {"value": movable * n_cpu / cpu_s, "unit": "body-steps/s", "cores": 1, "kind": "port",
                                "steps": n_cpu}}
    w.close()
    log(f"C2 pile: {out['c2']['ms_per_step']:.3f} ms/step on the GPU, oracle {cpu_s * 1e3 / n_cpu:.1f} ms/step")
    return out


def bench_serving(args, log):
    """C5: Poisson arrivals of world requests (bodies log-uniform in [16, 16384], horizons
    U[1, 600] steps, 70 % piles / 30 % jointed chains), admitted by HBM budget into one
    device-resident batch, continuous batching; body-steps/s and p50/p99 step latency per
    offered load."""
    from importlib import import_module
    serve = import_module("godot-avbd_amd.serve")
    params = pkg.default_params(iterations=args.iterations)
    rng = np.random.default_rng(5)
    chain_cache = {}

    def make(rid):
        n = int(round(math.exp(rng.uniform(math.log(16), math.log(16384)))))
        horizon = int(rng.integers(1, 601))
        if rng.uniform() < 0.7:
            b, j = pkg.scenes.pile(n, seed=50000 + rid, pitch=4.0, ground=max(40.0, 8.0 * math.sqrt(n)))
        else:
            nc = max(1, int(round(n / 33.0)))
            nc = max(1, int(round(2 ** (round(2 * math.log2(nc)) / 2))))   # a few cached sizes
            if nc not in chain_cache:
                chain_cache[nc] = pkg.scenes.chains(n_chains=nc, links=32, seed=51 + nc)
            b, j = chain_cache[nc]
        return serve.Request(rid, b, j, horizon, 0.0)

    out = []
    rid = 0
    for load in args.serve_loads:
        srv = serve.Server(params, args.serve_budget_gb * 1e9, max_worlds=args.serve_max_worlds,
                           joints_per_body=0.5)
        # start at the steady-state population of this load (mean horizon 300 ticks)
        for _ in range(int(load * 300)):
            r = make(rid); rid += 1
            r.horizon = int(rng.integers(1, r.horizon + 1))
            srv.submit(r)
        for _ in range(args.serve_warmup):
            srv.tick()
        srv.latencies.clear(); srv.admit_waits.clear(); srv.body_steps = 0
        resident = []
        a = time.perf_counter()
        for _ in range(args.serve_ticks):
            for _ in range(rng.poisson(load)):
                r = make(rid); rid += 1
                srv.submit(r)
            srv.tick()
            resident.append(len(srv.resident))
        el = time.perf_counter() - a
        lat = np.asarray(srv.latencies) * 1e3
        waits = np.asarray(srv.admit_waits) * 1e3 if srv.admit_waits else np.zeros(1)
        rec = {"offered_load_per_tick": load, "ticks": args.serve_ticks, "seconds": el,
               "value": srv.body_steps / el, "unit": "body-steps/s",
               "resident_worlds_mean": float(np.mean(resident)), "waiting_at_end": len(srv.waiting),
               "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99)),
               "p50_admission_wait_ms": float(np.percentile(waits, 50)),
               "p99_admission_wait_ms": float(np.percentile(waits, 99)),
               "hbm_budget_bytes": int(args.serve_budget_gb * 1e9), "device_bytes": srv.device_bytes,
               "body_capacity": srv.max_bodies}
        log(f"serving load {load}/tick: {rec['value'] / 1e6:.1f} M body-steps/s, p50 {rec['p50_step_ms']:.2f} ms, "
            f"p99 {rec['p99_step_ms']:.2f} ms, {rec['resident_worlds_mean']:.0f} resident worlds")
        srv.close()
        out.append(rec)
    return {"workload": "C5 serving: Poisson arrivals, bodies log-uniform [16, 16384], horizons U[1, 600], "
                        "70% piles / 30% chains, admission by HBM budget, continuous batching",
            "loads": out}


def bench_queries(args, log):
    """Ray casts per second on one world of `query_bodies` mixed boxes, spheres and cylinders
    after 120 steps: `query_rays` segments per call (a third straight down onto the pile), timed
    around avbd_world_raycast (host arrays in and out) and, with HIP events, around the kernel
    alone.  Each ray walks the step's LBVH refit to the bodies' current bounds and tests the exact
    shape of each body whose bounds its segment meets."""
    import ctypes as C
    abi = pkg.abi
    b, j = pkg.scenes.round_pile(args.query_bodies, seed=31, pitch=2.4)
    n = len(b["mass"])
    w = pkg.World(pkg.default_params(iterations=args.iterations), n, 0)
    w.upload(b, j)
    w.step(120)
    rng = np.random.default_rng(8)
    r = args.query_rays
    pos = w.download()["position"]
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    frm = rng.uniform(lo, hi, (r, 3)).astype(np.float32)
    frm[:, 2] = hi[2] + 2.0
    to = (frm + rng.normal(0, 6, (r, 3))).astype(np.float32)
    to[: r // 3, 2] = lo[2] - 2.0
    for _ in range(2):
        out = w.raycast(frm, to)
    times = []
    for _ in range(5):
        t = time.perf_counter()
        out = w.raycast(frm, to)
        times.append(time.perf_counter() - t)
    host_s = float(np.median(times))
    hits = int((out["body"] >= 0).sum())
    # kernel alone, HIP events on the library's stream through the profiling entry point
    prof = abi.QueryProfile()
    abi.check(w.lib.avbd_world_profile_raycast(w.h, r, frm.ctypes.data_as(abi.f32p), to.ctypes.data_as(abi.f32p), 5,
                                              C.byref(prof)))
    kern_s = prof.kernel_ms / 1e3
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle's serial restatement of the same query (the reference's server scans the
        # body list per query the same way), on a bounded sample of the rays, one core
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        o = orc.OracleWorld(pkg.default_params(iterations=args.iterations))
        st = w.download()
        o.upload({**b, **{k: st[k] for k in ("position", "rotation", "velocity", "angular_velocity",
                                                 "prev_velocity")}}, j)
        k = 4096
        t = time.perf_counter()
        ref = o.raycast(frm[:k], to[:k])
        cpu_s = time.perf_counter() - t
        assert np.array_equal(ref["body"], out["body"][:k])
        cpu = {"value": k / cpu_s, "unit": "rays/s", "cores": 1, "kind": "port",
               "sample": f"the first {k} rays of the same call, same world state"}
    line = {"metric": "rays/sec (segment casts: the refit LBVH, then each candidate's exact shape)",
            "value": r / kern_s, "unit": "rays/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
            "data": "synthetic", "config": {"workload": f"round_pile({args.query_bodies}) settled 120 steps",
                                            "bodies": n, "rays": r, "hits": hits},
            "kernel_ms": prof.kernel_ms, "bounds_ms": prof.bounds_ms, "host_call_ms": host_s * 1e3,
            "rays_per_s_host": r / host_s, "cpu_baseline": cpu}
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps after the settle phase and the autotuner")
    ap.add_argument("--settle", type=int, default=300,
                    help="steps that let the piles come to rest before anything is tuned or timed")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--min-bodies", type=int, default=8)
    ap.add_argument("--max-bodies", type=int, default=2048)
    ap.add_argument("--iterations", type=int, default=10)
    ap.add_argument("--plain-sample", type=int, default=48)
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-sample", type=int, default=160,
                    help="worlds (stratified by size) stepped by the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the CPU baseline (the host share of one GPU)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels one by one, batch and plain worlds (for tracing tools)")
    ap.add_argument("--tune-steps", type=int, default=2, help="autotuner steps per trial (0: keep defaults)")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3 chains world")
    ap.add_argument("--no-plain-configs", action="store_true",
                    help="skip C1 (406-cube pyramid with long-run statistics) and C2 (8192-box world)")
    ap.add_argument("--c3-warmup", type=int, default=240)
    ap.add_argument("--c3-steps", type=int, default=30)
    ap.add_argument("--no-serve", action="store_true", help="skip the C5 serving loop")
    ap.add_argument("--serve-loads", type=float, nargs="+", default=[0.5, 1.0, 2.0])
    ap.add_argument("--serve-ticks", type=int, default=120)
    ap.add_argument("--serve-warmup", type=int, default=20)
    ap.add_argument("--serve-budget-gb", type=float, default=32.0)
    ap.add_argument("--serve-max-worlds", type=int, default=2048)
    ap.add_argument("--summarize-rocprof", metavar="CSV",
                    help="print a per-kernel summary of a rocprofv3 CSV of a bench run and exit")
    ap.add_argument("--summary-steps", type=int, default=5)
    ap.add_argument("--summary-plain", action="store_true",
                    help="summarize the last plain world's steps (C3 runs last) instead of the batch's")
    ap.add_argument("--queries", action="store_true",
                    help="measure batched ray casts (the server's intersect_ray) on a settled mixed-shape pile and exit")
    ap.add_argument("--query-bodies", type=int, default=8192)
    ap.add_argument("--query-rays", type=int, default=262144)
    args = ap.parse_args()
    if args.summarize_rocprof:
        summarize_rocprof(args.summarize_rocprof, args.summary_steps, args.summary_plain)
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(worlds would shard across GPUs with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.queries:
        bench_queries(args, log)
        return

    params = pkg.default_params(iterations=args.iterations)
    t0 = time.time()
    worlds = pkg.scenes.batch_piles(args.worlds, args.min_bodies, args.max_bodies)
    sizes = np.array([len(b["mass"]) for b, _ in worlds])
    total_bodies = int(sizes.sum())
    log(f"built {len(worlds)} worlds, {total_bodies} bodies in {time.time() - t0:.1f}s")

    batch = pkg.Batch(params, len(worlds), total_bodies, 0, 0)
    ids = [batch.add(b, j) for b, j in worlds]
    if args.no_graph:
        batch.use_graph(False)
    log(f"uploaded in {time.time() - t0:.1f}s")

    for s in range(args.settle):
        batch.step()
        if (s + 1) % 30 == 0:
            log(f"settle step {s + 1}/{args.settle}")
    tuning = batch.autotune(args.tune_steps) if args.tune_steps > 0 else None
    if tuning:
        log(f"autotuned launch shapes ({8 * args.tune_steps} trial steps): {tuning}")
    for _ in range(args.warmup):
        batch.step()
    tot = batch.totals()
    movable = tot["movable_bodies"]

    # ---- timed region: K captured-graph steps of the whole batch
    lat = []
    pkg.abi.load_library()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        batch.step()
        lat.append((time.perf_counter() - a) * 1e3)
    elapsed = time.perf_counter() - t_start
    tot = batch.totals()
    lists = batch.list_stats()
    value = movable * args.steps / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    log(f"batched: {ms_per_step:.3f} ms/step, {value / 1e6:.1f} M body-steps/s, totals {tot}, lists {lists}")

    # ---- per-kernel durations (HIP events on the batch's stream) for the roofline
    prof = batch.profile(args.profile_steps)
    tot_p = batch.totals()
    iters = args.iterations
    dom = max(("primal", "dual", "narrow", "brake_levels", "traverse"),
              key=lambda k: prof["avg_ms"][k] * prof["launches"][k])
    nb, nm, nc, nj = tot_p["movable_bodies"], tot_p["manifolds"], tot_p["contacts"], tot_p["joints"]
    kernel_bytes = {"primal": primal_bytes(nb, nm, nc), "dual": dual_bytes(nm, nc, nj)}
    roof_kernel = dom if dom in kernel_bytes else max(kernel_bytes, key=lambda k: prof["avg_ms"][k])
    kb = kernel_bytes[roof_kernel]
    kms = prof["avg_ms"][roof_kernel]
    achieved = kb / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{roof_kernel}_batch.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    sb = step_bytes(nb, nm, nc, nj, iters)
    log(f"profile: {json.dumps(prof)}")

    # ---- plain single-world path on the same worlds (stratified sample, one after another)
    pick = sample_worlds(sizes, min(args.plain_sample, len(worlds)))
    plain_t, plain_n = [], []
    snaps = {}
    for w in pick:
        b, j = worlds[w]
        st = batch.download(ids[w])
        m = batch.manifolds(ids[w])
        start = {**b, **st}
        snaps[w] = (start, m)
        pw = pkg.World(params, len(b["mass"]), 0)
        pw.upload(start, j)
        pw.set_manifolds(m)
        if args.no_graph:
            pw.use_graph(False)
        pw.step()                   # first call pays one-time setup
        a = time.perf_counter()
        pw.step(args.plain_steps)
        plain_t.append((time.perf_counter() - a) / args.plain_steps)
        plain_n.append(int((b["mass"] > 0).sum()))
        pw.close()
    # per-world time model t(n) = c0 + c1 n, summed over every world of the batch
    A = np.stack([np.ones(len(plain_n)), np.asarray(plain_n, float)], axis=1)
    coef, *_ = np.linalg.lstsq(A, np.asarray(plain_t), rcond=None)
    movable_w = np.array([int((b["mass"] > 0).sum()) for b, _ in worlds])
    plain_step_s = float(np.sum(np.maximum(coef[0] + coef[1] * movable_w, min(plain_t))))
    plain_value = float(movable_w.sum()) / plain_step_s
    log(f"plain: sample {len(pick)} worlds, model {coef[0] * 1e3:.3f} ms + {coef[1] * 1e6:.3f} us/body; "
        f"{plain_value / 1e6:.3f} M body-steps/s")

    # ---- CPU baseline: the oracle (serial restatement of the reference step) on this host,
    # one core, then a larger stratified sample of worlds partitioned over the threads by
    # body count (longest first onto the least-loaded thread), each thread stepping its worlds
    # round-robin for the same wall time (ctypes releases the GIL inside the oracle's step)
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        from concurrent.futures import ThreadPoolExecutor
        cpu_pick = sample_worlds(sizes, min(args.cpu_sample, len(worlds)))
        order = sorted(cpu_pick, key=lambda w: -sizes[w])
        movable_of = {w: int((worlds[w][0]["mass"] > 0).sum()) for w in order}
        oworlds = {}
        for w in order:
            b, j = worlds[w]
            if w in snaps:
                start, m = snaps[w]
            else:
                start, m = {**b, **batch.download(ids[w])}, batch.manifolds(ids[w])
            o = oracle.OracleWorld(params)
            o.upload(start, j)
            o.set_manifolds(m)
            oworlds[w] = o

        def run(ws, seconds):
            done = 0
            end_t = time.perf_counter() + seconds
            while time.perf_counter() < end_t:
                for w in ws:
                    oworlds[w].step()
                    done += movable_of[w]
            return done

        one_s = args.cpu_seconds / 3.0
        a = time.perf_counter()
        one_done = run(order, one_s)
        one_value = one_done / (time.perf_counter() - a)
        try:
            affinity = len(os.sched_getaffinity(0))
        except AttributeError:
            affinity = os.cpu_count() or 1
        threads = max(1, min(affinity, args.cpu_threads, len(order)))
        parts, load = [[] for _ in range(threads)], [0] * threads
        for w in order:
            t = load.index(min(load))
            parts[t].append(w)
            load[t] += movable_of[w]
        a = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            multi_done = sum(ex.map(lambda ws: run(ws, args.cpu_seconds - one_s), parts))
        multi_value = multi_done / (time.perf_counter() - a)
        cpu_model = ""
        try:
            with open("/proc/cpuinfo") as f:
                cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": multi_value, "unit": "body-steps/s", "cores": threads, "kind": "port",
               "single_core_value": one_value, "host_cpus": os.cpu_count(), "host_cpus_usable": affinity,
               "cpu_model": cpu_model,
               "sample": f"{len(order)} worlds stratified by size (states copied from the GPU after the settle "
                         f"phase, {sum(movable_of.values())} movable bodies), stepped by the oracle (serial "
                         f"restatement of the reference step, {iters} iterations): {threads} threads (the host "
                         f"share of one GPU; the host has {os.cpu_count()} CPUs) with the worlds partitioned by "
                         f"body count for {args.cpu_seconds - one_s:.0f}s ({multi_done} body-steps), and 1 thread "
                         f"for {one_s:.0f}s ({one_done} body-steps)"}
        log(f"cpu baseline: {multi_value / 1e3:.1f} k body-steps/s on {threads} threads, "
            f"{one_value / 1e3:.1f} k on 1 ({cpu_model}, {os.cpu_count()} CPUs)")

    plain_cfg = None if args.no_plain_configs else bench_plain_configs(args, log)
    c3 = None if args.no_c3 else bench_chains(args, log)
    serving = None if args.no_serve else bench_serving(args, log)

    line = {
        "metric": "body-steps/sec @10 iters (batched vs plain, same GPU); p50/p99 step latency",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"C4 batched: {len(worlds)} independent worlds, bodies log-uniform "
                        f"[{args.min_bodies}, {args.max_bodies}], seeded random box piles, one captured graph per step",
            "worlds": len(worlds), "bodies": total_bodies, "movable_bodies": movable,
            "manifolds": tot["manifolds"], "contacts": tot["contacts"], "iterations": iters,
            "dt": "1/60", "parallelism": "single GPU", "settle_steps": args.settle,
        },
        "p50_step_ms": percentile(lat, 50),
        "p99_step_ms": percentile(lat, 99),
        "plain": {"value": plain_value, "unit": "body-steps/s", "sample_worlds": len(pick),
                  "model_ms": [coef[0] * 1e3, coef[1] * 1e3]},
        "batched_over_plain": value / plain_value,
        "roofline": {"bound": "hbm", "kernel": roof_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": kb, "avg_launch_ms": kms},
        "step_roofline": {"algorithmic_bytes_per_step": sb,
                          "achieved_GBs": sb / (ms_per_step * 1e-3) / 1e9,
                          "frac": sb / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "kernel_ms": prof["avg_ms"],
        "kernel_launches_per_step": {k: v / max(1, prof["steps"]) for k, v in prof["launches"].items()},
        "cpu_baseline": cpu,
        "tuning": tuning,
        "broadphase_lists": lists,
        "c1": None if plain_cfg is None else plain_cfg["c1"],
        "c2": None if plain_cfg is None else plain_cfg["c2"],
        "c3": c3,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
