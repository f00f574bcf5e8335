#!/usr/bin/env python3
"""Throughput bench: matched (query, molecule) pairs/s of SIGMo's hot path on one MI355X.

Workload (default): 100,000 synthetic molecules (seed 1, heavy-atom count
clamp(round(Normal(24,8)),6,64)) x 100 query graphs (k ~ U[4,20], 20% with one perturbed
label), plain per-phase form: signatures, filter, 6 x (signature refine + candidate refine),
GMCR, find-first join.  One step = one full pass of that pipeline on device-resident graphs
(upload and buffer allocation happen before the timed region, as in the reference driver).

Prints ONE JSON line with the throughput, a roofline entry for the dominant HBM-bound
kernel (algorithmic bytes / measured HIP-event time) and a CPU baseline: the CPU oracle
(a restatement of the reference kernels, OpenMP over all host cores) on a bounded molecule
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "matched (query,molecule) pairs/sec on 1 MI355X; serving p50/p99 latency"
HBM_KERNELS = ("refine", "filter", "gmcr", "sig_refine", "sig_generate", "pair_filter", "db_prepare")


# name: (seed, molecules, size law, queries, kmin, kmax, perturbed fraction, ring fraction)
CONFIGS = {
    1: (1, 100_000, "normal", 100, 4, 20, 0.2, -1.0),
    2: (2, 1_000_000, "lognormal", 1000, 4, 32, 0.0, 0.3),
    3: (3, 20_000_000, "lognormal", 2000, 4, 32, 0.0, 0.3),  # streamed in chunks
    4: (4, 10_000_000, "lognormal", 0, 4, 32, 0.0, 0.3),  # serving: queries come with the requests
}


def workload(args):
    from sigmo_amd import capi

    seed, _, law, _, kmin, kmax, perturb, ring = CONFIGS[args.config]
    dist = capi.SIZE_NORMAL if law == "normal" else capi.SIZE_LOGNORMAL
    qh = capi.HostGraphs.queries(seed, args.queries, kmin, kmax, dist, perturb_frac=perturb, ring_frac=ring)
    dh = capi.HostGraphs.molecules(seed, args.molecules, dist)
    return qh, dh


def describe(args) -> str:
    seed, _, law, _, kmin, kmax, perturb, ring = CONFIGS[args.config]
    sizes = "Normal(24,8) sizes" if law == "normal" else "LogNormal(ln 24,0.45) sizes"
    extra = f" {int(perturb * 100)}% perturbed" if perturb > 0 else ""
    extra += f" {int(ring * 100)}% with a ring" if ring >= 0 else ""
    mode = "find-first" if not args.find_all else "find-all"
    form = "plain per-phase form" if args.form == "plain" else "batched pair form"
    stream = ""
    if args.config == 3:
        stream = (f", streamed from host memory in chunks of {args.chunk_molecules} molecules"
                  f"{' (next chunk loaded on a side stream)' if not args.no_overlap else ''}")
    return (f"config {args.config} ({form}): {args.molecules} molecules seed {seed} {sizes}, "
            f"{args.queries} queries k~U[{kmin},{kmax}]{extra}, R={args.refinement_steps}, {mode}{stream}")


CPU_SAMPLE_PAIRS = 10_000_000  # ~3 s of oracle work on 16 cores at config 1 rates; whole workload for config 1


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(qh, dh, args) -> dict:
    """Oracle (CPU restatement of the reference kernels) on the first molecules of the workload."""
    from oracle import oracle as O

    threads = int(os.environ.get("SIGMO_CPU_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    count = args.cpu_molecules or max(1, CPU_SAMPLE_PAIRS // max(1, qh.num_graphs))
    sample = dh.slice(0, min(count, dh.num_graphs))
    qa, sa = qh.arrays(), sample.arrays()
    t0 = time.perf_counter()
    r = O.pipeline(qa, sa, args.refinement_steps, modes=[not args.find_all])
    dt = time.perf_counter() - t0
    return {
        "value": (r["total_first"] if not args.find_all else int((r["counts_all"] > 0).sum())) / dt,
        "unit": "pairs/s",
        "cores": O.threads(),
        "kind": "port",
        "sample": f"first {sample.num_graphs} of the {dh.num_graphs} molecules x all {qh.num_graphs} queries, "
                  f"R={args.refinement_steps}, {'find-first' if not args.find_all else 'find-all'}; {dt:.2f} s on "
                  f"{O.threads()} OpenMP threads ({_cpu_model()}, {os.cpu_count()} logical CPUs visible)",
    }


def _pct(x, q):
    import numpy as np

    return round(float(np.percentile(x, q)), 3) if len(x) else None


def _serve_online(capi, workers, qh, rgo, arr, **options) -> dict:
    """The same request stream through the online server: a client submits each request at its
    arrival time (sigmo_server_submit), then polls every result; latency is completion minus
    submission, measured by the server."""
    import numpy as np

    n = len(arr)
    parts = [qh.slice(int(rgo[r]), int(rgo[r + 1] - rgo[r])) for r in range(n)]  # prepared before the clock starts
    srv = capi.Server(workers, **options)
    t0 = time.perf_counter()
    ids = []
    for r in range(n):
        wait = arr[r] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ids.append(srv.submit(parts[r]))
    res = [srv.poll(i, -1.0) for i in ids]
    summary = srv.close()
    out = {k: np.array([x[k] for x in res]) for k in ("status", "latency_ms", "matched_pairs", "embeddings")}
    out["hash"] = np.array([x["hash"] for x in res], np.uint64)
    out["summary"] = summary
    return out


def serve_bench(args) -> int:
    """Serving (config 4): a resident database of LogNormal-size molecules; for each rate, a
    Poisson stream of requests carrying U[1,50] query graphs (k ~ U[4,32], 30% with a ring) goes
    through the continuous-batching loop.  Reports per-rate p50/p99 latency (completion minus
    arrival) and matched pairs/s, and checks the first requests of each rate against offline
    single-request runs."""
    import numpy as np

    from sigmo_amd import capi

    seed, _, law, _, kmin, kmax, _, ring = CONFIGS[args.config]
    dist = capi.SIZE_NORMAL if law == "normal" else capi.SIZE_LOGNORMAL
    rates = [float(x) for x in args.qps.split(",") if x]
    print(f"[serve] generating {args.molecules} molecules", file=sys.stderr, flush=True)
    dh = capi.HostGraphs.molecules(seed, args.molecules, dist)
    ctxs = [capi.Context(0) for _ in range(max(1, args.workers))]
    dd = capi.DeviceGraph(ctxs[0], dh)
    db = capi.Database(ctxs[0], dd, args.refinement_steps)
    ctxs[0].sync()
    workers = [capi.Batch(c, db) for c in ctxs]
    dev_mem = ctxs[0].device_info()[1]
    print(f"[serve] database resident: {dd.total_nodes} nodes, {db.device_bytes / 1e9:.2f} GB of level planes; "
          f"{dev_mem / 1e9:.0f} GB device", file=sys.stderr, flush=True)
    find_first = not args.find_all
    budget = int(args.budget_gb * 1e9)
    per_rate = []
    total_matched, total_wall = 0, 0.0
    for i, qps in enumerate(rates):
        arr, per = capi.arrivals(seed + i, args.requests, qps, 1, 50)
        rgo = np.concatenate([[0], np.cumsum(per)]).astype(np.uint32)
        qh = capi.HostGraphs.queries(seed * 1000 + i, int(rgo[-1]), kmin, kmax, dist, ring_frac=ring)
        # warm the buffers with the stream's largest request, untimed
        big = int(np.argmax(per))
        capi.serve(workers[:1], qh, rgo[big:big + 2], np.zeros(1), find_first=find_first)
        for c in ctxs:
            c.reset_phases()
        print(f"[serve] {qps:g} QPS: {args.requests} requests, {int(rgo[-1])} query graphs", file=sys.stderr, flush=True)
        if args.online:
            out = _serve_online(capi, workers, qh, rgo, arr, find_first=find_first, memory_budget=budget,
                                max_batch_requests=args.max_batch_requests)
        else:
            out = capi.serve(workers, qh, rgo, arr, find_first=find_first, memory_budget=budget,
                             max_batch_requests=args.max_batch_requests)
        s = out["summary"]
        done = out["status"] == capi.SERVE_DONE
        lat = out["latency_ms"][done]
        # offline check: the first requests alone, one batched run each
        offline_ok = True
        for r in range(min(4, args.requests)):
            if not done[r]:
                continue
            one = capi.HostGraphs.slice(qh, int(rgo[r]), int(rgo[r + 1] - rgo[r]))
            qd = capi.DeviceGraph(ctxs[0], one)
            workers[0].run(qd, find_first)
            st = workers[0].request_stats([0, one.num_graphs])[0]
            offline_ok &= (st["matched_pairs"], st["embeddings"], st["hash"]) == (
                int(out["matched_pairs"][r]), int(out["embeddings"][r]), int(out["hash"][r]))
        pairs_s = s["matched_pairs"] / s["wall_s"] if s["wall_s"] > 0 else 0.0
        total_matched += s["matched_pairs"]
        total_wall += s["wall_s"]
        kst = {}
        for c in ctxs:
            for name, (n, ms, by) in c.kernel_stats().items():
                a = kst.setdefault(name, [0, 0.0, 0.0])
                a[0] += n
                a[1] += ms
                a[2] += by
        entry = {
            "qps": qps, "requests": args.requests, "query_graphs": int(rgo[-1]), "done": s["done"], "rejected": s["rejected"],
            "expired": s["expired"], "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99), "max_ms": _pct(lat, 100),
            "offered_s": round(float(arr[-1]), 3), "wall_s": round(s["wall_s"], 3), "busy_frac": round(s["busy_s"] / s["wall_s"], 3),
            "matched_pairs": s["matched_pairs"], "pairs_per_s": round(pairs_s, 1),
            "query_graph_molecule_pairs_per_s": round(int(rgo[-1]) * dh.num_graphs / s["wall_s"], 1),
            "batches": s["batches"], "mean_batch_requests": round(s["done"] / max(1, s["batches"]), 2),
            "max_inflight_gb": round(s["max_inflight_bytes"] / 1e9, 2), "offline_equal": bool(offline_ok),
            "kernel_ms": {k: round(v[1], 2) for k, v in kst.items() if v[0]},
        }
        per_rate.append(entry)
        print(f"[serve] {json.dumps(entry)}", file=sys.stderr, flush=True)
    line = {
        "metric": METRIC,
        "value": round(total_matched / total_wall, 1) if total_wall else 0.0,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"config {args.config} serving: {args.molecules} molecules seed {seed} LogNormal(ln 24,0.45) sizes "
                        f"resident; Poisson requests of U[1,50] query graphs k~U[{kmin},{kmax}] {int(ring * 100)}% with a ring; "
                        f"R={args.refinement_steps}, {'find-first' if find_first else 'find-all'}; {len(ctxs)} workers, "
                        f"budget {args.budget_gb:g} GB" + ("; online submissions" if args.online else ""),
            "molecules": args.molecules,
            "data_nodes": dd.total_nodes,
        },
        "serving": per_rate,
    }
    print(json.dumps(line))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1, choices=sorted(CONFIGS))
    ap.add_argument("--molecules", type=int, default=None, help="override the config's molecule count")
    ap.add_argument("--queries", type=int, default=None, help="override the config's query count")
    ap.add_argument("--find-all", action="store_true", help="count every embedding instead of find-first")
    ap.add_argument("--refinement-steps", type=int, default=6)
    ap.add_argument("--cpu-molecules", type=int, default=0,
                    help="molecules in the CPU baseline sample (default: about 1e7 query x molecule pairs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 plain device copies of MIB MiB (the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--form", choices=("plain", "batched"), default="plain",
                    help="plain: per-phase kernels over the Q x D bitset; batched: per-pair evaluation from "
                         "per-molecule level planes (recomputed inside every step)")
    ap.add_argument("--serve", action="store_true",
                    help="serving mode (config 4): Poisson request streams against a resident database")
    ap.add_argument("--qps", default="50,200,800", help="serving: comma-separated request rates")
    ap.add_argument("--requests", type=int, default=200, help="serving: requests per rate")
    ap.add_argument("--workers", type=int, default=2, help="serving: batches in flight (one stream each)")
    ap.add_argument("--budget-gb", type=float, default=128.0, help="serving: device bytes for batches in flight")
    ap.add_argument("--max-batch-requests", type=int, default=256)
    ap.add_argument("--online", action="store_true",
                    help="serving: submit each request at its arrival time through the online server (sigmo_server_*) "
                         "instead of handing the loop the whole arrival schedule")
    ap.add_argument("--chunk-gb", type=float, default=128.0,
                    help="config 3: size chunks so the plain form's candidate bitset of one chunk fits this many GB")
    ap.add_argument("--chunk-molecules", type=int, default=0, help="config 3: molecules per chunk (overrides --chunk-gb)")
    ap.add_argument("--no-overlap", action="store_true", help="config 3: load chunks inline instead of on a side stream")
    ap.add_argument("--autotune", action="store_true", help="batched form: pick launch shapes with the autotuner before timing")
    ap.add_argument("--graph", action="store_true",
                    help="capture one whole step (the same kernel launches and work; the batched form's includes the "
                         "database refresh) as a hipGraph and time its replays; per-kernel figures then come from eager "
                         "steps run just before the capture")
    args = ap.parse_args()
    if args.serve and args.config == 1 and "--config" not in sys.argv:
        args.config = 4
    if args.molecules is None:
        args.molecules = CONFIGS[args.config][1]
    if args.queries is None:
        args.queries = CONFIGS[args.config][3]

    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * 1024 * 1024 // 4
        a = torch.ones(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 4, "copies": 3}))
        return 0

    if args.serve:
        if args.gpus != 1:
            print("bench.py measures one GPU; use --gpus 1.", file=sys.stderr)
            return 2
        return serve_bench(args)

    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs SIGMo's hot path on a single MI355X "
              "(molecule shards would run one per GPU with no communication); use --gpus 1.", file=sys.stderr)
        return 2

    from sigmo_amd import capi

    print(f"[bench] generating workload: {args.molecules} molecules, {args.queries} queries", file=sys.stderr, flush=True)
    qh, dh = workload(args)
    ctx = capi.Context(0)
    find_first = not args.find_all
    qd = capi.DeviceGraph(ctx, qh)
    if args.config == 3:
        # the shard stays in host memory; one step streams it through the device in chunks
        if not args.chunk_molecules:
            # the plain form's Q x D candidate bitset of one chunk within --chunk-gb; both forms use
            # the same chunks
            mean_n = dh.total_nodes / dh.num_graphs
            args.chunk_molecules = int(args.chunk_gb * 1e9 / (qd.total_nodes / 8.0) / mean_n)
        data_nodes = dh.total_nodes

        def step():
            return capi.stream_run(ctx, qd, dh, chunk_molecules=args.chunk_molecules, form=args.form,
                                   refinement_steps=args.refinement_steps, find_first=find_first, overlap=not args.no_overlap)
    elif args.form == "plain":
        dd = capi.DeviceGraph(ctx, dh)
        data_nodes = dd.total_nodes
        pipe = capi.Pipeline(ctx, qd, dd)

        def step():
            return pipe.execute(args.refinement_steps, find_first)

        if args.graph and args.config != 3:
            # per-kernel statistics from eager executes; the timed steps replay the captured execute
            ctx.reset_phases()
            for _ in range(max(1, args.warmup)):
                step()
            ctx.sync()
            graph_kstats, graph_phase_ms, graph_runs = ctx.kernel_stats(), ctx.phase_ms(), max(1, args.warmup)
            pipe.capture(args.refinement_steps, find_first)

            def step():
                return pipe.replay()
    else:
        dd = capi.DeviceGraph(ctx, dh)
        data_nodes = dd.total_nodes
        db = capi.Database(ctx, dd, args.refinement_steps)
        batch = capi.Batch(ctx, db)
        if args.autotune:
            print("[bench] autotuning", file=sys.stderr, flush=True)
            args.tuning, _ = batch.autotune(qd, find_first)
        def eager_step():
            db.refresh()  # the data side is recomputed every step, as the plain form does
            return batch.run(qd, find_first)

        step = eager_step
        if args.graph:
            # per-kernel statistics come from eager steps of the same kernels (a graph has no
            # per-kernel events); the timed steps are replays of the captured step
            ctx.reset_phases()
            for _ in range(max(1, args.warmup)):
                eager_step()
            ctx.sync()
            graph_kstats, graph_phase_ms, graph_runs = ctx.kernel_stats(), ctx.phase_ms(), max(1, args.warmup)
            batch.capture(qd, find_first, refresh_db=True)

            def step():
                return batch.replay()  # the captured step refreshes the data side too
    print(f"[bench] uploaded: Q={qd.total_nodes} query nodes, D={data_nodes} data nodes; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.reset_phases()
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        results.append(step())
    ctx.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    matched = results[-1]["matched_pairs"]
    assert all(r["matched_pairs"] == matched for r in results), "non-deterministic match count"
    value = matched / (ms_per_step / 1000.0)
    print(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step, {matched} matched pairs", file=sys.stderr, flush=True)

    kstats = ctx.kernel_stats()
    phase_ms = ctx.phase_ms()
    phase_ms = {k: round(phase_ms[k] / args.steps, 4) for k in capi.PHASE_NAMES}
    if args.graph and args.config != 3:
        # eager-step statistics, scaled to one step (launch counts per step stay integral)
        kstats = {k: (n * args.steps / graph_runs, ms * args.steps / graph_runs, by * args.steps / graph_runs)
                  for k, (n, ms, by) in graph_kstats.items()}
        phase_ms = {k: round(graph_phase_ms[k] / graph_runs, 4) for k in capi.PHASE_NAMES}
    kernels = {}
    for name, (n, ms, by) in kstats.items():
        if n:
            avg_ms = ms / n
            kernels[name] = {"launches_per_step": n / args.steps, "avg_ms": round(avg_ms, 5),
                             "gbs": round(by / n / (avg_ms / 1e3) / 1e9, 1) if by else None}
    hbm = {k: v for k, v in kstats.items() if k in HBM_KERNELS and v[0]}
    dom = max(hbm, key=lambda k: hbm[k][1])
    n, ms, by = hbm[dom]
    achieved = by / n / (ms / n / 1e3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_{dom}_config{args.config}_{args.form}.json"
    if tfile.exists() and args.molecules == CONFIGS[args.config][1] and args.queries == CONFIGS[args.config][3]:
        traffic = json.loads(tfile.read_text()).get("traffic_bytes_per_launch")
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": by / n,
        "avg_launch_ms": round(ms / n, 5),
    }
    overall = max(kstats, key=lambda k: kstats[k][1])

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(qh, dh, args)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": describe(args),
            "molecules": args.molecules,
            "queries": args.queries,
            "query_nodes": qd.total_nodes,
            "data_nodes": data_nodes,
            "refinement_steps": args.refinement_steps,
            "gmcr_pairs": results[-1]["gmcr_pairs"],
            "matched_pairs": matched,
            "embeddings": results[-1]["num_matches"] if args.find_all else None,
            "dfs_steps": results[-1]["dfs_steps"],
            "join_lane_use": (round(results[-1]["dfs_steps"] / (64.0 * results[-1]["join_wave_steps"]), 4)
                              if results[-1].get("join_wave_steps") else None),
            **({k: results[-1][k] for k in ("filter_wave_steps", "filter_lane_steps", "filter_scans")}
               if args.form == "batched" and args.config != 3 else {}),
            **({k: results[-1][k] for k in ("chunks", "wall_ms", "compute_ms", "load_ms", "load_wait_ms")}
               if args.config == 3 else {}),
            "form": args.form,
            **({"tuning": getattr(args, "tuning", None)} if args.form == "batched" else {}),
            "graph": bool(args.graph and args.config != 3),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phase_ms": phase_ms,
        "kernels": kernels,
        "dominant_kernel": overall,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
