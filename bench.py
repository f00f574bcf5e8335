#!/usr/bin/env python3
"""Throughput bench: matched (query, molecule) pairs/s of SIGMo's hot path on one MI355X.

Workload (default): one rank's shard of the multi-node run -- 20,000,000 synthetic molecules
(seed 3, heavy-atom count clamp(round(LogNormal(ln 24, 0.45)), 6, 128)) x 2,000 query graphs
(k ~ U[4,32], 30% with a ring), R = 6, find-first -- held in host memory and streamed through the
device in chunks, the next chunk uploaded on a side stream while the current one runs.  One
step = one pass over the whole shard.  The value is the batched persistent form's matched
(query, molecule) pairs/s; the plain per-phase form (signatures, filter, 6 x refine, GMCR, join
over a Q x D candidate bitset per chunk, chunks as large as the device holds) is timed in the
same run on the same shard, and the two must agree on every total.

Prints ONE JSON line with the throughput, a roofline entry for the dominant HBM-bound kernel of
each form (algorithmic bytes / measured HIP-event time) and a CPU baseline: the CPU oracle (a
restatement of the reference kernels, OpenMP over the host's cores) on a bounded molecule sample
of the same workload.  --config 1/2 time the device-resident configs instead; --serve runs the
serving workload.
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


def describe(args, plan=None) -> str:
    seed, _, law, _, kmin, kmax, perturb, ring = CONFIGS[args.config]
    sizes = "Normal(24,8) sizes" if law == "normal" else "LogNormal(ln 24,0.45) sizes"
    extra = f" {int(perturb * 100)}% perturbed" if perturb > 0 else ""
    extra += f" {int(ring * 100)}% with a ring" if ring >= 0 else ""
    mode = "find-first" if not args.find_all else "find-all"
    if args.form == "both":
        form = "batched persistent form; plain per-phase form timed in the same run"
    else:
        form = "plain per-phase form" if args.form == "plain" else "batched persistent form"
    stream = ""
    if plan:
        stream = (f", streamed from host memory in chunks (batched {plan['batched']} molecules, plain {plan['plain']} "
                  f"molecules = a {plan['bitset_gb']} GB candidate bitset per chunk)"
                  f"{'; next chunk loaded on a side stream' if not args.no_overlap else ''}")
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


def _cpu_threads() -> int:
    """Threads for the CPU baseline: OMP_NUM_THREADS when set (the host's share of cores for this
    process), otherwise every core the process may run on."""
    env = os.environ.get("SIGMO_CPU_THREADS") or os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(qh, dh, args) -> dict:
    """Oracle (CPU restatement of the reference kernels) on the first molecules of the workload,
    built with g++ -O3 -march=native -fopenmp for this host (the portable build when that fails)."""
    import tempfile

    from oracle import oracle as O

    build = "g++ -O3 -march=native -fopenmp"
    tmp = tempfile.TemporaryDirectory(prefix="sigmo_oracle_")
    try:
        O.use_library(O.build_native(tmp.name))
    except Exception as e:  # noqa: BLE001 -- fall back to the prebuilt portable oracle
        print(f"[bench] native oracle build failed ({e}); using the portable build", file=sys.stderr, flush=True)
        build = "g++ -O3 -march=x86-64-v2 -fopenmp (prebuilt)"
    threads = _cpu_threads()
    O.set_threads(threads)
    count = args.cpu_molecules or max(1, CPU_SAMPLE_PAIRS // max(1, qh.num_graphs))
    sample = dh.slice(0, min(count, dh.num_graphs))
    qa, sa = qh.arrays(), sample.arrays()
    t0 = time.perf_counter()
    r = O.pipeline(qa, sa, args.refinement_steps, modes=[not args.find_all])
    dt = time.perf_counter() - t0
    tmp.cleanup()
    return {
        "value": (r["total_first"] if not args.find_all else int((r["counts_all"] > 0).sum())) / dt,
        "unit": "pairs/s",
        "cores": O.threads(),
        "kind": "port",
        "sample": f"first {sample.num_graphs} of the {dh.num_graphs} molecules x all {qh.num_graphs} queries, "
                  f"R={args.refinement_steps}, {'find-first' if not args.find_all else 'find-all'}, whole pipeline in one "
                  f"pass (no chunking needed at this size); {dt:.2f} s on {O.threads()} OpenMP threads ({_cpu_model()}, "
                  f"{os.cpu_count()} logical CPUs visible, OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')}); "
                  f"built {build}",
    }


def calibrate_copy(mib: int) -> int:
    """Three device-to-device copies of `mib` MiB through the HIP runtime (its copy kernel), the
    known-size reference for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mib * 1024 * 1024
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, ctypes.c_size_t(n)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": n, "copies": 3}))
    return 0


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


def _device_free_bytes() -> int:
    """Free device memory as the HIP runtime reports it (hipMemGetInfo), without PyTorch."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return int(free.value)


def _even_chunks(G: int, max_chunk: int) -> int:
    """Molecules per chunk when G molecules are split into the fewest chunks of at most max_chunk."""
    n = max(1, -(-G // max(1, max_chunk)))
    return -(-G // n)


def chunk_plan(args, qd, dh) -> dict:
    """Config 3's chunk plans.  Plain form: the largest chunks whose Q x D candidate bitset, with
    the rest of one chunk's device buffers, fits the device (--chunk-gb caps the bitset instead).
    Batched form: chunks of at most --batched-chunk molecules (its candidate arena uses 32-bit
    offsets: the survivors' masks of 2 M molecules x 2,000 queries fit them, those of 4 M do not;
    2 M-molecule chunks run a config 3 pass 3% faster than 1 M ones)."""
    mean_n = dh.total_nodes / dh.num_graphs
    bitset_per_mol = qd.total_nodes * mean_n / 8.0
    if args.chunk_gb:
        budget = args.chunk_gb * 1e9
        source = f"--chunk-gb {args.chunk_gb:g}"
    else:
        free = _device_free_bytes()
        # per data node, besides the bitset: two alternating graph buffers (~70 B each), signature
        # state and planes (~32 B) and the GMCR arrays (pair bits and ~16 B per surviving pair)
        per_mol_other = mean_n * (2 * 70 + 32) + args.queries * (1 / 8.0 + 0.1 * 16)
        budget = (free - 8e9) * bitset_per_mol / (bitset_per_mol + per_mol_other)
        source = f"{free / 1e9:.1f} GB free on the device"
    plain = _even_chunks(dh.num_graphs, int(budget / bitset_per_mol))
    batched = _even_chunks(dh.num_graphs, args.batched_chunk)
    return {"plain": plain, "batched": batched, "bitset_gb": round(plain * bitset_per_mol / 1e9, 1), "source": source}


def time_form(args, ctx, step, steps: int, warmup: int) -> dict:
    """Warm up, then time `steps` calls of step() bracketed by device synchronisations; returns the
    per-step time, the last result and the context's per-kernel and per-phase figures."""
    from sigmo_amd import capi

    for _ in range(warmup):
        step()
    ctx.sync()
    ctx.reset_phases()
    results = []
    t0 = time.perf_counter()
    for _ in range(steps):
        results.append(step())
    ctx.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / steps
    matched = results[-1]["matched_pairs"]
    assert all(r["matched_pairs"] == matched for r in results), "non-deterministic match count"
    phase = ctx.phase_ms()
    return {"ms_per_step": ms, "matched": matched, "result": results[-1], "kstats": ctx.kernel_stats(),
            "phase_ms": {k: round(phase[k] / steps, 4) for k in capi.PHASE_NAMES}, "steps": steps}


def kernel_table(kstats: dict, steps: int) -> dict:
    out = {}
    for name, (n, ms, by) in kstats.items():
        if n:
            avg_ms = ms / n
            out[name] = {"launches_per_step": n / steps, "avg_ms": round(avg_ms, 5),
                         "gbs": round(by / n / (avg_ms / 1e3) / 1e9, 1) if by else None}
    return out


def roofline_of(kstats: dict, args, form: str) -> dict:
    """The form's dominant HBM-bound kernel (by time): algorithmic bytes per launch over its HIP-event
    launch time, with the rocprofv3-measured traffic per launch where a summary exists for it."""
    hbm = {k: v for k, v in kstats.items() if k in HBM_KERNELS and v[0]}
    dom = max(hbm, key=lambda k: hbm[k][1])
    n, ms, by = hbm[dom]
    achieved = by / n / (ms / n / 1e3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_{dom}_config{args.config}_{form}.json"
    if tfile.exists() and args.molecules == CONFIGS[args.config][1] and args.queries == CONFIGS[args.config][3]:
        traffic = json.loads(tfile.read_text()).get("traffic_bytes_per_launch")
    return {
        "bound": "hbm",
        "kernel": dom,
        "form": form,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": by / n,
        "avg_launch_ms": round(ms / n, 5),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=sorted(CONFIGS))
    ap.add_argument("--molecules", type=int, default=None, help="override the config's molecule count")
    ap.add_argument("--queries", type=int, default=None, help="override the config's query count")
    ap.add_argument("--find-all", action="store_true", help="count every embedding instead of find-first")
    ap.add_argument("--refinement-steps", type=int, default=6)
    ap.add_argument("--cpu-molecules", type=int, default=0,
                    help="molecules in the CPU baseline sample (default: about 1e7 query x molecule pairs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB through the HIP runtime (the known-size kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--form", choices=("plain", "batched", "both"), default="both",
                    help="both (default): the batched form gives the value and the plain form is timed in the same run "
                         "for the ratio; plain: per-phase kernels over the Q x D bitset; batched: per-pair evaluation from "
                         "per-molecule level planes (recomputed inside every step)")
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="timed steps of the plain form when --form both (default: 2 for config 3, --steps otherwise)")
    ap.add_argument("--plain-warmup", type=int, default=None,
                    help="untimed steps of the plain form when --form both (default: 1 for config 3, --warmup otherwise)")
    ap.add_argument("--serve", action="store_true",
                    help="serving mode (config 4): Poisson request streams against a resident database")
    ap.add_argument("--qps", default="50,200,800", help="serving: comma-separated request rates")
    ap.add_argument("--requests", type=int, default=200, help="serving: requests per rate")
    ap.add_argument("--workers", type=int, default=1,
                    help="serving: batches in flight (one stream each); at 50 QPS on 10 M molecules one worker gives "
                         "the lower tail (p99 156 ms against 254 ms with two)")
    ap.add_argument("--budget-gb", type=float, default=128.0, help="serving: device bytes for batches in flight")
    ap.add_argument("--max-batch-requests", type=int, default=256)
    ap.add_argument("--online", action="store_true",
                    help="serving: submit each request at its arrival time through the online server (sigmo_server_*) "
                         "instead of handing the loop the whole arrival schedule")
    ap.add_argument("--chunk-gb", type=float, default=0.0,
                    help="config 3: cap the plain form's candidate bitset of one chunk at this many GB (default: the "
                         "largest chunk that fits the device's free memory)")
    ap.add_argument("--batched-chunk", type=int, default=1 << 21, help="config 3: most molecules per chunk of the batched form")
    ap.add_argument("--no-overlap", action="store_true", help="config 3: load chunks inline instead of on a side stream")
    ap.add_argument("--stream-workers", type=int, default=3,
                    help="config 3, batched form: chunks run on this many contexts at once (each its own stream; "
                         "per-kernel times then overlap): 1 / 2 / 3 workers take 1.38 / 1.33 / 1.30 s per pass")
    ap.add_argument("--autotune", action="store_true", help="batched form: pick launch shapes with the autotuner before timing")
    ap.add_argument("--filter-mode", type=int, default=0, choices=(0, 1, 2),
                    help="configs 1-2, batched form: pair evaluation by molecule (1) or from the pre-tested pair list (2); "
                         "0 (default): the library's choice (list mode below 256 query graphs)")
    ap.add_argument("--graph", action="store_true",
                    help="configs 1-2: capture one whole step (the same kernel launches and work; the batched form's "
                         "includes the database refresh) as a hipGraph and time its replays; per-kernel figures then "
                         "come from eager steps run just before the capture")
    args = ap.parse_args()
    if args.serve and "--config" not in sys.argv:
        args.config = 4
    if args.molecules is None:
        args.molecules = CONFIGS[args.config][1]
    if args.queries is None:
        args.queries = CONFIGS[args.config][3]
    if args.plain_steps is None:
        args.plain_steps = 2 if args.config == 3 else args.steps
    if args.plain_warmup is None:
        args.plain_warmup = 1 if args.config == 3 else args.warmup

    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs SIGMo's hot path on a single MI355X "
              "(molecule shards would run one per GPU with no communication); use --gpus 1.", file=sys.stderr)
        return 2
    if args.serve:
        return serve_bench(args)
    if args.config == 4:
        print("config 4 is the serving workload: use --serve", file=sys.stderr)
        return 2

    from sigmo_amd import capi

    print(f"[bench] generating workload: {args.molecules} molecules, {args.queries} queries", file=sys.stderr, flush=True)
    qh, dh = workload(args)
    ctx = capi.Context(0)
    find_first = not args.find_all
    qd = capi.DeviceGraph(ctx, qh)
    forms = ["batched", "plain"] if args.form == "both" else [args.form]
    steps_of = {f: (args.plain_steps, args.plain_warmup) if (f == "plain" and args.form == "both") else (args.steps, args.warmup)
                for f in forms}
    timed, graph_stats = {}, {}
    plan = None
    if args.config == 3:
        # the shard stays in host memory; one step streams it through the device in chunks
        plan = chunk_plan(args, qd, dh)
        data_nodes = dh.total_nodes
        print(f"[bench] chunk plan: {plan}", file=sys.stderr, flush=True)
        for f in forms:
            # one session per form: its first pass (a warmup step) allocates the device buffers,
            # which later passes reuse, as a rank allocates once before its timed phases
            sess = capi.Stream(ctx, qd, dh, chunk_molecules=plan[f], form=f, refinement_steps=args.refinement_steps,
                               find_first=find_first, overlap=not args.no_overlap,
                               workers=args.stream_workers if f == "batched" else 1)
            print(f"[bench] {f}: warmup {steps_of[f][1]}, {steps_of[f][0]} timed steps", file=sys.stderr, flush=True)
            timed[f] = time_form(args, ctx, sess.pass_, max(1, steps_of[f][0]), max(1, steps_of[f][1]))
            sess.close()  # frees the form's device buffers before the next form allocates
            print(f"[bench] {f}: {timed[f]['ms_per_step']:.1f} ms/step, {timed[f]['matched']} matched pairs",
                  file=sys.stderr, flush=True)
    else:
        dd = capi.DeviceGraph(ctx, dh)
        data_nodes = dd.total_nodes
        print(f"[bench] uploaded: Q={qd.total_nodes} query nodes, D={data_nodes} data nodes", file=sys.stderr, flush=True)
        for f in forms:
            if f == "plain":
                pipe = capi.Pipeline(ctx, qd, dd)

                def step():
                    return pipe.execute(args.refinement_steps, find_first)

                if args.graph:
                    ctx.reset_phases()
                    for _ in range(max(1, args.warmup)):
                        step()
                    ctx.sync()
                    graph_stats[f] = (ctx.kernel_stats(), ctx.phase_ms(), max(1, args.warmup))
                    pipe.capture(args.refinement_steps, find_first)

                    def step():
                        return pipe.replay()
            else:
                db = capi.Database(ctx, dd, args.refinement_steps)
                batch = capi.Batch(ctx, db)
                if args.filter_mode:
                    batch.tuning = {"filter_mode": args.filter_mode}
                    args.tuning = batch.tuning
                if args.autotune:
                    print("[bench] autotuning", file=sys.stderr, flush=True)
                    args.tuning, _ = batch.autotune(qd, find_first)

                def step():
                    db.refresh()  # the data side is recomputed every step, as the plain form does
                    return batch.run(qd, find_first)

                if args.graph:
                    ctx.reset_phases()
                    for _ in range(max(1, args.warmup)):
                        step()
                    ctx.sync()
                    graph_stats[f] = (ctx.kernel_stats(), ctx.phase_ms(), max(1, args.warmup))
                    batch.capture(qd, find_first, refresh_db=True)

                    def step():
                        return batch.replay()  # the captured step refreshes the data side too
            timed[f] = time_form(args, ctx, step, *steps_of[f])
            print(f"[bench] {f}: {timed[f]['ms_per_step']:.3f} ms/step, {timed[f]['matched']} matched pairs",
                  file=sys.stderr, flush=True)
    if len(forms) == 2:
        assert timed["plain"]["matched"] == timed["batched"]["matched"], "the two forms disagree"
        assert timed["plain"]["result"]["gmcr_pairs"] == timed["batched"]["result"]["gmcr_pairs"], "the two forms disagree"
    for f, (ks, ph, runs) in graph_stats.items():
        # eager-step statistics of the same kernels, scaled to the timed steps
        n_steps = timed[f]["steps"]
        timed[f]["kstats"] = {k: (n * n_steps / runs, ms * n_steps / runs, by * n_steps / runs) for k, (n, ms, by) in ks.items()}
        timed[f]["phase_ms"] = {k: round(ph[k] / runs, 4) for k in capi.PHASE_NAMES}

    head = forms[0]  # the value: the batched form when both run
    main_t = timed[head]
    value = main_t["matched"] / (main_t["ms_per_step"] / 1000.0)
    res = main_t["result"]

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(qh, dh, args)

    per_form = {}
    for f in forms:
        t = timed[f]
        per_form[f] = {
            "ms_per_step": round(t["ms_per_step"], 4),
            "pairs_per_s": round(t["matched"] / (t["ms_per_step"] / 1000.0), 1),
            "steps": t["steps"],
            "roofline": roofline_of(t["kstats"], args, f),
            "phase_ms": t["phase_ms"],
            "kernels": kernel_table(t["kstats"], t["steps"]),
            "dominant_kernel": max(t["kstats"], key=lambda k: t["kstats"][k][1]),
            **({k: t["result"][k] for k in ("chunks", "wall_ms", "compute_ms", "load_ms", "load_wait_ms")}
               if args.config == 3 else {}),
            **({k: t["result"][k] for k in ("filter_wave_steps", "filter_lane_steps", "filter_scans")}
               if f == "batched" and args.config != 3 else {}),
            **({"chunk_molecules": plan[f]} if plan else {}),
        }
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_t["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": describe(args, plan),
            "molecules": args.molecules,
            "queries": args.queries,
            "query_nodes": qd.total_nodes,
            "data_nodes": data_nodes,
            "refinement_steps": args.refinement_steps,
            "gmcr_pairs": res["gmcr_pairs"],
            "matched_pairs": main_t["matched"],
            "embeddings": res["num_matches"] if args.find_all else None,
            "dfs_steps": res["dfs_steps"],
            "join_lane_use": (round(res["dfs_steps"] / (64.0 * res["join_wave_steps"]), 4) if res.get("join_wave_steps") else None),
            "form": head,
            **({"tuning": getattr(args, "tuning", None)} if "batched" in forms else {}),
            "graph": bool(args.graph and args.config != 3),
            **({"chunk_plan": plan} if plan else {}),
        },
        "roofline": per_form[head]["roofline"],
        "cpu_baseline": cpu,
        **({"batched_over_plain": round(timed["plain"]["ms_per_step"] / timed["batched"]["ms_per_step"], 3)}
           if len(forms) == 2 else {}),
        "forms": per_form,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
