"""One rank of the reference's multi-node driver (library/src/sigmo_mpi.cpp:81-240) on this
library.

Each rank reads its share of the data file as the reference's loadFileLinesMPI does
(sigmo_host_graphs_load_dat_shard: the rank's byte range, partial lines dropped, the n/e/l line
filter), loads every query graph (:117), runs the pipeline on its device (:160-207), and the ranks
sum their data-graph counts (:136-137) and matches (:221-222) and gather their times (:212-216)
over torch.distributed instead of MPI.  Shards are molecules only, so no rank exchanges anything
while it computes.  Sums travel as CPU tensors on a gloo group and as device tensors on an RCCL
("nccl") group.  Without an initialised process group a rank runs alone and its totals are its
own, which is how one GPU runs any single rank of a multi-rank job.
"""
from __future__ import annotations

import time

from . import capi


def _dist():
    import torch.distributed as dist

    return dist if dist.is_available() and dist.is_initialized() else None


def _device(dist, group):
    return "cuda" if dist.get_backend(group) == "nccl" else "cpu"


def reduce_totals(values, group=None) -> list[int]:
    """Sums of the integers `values` over the ranks of `group`, on every rank (MPI_Reduce with
    MPI_SUM, sigmo_mpi.cpp:136-137 and :221-222, delivers them to rank 0 only)."""
    dist = _dist()
    vals = [int(v) for v in values]
    if dist is None:
        return vals
    import torch

    t = torch.tensor(vals, dtype=torch.int64, device=_device(dist, group))
    dist.all_reduce(t, op=dist.ReduceOp.SUM, group=group)
    return [int(x) for x in t.cpu().tolist()]


def gather_times(ms: float, group=None) -> list[float]:
    """Every rank's time in rank order (MPI_Gather, sigmo_mpi.cpp:212-216)."""
    dist = _dist()
    if dist is None:
        return [float(ms)]
    import torch

    dev = _device(dist, group)
    mine = torch.tensor([float(ms)], dtype=torch.float64, device=dev)
    out = [torch.zeros(1, dtype=torch.float64, device=dev) for _ in range(dist.get_world_size(group))]
    dist.all_gather(out, mine, group=group)
    return [float(x.cpu().item()) for x in out]


def run_rank(ctx: capi.Context, query_path, data_path, max_data_graphs: int, refinement_steps: int = 6,
             find_all: bool = False, rank: int | None = None, nprocs: int | None = None, group=None) -> dict:
    """This rank's run.  rank and nprocs default to the process group's (0 and 1 without one).
    Returns this rank's counts and matches, the sums over all ranks, the reference's average
    (total // nprocs, sigmo_mpi.cpp:228) and every rank's time in ms."""
    dist = _dist()
    if rank is None:
        rank = dist.get_rank(group) if dist else 0
    if nprocs is None:
        nprocs = dist.get_world_size(group) if dist else 1
    data = capi.HostGraphs.load_dat_shard(data_path, rank, nprocs, max_data_graphs)
    query = capi.HostGraphs.load_dat(query_path)
    t0 = time.perf_counter()
    result = None
    matches = 0
    if data.num_graphs and query.num_graphs:  # an empty shard has nothing to match
        qd, dd = capi.DeviceGraph(ctx, query), capi.DeviceGraph(ctx, data)
        result = capi.pipeline_run(ctx, qd, dd, refinement_steps, find_first=not find_all)
        matches = int(result["num_matches"])
    rank_ms = (time.perf_counter() - t0) * 1e3
    total_graphs, total_matches = reduce_totals([data.num_graphs, matches], group)
    return {
        "rank": rank,
        "nprocs": nprocs,
        "query_graphs": query.num_graphs,
        "data_graphs": data.num_graphs,
        "total_data_graphs": total_graphs,
        "num_matches": matches,
        "total_matches": total_matches,
        "average_matches": total_matches // nprocs,
        "rank_ms": gather_times(rank_ms, group),
        "result": result,
    }
