"""GPU tests of the serving loop (sigmo_serve): per-request results are the same alone and in
a mixed continuous batch, equal the oracle's, and do not depend on the number of workers."""
from __future__ import annotations

import time

import numpy as np
import pytest

from sigmo_amd import capi
from oracle import oracle as O
from tests._util import request_stats

pytestmark = pytest.mark.gpu


def _requests(seed, count, qmin, qmax):
    """Request graph offsets for `count` requests of U[qmin, qmax] query graphs each."""
    _, per = capi.arrivals(seed, count, 100.0, qmin, qmax)
    return np.concatenate([[0], np.cumsum(per)]).astype(np.uint32)


@pytest.fixture(scope="module")
def served(ctx):
    dh = capi.HostGraphs.molecules(4, 4000, capi.SIZE_LOGNORMAL)
    rgo = _requests(4, 24, 1, 6)
    qh = capi.HostGraphs.queries(4, int(rgo[-1]), 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)
    dd = capi.DeviceGraph(ctx, dh)
    db = capi.Database(ctx, dd, 6)
    return dh, qh, rgo, db


def _stats_of(out):
    return list(zip(out["matched_pairs"].tolist(), out["embeddings"].tolist(), out["hash"].tolist()))


@pytest.mark.parametrize("find_first", [True, False])
def test_serve_alone_equals_mix_equals_oracle(ctx, served, find_first):
    dh, qh, rgo, db = served
    n = rgo.size - 1
    batch = capi.Batch(ctx, db)
    print(f"serve mix find_first={find_first}: {n} requests, {int(rgo[-1])} query graphs", flush=True)
    mix = capi.serve([batch], qh, rgo, np.zeros(n), find_first=find_first, max_batch_requests=8)
    assert (mix["status"] == capi.SERVE_DONE).all()
    assert mix["batch_requests"].max() > 1  # requests really shared batches
    print("serve alone", flush=True)
    alone = capi.serve([batch], qh, rgo, np.zeros(n), find_first=find_first, max_batch_requests=1)
    assert (alone["batch_requests"] == 1).all()
    assert _stats_of(mix) == _stats_of(alone)
    # the oracle over all query graphs at once, split by request
    key = "first" if find_first else "all"
    o = O.pipeline(qh.arrays(), dh.arrays(), 6, modes=[find_first])
    assert _stats_of(mix) == request_stats(o["offsets"], o["indices"], o[f"counts_{key}"], rgo)
    s = mix["summary"]
    assert s["matched_pairs"] == int((o[f"counts_{key}"] > 0).sum())
    assert s["gmcr_pairs"] == len(o["indices"])


def test_serve_two_workers_and_budget(ctx, served):
    dh, qh, rgo, db = served
    n = rgo.size - 1
    w = [capi.Batch(ctx2, db) for ctx2 in (capi.Context(0), capi.Context(0))]
    one = capi.serve([capi.Batch(ctx, db)], qh, rgo, np.zeros(n), max_batch_requests=5)
    per_graph = w[0].workspace_bytes(64, 64 * 18) // 64
    budget = per_graph * 5
    print(f"serve two workers, budget {budget} B", flush=True)
    two = capi.serve(w, qh, rgo, np.linspace(0.0, 0.05, n), max_batch_requests=5, memory_budget=budget,
                     bytes_per_graph=per_graph)
    done = two["status"] == capi.SERVE_DONE
    assert done.any()
    big = np.diff(rgo) * per_graph > budget
    assert (two["status"][big] == capi.SERVE_REJECTED).all() and done[~big].all()
    assert two["summary"]["max_inflight_bytes"] <= budget
    assert [s for s, d in zip(_stats_of(two), done) if d] == [s for s, d in zip(_stats_of(one), done) if d]


def test_request_stats_partition_the_batch(ctx, served):
    """Per-request sums over any split of a batch add up to the whole run's totals."""
    dh, qh, rgo, db = served
    batch = capi.Batch(ctx, db)
    qd = capi.DeviceGraph(ctx, qh)
    r = batch.run(qd, find_first=False)
    whole = batch.request_stats([0, qh.num_graphs])[0]
    assert whole["matched_pairs"] == r["matched_pairs"] and whole["embeddings"] == r["num_matches"]
    parts = batch.request_stats(rgo)
    assert sum(p["matched_pairs"] for p in parts) == r["matched_pairs"]
    assert sum(p["embeddings"] for p in parts) == r["num_matches"]
    off, idx, cnt = batch.pairs()
    assert [(p["matched_pairs"], p["embeddings"], p["hash"]) for p in parts] == request_stats(off, idx, cnt, rgo)
    with pytest.raises(capi.SigmoError):
        batch.request_stats([0, 1])  # does not cover the batch


def test_online_server_equals_offline(ctx, served):
    """Requests submitted while the server runs (two workers, one stream each) give the same
    per-request results as offline single-request runs."""
    dh, qh, rgo, db = served
    n = rgo.size - 1
    alone = capi.serve([capi.Batch(ctx, db)], qh, rgo, np.zeros(n), max_batch_requests=1)
    workers = [capi.Batch(c, db) for c in (capi.Context(0), capi.Context(0))]
    print(f"online server: {n} requests over two workers", flush=True)
    srv = capi.Server(workers, max_batch_requests=4)
    ids = []
    for r in range(n):
        ids.append(srv.submit(qh.slice(int(rgo[r]), int(rgo[r + 1] - rgo[r]))))
        if r % 6 == 5:
            time.sleep(0.002)
    res = [srv.poll(i, timeout_ms=-1) for i in ids]
    s = srv.close()
    assert ids == list(range(n))
    assert all(r["status"] == capi.SERVE_DONE for r in res)
    assert [(r["matched_pairs"], r["embeddings"], r["hash"]) for r in res] == _stats_of(alone)
    assert s["done"] == n and s["matched_pairs"] == int(alone["matched_pairs"].sum())
