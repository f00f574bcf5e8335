"""GPU runs of every configuration at its stated workload size, checked against the oracle's
committed digests where the oracle can run and, beyond that, against the other form and
through size-independent properties.

* Config 1 (100,000 molecules x 100 queries): the timed path, sigmo_pipeline_execute (fused
  filter-refine, deferred GMCR count), and the batched form, both join modes, per pair against
  the oracle's digests of the full workload (tests/golden/config1.json).
* Config 2, exact query set (1,000 queries, seed 2) against the first 20,000 molecules of seed 2:
  both forms per pair against the oracle's digests (tests/golden/config2_subset.json).
* Config 2 at full size (1,000,000 molecules x 1,000 queries, a 10^9-pair space): the batched
  form (arena and GMCR sized from the measured demand) equals the plain form (host-sized GMCR
  beyond the deferred-count limit) pair by pair in both modes, and find-first marks exactly the
  pairs find-all counts.
* Config 3 at full size (20,000,000 molecules x 2,000 queries, streamed): both forms over the
  same chunk plan give equal totals and equal per-chunk digests of (offsets, indices, counts).
* Config 4 (10,000,000 molecules resident): a Poisson stream of U[1,50]-graph requests served
  with continuous batching; every request is served and a sample of them equals its result when
  run alone.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from sigmo_amd import capi
from tests._util import GOLDEN, digest

pytestmark = pytest.mark.gpu


def _config2_queries():
    return capi.HostGraphs.queries(2, 1000, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)


def test_config1_timed_pipeline_against_golden(ctx):
    """100,000 molecules (seed 1) x 100 queries through the benched entry points."""
    gold = json.loads((GOLDEN / "config1.json").read_text())
    qh = capi.HostGraphs.queries(1, 100, 4, 20, capi.SIZE_NORMAL, perturb_frac=0.2)
    dh = capi.HostGraphs.molecules(1, 100_000)
    assert digest(*qh.arrays().values()) == gold["query_digest"]
    assert digest(*dh.arrays().values()) == gold["data_digest"]
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    pipe = capi.Pipeline(ctx, qd, dd)
    db = capi.Database(ctx, dd, 6)
    batch = capi.Batch(ctx, db)
    for mode, key in ((True, "first"), (False, "all")):
        for name, run, pairs in (("pipeline", lambda: pipe.execute(6, mode), pipe.pairs),
                                 ("batched", lambda: batch.run(qd, find_first=mode), batch.pairs)):
            print(f"config 1 full: {name} find_{key}", flush=True)
            r = run()
            off, idx, cnt = pairs()
            assert digest(off, idx) == gold["gmcr_digest"], name
            assert digest(cnt) == gold[f"counts_{key}_digest"], name
            assert r["num_matches"] == gold[f"matches_{key}"], name
            assert r["gmcr_pairs"] == gold["gmcr_pairs"], name


def test_config2_query_set_against_golden(ctx):
    """Config 2's exact 1,000 queries x the first 20,000 molecules of seed 2, both forms."""
    gold = json.loads((GOLDEN / "config2_subset.json").read_text())
    qh = _config2_queries()
    dh = capi.HostGraphs.molecules(2, gold["molecules"], capi.SIZE_LOGNORMAL)
    assert digest(*qh.arrays().values()) == gold["query_digest"]
    assert digest(*dh.arrays().values()) == gold["data_digest"]
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    pipe = capi.Pipeline(ctx, qd, dd)
    batch = capi.Batch(ctx, capi.Database(ctx, dd, 6))
    for mode, key in ((True, "first"), (False, "all")):
        for name, run, pairs in (("pipeline", lambda: pipe.execute(6, mode), pipe.pairs),
                                 ("batched", lambda: batch.run(qd, find_first=mode), batch.pairs)):
            print(f"config 2 query set x {gold['molecules']} molecules: {name} find_{key}", flush=True)
            r = run()
            off, idx, cnt = pairs()
            assert digest(off, idx) == gold["gmcr_digest"], name
            assert digest(cnt) == gold[f"counts_{key}_digest"], name
            assert r["num_matches"] == gold[f"matches_{key}"], name


def test_config2_full_batched_equals_plain(ctx):
    """1,000,000 molecules (seed 2) x config 2's 1,000 queries: the batched form equals the plain
    form pair by pair in both modes (no CPU oracle at this size)."""
    qh = _config2_queries()
    print("config 2 full: generating 1,000,000 molecules", flush=True)
    dh = capi.HostGraphs.molecules(2, 1_000_000, capi.SIZE_LOGNORMAL)
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    assert qd.num_graphs * dd.num_graphs > 1 << 25  # beyond the deferred-count and one-sync limits
    out = {}
    pipe = capi.Pipeline(ctx, qd, dd)
    for mode in (True, False):
        print(f"config 2 full: plain find_first={mode}", flush=True)
        r = pipe.execute(6, mode)
        out["plain", mode] = (r, *pipe.pairs())
    del pipe  # the 60 GB candidate bitset
    batch = capi.Batch(ctx, capi.Database(ctx, dd, 6))
    for mode in (True, False):
        print(f"config 2 full: batched find_first={mode}", flush=True)
        r = batch.run(qd, find_first=mode)
        assert r["one_sync"] == 0
        out["batched", mode] = (r, *batch.pairs())
    for mode in (True, False):
        pr, po, pi, pc = out["plain", mode]
        br, bo, bi, bc = out["batched", mode]
        np.testing.assert_array_equal(bo, po, err_msg="GMCR offsets")
        np.testing.assert_array_equal(bi, pi, err_msg="GMCR indices")
        np.testing.assert_array_equal(bc, pc, err_msg=f"per-pair counts, find_first={mode}")
        for k in ("num_matches", "matched_pairs", "gmcr_pairs", "dfs_steps"):
            assert pr[k] == br[k], k
    cf, ca = out["plain", True][3], out["plain", False][3]
    np.testing.assert_array_equal(cf, (ca > 0).astype(np.int32))
    assert out["plain", True][0]["gmcr_pairs"] > 1 << 25


def test_config3_full_stream_both_forms(ctx):
    """20,000,000 molecules (seed 3) x 2,000 queries streamed in 20 chunks of 1,000,000 molecules:
    the batched and plain forms give the same per-chunk GMCR and counts (digests) and totals."""
    qh = capi.HostGraphs.queries(3, 2000, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)
    print("config 3 full: generating 20,000,000 molecules", flush=True)
    dh = capi.HostGraphs.molecules(3, 20_000_000, capi.SIZE_LOGNORMAL)
    qd = capi.DeviceGraph(ctx, qh)
    runs = {}
    for form in ("batched", "plain"):
        print(f"config 3 full: {form} stream", flush=True)
        runs[form] = capi.stream_run(ctx, qd, dh, chunk_molecules=1_000_000, form=form, refinement_steps=6, find_first=True,
                                     digest=True)
        print(f"config 3 full: {form} {runs[form]['matched_pairs']} matched pairs, {runs[form]['wall_ms']:.0f} ms", flush=True)
    b, p = runs["batched"], runs["plain"]
    assert b["chunks"] == p["chunks"] == 20
    for k in ("num_matches", "matched_pairs", "gmcr_pairs", "dfs_steps"):
        assert b[k] == p[k], k
    assert b["chunk_digests"] == p["chunk_digests"]
    assert sum(c[2] for c in b["chunk_digests"]) == b["gmcr_pairs"]
    assert b["num_matches"] == b["matched_pairs"]  # find-first: one per matched pair


def test_config4_serving_at_10m(ctx):
    """10,000,000 molecules (seed 4) resident; 300 Poisson requests (50 QPS) of U[1,50] query graphs
    served by two workers; each sampled request equals its result when run alone."""
    print("config 4: generating 10,000,000 molecules", flush=True)
    dh = capi.HostGraphs.molecules(4, 10_000_000, capi.SIZE_LOGNORMAL)
    dd = capi.DeviceGraph(ctx, dh)
    db = capi.Database(ctx, dd, 6)
    ctx2 = capi.Context(0)
    workers = [capi.Batch(ctx, db), capi.Batch(ctx2, db)]
    arr, per = capi.arrivals(4, 300, 50.0, 1, 50)
    rgo = np.concatenate([[0], np.cumsum(per)]).astype(np.uint32)
    qh = capi.HostGraphs.queries(4000, int(rgo[-1]), 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)
    print(f"config 4: serving 300 requests, {int(rgo[-1])} query graphs", flush=True)
    out = capi.serve(workers, qh, rgo, arr, find_first=True, memory_budget=128 * 10**9)
    assert (out["status"] == capi.SERVE_DONE).all()
    assert out["summary"]["done"] == 300
    sample = [0, 1, 2, 3, 50, 149, 250, 299]
    for r in sample:
        one = qh.slice(int(rgo[r]), int(rgo[r + 1] - rgo[r]))
        print(f"config 4: request {r} alone ({one.num_graphs} query graphs)", flush=True)
        workers[0].run(capi.DeviceGraph(ctx, one), True)
        st = workers[0].request_stats([0, one.num_graphs])[0]
        assert (st["matched_pairs"], st["embeddings"], st["hash"]) == (
            int(out["matched_pairs"][r]), int(out["embeddings"][r]), int(out["hash"][r])), r


def test_serving_rejects_workers_sharing_a_context(ctx):
    dh = capi.HostGraphs.molecules(5, 300)
    dd = capi.DeviceGraph(ctx, dh)
    db = capi.Database(ctx, dd, 6)
    a, b = capi.Batch(ctx, db), capi.Batch(ctx, db)
    qh = capi.HostGraphs.queries(5, 4, 4, 10, capi.SIZE_NORMAL)
    with pytest.raises(capi.SigmoError, match="share a context"):
        capi.serve([a, b], qh, [0, 2, 4], [0.0, 0.0])
    with pytest.raises(capi.SigmoError, match="share a context"):
        capi.serve([a, a], qh, [0, 2, 4], [0.0, 0.0])
