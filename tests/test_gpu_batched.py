"""GPU parity of the batched form (sigmo_db_* / sigmo_batch_*) against the CPU oracle.

The batched form never materialises the candidate bitset: it evaluates each (query graph,
molecule) pair from level signatures prepared once per molecule.  Its GMCR arrays and per-pair
join counts must equal the oracle's restatement of the reference driver
(library/src/sigmo.cpp:102-159) bit for bit, in both join modes.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from sigmo_amd import capi
from oracle import oracle as O
from tests._util import GOLDEN, digest

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "cfg0":
        return (capi.HostGraphs.queries(0, 20, 4, 12, capi.SIZE_NORMAL), capi.HostGraphs.molecules(0, 2000), 6)
    if name == "cfg1_slice":
        return (capi.HostGraphs.queries(1, 100, 4, 20, capi.SIZE_NORMAL, perturb_frac=0.2),
                capi.HostGraphs.molecules(1, 5000), 6)
    if name == "lognormal_aliasing":  # molecules of 65..128 nodes use the second position word
        return (capi.HostGraphs.queries(2, 40, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3),
                capi.HostGraphs.molecules(2, 3000, capi.SIZE_LOGNORMAL), 6)
    if name == "reference_fixtures":
        return (capi.HostGraphs.load_dat(GOLDEN / "test_query.dat"), capi.HostGraphs.load_dat(GOLDEN / "test_data.dat"), 3)
    raise KeyError(name)


def _batched(ctx, qh, dh, R, modes=(True, False), db=None, batch=None, tuning=None):
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    db = db or capi.Database(ctx, dd, R)
    batch = batch or capi.Batch(ctx, db)
    if tuning is not None:
        batch.tuning = tuning
    out = {}
    for mode in modes:
        r = batch.run(qd, find_first=mode)
        off, idx, cnt = batch.pairs()
        key = "first" if mode else "all"
        out["offsets"], out["indices"] = off, idx
        out[f"counts_{key}"], out[f"total_{key}"] = cnt, r["num_matches"]
        out[f"result_{key}"] = r
    return out


def _assert_equal(g, o, modes=(True, False)):
    np.testing.assert_array_equal(g["offsets"], o["offsets"], err_msg="GMCR offsets")
    np.testing.assert_array_equal(g["indices"], o["indices"], err_msg="GMCR indices")
    for mode in modes:
        key = "first" if mode else "all"
        np.testing.assert_array_equal(g[f"counts_{key}"], o[f"counts_{key}"], err_msg=f"join counts ({key})")
        assert g[f"total_{key}"] == o[f"total_{key}"]
        r = g[f"result_{key}"]
        assert r["gmcr_pairs"] == len(o["indices"])
        assert r["matched_pairs"] == int((o[f"counts_{key}"] > 0).sum())


@pytest.mark.parametrize("mode", [1, 2], ids=["molecule_mode", "list_mode"])
@pytest.mark.parametrize("name", ["cfg0", "cfg1_slice", "lognormal_aliasing", "reference_fixtures"])
def test_batched_parity(ctx, name, mode):
    qh, dh, R = _case(name)
    print(f"batched case {name} filter_mode {mode}: queries {qh.info()} data {dh.info()} R={R}", flush=True)
    g = _batched(ctx, qh, dh, R, tuning={"filter_mode": mode})
    o = O.pipeline(qh.arrays(), dh.arrays(), R, modes=[True, False])
    _assert_equal(g, o)


@pytest.mark.parametrize("R", [1, 2, 8])
def test_batched_refinement_depths(ctx, R):
    qh, dh = capi.HostGraphs.queries(3, 30, 4, 16, capi.SIZE_NORMAL, perturb_frac=0.1), capi.HostGraphs.molecules(3, 1500)
    print(f"batched R={R}", flush=True)
    _assert_equal(_batched(ctx, qh, dh, R), O.pipeline(qh.arrays(), dh.arrays(), R, modes=[True, False]))


def test_batched_wildcards_overflow_and_replay(ctx):
    """Wildcard queries keep nearly every pair: the record and arena buffers overflow their first
    sizing and the run is replayed with larger ones; labels above 15 use the slow label path."""
    qs = [([255, 255], [(0, 1)])] * 40 + [([255, 8, 255], [(0, 1), (1, 2)])] * 20 + [([20, 8], [(0, 1)])] * 4
    mols = [([8, 0, 7, 8, 8, 9, 8], [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 6), (6, 0)]),
            ([20, 8, 8], [(0, 1), (1, 2)]), ([3], [])]
    extra = capi.HostGraphs.molecules(4, 1500)
    qh = capi.HostGraphs.from_edge_lists(qs)
    dh = capi.HostGraphs.from_edge_lists(mols).concat(extra)
    o = O.pipeline(qh.arrays(), dh.arrays(), 4, modes=[True, False])
    assert len(o["indices"]) > 65536  # beyond the initial record capacity
    for mode in (1, 2):
        print(f"batched wildcard case, filter_mode {mode}", flush=True)
        _assert_equal(_batched(ctx, qh, dh, 4, tuning={"filter_mode": mode}), o)


def test_batched_one_sync_arena_overflow_is_repeated(ctx):
    """A batch sized by one query set, then run on a wildcard-heavy set whose arena demand exceeds
    that size while its estimate does not: the second run takes the one-synchronisation path,
    overflows there, skips the join (no arena offset past the end is read) and repeats sized.
    Sizes: 200 molecules; 10,000 ordinary queries (k in 4..12, ~15% of 2 M pairs survive) size the
    arena at ~3.9 M words; 10,000 three-node wildcard paths keep all 2 M pairs (6 M words)."""
    dh = capi.HostGraphs.molecules(4, 200)
    q1 = capi.HostGraphs.queries(11, 10_000, 4, 12, capi.SIZE_NORMAL)
    q2 = capi.HostGraphs.from_edge_lists([([255, 255, 255], [(0, 1), (1, 2)])] * 10_000)
    dd = capi.DeviceGraph(ctx, dh)
    db = capi.Database(ctx, dd, 6)
    batch = capi.Batch(ctx, db)
    batch.tuning = {"filter_blocks_per_cu": 1}
    print("one-sync overflow: sizing run", flush=True)
    first = batch.run(capi.DeviceGraph(ctx, q1), find_first=True)
    assert first["one_sync"] == 0
    o = O.pipeline(q2.arrays(), dh.arrays(), 6, modes=[True, False])
    print("one-sync overflow: wildcard run", flush=True)
    g = _batched(ctx, q2, dh, 6, db=db, batch=batch)
    r = g["result_first"]
    assert r["one_sync"] == 1 and r["arena_repeats"] >= 1, r
    _assert_equal(g, o)


def test_batched_more_than_1024_query_graphs(ctx):
    """2,000 query graphs (config 3's set, k in 4..32, 30% with a ring) x 1,500 molecules of seed 3:
    the query preparation, the per-class pre-test maxima and both evaluation modes past 1,024
    query graphs (32 bitmask words per molecule) against the oracle."""
    qh = capi.HostGraphs.queries(3, 2000, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)
    dh = capi.HostGraphs.molecules(3, 1500, capi.SIZE_LOGNORMAL)
    o = O.pipeline(qh.arrays(), dh.arrays(), 6, modes=[True, False])
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    db = capi.Database(ctx, dd, 6)
    for mode in (1, 2):
        batch = capi.Batch(ctx, db)
        for rep in range(2):
            print(f"2,000 query graphs x 1,500 molecules: filter_mode {mode}, run {rep}", flush=True)
            _assert_equal(_batched(ctx, qh, dh, 6, db=db, batch=batch, tuning={"filter_mode": mode}), o)


def test_batched_reuse_across_query_sets(ctx):
    """One database and one batch object serve successive query sets, larger then smaller."""
    dh = capi.HostGraphs.molecules(6, 2500)
    dd = capi.DeviceGraph(ctx, dh)
    db = capi.Database(ctx, dd, 6)
    batch = capi.Batch(ctx, db)
    for seed, count in ((6, 60), (7, 9), (8, 25)):
        qh = capi.HostGraphs.queries(seed, count, 4, 20, capi.SIZE_NORMAL, perturb_frac=0.2)
        print(f"batched reuse seed {seed}", flush=True)
        g = _batched(ctx, qh, dh, 6, db=db, batch=batch)
        _assert_equal(g, O.pipeline(qh.arrays(), dh.arrays(), 6, modes=[True, False]))
    assert db.device_bytes > 0


def test_batched_matches_plain_pipeline(ctx):
    qh, dh, R = _case("cfg1_slice")
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    pipe = capi.Pipeline(ctx, qd, dd)
    batch = capi.Batch(ctx, capi.Database(ctx, dd, R))
    for mode in (True, False):
        p = pipe.execute(R, mode)
        b = batch.run(qd, find_first=mode)
        assert (p["num_matches"], p["gmcr_pairs"], p["matched_pairs"]) == (b["num_matches"], b["gmcr_pairs"], b["matched_pairs"])
        assert p["dfs_steps"] == b["dfs_steps"]
        po, pi, pc = pipe.pairs()
        bo, bi, bc = batch.pairs()
        np.testing.assert_array_equal(po, bo)
        np.testing.assert_array_equal(pi, bi)
        np.testing.assert_array_equal(pc, bc)


def test_batched_config1_full_against_golden(ctx):
    """100,000 molecules x 100 queries (seed 1) at full size against the oracle's committed digests."""
    gold = json.loads((GOLDEN / "config1.json").read_text())
    qh = capi.HostGraphs.queries(1, 100, 4, 20, capi.SIZE_NORMAL, perturb_frac=0.2)
    dh = capi.HostGraphs.molecules(1, 100_000)
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    batch = capi.Batch(ctx, capi.Database(ctx, dd, 6))
    first = batch.run(qd, find_first=True)
    off, idx, cf = batch.pairs()
    assert digest(off, idx) == gold["gmcr_digest"]
    assert first["num_matches"] == gold["matches_first"]
    assert digest(cf) == gold["counts_first_digest"]
    allr = batch.run(qd, find_first=False)
    _, _, ca = batch.pairs()
    assert allr["num_matches"] == gold["matches_all"]
    assert digest(ca) == gold["counts_all_digest"]


def test_batched_rejects_unsupported_inputs(ctx):
    dh = capi.HostGraphs.molecules(9, 20)
    dd = capi.DeviceGraph(ctx, dh)
    with pytest.raises(capi.SigmoError):
        capi.Database(ctx, dd, 0)
    with pytest.raises(capi.SigmoError):
        capi.Database(ctx, dd, 9)
    batch = capi.Batch(ctx, capi.Database(ctx, dd, 2))
    single = capi.DeviceGraph(ctx, capi.HostGraphs.from_edge_lists([([8], [])]))
    with pytest.raises(capi.SigmoError):
        batch.run(single)


def test_batched_node_with_every_label_at_one_level(ctx):
    """A query node whose level-1 sphere holds all 16 labels has 16 requirements at R=1: its token
    segment is full, the reader's lookahead ends exactly at the spare word."""
    star = (list(range(16)) + [3], [(16, i) for i in range(16)])
    path = ([3, 7, 3], [(0, 1), (1, 2)])
    mol_star = (list(range(16)) + [3, 9], [(16, i) for i in range(16)] + [(16, 17)])
    qh = capi.HostGraphs.from_edge_lists([path, star])  # the star is the last query graph
    dh = capi.HostGraphs.from_edge_lists([mol_star, path, mol_star]).concat(capi.HostGraphs.molecules(10, 200))
    for R in (1, 2):
        print(f"batched full segment R={R}", flush=True)
        g = _batched(ctx, qh, dh, R)
        o = O.pipeline(qh.arrays(), dh.arrays(), R, modes=[True, False])
        _assert_equal(g, o)
        assert o["counts_all"].sum() > 0


@pytest.mark.parametrize("name", ["cfg1_slice", "lognormal_aliasing"])
def test_batched_graph_replay_equals_eager(ctx, name):
    """One captured hipGraph per mode, replayed three times with other eager work and
    synchronisations between the replays; every replay equals the eager run and the oracle."""
    qh, dh, R = _case(name)
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    db = capi.Database(ctx, dd, R)
    batch = capi.Batch(ctx, db)
    other = capi.Batch(ctx, db)
    other_q = capi.DeviceGraph(ctx, capi.HostGraphs.queries(11, 15, 4, 12, capi.SIZE_NORMAL))
    o = O.pipeline(qh.arrays(), dh.arrays(), R, modes=[True, False])
    for mode in (True, False):
        key = "first" if mode else "all"
        eager = batch.run(qd, find_first=mode)
        e_off, e_idx, e_cnt = batch.pairs()
        print(f"graph capture {name} find_first={mode}", flush=True)
        batch.capture(qd, find_first=mode, refresh_db=True)
        for rep in range(3):
            other.run(other_q, find_first=not mode)  # eager work on the same stream between replays
            ctx.sync()
            r = batch.replay()
            off, idx, cnt = batch.pairs()
            np.testing.assert_array_equal(off, e_off)
            np.testing.assert_array_equal(idx, e_idx)
            np.testing.assert_array_equal(cnt, e_cnt, err_msg=f"replay {rep}")
            assert (r["num_matches"], r["gmcr_pairs"], r["matched_pairs"]) == (eager["num_matches"], eager["gmcr_pairs"],
                                                                               eager["matched_pairs"])
        np.testing.assert_array_equal(cnt, o[f"counts_{key}"])


def test_batched_graph_dropped_when_buffers_grow(ctx):
    dh = capi.HostGraphs.molecules(12, 800)
    dd = capi.DeviceGraph(ctx, dh)
    batch = capi.Batch(ctx, capi.Database(ctx, dd, 4))
    small = capi.DeviceGraph(ctx, capi.HostGraphs.queries(12, 5, 4, 8, capi.SIZE_NORMAL))
    big = capi.DeviceGraph(ctx, capi.HostGraphs.queries(13, 120, 4, 16, capi.SIZE_NORMAL))
    print("graph invalidation", flush=True)
    batch.capture(small)
    batch.replay()
    batch.run(big)  # grows the per-pair buffers: the graph's addresses are gone
    with pytest.raises(capi.SigmoError):
        batch.replay()


def test_batched_tuning_and_autotune_keep_results(ctx):
    qh, dh, R = _case("lognormal_aliasing")
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    batch = capi.Batch(ctx, capi.Database(ctx, dd, R))
    base = batch.run(qd, find_first=False)
    ref = batch.pairs()
    for t in ({"filter_blocks_per_cu": 1, "filter_claim": 1, "join_blocks_per_cu": 1, "join_chunk": 1},
              {"filter_blocks_per_cu": 8, "filter_claim": 64, "join_blocks_per_cu": 4, "join_chunk": 4096},
              {"filter_mode": 1}, {"filter_mode": 2}, {"filter_mode": 1, "filter_claim": 64},
              {"filter_mode": 2, "filter_claim": 1, "filter_blocks_per_cu": 1}):
        print(f"tuning {t}", flush=True)
        batch.tuning = t
        r = batch.run(qd, find_first=False)
        for a, b in zip(batch.pairs(), ref):
            np.testing.assert_array_equal(a, b)
        assert r["num_matches"] == base["num_matches"]
    with pytest.raises(capi.SigmoError):
        batch.tuning = {"join_blocks_per_cu": 5}
    with pytest.raises(capi.SigmoError):
        batch.tuning = {"filter_mode": 3}
    batch.tuning = {}
    chosen, ms = batch.autotune(qd, find_first=False)
    assert ms > 0 and batch.tuning == chosen
    batch.run(qd, find_first=False)
    for a, b in zip(batch.pairs(), ref):
        np.testing.assert_array_equal(a, b)
