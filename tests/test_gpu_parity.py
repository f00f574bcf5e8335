"""GPU parity: every phase of the HIP path against the CPU oracle, bit for bit.

Each case runs signatures (every level), filter, refine, GMCR and the join (find-first and
find-all) through the C ABI and compares each downloaded intermediate with the oracle's
restatement of the reference kernels on the same seeded input.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from sigmo_amd import capi
from oracle import oracle as O
from tests._util import GOLDEN, assert_phases_equal, digest, gpu_phases, oracle_phases, witness_graphs

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "cfg0":  # 2,000 molecules (seed 0), 20 queries k in U[4,12]
        return (capi.HostGraphs.queries(0, 20, 4, 12, capi.SIZE_NORMAL), capi.HostGraphs.molecules(0, 2000), 6)
    if name == "cfg1_slice":  # the 100,000-molecule workload's law on 5,000 molecules, 100 queries, 20% perturbed
        return (capi.HostGraphs.queries(1, 100, 4, 20, capi.SIZE_NORMAL, perturb_frac=0.2),
                capi.HostGraphs.molecules(1, 5000), 6)
    if name == "lognormal_aliasing":  # molecules up to 128 nodes (positions alias mod 64), k up to 32
        return (capi.HostGraphs.queries(2, 40, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3),
                capi.HostGraphs.molecules(2, 3000, capi.SIZE_LOGNORMAL), 6)
    if name == "reference_fixtures":
        return (capi.HostGraphs.load_dat(GOLDEN / "test_query.dat"), capi.HostGraphs.load_dat(GOLDEN / "test_data.dat"), 3)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["cfg0", "cfg1_slice", "lognormal_aliasing", "reference_fixtures"])
def test_phase_parity(ctx, name):
    qh, dh, R = _case(name)
    print(f"case {name}: queries {qh.info()} data {dh.info()} R={R}", flush=True)
    g = gpu_phases(ctx, qh, dh, R)
    o = oracle_phases(qh, dh, R)
    assert_phases_equal(g, o)
    assert g["total_first"] == int((g["counts_all"] > 0).sum())


def test_aliasing_case_has_large_molecules():
    _, dh, _ = _case("lognormal_aliasing")
    n = np.diff(dh.arrays()["graph_offsets"].astype(np.int64))
    assert n.max() > 64  # the case really exercises the mod-64 positions


@pytest.mark.parametrize("idx", [0, 1])
def test_dfs_quirk_witnesses(ctx, idx):
    q, mol, expected = witness_graphs()[idx]
    qh, dh = capi.HostGraphs.from_edge_lists([q]), capi.HostGraphs.from_edge_lists([mol])
    g = gpu_phases(ctx, qh, dh, 0)
    o = oracle_phases(qh, dh, 0)
    assert_phases_equal(g, o)
    assert g["total_all"] == expected


def test_odd_reference_stride_and_wildcard(ctx):
    # 3 molecules of 7 nodes: D = 21, one uint32 word per row (odd stride); one wildcard query node
    mols = [([8, 0, 7, 8, 8, 9, 8], [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 6), (6, 0)])] * 3
    qs = [([8, 255, 7], [(0, 1), (1, 2)]), ([8, 8], [(0, 1)])]
    qh, dh = capi.HostGraphs.from_edge_lists(qs), capi.HostGraphs.from_edge_lists(mols)
    g = gpu_phases(ctx, qh, dh, 2)
    o = oracle_phases(qh, dh, 2)
    assert_phases_equal(g, o)
    assert g["filtered"].size == 5  # 5 query rows x ceil(21/32) words


def test_repeated_view_size_accumulates_reachables(ctx):
    # refine(view_size=1) twice: the second call sees every neighbour in the old reachables
    qh, dh = capi.HostGraphs.queries(5, 4, 4, 8), capi.HostGraphs.molecules(5, 50)
    qa, da = qh.arrays(), dh.arrays()
    dd = capi.DeviceGraph(ctx, dh)
    sig = capi.Signature(ctx, dd.total_nodes, qh.total_nodes)
    D = dd.total_nodes
    osig, oreach = np.zeros(D, np.uint64), np.zeros(D, np.uint64)
    for view in (1, 1, 3, 2):
        sig.refine_data(dd, view)
        O.refine_signatures(da, view, osig, oreach)
        np.testing.assert_array_equal(sig.download(capi.SCOPE_DATA), osig, err_msg=f"view {view}")
    del qa


def test_generate_accumulates(ctx):
    dh = capi.HostGraphs.molecules(7, 30)
    da = dh.arrays()
    dd = capi.DeviceGraph(ctx, dh)
    sig = capi.Signature(ctx, dd.total_nodes, 1)
    osig = np.zeros(dd.total_nodes, np.uint64)
    for _ in range(3):  # counts keep saturating at 15 as in signature.hpp:44-51
        sig.generate_data(dd)
        O.generate_signatures(da, osig)
        np.testing.assert_array_equal(sig.download(capi.SCOPE_DATA), osig)


def test_candidate_counts_match_reference_rows(ctx):
    qh, dh, R = _case("cfg0")
    g = gpu_phases(ctx, qh, dh, R, modes=(True,))
    Q, D = qh.total_nodes, dh.total_nodes
    np.testing.assert_array_equal(g["row_counts"], O.candidate_counts(Q, D, g["words"]))


def test_pipeline_entry_point_matches_phases(ctx):
    qh, dh, R = _case("cfg0")
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    o = O.pipeline(qh.arrays(), dh.arrays(), R, modes=[True, False])
    r1 = capi.pipeline_run(ctx, qd, dd, R, find_first=True)
    r2 = capi.pipeline_run(ctx, qd, dd, R, find_first=False)
    assert r1["num_matches"] == o["total_first"] == r1["matched_pairs"]
    assert r2["num_matches"] == o["total_all"]
    assert r1["gmcr_pairs"] == len(o["indices"])
    assert r1["total_ms"] > 0


@pytest.mark.parametrize("name", ["cfg1_slice", "lognormal_aliasing"])
@pytest.mark.parametrize("R", [0, 1, 2])
def test_pipeline_few_refinements(ctx, name, R):
    """The pipeline runs the filter fused with its first candidate refinement (after that level's
    signatures): with R = 1 that fused launch is the only refinement, with R = 0 the filter runs
    alone.  GMCR pairs and per-pair counts equal the oracle's, which filters and refines apart."""
    qh, dh, _ = _case(name)
    print(f"pipeline {name} R={R}", flush=True)
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    pipe = capi.Pipeline(ctx, qd, dd)
    o = O.pipeline(qh.arrays(), dh.arrays(), R, modes=[True, False])
    for mode in (True, False):
        key = "first" if mode else "all"
        r = pipe.execute(R, mode)
        off, idx, cnt = pipe.pairs()
        np.testing.assert_array_equal(off, o["offsets"])
        np.testing.assert_array_equal(idx, o["indices"])
        np.testing.assert_array_equal(cnt, o[f"counts_{key}"])
        assert r["num_matches"] == o[f"total_{key}"]


def test_phase_and_kernel_timing(ctx):
    """Phase and per-kernel times of executes: the pipeline's phases share their boundary events,
    so the phases tile the execute (their sum stays within its total), every launch is counted
    once and every kernel time is positive."""
    qh, dh, R = _case("cfg1_slice")
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    pipe = capi.Pipeline(ctx, qd, dd)
    pipe.execute(R, True)
    ctx.sync()
    ctx.reset_phases()
    runs = [pipe.execute(R, True) for _ in range(3)]
    ctx.sync()
    phases = ctx.phase_ms()
    ks = ctx.kernel_stats()
    total = sum(r["total_ms"] for r in runs)
    assert all(v >= 0 for v in phases.values())
    assert 0.5 * total <= sum(phases.values()) <= 1.02 * total + 0.05
    n = len(runs)
    # the first candidate refinement runs fused into the filter's launch
    expect = {"sig_generate": n, "sig_refine": n * R, "filter": n, "refine": n * (R - 1), "gmcr": n, "join": n}
    for name, launches in expect.items():
        got, ms, by = ks[name]
        assert got == launches, name
        assert ms > 0, name
    for r in runs:
        assert sum(r["phase_ms"].values()) <= 1.02 * r["total_ms"] + 0.05


def test_config1_full_against_golden(ctx):
    """100,000 molecules (seed 1) x 100 queries at full size against the oracle's committed digest."""
    path = GOLDEN / "config1.json"
    gold = json.loads(path.read_text())
    qh = capi.HostGraphs.queries(1, 100, 4, 20, capi.SIZE_NORMAL, perturb_frac=0.2)
    dh = capi.HostGraphs.molecules(1, 100_000)
    assert digest(*qh.arrays().values()) == gold["query_digest"]
    assert digest(*dh.arrays().values()) == gold["data_digest"]
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    sig = capi.Signature(ctx, dd.total_nodes, qd.total_nodes)
    cand = capi.Candidates(ctx, qd.total_nodes, dd.total_nodes)
    sig.generate_data(dd)
    sig.generate_query(qd)
    capi.filter_candidates(ctx, qd, dd, sig, cand)
    for r in range(1, 7):
        sig.refine_data(dd, r)
        sig.refine_query(qd, r)
        capi.refine_candidates(ctx, qd, dd, sig, cand)
    assert digest(cand.counts()) == gold["row_counts_digest"]
    assert digest(cand.download()) == gold["words_digest"]
    gm = capi.GMCR(ctx)
    gm.generate(qd, dd, cand)
    off, idx = gm.download()
    assert digest(off, idx) == gold["gmcr_digest"]
    first, cf = capi.join_candidates(ctx, qd, dd, cand, gm, find_first=True)
    allc, ca = capi.join_candidates(ctx, qd, dd, cand, gm, find_first=False)
    assert first == gold["matches_first"]
    assert allc == gold["matches_all"]
    assert digest(cf) == gold["counts_first_digest"]
    assert digest(ca) == gold["counts_all_digest"]
    # size-independent property: find-first marks exactly the pairs find-all counts
    np.testing.assert_array_equal(cf, (ca > 0).astype(np.int32))


@pytest.mark.parametrize("views", [(1, 2, 3, 4, 5, 6), (1, 2, 4, 5, 5, 6, 2, 7), (3, 4, 1, 5), (2, 2, 3)])
def test_refine_view_sequences(ctx, views):
    """Sequential views take the one-layer path, jumps and repeats the full BFS; both must
    equal the reference restatement call by call (molecules up to 128 nodes)."""
    dh = capi.HostGraphs.molecules(8, 400, capi.SIZE_LOGNORMAL)
    da = dh.arrays()
    dd = capi.DeviceGraph(ctx, dh)
    sig = capi.Signature(ctx, dd.total_nodes, 1)
    osig, oreach = np.zeros(dd.total_nodes, np.uint64), np.zeros(dd.total_nodes, np.uint64)
    for view in views:
        sig.refine_data(dd, view)
        O.refine_signatures(da, view, osig, oreach)
        np.testing.assert_array_equal(sig.download(capi.SCOPE_DATA), osig, err_msg=f"views {views} at {view}")


def test_refine_with_two_graphs_of_equal_size(ctx):
    """Reachables are per node id, not per graph: refining graph B after graph A combines
    them exactly as the reference does (the one-layer shortcut must not be taken)."""
    a = capi.HostGraphs.molecules(9, 200)
    n = a.total_nodes
    # graph B: same node count, every molecule's edges reversed in id order
    arr = a.arrays()
    go = arr["graph_offsets"]
    graphs = []
    for g in range(len(go) - 1):
        lo, hi = int(go[g]), int(go[g + 1])
        m = hi - lo
        edges = set()
        for v in range(lo, hi):
            for w in arr["column_indices"][arr["row_offsets"][v]:arr["row_offsets"][v + 1]]:
                u1, u2 = m - 1 - (v - lo), m - 1 - (int(w) - lo)
                edges.add((min(u1, u2), max(u1, u2)))
        graphs.append((arr["node_labels"][lo:hi].tolist(), sorted(edges)))
    b = capi.HostGraphs.from_edge_lists(graphs)
    assert b.total_nodes == n
    da, db = a.arrays(), b.arrays()
    ga, gb = capi.DeviceGraph(ctx, a), capi.DeviceGraph(ctx, b)
    sig = capi.Signature(ctx, n, 1)
    osig, oreach = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    for graph, arrays, view in ((ga, da, 1), (ga, da, 2), (gb, db, 3), (gb, db, 4), (ga, da, 5)):
        sig.refine_data(graph, view)
        O.refine_signatures(arrays, view, osig, oreach)
        np.testing.assert_array_equal(sig.download(capi.SCOPE_DATA), osig, err_msg=f"view {view}")


@pytest.mark.parametrize("name", ["cfg1_slice", "lognormal_aliasing"])
def test_pipeline_graph_replay_equals_execute(ctx, name):
    """The plain pipeline captured as one hipGraph: three replays per mode, with other eager
    work and synchronisations between them, each equal to the eager execute and the oracle."""
    qh, dh, R = _case(name)
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    pipe = capi.Pipeline(ctx, qd, dd)
    other_q = capi.DeviceGraph(ctx, capi.HostGraphs.queries(21, 12, 4, 10, capi.SIZE_NORMAL))
    other = capi.Pipeline(ctx, other_q, dd)
    o = O.pipeline(qh.arrays(), dh.arrays(), R, modes=[True, False])
    for mode in (True, False):
        key = "first" if mode else "all"
        eager = pipe.execute(R, mode)
        e_off, e_idx, e_cnt = pipe.pairs()
        print(f"pipeline capture {name} find_first={mode}", flush=True)
        pipe.capture(R, mode)
        for rep in range(3):
            other.execute(2, not mode)  # eager work on the same stream between replays
            ctx.sync()
            r = pipe.replay()
            off, idx, cnt = pipe.pairs()
            np.testing.assert_array_equal(off, e_off)
            np.testing.assert_array_equal(idx, e_idx)
            np.testing.assert_array_equal(cnt, e_cnt, err_msg=f"replay {rep}")
            assert (r["num_matches"], r["gmcr_pairs"], r["matched_pairs"], r["dfs_steps"]) == (
                eager["num_matches"], eager["gmcr_pairs"], eager["matched_pairs"], eager["dfs_steps"])
        np.testing.assert_array_equal(cnt, o[f"counts_{key}"])
