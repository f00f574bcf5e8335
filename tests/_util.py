"""Shared helpers for the test suite: workload builders and GPU/oracle phase runners."""
from __future__ import annotations

import hashlib
from pathlib import Path

import numpy as np

from sigmo_amd import capi
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def witness_graphs():
    """Hand-traced witnesses of the reference DFS's stale-state quirks (isomorphism.hpp:319-371):
    (queries, molecules, expected find-all per pair)."""
    # 1. stale sibling in visited: molecule path 0-2-1, query star centred on q0.
    mol1 = ([8, 8, 8], [(0, 2), (1, 2)])
    q1 = ([8, 8, 8], [(0, 1), (0, 2)])
    # 2. stale unset: molecule path 0-2-3 plus isolated node 1; query path q0-q2-q1.
    mol2 = ([8, 8, 8, 8], [(0, 2), (2, 3)])
    q2 = ([8, 8, 8], [(0, 2), (1, 2)])
    return (q1, mol1, 1), (q2, mol2, 4)


def gpu_phases(ctx, qh: capi.HostGraphs, dh: capi.HostGraphs, R: int, modes=(True, False)) -> dict:
    """Run every phase through the C ABI and download each intermediate result."""
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    Q, D = qd.total_nodes, dd.total_nodes
    sig = capi.Signature(ctx, D, Q)
    cand = capi.Candidates(ctx, Q, D)
    out = {"levels": []}
    sig.generate_data(dd)
    sig.generate_query(qd)
    out["levels"].append((sig.download(capi.SCOPE_DATA), sig.download(capi.SCOPE_QUERY)))
    capi.filter_candidates(ctx, qd, dd, sig, cand)
    out["filtered"] = cand.download()
    for r in range(1, R + 1):
        sig.refine_data(dd, r)
        sig.refine_query(qd, r)
        capi.refine_candidates(ctx, qd, dd, sig, cand)
        out["levels"].append((sig.download(capi.SCOPE_DATA), sig.download(capi.SCOPE_QUERY)))
    out["words"] = cand.download()
    out["row_counts"] = cand.counts()
    gm = capi.GMCR(ctx)
    gm.generate(qd, dd, cand)
    out["offsets"], out["indices"] = gm.download()
    for mode in modes:
        total, counts = capi.join_candidates(ctx, qd, dd, cand, gm, find_first=mode)
        key = "first" if mode else "all"
        out[f"total_{key}"] = total
        out[f"counts_{key}"] = counts
    return out


def assert_phases_equal(g: dict, o: dict, modes=(True, False)) -> None:
    assert len(g["levels"]) == len(o["levels"])
    for r, ((gd, gq), (od, oq)) in enumerate(zip(g["levels"], o["levels"])):
        np.testing.assert_array_equal(gd, od, err_msg=f"data signatures, level {r}")
        np.testing.assert_array_equal(gq, oq, err_msg=f"query signatures, level {r}")
    np.testing.assert_array_equal(g["filtered"], o["filtered"], err_msg="filtered candidates")
    np.testing.assert_array_equal(g["words"], o["words"], err_msg="refined candidates")
    np.testing.assert_array_equal(g["offsets"], o["offsets"], err_msg="GMCR offsets")
    np.testing.assert_array_equal(g["indices"], o["indices"], err_msg="GMCR indices")
    for mode in modes:
        key = "first" if mode else "all"
        np.testing.assert_array_equal(g[f"counts_{key}"], o[f"counts_{key}"], err_msg=f"join counts ({key})")
        assert g[f"total_{key}"] == o[f"total_{key}"]


def oracle_phases(qh: capi.HostGraphs, dh: capi.HostGraphs, R: int, modes=(True, False)) -> dict:
    return O.pipeline(qh.arrays(), dh.arrays(), R, keep_levels=True, modes=list(modes))


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 output function on uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def request_stats(offsets, indices, counts, request_graph_offsets) -> list[tuple[int, int, int]]:
    """Per-request (matched pairs, embeddings, hash) of a GMCR + counts result, the statistics
    sigmo_batch_request_stats reports: request r owns query graphs [rgo[r], rgo[r+1]), and a pair
    (molecule m, query graph g, count c) adds mix64(mix64(((g - rgo[r]) << 32) | m) + c)."""
    rgo = np.asarray(request_graph_offsets, np.int64)
    mol = np.repeat(np.arange(len(offsets) - 1, dtype=np.uint64), np.diff(np.asarray(offsets, np.int64)))
    g = np.asarray(indices, np.int64)
    c = np.asarray(counts, np.int64)
    req = np.searchsorted(rgo, g, side="right") - 1
    with np.errstate(over="ignore"):
        key = ((g - rgo[req]).astype(np.uint64) << np.uint64(32)) | mol
        h = _mix64(_mix64(key) + c.astype(np.uint32).astype(np.uint64))
    out = []
    for r in range(len(rgo) - 1):
        m = req == r
        with np.errstate(over="ignore"):
            hs = int(np.sum(h[m], dtype=np.uint64)) if m.any() else 0
        out.append((int((c[m] > 0).sum()), int(c[m].sum()), hs))
    return out
