"""GPU parity of the wildcard/edge-label join (sigmo_join_wildcard, the reference's
joinWildcardCandidates, library/include/isomorphism.hpp:383-556) against the oracle's pinned
restatement: per-pair counts and totals, both modes, on candidates produced by the GPU phases."""
from __future__ import annotations

import numpy as np
import pytest

from sigmo_amd import capi
from oracle import oracle as O
from tests._util import GOLDEN
from tests.test_oracle_wildcard import batch_el, random_case

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "cfg0":
        return capi.HostGraphs.queries(0, 20, 4, 12, capi.SIZE_NORMAL), capi.HostGraphs.molecules(0, 2000), 6
    if name == "lognormal_wide":  # molecules of 65..128 nodes: two-word positions, mod-64 visited
        return (capi.HostGraphs.queries(2, 30, 4, 24, capi.SIZE_LOGNORMAL, ring_frac=0.3),
                capi.HostGraphs.molecules(2, 1500, capi.SIZE_LOGNORMAL), 4)
    if name == "reference_fixtures":
        return capi.HostGraphs.load_dat(GOLDEN / "test_query.dat"), capi.HostGraphs.load_dat(GOLDEN / "test_data.dat"), 3
    if name == "edge_labels_small":
        queries, mols = random_case(7, n_mol=60)
        q, d = batch_el(queries), batch_el(mols)
        return (capi.HostGraphs.from_arrays(q["graph_offsets"], q["row_offsets"], q["column_indices"], q["node_labels"], q["edge_labels"]),
                capi.HostGraphs.from_arrays(d["graph_offsets"], d["row_offsets"], d["column_indices"], d["node_labels"], d["edge_labels"]), 0)
    raise KeyError(name)


def _gpu_wildcard(ctx, qh, dh, R, find_first):
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    sig = capi.Signature(ctx, dd.total_nodes, qd.total_nodes)
    cand = capi.Candidates(ctx, qd.total_nodes, dd.total_nodes)
    sig.generate_data(dd)
    sig.generate_query(qd)
    capi.filter_candidates(ctx, qd, dd, sig, cand)
    for r in range(1, R + 1):
        sig.refine_data(dd, r)
        sig.refine_query(qd, r)
        capi.refine_candidates(ctx, qd, dd, sig, cand)
    gm = capi.GMCR(ctx)
    gm.generate(qd, dd, cand)
    off, idx = gm.download()
    total, counts = capi.join_wildcard(ctx, qd, dd, cand, gm, find_first)
    return cand.download(), off, idx, total, counts


@pytest.mark.parametrize("name", ["cfg0", "lognormal_wide", "reference_fixtures", "edge_labels_small"])
def test_wildcard_join_parity(ctx, name):
    qh, dh, R = _case(name)
    q, d = qh.arrays(), dh.arrays()
    for find_first in (True, False):
        print(f"wildcard join {name} find_first={find_first}", flush=True)
        words, off, idx, total, counts = _gpu_wildcard(ctx, qh, dh, R, find_first)
        o_off, o_idx = O.gmcr(q, d, words)
        np.testing.assert_array_equal(off, o_off)
        np.testing.assert_array_equal(idx, o_idx)
        o_total, o_counts = O.join_wildcard(q, d, words, o_off, o_idx, find_first)
        np.testing.assert_array_equal(counts, o_counts, err_msg=f"{name} find_first={find_first}")
        assert total == o_total
        assert (counts > 0).sum() > 0
