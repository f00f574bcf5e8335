"""GPU test of sigmo_join_noalloc (the join with its counters and per-pair counts in the library's
static device scratch instead of allocated buffers): its totals and per-pair counts equal
sigmo_join's on the same GMCR, in both modes, on two calls in a row -- the second with a smaller
input, so whatever the first left in the scratch must not leak into it -- and a GMCR beyond the
scratch's capacity is rejected before any launch."""
from __future__ import annotations

import numpy as np
import pytest

from sigmo_amd import capi

pytestmark = pytest.mark.gpu


def _phases(ctx, qh, dh, R=6):
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
    return qd, dd, cand, gm


def test_join_noalloc_equals_join_on_two_calls(ctx):
    cases = [
        # 3,000 LogNormal-size molecules (some of 65..128 nodes: the wide-pair list is used)
        ("larger", capi.HostGraphs.queries(7, 40, 4, 24, capi.SIZE_LOGNORMAL, ring_frac=0.3),
         capi.HostGraphs.molecules(7, 3000, capi.SIZE_LOGNORMAL)),
        # then a smaller input: fewer pairs than the first call left in the scratch
        ("smaller", capi.HostGraphs.queries(8, 6, 4, 10, capi.SIZE_NORMAL), capi.HostGraphs.molecules(8, 400)),
    ]
    sizes = []
    for name, qh, dh in cases:
        qd, dd, cand, gm = _phases(ctx, qh, dh)
        sizes.append(gm.total)
        for find_first in (True, False):
            print(f"join_noalloc {name}: {gm.total} pairs, find_first={find_first}", flush=True)
            t0, c0 = capi.join_candidates(ctx, qd, dd, cand, gm, find_first=find_first)
            t1, c1 = capi.join_candidates_noalloc(ctx, qd, dd, cand, gm, find_first=find_first)
            assert t1 == t0
            np.testing.assert_array_equal(c1, c0)
    assert 0 < sizes[1] < sizes[0]
    assert max(dh.arrays()["graph_offsets"][1:] - dh.arrays()["graph_offsets"][:-1]) <= 64  # second input: no wide list


def test_join_noalloc_rejects_more_pairs_than_its_scratch(ctx):
    # every one of 200 molecules survives each of 11,000 two-node wildcard queries: 2.2 M pairs
    qh = capi.HostGraphs.from_edge_lists([([255, 255], [(0, 1)])] * 11_000)
    dh = capi.HostGraphs.molecules(9, 200)
    qd, dd, cand, gm = _phases(ctx, qh, dh, R=1)
    assert gm.total > capi.JOIN_NOALLOC_MAX_PAIRS
    print(f"join_noalloc over capacity: {gm.total} pairs", flush=True)
    with pytest.raises(capi.SigmoError, match="at most"):
        capi.join_candidates_noalloc(ctx, qd, dd, cand, gm, find_first=True)
