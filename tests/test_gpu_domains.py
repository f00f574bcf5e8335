"""GPU parity of the alternative domains: ViewBased signature refinement
(sigmo_signatures_refine_view_based, signature.hpp:298-334) and the Data-domain filter
(sigmo_candidates_filter_data, isomorphism.hpp:49-52), against the pinned oracle."""
from __future__ import annotations

import numpy as np
import pytest

from sigmo_amd import capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size", [capi.SIZE_NORMAL, capi.SIZE_LOGNORMAL])
def test_view_based_refinement_parity(ctx, size):
    qh = capi.HostGraphs.queries(5, 25, 4, 20, size)
    dh = capi.HostGraphs.molecules(5, 1500, size)
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    q, d = qh.arrays(), dh.arrays()
    sig = capi.Signature(ctx, dd.total_nodes, qd.total_nodes)
    sig.generate_data(dd)
    sig.generate_query(qd)
    dsig, qsig = O.generate_signatures(d), O.generate_signatures(q)
    words = O.filter_candidates(q, d)
    cand = capi.Candidates(ctx, qd.total_nodes, dd.total_nodes)
    capi.filter_candidates(ctx, qd, dd, sig, cand)
    for view in (1, 2, 3):  # the driver passes view_size = refinement step
        print(f"view-based refine size={size} view={view}", flush=True)
        sig.refine_view_based(dd, capi.SCOPE_DATA, view)
        sig.refine_view_based(qd, capi.SCOPE_QUERY, view)
        O.refine_view_based(d, view, dsig)
        O.refine_view_based(q, view, qsig)
        np.testing.assert_array_equal(sig.download(capi.SCOPE_DATA), dsig)
        np.testing.assert_array_equal(sig.download(capi.SCOPE_QUERY), qsig)
        # the bit-sliced data planes follow the signatures: a candidate refine over them
        # equals the oracle's refine over the downloaded signatures
        capi.refine_candidates(ctx, qd, dd, sig, cand)
        O.refine_candidates(qsig, dsig, words)
        np.testing.assert_array_equal(cand.download(), words)


def test_filter_data_domain_parity(ctx):
    qs = [([255, 8], [(0, 1)])] + [([7, 8, 0], [(0, 1), (1, 2)])] * 3
    qh = capi.HostGraphs.from_edge_lists(qs).concat(capi.HostGraphs.queries(6, 40, 4, 12, capi.SIZE_NORMAL))
    dh = capi.HostGraphs.molecules(6, 700)
    qd, dd = capi.DeviceGraph(ctx, qh), capi.DeviceGraph(ctx, dh)
    cand = capi.Candidates(ctx, dd.total_nodes, qd.total_nodes)  # rows = data nodes
    print("data-domain filter", flush=True)
    capi.filter_candidates_data(ctx, qd, dd, cand)
    np.testing.assert_array_equal(cand.download(), O.filter_candidates_data(qh.arrays(), dh.arrays()))
    wrong = capi.Candidates(ctx, qd.total_nodes, dd.total_nodes)
    with pytest.raises(capi.SigmoError):
        capi.filter_candidates_data(ctx, qd, dd, wrong)
