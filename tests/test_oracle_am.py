"""Pinning the oracle's adjacency-matrix (AM) query-graph path: am_build (toAMGraph and
createDeviceAMGraph, library/include/graph.hpp:208-214, 345-392), am_generate
(generateAMSignatures, signature.hpp:100-120), am_refine_power_graph
(refineAMSignatures<PowerGraph>, signature.hpp:250-296) and am_refine_view_based
(refineAMSignatures<ViewBased>, signature.hpp:214-248).

* Layout: the row widths the reference's float square root gives (utils.hpp:25-28), tabulated by
  hand for the sizes where they differ from n, and the reference fixture's first graph's bits.
* Generation and ViewBased refinement: on graphs of degree <= 4 the AM neighbour lists are the
  CSR ones, so both equal the CSR restatements (themselves pinned by closed forms in
  test_oracle_pinning.py / test_oracle_domains.py); a degree-5 star shows the fifth neighbour
  dropped (MAX_NEIGHBORS = 4, types.hpp:28).
* PowerGraph: the label histogram of networkx's BFS ball of radius view_size minus the node
  itself (saturating at 15); a planted mistake (the sphere instead of the ball) fails; repeated
  calls give the same result (each call rebuilds the signature); a hand-traced 40-node path shows
  positions >= 32 aliasing mod 32 through Bitset<uint32_t>.
"""
from __future__ import annotations

import networkx as nx
import numpy as np
import pytest

from oracle import oracle as O
from sigmo_amd import capi
from tests._util import GOLDEN
from tests.test_oracle_pinning import batch_from_graphs


def sig_matrix(sigs):
    return np.array([[(int(s) >> (4 * l)) & 15 for l in range(16)] for s in sigs], np.int64)


def queries(seed=21, count=60):
    return capi.HostGraphs.queries(seed, count, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3).arrays()


def test_am_row_width_and_words_by_hand():
    # (n, words, W): words = ceil(n^2 / 64); W = floor(sqrt(64 * words)), hand-computed
    table = [(1, 1, 8), (3, 1, 8), (5, 1, 8), (8, 1, 8), (9, 2, 11), (10, 2, 11), (11, 2, 11), (12, 3, 13),
             (16, 4, 16), (20, 7, 21), (31, 16, 32), (32, 16, 32), (33, 18, 33), (40, 25, 40), (64, 64, 64)]
    for n, words, W in table:
        g = batch_from_graphs([([0] * n, [(0, n - 1)] if n > 1 else [])])
        am = O.am_build(g)
        assert len(am["adjacency"]) == words, n
        if n > 1:  # the edge (0, n-1) sits at bits 0*W + (n-1) and (n-1)*W + 0
            bits = {i for i in range(64 * words) if (int(am["adjacency"][i // 64]) >> (i % 64)) & 1}
            assert bits == {n - 1, (n - 1) * W}, (n, bits)


def test_am_layout_of_the_reference_query_fixture():
    qh = capi.HostGraphs.load_dat(GOLDEN / "test_query.dat")
    am = O.am_build(qh.arrays())
    # test_query.dat: 4 graphs of 3, 4, 5, 4 nodes, one 64-bit word each (W = 8)
    assert am["G"] == 4 and am["N"] == 16
    assert am["num_nodes"].tolist() == [3, 7, 12, 16]
    assert am["graph_offsets"].tolist() == [0, 1, 2, 3]
    # graph 0 ("e#2 0 1 1 0 2 1"): edges 0-1, 0-2 -> bits 0*8+1, 1*8+0, 0*8+2, 2*8+0
    assert int(am["adjacency"][0]) == (1 << 1) | (1 << 8) | (1 << 2) | (1 << 16)
    # graph 2: the 5-cycle 0-1-2-3-4-0
    expect = 0
    for u, v in [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]:
        expect |= (1 << (u * 8 + v)) | (1 << (v * 8 + u))
    assert int(am["adjacency"][2]) == expect


def test_am_generate_equals_csr_generate_on_degree_four_graphs():
    g = queries()
    assert np.diff(g["row_offsets"].astype(np.int64)).max() <= 4
    am = O.am_build(g)
    np.testing.assert_array_equal(O.am_generate(am), O.generate_signatures(g))
    base = np.arange(len(g["node_labels"]), dtype=np.uint64) * np.uint64(0x1000)  # onto existing values
    np.testing.assert_array_equal(O.am_generate(am, base.copy()), O.generate_signatures(g, base.copy()))


def test_am_fifth_neighbour_is_dropped():
    # star: centre 0 (label 1) with leaves 1..5 labelled 2, 3, 4, 5, 6
    g = batch_from_graphs([([1, 2, 3, 4, 5, 6], [(0, i) for i in range(1, 6)])])
    am = O.am_build(g)
    m = sig_matrix(O.am_generate(am))
    assert m[0].tolist() == [0, 0, 1, 1, 1, 1, 0] + [0] * 9  # labels 2..5 counted, 6 dropped
    assert all(m[i, 1] == 1 for i in range(1, 6))
    vb = O.am_generate(am)
    O.am_refine_view_based(am, 1, vb)
    csr = O.generate_signatures(g)
    O.refine_view_based(g, 1, csr)
    assert sig_matrix(vb)[0].tolist() != sig_matrix(csr)[0].tolist()  # the centre's fifth leaf is missing


def _ball_histograms(g, radius):
    go = g["graph_offsets"].astype(np.int64)
    out = np.zeros((len(g["node_labels"]), 16), np.int64)
    for gi in range(len(go) - 1):
        s, e = go[gi], go[gi + 1]
        G = nx.Graph()
        G.add_nodes_from(range(s, e))
        for v in range(s, e):
            for i in range(g["row_offsets"][v], g["row_offsets"][v + 1]):
                G.add_edge(v, int(g["column_indices"][i]))
        for v in range(s, e):
            ball = nx.single_source_shortest_path_length(G, v, cutoff=radius)
            for u in ball:
                if u != v and g["node_labels"][u] < 16:
                    out[v, g["node_labels"][u]] += 1
    return np.minimum(out, 15)


@pytest.mark.parametrize("view", [1, 2, 3, 4])
def test_am_power_graph_is_the_ball_histogram(view):
    g = queries(seed=22, count=40)
    assert (np.diff(g["graph_offsets"].astype(np.int64)) <= 32).all()
    am = O.am_build(g)
    sigs = O.am_generate(am)
    O.am_refine_power_graph(am, view, sigs)
    np.testing.assert_array_equal(sig_matrix(sigs), _ball_histograms(g, view))
    again = sigs.copy()
    O.am_refine_power_graph(am, view, again)  # each call clears and rebuilds
    np.testing.assert_array_equal(again, sigs)
    if view > 1:  # the planted sphere-only histogram is caught
        bad = O.am_generate(am)
        O.am_refine_power_graph(am, view, bad, mutation=1)
        assert not np.array_equal(sig_matrix(bad), _ball_histograms(g, view))


def test_am_power_graph_aliases_positions_mod_32():
    # a 40-node path; node 33's neighbours 32 and 34 land on bits 0 and 2 of a uint32 set, so
    # its view-1 signature counts the labels of nodes 0 and 2 (labels 1 and 2), not 32 and 34 (3)
    labels = [0] * 40
    labels[0], labels[2], labels[32], labels[34] = 1, 2, 3, 3
    g = batch_from_graphs([(labels, [(i, i + 1) for i in range(39)])])
    am = O.am_build(g)
    sigs = O.am_generate(am)
    O.am_refine_power_graph(am, 1, sigs)
    m = sig_matrix(sigs)
    assert m[33].tolist()[:4] == [0, 1, 1, 0]
    assert m[5].tolist()[:4] == [2, 0, 0, 0]  # an unaliased node: two label-0 neighbours


@pytest.mark.parametrize("view", [1, 2])
def test_am_view_based_equals_csr_view_based(view):
    g = queries(seed=23, count=50)
    am = O.am_build(g)
    a, c = O.am_generate(am), O.generate_signatures(g)
    for v in range(1, view + 1):
        O.am_refine_view_based(am, v, a)
        O.refine_view_based(g, v, c)
        np.testing.assert_array_equal(a, c)
