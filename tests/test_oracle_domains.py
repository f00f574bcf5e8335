"""Pinning the oracle's ViewBased refinement (oracle_refine_view_based, after
library/include/signature.hpp:298-334) and Data-domain filter (oracle_filter_data, after
isomorphism.hpp:22-62 with CandidatesDomain::Data).

* ViewBased, view_size 1, first pass on graphs whose counts stay small: the closed form
  H1 = H0 + A.H0 - deg(v) * onehot(label(v)) (numpy, from the adjacency matrix).
* Hand-traced cases of the uint8 arithmetic: a neighbour count below view_size wraps and then
  decrements through incrementLabelCount; an add that would pass 15 is dropped, not clamped.
* Data-domain filter: numpy label equality / wildcard, and the transpose of the Query-domain
  filter.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_pinning import batch_from_graphs


def sig_matrix(sigs):
    return np.array([[(int(s) >> (4 * l)) & 15 for l in range(16)] for s in sigs], np.int64)


def adjacency(g):
    n = len(g["node_labels"])
    A = np.zeros((n, n), np.int64)
    ro, ci = g["row_offsets"], g["column_indices"]
    for v in range(n):
        for i in range(ro[v], ro[v + 1]):
            A[v, ci[i]] += 1
    return A


@pytest.mark.parametrize("seed", [11, 12])
def test_view_based_first_pass_closed_form(seed):
    rng = np.random.default_rng(seed)
    graphs = []
    for _ in range(20):  # trees of degree <= 3 over 8 labels: every sum stays below 16
        n = int(rng.integers(3, 12))
        edges, deg = [], [0] * n
        for i in range(1, n):
            cands = [j for j in range(i) if deg[j] < 3]
            p = int(rng.choice(cands))
            edges.append((p, i))
            deg[p] += 1
            deg[i] += 1
        graphs.append((rng.integers(0, 8, n).tolist(), edges))
    g = batch_from_graphs(graphs)
    sigs = O.generate_signatures(g)
    H0 = sig_matrix(sigs)
    A = adjacency(g)
    onehot = np.zeros((len(H0), 16), np.int64)
    onehot[np.arange(len(H0)), g["node_labels"]] = 1
    expected = H0 + A @ H0 - A.sum(axis=1)[:, None] * onehot
    assert expected.max() < 16
    O.refine_view_based(g, 1, sigs)
    np.testing.assert_array_equal(sig_matrix(sigs), expected)


def test_view_based_uint8_wrap_and_dropped_add():
    # path a-b-c, all label 0: H0 = a{0:1} b{0:2} c{0:1}.  With view_size 2, b sees
    # (1 - 2) mod 256 = 255 from a and from c: 2 + 255 = 1 (mod 256), then 1 + 255 = 0.
    g = batch_from_graphs([([0, 0, 0], [(0, 1), (1, 2)])])
    sigs = O.generate_signatures(g)
    O.refine_view_based(g, 2, sigs)
    assert sig_matrix(sigs)[:, 0].tolist() == [1, 0, 1]
    # the planted mistake (own-label count not reduced) adds 2 to a and c and 1 + 1 to b
    sigs = O.generate_signatures(g)
    O.refine_view_based(g, 2, sigs, mutation=1)
    assert sig_matrix(sigs)[:, 0].tolist() == [3, 4, 3]
    # star: centre 0 (label 1) with 14 leaves of label 2, and a second centre joined to it
    # whose own 3 leaves make its label-2 count 3.  View 1: the centre holds 14 for label 2
    # and adds 3 -> 17 >= 16, so the add is dropped (stays 14, not 15).
    labels = [1] + [2] * 14 + [3] + [2] * 3
    edges = [(0, i) for i in range(1, 15)] + [(0, 15)] + [(15, 16 + i) for i in range(3)]
    g = batch_from_graphs([(labels, edges)])
    sigs = O.generate_signatures(g)
    assert sig_matrix(sigs)[0, 2] == 14
    O.refine_view_based(g, 1, sigs)
    assert sig_matrix(sigs)[0, 2] == 14


def test_filter_data_domain_closed_form_and_transpose():
    rng = np.random.default_rng(21)
    q = batch_from_graphs([(rng.integers(0, 5, 7).tolist() + [255], []) for _ in range(9)])
    d = batch_from_graphs([(rng.integers(0, 5, 30).tolist(), []) for _ in range(5)])
    Q, D = len(q["node_labels"]), len(d["node_labels"])
    words = O.filter_candidates_data(q, d)
    S = (Q + 31) // 32
    bits = np.unpackbits(words.view(np.uint8), bitorder="little").reshape(D, S * 32)[:, :Q]
    ql, dl = q["node_labels"].astype(int), d["node_labels"].astype(int)
    expected = (ql[None, :] == dl[:, None]) | (ql[None, :] == 255)
    np.testing.assert_array_equal(bits.astype(bool), expected)
    qwords = O.filter_candidates(q, d)
    SQ = (D + 31) // 32
    qbits = np.unpackbits(qwords.view(np.uint8), bitorder="little").reshape(Q, SQ * 32)[:, :D]
    np.testing.assert_array_equal(qbits.T.astype(bool), bits.astype(bool))
