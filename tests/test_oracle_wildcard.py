"""Pinning the oracle's restatement of the wildcard/edge-label join (oracle_join_wildcard,
after library/include/isomorphism.hpp:383-556) before it is used as the GPU kernel's checker.

Checks that do not restate the oracle's loop:
* brute force over all injective maps on molecules whose query nodes have pairwise distinct
  labels (candidate sets are then disjoint, so the reference's stale-unset and kept-sibling
  bookkeeping cannot change a result): find-all = number of induced embeddings whose adjacent
  pairs carry equal edge labels; find-first = number of distinct images of the first node of
  the matching order (the node of most candidates, ties to the larger index);
* planted mistakes (edge labels ignored, matching order = query order, ties to the smaller
  index) must each contradict the brute force somewhere.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import oracle as O


def batch_el(graphs):
    """graphs: list of (labels, [(u, v, edge_label)]) -> batched CSR with edge labels (each
    undirected edge appended to both rows in input order, as graph.hpp:216-247 does)."""
    go, ro, ci, el, nl = [0], [0], [], [], []
    for labels, edges in graphs:
        n = len(labels)
        base = go[-1]
        rows = [[] for _ in range(n)]
        for u, v, l in edges:
            rows[u].append((v, l))
            rows[v].append((u, l))
        for r in rows:
            for w, l in r:
                ci.append(base + w)
                el.append(l)
            ro.append(len(ci))
        nl.extend(labels)
        go.append(base + n)
    return {"graph_offsets": np.array(go, np.uint32), "row_offsets": np.array(ro, np.uint32),
            "column_indices": np.array(ci, np.uint32), "node_labels": np.array(nl, np.uint8),
            "edge_labels": np.array(el, np.uint8)}


def random_case(seed, n_mol=25):
    rng = np.random.default_rng(seed)
    mols = []
    for _ in range(n_mol):
        n = int(rng.integers(5, 13))
        labels = rng.integers(0, 4, n).tolist()
        edges = [(int(rng.integers(0, i)), i, int(rng.integers(1, 4))) for i in range(1, n)]
        for _ in range(int(rng.integers(0, 4))):  # a few extra edges (rings)
            a, b = sorted(int(x) for x in rng.choice(n, 2, replace=False))
            if all({a, b} != {u, v} for u, v, _ in edges):
                edges.append((a, b, int(rng.integers(1, 4))))
        mols.append((labels, edges))
    queries = []
    for _ in range(12):  # pairwise distinct labels, random connected shapes and edge labels
        k = int(rng.integers(2, 5))
        labels = rng.permutation(4)[:k].tolist()
        edges = [(int(rng.integers(0, i)), i, int(rng.integers(1, 4))) for i in range(1, k)]
        if k >= 3 and rng.random() < 0.4 and all({0, k - 1} != {u, v} for u, v, _ in edges):
            edges.append((0, k - 1, int(rng.integers(1, 4))))
        queries.append((labels, edges))
    return queries, mols


def brute(query, mol):
    """All induced embeddings with equal edge labels on query edges (maps as tuples)."""
    ql, qe = query
    ml, me = mol
    qadj = {}
    for u, v, l in qe:
        qadj.setdefault((min(u, v), max(u, v)), l)
    madj = {}
    for u, v, l in me:
        madj.setdefault((min(u, v), max(u, v)), l)
    cands = [[p for p in range(len(ml)) if ml[p] == ql[i]] for i in range(len(ql))]
    out = []
    for f in itertools.product(*cands):
        if len(set(f)) < len(f):
            continue
        ok = True
        for a, b in itertools.combinations(range(len(ql)), 2):
            qk, mk = (a, b), (min(f[a], f[b]), max(f[a], f[b]))
            if (qk in qadj) != (mk in madj) or (qk in qadj and qadj[qk] != madj[mk]):
                ok = False
                break
        if ok:
            out.append(f)
    return out, [len(c) for c in cands]


def first_node(ccounts):
    m = max(ccounts)
    return max(i for i, c in enumerate(ccounts) if c == m)


def wildcard_table(queries, mols, find_first, mutation=0):
    q, d = batch_el(queries), batch_el(mols)
    words = O.filter_candidates(q, d)  # label candidates only, no refinement
    offsets, indices = O.gmcr(q, d, words)
    _, counts = O.join_wildcard(q, d, words, offsets, indices, find_first, mutation=mutation)
    return O.pair_table(offsets, indices, counts)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_wildcard_join_equals_brute_force(seed):
    queries, mols = random_case(seed)
    for find_first in (False, True):
        table = wildcard_table(queries, mols, find_first)
        assert len(table) > 10
        nonzero = 0
        for qg, dg, cnt in table:
            emb, cc = brute(queries[qg], mols[dg])
            if find_first:
                want = len({f[first_node(cc)] for f in emb})
            else:
                want = len(emb)
            assert cnt == want, (seed, find_first, qg, dg, cnt, want)
            nonzero += cnt > 0
        assert nonzero >= 5


@pytest.mark.parametrize("mutation,find_first", [(1, False), (2, True), (3, True)])
def test_planted_wildcard_mistakes_contradict_brute_force(mutation, find_first):
    mismatches = 0
    for seed in (1, 2, 3, 4, 5):
        queries, mols = random_case(seed)
        table = wildcard_table(queries, mols, find_first, mutation=mutation)
        for qg, dg, cnt in table:
            emb, cc = brute(queries[qg], mols[dg])
            want = len({f[first_node(cc)] for f in emb}) if find_first else len(emb)
            mismatches += cnt != want
    assert mismatches > 0


def test_matching_order_ties_and_max():
    """Hand case: query B0 - A1 on a star with centre B and three A leaves.  A1 has the most
    candidates (3), so it goes first and find-first counts its 3 images; starting from B0 would
    count 1.  With two A nodes tied, the larger index goes first."""
    star = ([1, 0, 0, 0], [(0, 1, 1), (0, 2, 1), (0, 3, 1)])
    q = ([1, 0], [(0, 1, 1)])
    assert wildcard_table([q], [star], True)[0, 2] == 3
    assert wildcard_table([q], [star], True, mutation=2)[0, 2] == 1
    assert wildcard_table([q], [star], False)[0, 2] == 3
    # A0 - B1 - A2 - C3 path query on A-B-A-C plus a pendant A on the B: A0 and A2 tie (3 each)
    mol = ([0, 1, 0, 2, 0], [(0, 1, 1), (1, 2, 1), (2, 3, 1), (1, 4, 1)])
    q2 = ([0, 1, 0, 2], [(0, 1, 1), (1, 2, 1), (2, 3, 1)])
    emb, cc = brute(q2, mol)
    assert cc[0] == cc[2] == 3 and first_node(cc) == 2
    assert len({f[2] for f in emb}) != len({f[0] for f in emb})  # the tie rule is observable
    assert wildcard_table([q2], [mol], True)[0, 2] == len({f[2] for f in emb})
    assert wildcard_table([q2], [mol], True, mutation=3)[0, 2] == len({f[0] for f in emb})


def test_edge_labels_must_match():
    mol = ([0, 1], [(0, 1, 2)])
    assert wildcard_table([([0, 1], [(0, 1, 2)])], [mol], False)[0, 2] == 1
    assert wildcard_table([([0, 1], [(0, 1, 3)])], [mol], False)[0, 2] == 0
    assert wildcard_table([([0, 1], [(0, 1, 3)])], [mol], False, mutation=1)[0, 2] == 1
