"""Pinning the CPU oracle (oracle/oracle.cpp) before it is trusted as the parity reference.

Each oracle function is checked against something that does not restate its own formula:

======================  ===================================================================
oracle function          pinned by
======================  ===================================================================
sig_set / sig_inc        reference known-answer test library/test/candidates.cpp:13-32
cand insert/remove/count reference known-answer test library/test/candidates.cpp:119-143
                         (64-bit word image) + numpy popcounts over random ranges
generate_signatures      closed form min(15, #neighbours with label l), computed in Python
                         from the reference fixtures (test_data.dat, test_pool.bin) and
                         generated molecules
refine_signatures        networkx BFS spheres (n <= 64); hand-traced positions of a 70-node
                         path for the mod-64 aliasing of Bitset<uint64_t>
filter_candidates        numpy label equality / wildcard closed form; the superset property
                         of library/test/filter.cpp:14-69
refine_candidates        brute force: label match and sphere-histogram dominance at every
                         level, from networkx distances
gmcr                     brute force over all (query graph, molecule) pairs
join                     hand-traced DFS-quirk witnesses (1 and 4 embeddings); networkx
                         induced-subgraph counts where the DFS quirks cannot fire (k = 2,
                         unique-label molecules); the oracle's textbook mode equals networkx
dat_shard_lines          byte ranges worked out by hand on files of 40-byte lines (a rank that
                         starts on a line still drops it, the line limit counts lines before
                         the n/e/l filter, the NUL run past the end): tests/test_shard.py
======================  ===================================================================

Planted mistakes (the oracle's `mutation` switches) must make the matching check fail.
"""
from __future__ import annotations

import networkx as nx
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import GOLDEN, witness_graphs


# ---- helpers independent of the product library -------------------------------------------------

def batch_from_graphs(graphs):
    """graphs: list of (labels, edges) -> batched CSR dict built here (reference counting-sort order)."""
    go, ro, ci, nl = [0], [0], [], []
    for labels, edges in graphs:
        n = len(labels)
        base = go[-1]
        rows = [[] for _ in range(n)]
        for u, v in edges:
            rows[u].append(v)
            rows[v].append(u)
        for r in rows:
            ci.extend(base + w for w in r)
            ro.append(len(ci))
        nl.extend(labels)
        go.append(base + n)
    return {
        "graph_offsets": np.array(go, np.uint32),
        "row_offsets": np.array(ro, np.uint32),
        "column_indices": np.array(ci, np.uint32),
        "node_labels": np.array(nl, np.uint8),
        "edge_labels": np.zeros(len(ci), np.uint8),
    }


def parse_dat(path):
    graphs = []
    for line in open(path):
        t = line.split()
        if not t:
            continue
        n = int(t[0][2:])
        labels = [0] * n
        i = 2
        for _ in range(n):
            labels[int(t[i])] = int(t[i + 1])
            i += 2
        ne = int(t[i][2:])
        rest = t[i + 1:]
        step = 3 if len(rest) == 3 * ne else 2
        edges = [(int(rest[j * step]), int(rest[j * step + 1])) for j in range(ne)]
        graphs.append((labels, edges))
    return graphs


def decode_pool(path):
    """library/test/data/test_pool.bin: u32 G, then per graph u32 n, (n+1) u32 row offsets,
    E u32 columns, n u8 labels."""
    b = open(path, "rb").read()
    G = int(np.frombuffer(b[:4], np.uint32)[0])
    off, graphs = 4, []
    for _ in range(G):
        n = int(np.frombuffer(b[off:off + 4], np.uint32)[0])
        off += 4
        ro = np.frombuffer(b[off:off + 4 * (n + 1)], np.uint32)
        off += 4 * (n + 1)
        E = int(ro[-1])
        ci = np.frombuffer(b[off:off + 4 * E], np.uint32)
        off += 4 * E
        lab = np.frombuffer(b[off:off + n], np.uint8)
        off += n
        edges = sorted({(min(u, int(v)), max(u, int(v))) for u in range(n) for v in ci[ro[u]:ro[u + 1]]})
        graphs.append((lab.tolist(), edges))
    return graphs


def random_molecules(seed, count, nmax=40):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(3, nmax))
        labels = rng.choice([8, 8, 8, 0, 7, 9, 10, 1], size=n).tolist()
        edges = [(int(rng.integers(0, i)), i) for i in range(1, n)]
        for _ in range(int(rng.integers(0, 4))):
            a, c = sorted(rng.choice(n, 2, replace=False).tolist())
            if (a, c) not in edges:
                edges.append((a, c))
        out.append((labels, edges))
    return out


def nx_graph(labels, edges):
    g = nx.Graph()
    g.add_nodes_from((i, {"l": l}) for i, l in enumerate(labels))
    g.add_edges_from(edges)
    return g


def sphere_hist(graph, r):
    """Per node: 16 saturated label counts of the nodes at BFS distance exactly r."""
    labels, edges = graph
    g = nx_graph(labels, edges)
    out = []
    for v in range(len(labels)):
        d = nx.single_source_shortest_path_length(g, v, cutoff=r)
        h = [0] * 16
        for u, du in d.items():
            if du == r and r > 0 and labels[u] < 16:
                h[labels[u]] += 1
        out.append([min(c, 15) for c in h])
    return out


def unpack_sig(s):
    return [(int(s) >> (4 * l)) & 15 for l in range(16)]


def words_to_bits(words, Q, D):
    stride = (D + 31) // 32
    w = words.reshape(Q, stride).astype("<u4")
    bits = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")
    return bits[:, :D].astype(bool)


# ---- SignatureDevice (signature.hpp:23-54) ---------------------------------------------------------

def test_signature_known_answers():
    # library/test/candidates.cpp:13-32
    s = 0
    for label, count in ((0, 1), (2, 3), (5, 2), (15, 1)):
        s = O.sig_set(s, label, count)
    assert s == 0b0001000000000000000000000000000000000000001000000000001100000001
    assert O.sig_get(s, 15) == 1
    s = O.sig_inc(s, 15, 1)
    assert s == 0b0010000000000000000000000000000000000000001000000000001100000001
    s = O.sig_inc(s, 14, 1)
    assert s == 0b0010000100000000000000000000000000000000001000000000001100000001
    s = O.sig_inc(s, 14, 14)
    assert s == 0b0010111100000000000000000000000000000000001000000000001100000001


def test_signature_saturation_edges():
    s = O.sig_set(0, 3, 15)
    assert O.sig_inc(s, 3, 1) == s  # full nibble: no change
    s = O.sig_set(0, 3, 10)
    assert O.sig_inc(s, 3, 10) == s  # 10 + 10 does not fit four bits: the set is skipped
    assert O.sig_set(0, 16, 1) == 0 and O.sig_get(0xFFFFFFFFFFFFFFFF, 16) == 0  # labels >= 16 ignored
    assert O.sig_inc(0, 255, 1) == 0


# ---- candidates (candidates.hpp:16-162) --------------------------------------------------------------

def test_candidates_known_answers():
    # library/test/candidates.cpp:119-143, which states the image in 64-bit words
    c = O.RefCandidates(2, 128)
    for d in (0, 31, 32, 124):
        c.insert(0, d)
    w64 = c.words.view(np.uint64)
    assert w64[0] == 0b0000000000000000000000000000000110000000000000000000000000000001
    assert w64[1] == 0b0001000000000000000000000000000000000000000000000000000000000000
    assert c.count(0) == 4
    c.remove(0, 32)
    assert w64[0] == 0b0000000000000000000000000000000010000000000000000000000000000001
    assert c.count(0) == 3
    c.insert(1, 0)
    assert w64[2] == 1 and w64[3] == 0


def test_candidate_range_count_and_select():
    rng = np.random.default_rng(3)
    D = 300
    c = O.RefCandidates(3, D)
    dense = rng.random((3, D)) < 0.4
    for q, d in zip(*np.nonzero(dense)):
        c.insert(int(q), int(d))
    cases = [(0, 32), (32, 64), (5, 37), (31, 33), (60, 64), (64, 65), (1, 300), (288, 300), (100, 101), (7, 9)]
    cases += [tuple(sorted(rng.choice(D + 1, 2, replace=False).tolist())) for _ in range(200)]
    for q in range(3):
        for s, e in cases:
            if s == e:
                continue
            members = np.nonzero(dense[q, s:e])[0] + s
            assert c.count(q, s, e) == len(members), (q, s, e)
            for i in range(len(members) + 2):
                want = int(members[i]) if i < len(members) else 0xFFFFFFFF
                assert c.at(q, i, s, e) == want, (q, s, e, i)


# ---- signatures -------------------------------------------------------------------------------------

def fixture_graph_sets():
    return {
        "test_data.dat": parse_dat(GOLDEN / "test_data.dat"),
        "test_query.dat": parse_dat(GOLDEN / "test_query.dat"),
        "test_pool.bin": decode_pool(GOLDEN / "test_pool.bin"),
        "random": random_molecules(11, 60),
        "shapes": [
            ([8] * 20, [(i, i + 1) for i in range(19)]),                      # path
            ([8, 7] * 6, [(i, (i + 1) % 12) for i in range(12)]),              # ring
            ([0] + [8] * 12, [(0, i) for i in range(1, 13)]),                  # star (saturation at 12 < 15)
            ([9] * 20, [(a, b) for a in range(20) for b in range(a + 1, 20)]),  # K20: 19 neighbours -> 15
        ],
    }


@pytest.mark.parametrize("name", list(fixture_graph_sets()))
def test_level0_signatures_closed_form(name):
    graphs = fixture_graph_sets()[name]
    sigs = O.generate_signatures(batch_from_graphs(graphs))
    i = 0
    for labels, edges in graphs:
        nb = [[] for _ in labels]
        for u, v in edges:
            nb[u].append(v)
            nb[v].append(u)
        for v in range(len(labels)):
            want = [0] * 16
            for u in nb[v]:
                if labels[u] < 16:
                    want[labels[u]] += 1
            assert unpack_sig(sigs[i]) == [min(x, 15) for x in want], (name, v)
            i += 1


@pytest.mark.parametrize("name", list(fixture_graph_sets()))
@pytest.mark.parametrize("mutation", [0, 2])
def test_levels_are_bfs_spheres(name, mutation):
    graphs = fixture_graph_sets()[name]
    b = batch_from_graphs(graphs)
    N = len(b["node_labels"])
    sigs, reach = np.zeros(N, np.uint64), np.zeros(N, np.uint64)
    ok = True
    for r in range(1, 7):
        O.refine_signatures(b, r, sigs, reach, mutation)
        want = [h for g in graphs for h in sphere_hist(g, r)]
        got = [unpack_sig(s) for s in sigs]
        ok = ok and got == want
        if mutation == 0:
            assert got == want, (name, r)
    if mutation == 2:  # planted: the start node stays in the sphere
        assert not ok


def test_aliasing_positions_mod_64():
    """Hand-traced Bitset<uint64_t> BFS on a 70-node path (labels i % 11).

    Node 64: its start position is 64 mod 64 = 0 and position 0 expands node 0, whose only
    neighbour is node 1: level 1 = {label(1)} = {1}.
    Node 63: neighbours 62 and 64; 64 lands on position 0, whose label is read from node 0:
    level 1 = {label(62)=7, label(0)=0}.  Level 2 expands position 62 (node 62 -> 61, 63;
    63 is the start) and position 0 (node 0 -> 1): {label(61)=6, label(1)=1}.
    Node 1: an ordinary node: level 1 = {label(0)=0, label(2)=2}.
    """
    n = 70
    b = batch_from_graphs([([i % 11 for i in range(n)], [(i, i + 1) for i in range(n - 1)])])
    for mutation, expect_ok in ((0, True), (1, False)):
        sigs, reach = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
        O.refine_signatures(b, 1, sigs, reach, mutation)
        lvl1 = sigs.copy()
        O.refine_signatures(b, 2, sigs, reach, mutation)
        lvl2 = sigs.copy()

        def h(s):
            return {l: c for l, c in enumerate(unpack_sig(s)) if c}

        ok = (h(lvl1[64]) == {1: 1} and h(lvl1[63]) == {7: 1, 0: 1} and h(lvl2[63]) == {6: 1, 1: 1}
              and h(lvl1[1]) == {0: 1, 2: 1})
        assert ok == expect_ok, mutation


# ---- filter / refine ----------------------------------------------------------------------------------

def test_filter_closed_form_and_superset():
    queries = parse_dat(GOLDEN / "test_query.dat") + [([8, 255, 7, 0], [(0, 1), (1, 2), (2, 3)])]
    data = parse_dat(GOLDEN / "test_data.dat") + random_molecules(5, 30)
    q, d = batch_from_graphs(queries), batch_from_graphs(data)
    Q, D = len(q["node_labels"]), len(d["node_labels"])
    words = O.filter_candidates(q, d)
    ql, dl = q["node_labels"][:, None], d["node_labels"][None, :]
    want = (ql == dl) | (ql == 255)
    np.testing.assert_array_equal(words_to_bits(words, Q, D), want)
    # filter.cpp:14-69: every label-equal node passing the level-0 signature test is present
    dsig, qsig = O.generate_signatures(d), O.generate_signatures(q)
    got = words_to_bits(words, Q, D)
    for i in range(Q):
        for j in range(D):
            if q["node_labels"][i] == d["node_labels"][j] and all(
                    a <= c for a, c in zip(unpack_sig(qsig[i]), unpack_sig(dsig[j]))):
                assert got[i, j]


@pytest.mark.parametrize("mutation", [0, 1])
def test_refine_is_label_match_and_sphere_dominance(mutation):
    queries = parse_dat(GOLDEN / "test_query.dat") + random_molecules(21, 8, nmax=10)
    data = parse_dat(GOLDEN / "test_data.dat") + random_molecules(22, 25)
    q, d = batch_from_graphs(queries), batch_from_graphs(data)
    Q, D = len(q["node_labels"]), len(d["node_labels"])
    R = 4
    words = O.filter_candidates(q, d)
    qsig, dsig = np.zeros(Q, np.uint64), np.zeros(D, np.uint64)
    qre, dre = np.zeros(Q, np.uint64), np.zeros(D, np.uint64)
    for r in range(1, R + 1):
        O.refine_signatures(d, r, dsig, dre)
        O.refine_signatures(q, r, qsig, qre)
        O.refine_candidates(qsig, dsig, words, mutation)
    got = words_to_bits(words, Q, D)
    want = q["node_labels"][:, None] == d["node_labels"][None, :]
    for r in range(1, R + 1):
        hq = np.array([h for g in queries for h in sphere_hist(g, r)])
        hd = np.array([h for g in data for h in sphere_hist(g, r)])
        want &= (hq[:, None, :] <= hd[None, :, :]).all(axis=2)
    if mutation == 0:
        np.testing.assert_array_equal(got, want)
    else:  # planted: labels 8..15 never compared
        assert not np.array_equal(got, want)


# ---- GMCR -----------------------------------------------------------------------------------------------

def test_gmcr_brute_force():
    queries = random_molecules(31, 12, nmax=8)
    data = random_molecules(32, 80)
    q, d = batch_from_graphs(queries), batch_from_graphs(data)
    Q, D = len(q["node_labels"]), len(d["node_labels"])
    words = O.filter_candidates(q, d)
    offsets, indices = O.gmcr(q, d, words)
    bits = words_to_bits(words, Q, D)
    qgo, dgo = q["graph_offsets"], d["graph_offsets"]
    want_off, want_idx = [0], []
    for dg in range(len(dgo) - 1):
        for qg in range(len(qgo) - 1):
            if all(bits[i, dgo[dg]:dgo[dg + 1]].any() for i in range(qgo[qg], qgo[qg + 1])):
                want_idx.append(qg)
        want_off.append(len(want_idx))
    np.testing.assert_array_equal(offsets, want_off)
    np.testing.assert_array_equal(indices, want_idx)


def test_gmcr_rejects_single_node_queries():
    q = batch_from_graphs([([8], [])])
    d = batch_from_graphs([([8, 8], [(0, 1)])])
    with pytest.raises(ValueError):
        O.gmcr(q, d, O.filter_candidates(q, d))


# ---- join -------------------------------------------------------------------------------------------------

def join_counts(queries, data, find_first, mutation=0, R=0):
    q, d = batch_from_graphs(queries), batch_from_graphs(data)
    r = O.pipeline(q, d, R, modes=[find_first], join_mutation=mutation)
    key = "first" if find_first else "all"
    return O.pair_table(r["offsets"], r["indices"], r[f"counts_{key}"])


def nx_induced_count(qgraph, mgraph):
    gq, gm = nx_graph(*qgraph), nx_graph(*mgraph)
    gmatch = nx.algorithms.isomorphism.GraphMatcher(gm, gq, node_match=lambda a, b: a["l"] == b["l"])
    return sum(1 for _ in gmatch.subgraph_isomorphisms_iter())


@pytest.mark.parametrize("idx", [0, 1])
def test_dfs_quirk_witnesses(idx):
    q, mol, expected = witness_graphs()[idx]
    assert join_counts([q], [mol], False)[0, 2] == expected
    assert join_counts([q], [mol], True)[0, 2] == 1
    # the textbook DFS (planted mistake) counts the true embeddings instead
    assert join_counts([q], [mol], False, mutation=1)[0, 2] == 2 == nx_induced_count(q, mol)


def test_join_two_node_queries_equal_induced_counts():
    data = random_molecules(41, 40)
    queries = [([8, 8], [(0, 1)]), ([8, 8], []), ([8, 7], [(0, 1)]), ([0, 8], []), ([9, 8], [(0, 1)])]
    table = join_counts(queries, data, False)
    assert len(table) > 0
    for qg, dg, cnt in table:
        assert cnt == nx_induced_count(queries[qg], data[dg]), (qg, dg)


def test_join_unique_label_molecules_equal_induced_counts():
    rng = np.random.default_rng(51)
    data, queries = [], []
    for _ in range(30):
        n = int(rng.integers(6, 15))
        labels = rng.permutation(16)[:n].tolist()
        edges = [(int(rng.integers(0, i)), i) for i in range(1, n)]
        data.append((labels, edges))
    for _ in range(15):  # connected induced subgraphs of random molecules
        labels, edges = data[int(rng.integers(0, len(data)))]
        g = nx_graph(labels, edges)
        k = int(rng.integers(2, 6))
        nodes = list(nx.bfs_tree(g, int(rng.integers(0, len(labels)))).nodes())[:k]
        sub = g.subgraph(nodes)
        pos = {v: i for i, v in enumerate(nodes)}
        queries.append(([labels[v] for v in nodes], [(pos[a], pos[b]) for a, b in sub.edges()]))
    table = join_counts(queries, data, False)
    assert (table[:, 2] > 0).sum() >= 15
    for qg, dg, cnt in table:
        assert cnt == nx_induced_count(queries[qg], data[dg]), (qg, dg)


def induced_bfs_queries(data, seed, count, kmin, kmax):
    rng = np.random.default_rng(seed)
    queries = []
    for _ in range(count):
        labels, edges = data[int(rng.integers(0, len(data)))]
        g = nx_graph(labels, edges)
        k = int(rng.integers(kmin, kmax + 1))
        nodes = list(nx.bfs_tree(g, int(rng.integers(0, len(labels)))).nodes())[:k]
        pos = {v: i for i, v in enumerate(nodes)}
        queries.append(([labels[v] for v in nodes], [(pos[a], pos[b]) for a, b in g.subgraph(nodes).edges()]))
    return queries


def test_textbook_mode_matches_networkx_and_faithful_mode_differs():
    data = random_molecules(61, 40, nmax=20)
    queries = induced_bfs_queries(data, 62, 12, 3, 6)
    faithful = join_counts(queries, data, False)
    textbook = join_counts(queries, data, False, mutation=1)
    np.testing.assert_array_equal(faithful[:, :2], textbook[:, :2])
    for qg, dg, cnt in textbook:
        assert cnt == nx_induced_count(queries[qg], data[dg])
    assert (faithful[:, 2] != textbook[:, 2]).any()  # the reference bookkeeping is not textbook


@pytest.mark.parametrize("mutation", [2, 3])
def test_planted_join_mistakes_change_results(mutation):
    from sigmo_amd import capi

    qh = capi.HostGraphs.queries(0, 20, 4, 12, capi.SIZE_NORMAL)
    dh = capi.HostGraphs.molecules(0, 400)
    q, d = qh.arrays(), dh.arrays()
    base = O.pipeline(q, d, 2, modes=[False])
    bad = O.pipeline(q, d, 2, modes=[False], join_mutation=mutation)
    assert not np.array_equal(base["counts_all"], bad["counts_all"])


def test_find_first_marks_exactly_the_pairs_find_all_counts():
    import json

    gold = json.loads((GOLDEN / "config0.json").read_text())
    first, alls = np.array(gold["counts_first"]), np.array(gold["counts_all"])
    np.testing.assert_array_equal(first, (alls > 0).astype(first.dtype))
    assert gold["matches_first"] == int(first.sum()) and gold["matches_all"] == int(alls.sum())


def test_config0_golden_matches_oracle():
    import json

    from sigmo_amd import capi
    from tests._util import digest

    gold = json.loads((GOLDEN / "config0.json").read_text())
    qh = capi.HostGraphs.queries(0, 20, 4, 12, capi.SIZE_NORMAL)
    dh = capi.HostGraphs.molecules(0, 2000)
    assert digest(*qh.arrays().values()) == gold["query_digest"]
    assert digest(*dh.arrays().values()) == gold["data_digest"]
    r = O.pipeline(qh.arrays(), dh.arrays(), 6, keep_levels=True, modes=[True, False])
    assert [[digest(a), digest(b)] for a, b in r["levels"]] == gold["level_digests"]
    assert digest(r["words"]) == gold["words_digest"]
    assert r["offsets"].tolist() == gold["offsets"] and r["indices"].tolist() == gold["indices"]
    assert r["counts_first"].tolist() == gold["counts_first"]
    assert r["counts_all"].tolist() == gold["counts_all"]
