"""CPU tests of the host side of the product library: C-ABI exports, generator, .dat I/O,
batch slicing and input validation (no GPU needed)."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

from sigmo_amd import capi
from tests._util import GOLDEN, digest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "sigmo_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sigmo_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 40
    L = capi.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(capi.EXPORTED) == names


def test_device_entry_points_fail_cleanly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.SigmoError):
        capi.Context(0)


# ---- generator ------------------------------------------------------------------------------------

def test_generator_is_deterministic_and_slice_invariant():
    a = capi.HostGraphs.molecules(9, 300)
    b = capi.HostGraphs.molecules(9, 300)
    assert digest(*a.arrays().values()) == digest(*b.arrays().values())
    part = capi.HostGraphs.molecules(9, 120, first=150)
    assert digest(*part.arrays().values()) == digest(*a.slice(150, 120).arrays().values())
    threaded = capi.HostGraphs.molecules(9, 9000, threads=7)
    single = capi.HostGraphs.molecules(9, 9000, threads=1)
    assert digest(*threaded.arrays().values()) == digest(*single.arrays().values())
    other = capi.HostGraphs.molecules(10, 300)
    assert digest(*other.arrays().values()) != digest(*a.arrays().values())


def _degrees_ok(arr):
    cap = {8: 4, 0: 3, 7: 2, 9: 2, 5: 3, 10: 1, 1: 1, 3: 1}
    deg = np.diff(arr["row_offsets"].astype(np.int64))
    caps = np.array([cap[int(l)] for l in arr["node_labels"]])
    return bool((deg <= caps).all())


@pytest.mark.parametrize("dist,lo,hi,mean", [(capi.SIZE_NORMAL, 6, 64, 24.0), (capi.SIZE_LOGNORMAL, 6, 128, 26.5)])
def test_molecule_law(dist, lo, hi, mean):
    arr = capi.HostGraphs.molecules(1, 20000, dist).arrays()
    n = np.diff(arr["graph_offsets"].astype(np.int64))
    assert n.max() <= hi
    assert (n >= lo).mean() > 0.995  # only molecules whose tree ran out of free valence end early
    assert abs(n.mean() - mean) < 1.0
    labels = np.bincount(arr["node_labels"], minlength=11) / len(arr["node_labels"])
    for code, p in ((8, 0.72), (0, 0.11), (7, 0.11), (9, 0.02)):
        assert abs(labels[code] - p) < 0.02, (code, labels[code])
    assert _degrees_ok(arr)
    # undirected CSR without self loops or duplicate edges
    go, ro, ci = arr["graph_offsets"], arr["row_offsets"], arr["column_indices"]
    for v in range(0, len(arr["node_labels"]), 997):
        row = ci[ro[v]:ro[v + 1]]
        assert v not in row and len(set(row.tolist())) == len(row)
    edges = len(ci) // 2
    assert edges > len(arr["node_labels"]) - len(go) + 1  # ring closures were added
    el = arr["edge_labels"]
    assert set(np.unique(el).tolist()) <= {1, 2, 3, 4} and (el == 4).any()


def _is_connected(labels, ro, ci, base):
    n = len(labels)
    seen, stack = {0}, [0]
    while stack:
        u = stack.pop()
        for w in ci[ro[base + u]:ro[base + u + 1]] - base:
            if int(w) not in seen:
                seen.add(int(w))
                stack.append(int(w))
    return len(seen) == n


def test_queries_are_connected_induced_and_sized():
    qh = capi.HostGraphs.queries(2, 200, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)
    arr = qh.arrays()
    go, ro, ci = arr["graph_offsets"], arr["row_offsets"], arr["column_indices"]
    k = np.diff(go.astype(np.int64))
    assert k.min() >= 4 and k.max() <= 32
    cyclic = 0
    for g in range(len(k)):
        b = int(go[g])
        assert _is_connected(arr["node_labels"][b:b + k[g]], ro, ci, b)
        e = (ro[b + k[g]] - ro[b]) // 2
        cyclic += int(e >= k[g])
    assert cyclic == 60  # exactly 30% contain a cycle


def test_query_perturbation_fraction():
    base = capi.HostGraphs.queries(1, 100, 4, 20, capi.SIZE_NORMAL).arrays()
    pert = capi.HostGraphs.queries(1, 100, 4, 20, capi.SIZE_NORMAL, perturb_frac=0.2).arrays()
    np.testing.assert_array_equal(base["graph_offsets"], pert["graph_offsets"])
    diff = base["node_labels"] != pert["node_labels"]
    changed = {int(np.searchsorted(base["graph_offsets"], i, side="right") - 1) for i in np.nonzero(diff)[0]}
    assert len(changed) == 20 and diff.sum() == 20


def test_arrivals():
    t, q = capi.arrivals(4, 20000, 200.0, 1, 50)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert (gaps > 0).all() and abs(gaps.mean() - 1 / 200.0) < 2e-4
    assert q.min() == 1 and q.max() == 50


# ---- .dat I/O (graph.hpp:151-247, io.hpp:37-55) --------------------------------------------------------

def test_reference_fixture_files_parse():
    q = capi.HostGraphs.load_dat(GOLDEN / "test_query.dat")
    d = capi.HostGraphs.load_dat(GOLDEN / "test_data.dat")
    assert q.info() == (4, 16, 2 * (2 + 3 + 5 + 3))
    assert d.info() == (2, 22, 2 * (13 + 10))
    a = d.arrays()
    # first molecule: node 0 has the single edge 0-2 (column ids are global)
    assert a["column_indices"][a["row_offsets"][0]:a["row_offsets"][1]].tolist() == [2]
    # second molecule's edge "0 2" lands at global ids 13 and 15
    assert 15 in a["column_indices"][a["row_offsets"][13]:a["row_offsets"][14]].tolist()


def test_dat_round_trip(tmp_path):
    g = capi.HostGraphs.molecules(3, 200)
    p = tmp_path / "m.dat"
    g.save_dat(p)
    h = capi.HostGraphs.load_dat(p)
    a, b = g.arrays(), h.arrays()
    np.testing.assert_array_equal(a["graph_offsets"], b["graph_offsets"])
    np.testing.assert_array_equal(a["node_labels"], b["node_labels"])
    # CSR neighbour order may differ; compare neighbour sets
    for v in range(0, len(a["node_labels"]), 7):
        ra = sorted(a["column_indices"][a["row_offsets"][v]:a["row_offsets"][v + 1]].tolist())
        rb = sorted(b["column_indices"][b["row_offsets"][v]:b["row_offsets"][v + 1]].tolist())
        assert ra == rb


@pytest.mark.parametrize("line", [
    "n#3 l#2 0 1 1 1 2 1 e#2 0 1 1 2 1",     # token count neither 2E nor 3E
    "x#3 l#2 0 1 1 1 2 1 e#0",               # bad header
    "n#2 l#2 0 1 5 1 e#0",                   # node id out of range
    "n#2 l#2 0 1 1 1 e#1 0 9",               # edge endpoint out of range
])
def test_malformed_dat_is_rejected(line):
    with pytest.raises(capi.SigmoError, match="Invalid graph format"):
        capi.HostGraphs.parse_dat(line + "\n")


def test_unlabelled_edges_default_to_zero():
    g = capi.HostGraphs.parse_dat("n#3 l#2 0 1 1 1 2 1 e#2 0 1 1 2\n")
    assert g.arrays()["edge_labels"].tolist() == [0, 0, 0, 0]


def test_from_arrays_validates_graph_membership():
    ok = dict(graph_offsets=[0, 2, 4], row_offsets=[0, 1, 2, 3, 4], column_indices=[1, 0, 3, 2], node_labels=[8, 8, 8, 8])
    capi.HostGraphs.from_arrays(**ok)
    bad = dict(ok, column_indices=[2, 0, 3, 2])  # node 0 points into the second graph
    with pytest.raises(capi.SigmoError, match="leaves its graph"):
        capi.HostGraphs.from_arrays(**bad)


def test_from_arrays_rejects_asymmetric_adjacency():
    # a path 0-1-2 listed in both directions is accepted; dropping 2 -> 1 leaves 1 -> 2 one-way,
    # which the reference's CSR (every edge appended in both directions) never produces
    ok = dict(graph_offsets=[0, 3], row_offsets=[0, 1, 3, 4], column_indices=[1, 0, 2, 1], node_labels=[8, 7, 8])
    capi.HostGraphs.from_arrays(**ok)
    one_way = dict(graph_offsets=[0, 3], row_offsets=[0, 1, 3, 3], column_indices=[1, 0, 2], node_labels=[8, 7, 8])
    with pytest.raises(capi.SigmoError, match="not symmetric"):
        capi.HostGraphs.from_arrays(**one_way)


def test_slice_and_concat():
    g = capi.HostGraphs.molecules(4, 50)
    a, b = g.slice(0, 20), g.slice(20, 30)
    c = a.concat(b)
    assert digest(*c.arrays().values()) == digest(*g.arrays().values())
    with pytest.raises(capi.SigmoError):
        g.slice(40, 20)
