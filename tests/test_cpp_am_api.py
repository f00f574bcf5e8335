"""The host half of include/sigmo.hpp's adjacency-matrix API (graph.hpp:20-35, io.hpp:17-35,
utils.hpp:17-62), compiled with g++ against the header and run on the CPU: the matrices
loadAMGraphsFromFile builds equal the oracle's AM layout, getNeighbors lists the first four
row bits in ascending order (NULL_NODE-terminated) and isNeighbor reports exactly the edges."""
from __future__ import annotations

import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from sigmo_amd import build as B
from sigmo_amd import capi
from tests._util import GOLDEN


@pytest.fixture(scope="module")
def am_api(tmp_path_factory):
    if not B.LIB.exists() or shutil.which(B.CXX) is None:
        pytest.skip("library or host compiler missing")
    exe = tmp_path_factory.mktemp("am") / "am_api"
    subprocess.run([B.CXX, "-std=c++17", "-O1", "-I", str(B.ROOT / "include"), str(B.ROOT / "tests" / "cpp" / "am_api.cpp"),
                    "-L", str(B.PKG), "-lsigmo_hip", f"-Wl,-rpath,{B.PKG}", "-o", str(exe)], check=True)
    return exe


def _parse(lines):
    graphs = []
    for line in lines:
        tag, *rest = line.split()
        if tag == "g":
            graphs.append({"n": int(rest[0]), "words": [int(x, 16) for x in rest[2:]], "nodes": []})
        elif tag == "n":
            graphs[-1]["nodes"].append({"first": [int(x) for x in rest], "adj": []})
        else:
            graphs[-1]["nodes"][int(rest[0])]["adj"].append(int(rest[1]))
    return graphs


def _check(exe, path, host):
    out = subprocess.run([str(exe), str(path)], check=True, capture_output=True, text=True).stdout.splitlines()
    got = _parse(out)
    g = host.arrays()
    ref = O.am_build(g)
    go = g["graph_offsets"].astype(np.int64)
    ro, ci = g["row_offsets"].astype(np.int64), g["column_indices"].astype(np.int64)
    assert len(got) == len(go) - 1
    for gi, h in enumerate(got):
        n = go[gi + 1] - go[gi]
        assert h["n"] == n and len(h["nodes"]) == n
        lo = int(ref["graph_offsets"][gi])
        assert h["words"] == [int(w) for w in ref["adjacency"][lo:lo + len(h["words"])]], gi
        end = int(ref["graph_offsets"][gi + 1]) if gi + 1 < len(got) else len(ref["adjacency"])
        assert len(h["words"]) == end - lo
        for u, node in enumerate(h["nodes"]):
            nbrs = sorted({int(c) - int(go[gi]) for c in ci[ro[go[gi] + u]:ro[go[gi] + u + 1]]})
            assert node["first"] == nbrs[:4], (gi, u)
            assert node["adj"] == nbrs, (gi, u)


def test_am_api_on_the_reference_query_fixture(am_api):
    _check(am_api, GOLDEN / "test_query.dat", capi.HostGraphs.load_dat(GOLDEN / "test_query.dat"))


def test_am_api_on_dense_graphs(am_api, tmp_path):
    rng = np.random.default_rng(5)
    graphs = []
    for n in (1, 2, 9, 12, 20, 33, 40, 64):
        edges = sorted({(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < 0.3})
        graphs.append((rng.integers(0, 20, n).tolist(), edges))
    host = capi.HostGraphs.from_edge_lists(graphs)
    path = tmp_path / "dense.dat"
    host.save_dat(path)
    _check(am_api, path, capi.HostGraphs.load_dat(path))
