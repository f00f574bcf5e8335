"""GPU parity of the adjacency-matrix (AM) graph path against the oracle (itself pinned in
test_oracle_am.py): the device image (createDeviceAMGraph, graph.hpp:345-392), generateAMSignatures
(signature.hpp:100-120), refineAMSignatures<PowerGraph> (signature.hpp:250-296) and
refineAMSignatures<ViewBased> (signature.hpp:214-248), bit-exact on both scopes."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from sigmo_amd import capi
from tests._util import GOLDEN

pytestmark = pytest.mark.gpu


def _random_graphs(seed, count, nmax):
    """Random graphs of 1..nmax nodes with degrees up to ~10 (so the four-neighbour cut matters)
    and labels 0..20 (labels >= 16 are not counted)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(1, nmax + 1))
        labels = rng.integers(0, 21, n).tolist()
        m = int(rng.integers(0, 3 * n + 1)) if n > 1 else 0
        edges = set()
        for _ in range(m):
            a, b = (int(x) for x in rng.integers(0, n, 2))
            if a != b:
                edges.add((min(a, b), max(a, b)))
        out.append((labels, sorted(edges)))
    return out


def _batch():
    star = ([1, 2, 3, 4, 5, 6, 7], [(0, i) for i in range(1, 7)])
    path40 = ([i % 5 for i in range(40)], [(i, i + 1) for i in range(39)])
    clique64 = ([i % 16 for i in range(64)], [(a, b) for a in range(64) for b in range(a + 1, 64) if (a * 7 + b) % 3])
    hand = capi.HostGraphs.from_edge_lists([star, path40, clique64] + _random_graphs(31, 400, 64))
    fixture = capi.HostGraphs.load_dat(GOLDEN / "test_query.dat")
    gen = capi.HostGraphs.queries(31, 1500, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)
    return fixture.concat(hand).concat(gen)


@pytest.fixture(scope="module")
def batch():
    return _batch()


def test_am_device_image_matches_oracle(ctx, batch):
    print("AM upload", flush=True)
    am = capi.DeviceAMGraph(ctx, batch)
    ref = O.am_build(batch.arrays())
    assert (am.num_graphs, am.total_nodes, am.adjacency_words) == (ref["G"], ref["N"], len(ref["adjacency"]))
    got = am.download()
    for k in ("graph_offsets", "num_nodes", "adjacency", "node_labels"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


@pytest.mark.parametrize("scope", [capi.SCOPE_QUERY, capi.SCOPE_DATA])
def test_am_signatures_parity(ctx, batch, scope):
    am = capi.DeviceAMGraph(ctx, batch)
    ref = O.am_build(batch.arrays())
    n = am.total_nodes
    sig = capi.Signature(ctx, n, n)
    print(f"AM generate scope={scope}", flush=True)
    sig.generate_am(am, scope=scope)
    want = O.am_generate(ref)
    np.testing.assert_array_equal(sig.download(scope), want)
    sig.generate_am(am, scope=scope)  # onto the existing values
    O.am_generate(ref, want)
    np.testing.assert_array_equal(sig.download(scope), want)
    for view in (1, 2, 3, 4, 6):
        print(f"AM PowerGraph refine scope={scope} view={view}", flush=True)
        sig.refine_am(am, view, scope=scope)
        O.am_refine_power_graph(ref, view, want)
        np.testing.assert_array_equal(sig.download(scope), want)
    vb = capi.Signature(ctx, n, n)
    vb.generate_am(am, scope=scope)
    want = O.am_generate(ref)
    for view in (1, 2, 3):
        print(f"AM ViewBased refine scope={scope} view={view}", flush=True)
        vb.refine_am_view_based(am, view, scope=scope)
        O.am_refine_view_based(ref, view, want)
        np.testing.assert_array_equal(vb.download(scope), want)


def test_am_data_planes_follow_the_signatures(ctx, batch):
    # The data scope also writes the bit-sliced planes the candidate refine reads: a refine
    # over them equals the oracle's refine over the downloaded signatures.
    batch = batch.slice(0, 300)
    am = capi.DeviceAMGraph(ctx, batch)
    csr = capi.DeviceGraph(ctx, batch)
    g = batch.arrays()
    n = am.total_nodes
    sig = capi.Signature(ctx, n, n)
    sig.generate_query(csr)
    cand = capi.Candidates(ctx, n, n)
    capi.filter_candidates(ctx, csr, csr, sig, cand)
    words = O.filter_candidates(g, g)
    qsig = O.generate_signatures(g)
    ref = O.am_build(g)
    dsig = O.am_generate(ref)
    print("AM data planes", flush=True)
    sig.generate_am(am, scope=capi.SCOPE_DATA)
    sig.refine_am(am, 2, scope=capi.SCOPE_DATA)
    O.am_refine_power_graph(ref, 2, dsig)
    np.testing.assert_array_equal(sig.download(capi.SCOPE_DATA), dsig)
    capi.refine_candidates(ctx, csr, csr, sig, cand)
    O.refine_candidates(qsig, dsig, words)
    np.testing.assert_array_equal(cand.download(), words)


def test_am_rejects_large_graphs_and_mismatched_signatures(ctx):
    big = capi.HostGraphs.from_edge_lists([([0] * 65, [(i, i + 1) for i in range(64)])])
    with pytest.raises(capi.SigmoError):
        capi.DeviceAMGraph(ctx, big)
    small = capi.HostGraphs.from_edge_lists([([0, 1, 2], [(0, 1), (1, 2)])])
    am = capi.DeviceAMGraph(ctx, small)
    sig = capi.Signature(ctx, 5, 4)
    with pytest.raises(capi.SigmoError):
        sig.generate_am(am)
    with pytest.raises(capi.SigmoError):
        sig.refine_am(am, 1, scope=capi.SCOPE_DATA)
