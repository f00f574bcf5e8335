"""GPU tests of streaming (sigmo_stream_run): a molecule shard run in chunks, with and without
the overlapped loader, gives the GMCR and per-pair counts of one run over the whole shard, in
both forms and both join modes, and those equal the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from sigmo_amd import capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shard(ctx):
    # LogNormal sizes: some molecules exceed 64 nodes; chunk sizes below leave a ragged last chunk
    dh = capi.HostGraphs.molecules(3, 3000, capi.SIZE_LOGNORMAL)
    qh = capi.HostGraphs.queries(3, 40, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)
    qd = capi.DeviceGraph(ctx, qh)
    o = O.pipeline(qh.arrays(), dh.arrays(), 6, modes=[True, False])
    return dh, qh, qd, o


@pytest.mark.parametrize("form", ["plain", "batched"])
@pytest.mark.parametrize("chunk,overlap", [(0, True), (997, True), (997, False), (400, True)])
def test_stream_chunks_equal_one_run(ctx, shard, form, chunk, overlap):
    dh, qh, qd, o = shard
    for find_first in (True, False):
        print(f"stream form={form} chunk={chunk} overlap={overlap} find_first={find_first}", flush=True)
        r = capi.stream_run(ctx, qd, dh, chunk_molecules=chunk, form=form, find_first=find_first, overlap=overlap, collect=True)
        key = "first" if find_first else "all"
        np.testing.assert_array_equal(r["offsets"], o["offsets"], err_msg="GMCR offsets")
        np.testing.assert_array_equal(r["indices"], o["indices"], err_msg="GMCR indices")
        np.testing.assert_array_equal(r["counts"], o[f"counts_{key}"], err_msg=f"counts ({key})")
        assert r["num_matches"] == o[f"total_{key}"]
        assert r["gmcr_pairs"] == len(o["indices"])
        assert r["matched_pairs"] == int((o[f"counts_{key}"] > 0).sum())
        assert r["chunks"] == (1 if chunk == 0 else -(-dh.num_graphs // chunk))


def test_stream_growing_chunks_and_forms_agree(ctx, shard):
    """Chunks whose node counts grow make the loader's buffers and the per-chunk pipeline and
    database grow between chunks; both forms still agree on every total."""
    _, qh, qd, _ = shard
    small = capi.HostGraphs.molecules(5, 300, capi.SIZE_NORMAL)
    big = capi.HostGraphs.molecules(5, 600, capi.SIZE_LOGNORMAL, first=300)
    dh = small.concat(big)
    res = {}
    for form in ("plain", "batched"):
        print(f"stream growing form={form}", flush=True)
        res[form] = capi.stream_run(ctx, qd, dh, chunk_molecules=300, form=form, find_first=False, collect=True)
    for k in ("offsets", "indices", "counts"):
        np.testing.assert_array_equal(res["plain"][k], res["batched"][k])
    o = O.pipeline(qh.arrays(), dh.arrays(), 6, modes=[False])
    np.testing.assert_array_equal(res["batched"]["counts"], o["counts_all"])


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_stream_session_passes_repeat_the_result(ctx, shard, form):
    """A session (sigmo_stream_open / pass / close) keeps its device buffers between passes; every
    pass gives the oracle's GMCR and counts."""
    dh, qh, qd, o = shard
    sess = capi.Stream(ctx, qd, dh, chunk_molecules=997, form=form, find_first=False)
    for p in range(3):
        print(f"stream session form={form} pass {p}", flush=True)
        r = sess.pass_(collect=True)
        np.testing.assert_array_equal(r["offsets"], o["offsets"], err_msg="GMCR offsets")
        np.testing.assert_array_equal(r["indices"], o["indices"], err_msg="GMCR indices")
        np.testing.assert_array_equal(r["counts"], o["counts_all"], err_msg="counts")
        assert r["chunks"] == 4
    sess.close()


@pytest.mark.parametrize("workers", [2, 3])
def test_stream_workers_equal_one_run(ctx, shard, workers):
    # chunks on several contexts at once (batched form): results in chunk order, equal to the oracle
    dh, qh, qd, o = shard
    for find_first in (True, False):
        print(f"stream workers={workers} find_first={find_first}", flush=True)
        r = capi.stream_run(ctx, qd, dh, chunk_molecules=400, form="batched", find_first=find_first, collect=True, workers=workers)
        key = "first" if find_first else "all"
        np.testing.assert_array_equal(r["offsets"], o["offsets"], err_msg="GMCR offsets")
        np.testing.assert_array_equal(r["indices"], o["indices"], err_msg="GMCR indices")
        np.testing.assert_array_equal(r["counts"], o[f"counts_{key}"], err_msg=f"counts ({key})")
        assert r["num_matches"] == o[f"total_{key}"]
        assert r["chunks"] == -(-dh.num_graphs // 400)
    # a session keeps its workers' buffers: a second pass gives the same digests
    s = capi.Stream(ctx, qd, dh, chunk_molecules=400, form="batched", workers=workers)
    a, b = s.pass_(digest=True), s.pass_(digest=True)
    s.close()
    assert a["chunk_digests"] == b["chunk_digests"]
