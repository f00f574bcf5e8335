"""The serving loop's scheduler (sigmo_serve_schedule) on the host, with a Python executor that
stands in for the device: admission order, the memory budget, batch limits, deadlines,
rejection, overlap of workers and error propagation.  No GPU is needed; the device executor
(sigmo_serve) is covered by tests/test_gpu_serve.py."""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from sigmo_amd import capi


def _requests(graphs_per_request, gap_s):
    rgo = np.concatenate([[0], np.cumsum(graphs_per_request)]).astype(np.uint32)
    arr = np.arange(len(graphs_per_request), dtype=np.float64) * gap_s
    return rgo, arr


class Recorder:
    """Executor that sleeps `run_s` per batch and reports request ids as statistics."""

    def __init__(self, run_s=0.0, fail_on_batch=None):
        self.run_s = run_s
        self.batches = []
        self.lock = threading.Lock()
        self.active = 0
        self.max_active = 0
        self.fail_on_batch = fail_on_batch

    def __call__(self, worker, reqs):
        with self.lock:
            self.batches.append((worker, list(reqs)))
            idx = len(self.batches) - 1
            self.active += 1
            self.max_active = max(self.max_active, self.active)
        try:
            if self.fail_on_batch is not None and idx == self.fail_on_batch:
                raise RuntimeError("planted executor failure")
            time.sleep(self.run_s)
            return [(r, 10 * r, 1000 + r) for r in reqs], 7 * len(reqs)
        finally:
            with self.lock:
                self.active -= 1


def test_every_request_served_in_arrival_order():
    rgo, arr = _requests([3, 1, 4, 1, 5, 9, 2, 6], 0.002)
    rec = Recorder(run_s=0.005)
    out = capi.serve_schedule(1, rgo, arr, rec)
    assert (out["status"] == capi.SERVE_DONE).all()
    served = [r for _, reqs in rec.batches for r in reqs]
    assert served == list(range(8))  # FIFO, each request exactly once
    np.testing.assert_array_equal(out["matched_pairs"], np.arange(8))
    np.testing.assert_array_equal(out["embeddings"], 10 * np.arange(8))
    np.testing.assert_array_equal(out["hash"], 1000 + np.arange(8, dtype=np.uint64))
    assert (out["admit_s"] >= arr - 1e-9).all()
    assert (out["done_s"] >= out["admit_s"]).all()
    np.testing.assert_allclose(out["latency_ms"], (out["done_s"] - arr) * 1e3, rtol=0, atol=1e-6)
    s = out["summary"]
    assert s["done"] == 8 and s["batches"] == len(rec.batches)
    assert s["matched_pairs"] == sum(range(8)) and s["gmcr_pairs"] == 7 * 8


def test_requests_arriving_during_a_batch_join_the_next_one():
    """Continuous batching: a batch takes every request that arrived while the previous ran."""
    rgo, arr = _requests([1] * 12, 0.004)
    rec = Recorder(run_s=0.02)
    out = capi.serve_schedule(1, rgo, arr, rec)
    assert (out["status"] == capi.SERVE_DONE).all()
    sizes = [len(reqs) for _, reqs in rec.batches]
    assert sizes[0] == 1 and max(sizes) >= 3  # later batches carry the backlog
    for b, (_, reqs) in enumerate(rec.batches):
        assert (out["batch"][reqs] == b).all()
        assert (out["batch_requests"][reqs] == len(reqs)).all()


def test_batch_limits():
    rgo, arr = _requests([2, 2, 2, 2, 2, 2, 2, 2, 7], 0.0)
    rec = Recorder()
    out = capi.serve_schedule(1, rgo, arr, rec, max_batch_graphs=5, max_batch_requests=3)
    assert (out["status"] == capi.SERVE_DONE).all()
    for _, reqs in rec.batches:
        g = int(sum(rgo[r + 1] - rgo[r] for r in reqs))
        assert len(reqs) <= 3
        assert g <= 5 or len(reqs) == 1  # a larger request still runs, alone
    assert [8] in [reqs for _, reqs in rec.batches]


def test_memory_budget_bounds_bytes_in_flight_and_rejects_oversized_requests():
    rgo, arr = _requests([4, 4, 4, 20, 4, 4, 4, 4], 0.0)
    rec = Recorder(run_s=0.01)
    budget, per_graph, per_batch = 10_000, 1_000, 500
    out = capi.serve_schedule(3, rgo, arr, rec, memory_budget=budget, bytes_per_graph=per_graph, bytes_per_batch=per_batch)
    status = out["status"]
    assert status[3] == capi.SERVE_REJECTED  # 20 graphs cost 20,500 B > budget
    assert (np.delete(status, 3) == capi.SERVE_DONE).all()
    assert out["summary"]["rejected"] == 1 and out["summary"]["done"] == 7
    assert 0 < out["summary"]["max_inflight_bytes"] <= budget
    for _, reqs in rec.batches:
        assert 3 not in reqs
        assert per_batch + per_graph * sum(int(rgo[r + 1] - rgo[r]) for r in reqs) <= budget


def test_workers_overlap():
    rgo, arr = _requests([1] * 6, 0.0)
    rec = Recorder(run_s=0.05)
    out = capi.serve_schedule(2, rgo, arr, rec, max_batch_requests=1)
    assert (out["status"] == capi.SERVE_DONE).all()
    assert rec.max_active == 2
    assert set(out["worker"].tolist()) == {0, 1}
    # busy time counts overlapping batches once
    s = out["summary"]
    assert s["busy_s"] <= s["wall_s"] + 1e-6
    assert s["busy_s"] < 6 * 0.05


def test_deadline_expires_waiting_requests():
    rgo, arr = _requests([1] * 6, 0.0)
    rec = Recorder(run_s=0.08)
    out = capi.serve_schedule(1, rgo, arr, rec, max_batch_requests=1, deadline_ms=50.0)
    status = out["status"]
    assert status[0] == capi.SERVE_DONE
    assert (status[1:] == capi.SERVE_EXPIRED).all()  # each waited > 50 ms behind the first
    assert out["summary"]["expired"] == 5
    assert [reqs for _, reqs in rec.batches] == [[0]]


def test_executor_failure_stops_the_loop():
    rgo, arr = _requests([1] * 5, 0.0)
    rec = Recorder(fail_on_batch=1)
    with pytest.raises(capi.SigmoError, match="serving batch 1"):
        capi.serve_schedule(1, rgo, arr, rec, max_batch_requests=1)


@pytest.mark.parametrize("rgo,arr", [
    ([0, 1, 1], [0.0, 0.1]),       # a request with no query graph
    ([0, 1, 2], [0.2, 0.1]),       # arrivals out of order
    ([0, 1, 2], [-0.1, 0.1]),      # negative arrival
])
def test_invalid_requests(rgo, arr):
    with pytest.raises(capi.SigmoError):
        capi.serve_schedule(1, rgo, arr, Recorder())


def test_no_requests():
    out = capi.serve_schedule(2, [0], [], Recorder())
    assert out["summary"]["batches"] == 0


# ---- online server (sigmo_server_open_schedule): requests submitted while the loop runs ------------

def test_online_server_serves_submissions_as_they_come():
    rec = Recorder(run_s=0.01)
    srv = capi.Server(num_workers=1, executor=rec, max_batch_requests=16)
    ids = []
    for i in range(10):
        ids.append(srv.submit(1 + i % 3))
        time.sleep(0.003)
    assert ids == list(range(10))
    res = [srv.poll(i, timeout_ms=-1) for i in ids]
    s = srv.close()
    assert all(r["status"] == capi.SERVE_DONE for r in res)
    assert [r["matched_pairs"] for r in res] == ids and [r["hash"] for r in res] == [1000 + i for i in ids]
    served = [r for _, reqs in rec.batches for r in reqs]
    assert served == ids  # FIFO, each request once
    sizes = [len(reqs) for _, reqs in rec.batches]
    assert max(sizes) >= 2  # submissions during a batch joined the next one
    assert all(r["done_s"] >= r["admit_s"] >= 0 for r in res)
    assert s["done"] == 10 and s["batches"] == len(rec.batches) and s["matched_pairs"] == sum(ids)


def test_online_server_poll_without_waiting_sees_pending():
    gate = threading.Event()

    def slow(worker, reqs):
        gate.wait(5)
        return [(r, 0, 0) for r in reqs], 0

    srv = capi.Server(num_workers=1, executor=slow)
    rid = srv.submit(2)
    assert srv.poll(rid, timeout_ms=0)["status"] == capi.SERVE_PENDING
    assert srv.poll(rid, timeout_ms=20)["status"] == capi.SERVE_PENDING  # timed out
    gate.set()
    assert srv.poll(rid, timeout_ms=-1)["status"] == capi.SERVE_DONE
    with pytest.raises(capi.SigmoError):
        srv.poll(rid + 1, timeout_ms=0)  # unknown id
    srv.close()


def test_online_server_threads_budget_and_close_drains_the_queue():
    rec = Recorder(run_s=0.01)
    srv = capi.Server(num_workers=2, executor=rec, memory_budget=1000, bytes_per_graph=100, max_batch_requests=4)

    def client(k):
        return [srv.submit(1 + (k + i) % 4) for i in range(6)]

    with __import__("concurrent.futures").futures.ThreadPoolExecutor(3) as ex:
        ids = sorted(i for part in ex.map(client, range(3)) for i in part)
    big = srv.submit(11)  # 1100 bytes > the budget: rejected
    s = srv.close()  # serves everything already queued
    assert ids == list(range(18))
    assert s["done"] == 18 and s["rejected"] == 1
    assert s["max_inflight_bytes"] <= 1000
    served = sorted(r for _, reqs in rec.batches for r in reqs)
    assert served == ids and big not in served


def test_online_server_executor_failure_is_reported():
    rec = Recorder(fail_on_batch=0)
    srv = capi.Server(num_workers=1, executor=rec)
    rid = srv.submit(1)
    with pytest.raises(capi.SigmoError, match="serving batch 0"):
        srv.poll(rid, timeout_ms=-1)
    with pytest.raises(capi.SigmoError):
        srv.submit(1)
    with pytest.raises(capi.SigmoError):
        srv.close()


def test_online_server_argument_checks():
    srv = capi.Server(num_workers=1, executor=Recorder())
    with pytest.raises(capi.SigmoError):
        srv.submit(0)  # no query graph
    with pytest.raises(capi.SigmoError):
        srv.submit(capi.HostGraphs.queries(1, 2, 4, 8))  # a scheduling server takes counts
    srv.close()
    with pytest.raises(capi.SigmoError):
        capi.Server(num_workers=0, executor=Recorder())
