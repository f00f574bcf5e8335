"""ctypes binding of ``libsigmo_hip.so`` (the C ABI declared in ``include/sigmo_hip.h``).

Thin by design: every call goes straight to the native library, whose entry points mirror
the reference classes (paths relative to the SIGMo repository):

=====================================  ==========================================================
Python                                 reference
=====================================  ==========================================================
``HostGraphs.load_dat``                ``io::loadCSRGraphsFromFile`` (library/include/io.hpp:46-55)
``dat_shard_lines``                    ``loadFileLinesMPI`` (library/src/sigmo_mpi.cpp:18-80)
``HostGraphs.load_dat_shard``          that, then ``io::loadCSRGraphsFromLines`` (io.hpp:37-44)
``Context``                            the ``sycl::queue`` (library/src/sigmo.cpp:19)
``DeviceGraph``                        ``createDeviceCSRGraph`` (library/include/graph.hpp:255-314)
``Signature``                          ``signature::Signature<>`` (library/include/signature.hpp:20-212)
``Candidates``                         ``candidates::Candidates`` (library/include/candidates.hpp:14-204)
``filter_candidates``                  ``isomorphism::filter::filterCandidates`` (isomorphism.hpp:22-62)
``refine_candidates``                  ``isomorphism::filter::refineCandidates`` (isomorphism.hpp:64-127)
``GMCR``                               ``isomorphism::mapping::GMCR`` (library/include/gmcr.hpp:18-135)
``join_candidates``                    ``isomorphism::join::joinCandidates`` (isomorphism.hpp:270-381)
=====================================  ==========================================================

There is no CPU fallback: importing works without a GPU (so the host-side generator and
.dat I/O stay usable), but every device entry point fails with :class:`SigmoError` if the
library is missing or no HIP device is present.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("SIGMO_HIP_LIB", _PKG / "libsigmo_hip.so"))

OK = 0
ERR_INVALID, ERR_NOMEM, ERR_HIP, ERR_UNSUPPORTED, ERR_OVERFLOW, ERR_IO = -1, -2, -3, -4, -5, -6
SCOPE_DATA, SCOPE_QUERY = 0, 1
PHASE_SIG_DATA, PHASE_SIG_QUERY, PHASE_FILTER, PHASE_GMCR, PHASE_JOIN = range(5)
PHASE_NAMES = ("sig_data", "sig_query", "filter", "gmcr", "join")
SIZE_NORMAL, SIZE_LOGNORMAL = 0, 1

# Every exported symbol of include/sigmo_hip.h, for the load check.
EXPORTED = (
    "sigmo_last_error", "sigmo_ctx_create", "sigmo_ctx_destroy", "sigmo_ctx_sync", "sigmo_phase_ms",
    "sigmo_phase_reset", "sigmo_device_info", "sigmo_host_graphs_from_arrays", "sigmo_host_graphs_load_dat",
    "sigmo_host_graphs_parse_dat", "sigmo_dat_shard_lines", "sigmo_host_graphs_load_dat_shard", "sigmo_host_graphs_save_dat", "sigmo_host_graphs_info", "sigmo_host_graphs_arrays",
    "sigmo_host_graphs_slice", "sigmo_host_graphs_concat", "sigmo_host_graphs_free", "sigmo_gen_molecules",
    "sigmo_gen_queries", "sigmo_gen_arrivals", "sigmo_graph_upload", "sigmo_graph_free", "sigmo_graph_info",
    "sigmo_graph_alloc_size", "sigmo_signatures_create", "sigmo_signatures_free", "sigmo_signatures_generate",
    "sigmo_signatures_refine", "sigmo_signatures_download", "sigmo_signatures_refine_view_based", "sigmo_candidates_filter_data", "sigmo_signatures_alloc_size", "sigmo_candidates_create",
    "sigmo_am_graph_upload", "sigmo_am_graph_free", "sigmo_am_graph_info", "sigmo_am_graph_download",
    "sigmo_signatures_generate_am", "sigmo_signatures_refine_am", "sigmo_signatures_refine_am_view_based",
    "sigmo_candidates_free", "sigmo_candidates_filter", "sigmo_candidates_refine", "sigmo_candidates_download",
    "sigmo_candidates_counts", "sigmo_candidates_alloc_size", "sigmo_gmcr_create", "sigmo_gmcr_free",
    "sigmo_gmcr_generate", "sigmo_gmcr_total", "sigmo_gmcr_download", "sigmo_join", "sigmo_join_wildcard", "sigmo_pipeline_run",
    "sigmo_pipeline_create", "sigmo_pipeline_execute", "sigmo_pipeline_pairs", "sigmo_pipeline_free",
    "sigmo_pipeline_capture", "sigmo_pipeline_replay",
    "sigmo_kernel_stats", "sigmo_event_mark", "sigmo_event_elapsed", "sigmo_db_prepare", "sigmo_db_free",
    "sigmo_db_alloc_size", "sigmo_db_refresh", "sigmo_batch_create", "sigmo_batch_run", "sigmo_batch_pairs", "sigmo_batch_free",
    "sigmo_batch_request_stats", "sigmo_batch_workspace_bytes", "sigmo_serve_schedule", "sigmo_serve", "sigmo_stream_run",
    "sigmo_server_open", "sigmo_server_submit", "sigmo_server_poll", "sigmo_server_close", "sigmo_server_open_schedule",
    "sigmo_server_submit_count",
    "sigmo_batch_set_tuning", "sigmo_batch_get_tuning", "sigmo_batch_autotune", "sigmo_batch_capture", "sigmo_batch_replay",
    "sigmo_stream_open", "sigmo_stream_pass", "sigmo_stream_close", "sigmo_join_noalloc",
)
KERNEL_NAMES = ("sig_generate", "sig_refine", "filter", "refine", "gmcr", "join", "pair_filter", "db_prepare")


class SigmoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class PipelineResult(C.Structure):
    _fields_ = [
        ("num_matches", C.c_uint64),
        ("gmcr_pairs", C.c_uint64),
        ("matched_pairs", C.c_uint64),
        ("phase_ms", C.c_double * 5),
        ("total_ms", C.c_double),
        ("dfs_steps", C.c_uint64),
        ("join_wave_steps", C.c_uint64),
    ]


class BatchResult(C.Structure):
    _fields_ = [
        ("num_matches", C.c_uint64),
        ("matched_pairs", C.c_uint64),
        ("gmcr_pairs", C.c_uint64),
        ("dfs_steps", C.c_uint64),
        ("filter_wave_steps", C.c_uint64),
        ("filter_lane_steps", C.c_uint64),
        ("filter_scans", C.c_uint64),
        ("join_wave_steps", C.c_uint64),
        ("one_sync", C.c_uint64),
        ("arena_repeats", C.c_uint64),
    ]


class BatchTuning(C.Structure):
    _fields_ = [("filter_blocks_per_cu", C.c_uint32), ("filter_claim", C.c_uint32), ("join_blocks_per_cu", C.c_uint32),
                ("join_chunk", C.c_uint32), ("filter_mode", C.c_uint32)]


class RequestStats(C.Structure):
    _fields_ = [("matched_pairs", C.c_uint64), ("embeddings", C.c_uint64), ("hash", C.c_uint64)]


class ServeOptions(C.Structure):
    _fields_ = [
        ("max_batch_requests", C.c_uint32),
        ("max_batch_graphs", C.c_uint32),
        ("memory_budget", C.c_uint64),
        ("bytes_per_graph", C.c_uint64),
        ("bytes_per_batch", C.c_uint64),
        ("deadline_ms", C.c_double),
        ("find_first", C.c_int),
    ]


class ServeResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("batch", C.c_uint32),
        ("batch_requests", C.c_uint32),
        ("worker", C.c_uint32),
        ("admit_s", C.c_double),
        ("done_s", C.c_double),
        ("latency_ms", C.c_double),
        ("matched_pairs", C.c_uint64),
        ("embeddings", C.c_uint64),
        ("hash", C.c_uint64),
    ]


class ServeSummary(C.Structure):
    _fields_ = [
        ("batches", C.c_uint32),
        ("done", C.c_uint32),
        ("rejected", C.c_uint32),
        ("expired", C.c_uint32),
        ("wall_s", C.c_double),
        ("busy_s", C.c_double),
        ("max_inflight_bytes", C.c_uint64),
        ("matched_pairs", C.c_uint64),
        ("embeddings", C.c_uint64),
        ("gmcr_pairs", C.c_uint64),
    ]


SERVE_EXECUTOR = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(RequestStats),
                             C.POINTER(C.c_uint64))
SERVE_DONE, SERVE_REJECTED, SERVE_EXPIRED, SERVE_PENDING = 0, 1, 2, 3


class StreamOptions(C.Structure):
    _fields_ = [("chunk_molecules", C.c_uint32), ("refinement_steps", C.c_uint32), ("find_first", C.c_int), ("form", C.c_int),
                ("overlap", C.c_int), ("workers", C.c_uint32)]


class StreamResult(C.Structure):
    _fields_ = [
        ("num_matches", C.c_uint64),
        ("matched_pairs", C.c_uint64),
        ("gmcr_pairs", C.c_uint64),
        ("dfs_steps", C.c_uint64),
        ("chunks", C.c_uint32),
        ("wall_ms", C.c_double),
        ("compute_ms", C.c_double),
        ("load_ms", C.c_double),
        ("load_wait_ms", C.c_double),
    ]


STREAM_CHUNK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                              C.POINTER(C.c_int32), C.c_uint64)
FORM_PLAIN, FORM_BATCHED = 0, 1

_lib = None


def lib() -> C.CDLL:
    """Load the native library (raises SigmoError if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SigmoError(ERR_INVALID, f"{LIB_PATH} is missing: run `python -m sigmo_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        vp, u32, u64, i32, dbl = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
        pvp = C.POINTER(C.c_void_p)
        sig = {
            "sigmo_last_error": ([], C.c_char_p),
            "sigmo_ctx_create": ([i32, pvp], i32),
            "sigmo_ctx_destroy": ([vp], i32),
            "sigmo_ctx_sync": ([vp], i32),
            "sigmo_phase_ms": ([vp, i32, C.POINTER(dbl)], i32),
            "sigmo_phase_reset": ([vp], i32),
            "sigmo_device_info": ([vp, C.c_char_p, C.c_size_t, C.POINTER(u64)], i32),
            "sigmo_host_graphs_from_arrays": ([u32, vp, vp, vp, vp, vp, pvp], i32),
            "sigmo_host_graphs_load_dat": ([C.c_char_p, pvp], i32),
            "sigmo_host_graphs_parse_dat": ([C.c_char_p, C.c_size_t, pvp], i32),
            "sigmo_dat_shard_lines": ([C.c_char_p, u32, u32, u32, vp, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(u32)], i32),
            "sigmo_host_graphs_load_dat_shard": ([C.c_char_p, u32, u32, u32, pvp], i32),
            "sigmo_host_graphs_save_dat": ([vp, C.c_char_p], i32),
            "sigmo_host_graphs_info": ([vp, C.POINTER(u32), C.POINTER(u64), C.POINTER(u64)], i32),
            "sigmo_host_graphs_arrays": ([vp, pvp, pvp, pvp, pvp, pvp], i32),
            "sigmo_host_graphs_slice": ([vp, u32, u32, pvp], i32),
            "sigmo_host_graphs_concat": ([vp, vp, pvp], i32),
            "sigmo_host_graphs_free": ([vp], i32),
            "sigmo_gen_molecules": ([u64, u64, u64, i32, i32, pvp], i32),
            "sigmo_gen_queries": ([u64, u32, u32, u32, i32, dbl, dbl, pvp], i32),
            "sigmo_gen_arrivals": ([u64, u32, dbl, u32, u32, vp, vp], i32),
            "sigmo_graph_upload": ([vp, vp, pvp], i32),
            "sigmo_graph_free": ([vp], i32),
            "sigmo_graph_info": ([vp, C.POINTER(u32), C.POINTER(u64), C.POINTER(u64)], i32),
            "sigmo_graph_alloc_size": ([vp, C.POINTER(u64), C.POINTER(u64)], i32),
            "sigmo_signatures_create": ([vp, u64, u64, pvp], i32),
            "sigmo_signatures_free": ([vp], i32),
            "sigmo_signatures_generate": ([vp, vp, vp, i32], i32),
            "sigmo_signatures_refine": ([vp, vp, vp, i32, u32], i32),
            "sigmo_signatures_download": ([vp, vp, i32, vp, u64], i32),
            "sigmo_signatures_refine_view_based": ([vp, vp, vp, i32, u32], i32),
            "sigmo_am_graph_upload": ([vp, vp, pvp], i32),
            "sigmo_am_graph_free": ([vp], i32),
            "sigmo_am_graph_info": ([vp, C.POINTER(u32), C.POINTER(u64), C.POINTER(u64)], i32),
            "sigmo_am_graph_download": ([vp, vp, vp, vp, vp, vp], i32),
            "sigmo_signatures_generate_am": ([vp, vp, vp, i32], i32),
            "sigmo_signatures_refine_am": ([vp, vp, vp, i32, u32], i32),
            "sigmo_signatures_refine_am_view_based": ([vp, vp, vp, i32, u32], i32),
            "sigmo_candidates_filter_data": ([vp, vp, vp, vp], i32),
            "sigmo_signatures_alloc_size": ([vp, i32, C.POINTER(u64)], i32),
            "sigmo_candidates_create": ([vp, u64, u64, pvp], i32),
            "sigmo_candidates_free": ([vp], i32),
            "sigmo_candidates_filter": ([vp, vp, vp, vp, vp], i32),
            "sigmo_candidates_refine": ([vp, vp, vp, vp, vp], i32),
            "sigmo_candidates_download": ([vp, vp, vp, u64], i32),
            "sigmo_candidates_counts": ([vp, vp, vp, u64], i32),
            "sigmo_candidates_alloc_size": ([vp, C.POINTER(u64), C.POINTER(u64)], i32),
            "sigmo_gmcr_create": ([vp, pvp], i32),
            "sigmo_gmcr_free": ([vp], i32),
            "sigmo_gmcr_generate": ([vp, vp, vp, vp, vp], i32),
            "sigmo_gmcr_total": ([vp, C.POINTER(u64)], i32),
            "sigmo_gmcr_download": ([vp, vp, vp, u64, vp, u64], i32),
            "sigmo_join": ([vp, vp, vp, vp, vp, i32, C.POINTER(u64), vp], i32),
            "sigmo_join_noalloc": ([vp, vp, vp, vp, vp, i32, C.POINTER(u64), vp], i32),
            "sigmo_join_wildcard": ([vp, vp, vp, vp, vp, i32, C.POINTER(u64), vp], i32),
            "sigmo_pipeline_run": ([vp, vp, vp, u32, i32, C.POINTER(PipelineResult)], i32),
            "sigmo_pipeline_create": ([vp, vp, vp, pvp], i32),
            "sigmo_pipeline_execute": ([vp, u32, i32, C.POINTER(PipelineResult)], i32),
            "sigmo_pipeline_pairs": ([vp, vp, u64, vp, vp, u64], i32),
            "sigmo_pipeline_capture": ([vp, u32, i32], i32),
            "sigmo_pipeline_replay": ([vp, C.POINTER(PipelineResult)], i32),
            "sigmo_pipeline_free": ([vp], i32),
            "sigmo_kernel_stats": ([vp, i32, C.POINTER(u64), C.POINTER(dbl), C.POINTER(dbl)], i32),
            "sigmo_event_mark": ([vp, C.POINTER(u64)], i32),
            "sigmo_event_elapsed": ([vp, u64, u64, C.POINTER(dbl)], i32),
            "sigmo_db_prepare": ([vp, vp, u32, pvp], i32),
            "sigmo_db_free": ([vp], i32),
            "sigmo_db_refresh": ([vp], i32),
            "sigmo_db_alloc_size": ([vp, C.POINTER(u64)], i32),
            "sigmo_batch_create": ([vp, vp, pvp], i32),
            "sigmo_batch_run": ([vp, vp, i32, C.POINTER(BatchResult)], i32),
            "sigmo_batch_pairs": ([vp, vp, u64, vp, vp, u64], i32),
            "sigmo_batch_free": ([vp], i32),
            "sigmo_batch_request_stats": ([vp, vp, u32, C.POINTER(RequestStats)], i32),
            "sigmo_batch_set_tuning": ([vp, C.POINTER(BatchTuning)], i32),
            "sigmo_batch_get_tuning": ([vp, C.POINTER(BatchTuning)], i32),
            "sigmo_batch_autotune": ([vp, vp, i32, u32, C.POINTER(BatchTuning), C.POINTER(dbl)], i32),
            "sigmo_batch_capture": ([vp, vp, i32, i32], i32),
            "sigmo_batch_replay": ([vp, C.POINTER(BatchResult)], i32),
            "sigmo_batch_workspace_bytes": ([vp, u32, u64, C.POINTER(u64)], i32),
            "sigmo_serve_schedule": ([u32, vp, vp, u32, C.POINTER(ServeOptions), SERVE_EXECUTOR, vp, C.POINTER(ServeResult),
                                      C.POINTER(ServeSummary)], i32),
            "sigmo_serve": ([vp, u32, vp, vp, vp, u32, C.POINTER(ServeOptions), C.POINTER(ServeResult), C.POINTER(ServeSummary)], i32),
            "sigmo_stream_run": ([vp, vp, vp, C.POINTER(StreamOptions), STREAM_CHUNK_FN, vp, C.POINTER(StreamResult)], i32),
            "sigmo_stream_open": ([vp, vp, vp, C.POINTER(StreamOptions), C.POINTER(vp)], i32),
            "sigmo_stream_pass": ([vp, STREAM_CHUNK_FN, vp, C.POINTER(StreamResult)], i32),
            "sigmo_stream_close": ([vp], i32),
            "sigmo_server_open": ([vp, u32, C.POINTER(ServeOptions), C.POINTER(vp)], i32),
            "sigmo_server_submit": ([vp, vp, C.POINTER(u64)], i32),
            "sigmo_server_poll": ([vp, u64, dbl, C.POINTER(ServeResult)], i32),
            "sigmo_server_close": ([vp, C.POINTER(ServeSummary)], i32),
            "sigmo_server_open_schedule": ([u32, C.POINTER(ServeOptions), SERVE_EXECUTOR, vp, C.POINTER(vp)], i32),
            "sigmo_server_submit_count": ([vp, u32, C.POINTER(u64)], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != OK:
        raise SigmoError(status, lib().sigmo_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


# ---- host graphs ---------------------------------------------------------------------------------


class HostGraphs:
    """Host batched CSR (graph_offsets, row_offsets, column_indices, node_labels, edge_labels)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.sigmo_host_graphs_free(self._h)
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @classmethod
    def from_arrays(cls, graph_offsets, row_offsets, column_indices, node_labels, edge_labels=None) -> "HostGraphs":
        go = np.ascontiguousarray(graph_offsets, dtype=np.uint32)
        ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
        ci = np.ascontiguousarray(column_indices, dtype=np.uint32)
        nl = np.ascontiguousarray(node_labels, dtype=np.uint8)
        el = np.zeros(len(ci), np.uint8) if edge_labels is None else np.ascontiguousarray(edge_labels, dtype=np.uint8)
        h = C.c_void_p()
        check(lib().sigmo_host_graphs_from_arrays(len(go) - 1, _ptr(go), _ptr(ro), _ptr(ci) if len(ci) else None,
                                                  _ptr(nl) if len(nl) else None, _ptr(el) if len(el) else None, C.byref(h)))
        return cls(h)

    @classmethod
    def from_edge_lists(cls, graphs) -> "HostGraphs":
        """graphs: iterable of (labels, edges[, edge_labels]) with local ids."""
        lines = []
        for g in graphs:
            labels, edges = g[0], g[1]
            el = g[2] if len(g) > 2 else [1] * len(edges)
            toks = [f"n#{len(labels)}", "l#19"]
            toks += [f"{i} {int(l)}" for i, l in enumerate(labels)]
            toks.append(f"e#{len(edges)}")
            toks += [f"{u} {v} {int(l)}" for (u, v), l in zip(edges, el)]
            lines.append(" ".join(toks))
        return cls.parse_dat("\n".join(lines) + "\n")

    @classmethod
    def load_dat(cls, path) -> "HostGraphs":
        h = C.c_void_p()
        check(lib().sigmo_host_graphs_load_dat(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def load_dat_shard(cls, path, rank: int, nprocs: int, max_lines: int) -> "HostGraphs":
        """Rank `rank` of `nprocs`'s share of a .dat file, read as the reference's multi-node
        driver reads it (sigmo_mpi.cpp:18-80 then io.hpp:37-44); see dat_shard_lines."""
        h = C.c_void_p()
        check(lib().sigmo_host_graphs_load_dat_shard(str(path).encode(), rank, nprocs, max_lines, C.byref(h)))
        return cls(h)

    @classmethod
    def parse_dat(cls, text: str) -> "HostGraphs":
        b = text.encode()
        h = C.c_void_p()
        check(lib().sigmo_host_graphs_parse_dat(b, len(b), C.byref(h)))
        return cls(h)

    @classmethod
    def molecules(cls, seed: int, count: int, size_dist: int = SIZE_NORMAL, first: int = 0, threads: int = 0) -> "HostGraphs":
        h = C.c_void_p()
        check(lib().sigmo_gen_molecules(seed, first, count, size_dist, threads, C.byref(h)))
        return cls(h)

    @classmethod
    def queries(cls, seed: int, count: int, kmin: int, kmax: int, size_dist: int = SIZE_NORMAL,
                perturb_frac: float = 0.0, ring_frac: float = -1.0) -> "HostGraphs":
        h = C.c_void_p()
        check(lib().sigmo_gen_queries(seed, count, kmin, kmax, size_dist, perturb_frac, ring_frac, C.byref(h)))
        return cls(h)

    def save_dat(self, path) -> None:
        check(lib().sigmo_host_graphs_save_dat(self._h, str(path).encode()))

    def info(self):
        g, n, e = C.c_uint32(), C.c_uint64(), C.c_uint64()
        check(lib().sigmo_host_graphs_info(self._h, C.byref(g), C.byref(n), C.byref(e)))
        return g.value, n.value, e.value

    @property
    def num_graphs(self) -> int:
        return self.info()[0]

    @property
    def total_nodes(self) -> int:
        return self.info()[1]

    @property
    def total_edges(self) -> int:
        return self.info()[2]

    def arrays(self) -> dict:
        """Copies of the five batched-CSR arrays as numpy arrays."""
        G, N, E = self.info()
        p = [C.c_void_p() for _ in range(5)]
        check(lib().sigmo_host_graphs_arrays(self._h, *[C.byref(x) for x in p]))

        def view(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(ptr.value)
            return np.frombuffer(buf, dtype=dt).copy()

        return {
            "graph_offsets": view(p[0], G + 1, np.uint32),
            "row_offsets": view(p[1], N + 1, np.uint32),
            "column_indices": view(p[2], E, np.uint32),
            "node_labels": view(p[3], N, np.uint8),
            "edge_labels": view(p[4], E, np.uint8),
        }

    def slice(self, first: int, count: int) -> "HostGraphs":
        h = C.c_void_p()
        check(lib().sigmo_host_graphs_slice(self._h, first, count, C.byref(h)))
        return HostGraphs(h)

    def concat(self, other: "HostGraphs") -> "HostGraphs":
        h = C.c_void_p()
        check(lib().sigmo_host_graphs_concat(self._h, other._h, C.byref(h)))
        return HostGraphs(h)


def arrivals(seed: int, count: int, qps: float, qmin: int = 1, qmax: int = 50):
    t = np.zeros(count, np.float64)
    q = np.zeros(count, np.uint32)
    check(lib().sigmo_gen_arrivals(seed, count, qps, qmin, qmax, _ptr(t), _ptr(q)))
    return t, q


def dat_shard_lines(path, rank: int, nprocs: int, max_lines: int) -> list[bytes]:
    """The .dat lines rank `rank` of `nprocs` reads in the reference's multi-node driver
    (loadFileLinesMPI, sigmo_mpi.cpp:18-80): its byte range [rank * (size // nprocs), + 2048 *
    max_lines + 1024), the partial first line dropped on ranks > 0 and the partial last line on
    ranks < nprocs - 1, at most max_lines lines, then only those with exactly one 'n', 'e' and
    'l' (:72-77).  A NUL run past the end of the file comes back as one NUL byte."""
    n = C.c_size_t()
    k = C.c_uint32()
    check(lib().sigmo_dat_shard_lines(str(path).encode(), rank, nprocs, max_lines, None, 0, C.byref(n), C.byref(k)))
    buf = C.create_string_buffer(max(1, n.value))
    check(lib().sigmo_dat_shard_lines(str(path).encode(), rank, nprocs, max_lines, buf, n.value, C.byref(n), C.byref(k)))
    text = buf.raw[: n.value]
    lines = text.split(b"\n")[:-1] if text else []
    if len(lines) != k.value:
        raise SigmoError(ERR_IO, f"{path} changed between the two reads of its shard")
    return lines


# ---- device objects ---------------------------------------------------------------------------------


class Context:
    """One device + one HIP stream (the reference's sycl::queue)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().sigmo_ctx_create(device, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().sigmo_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    @property
    def handle(self):
        return self._h

    def sync(self):
        check(lib().sigmo_ctx_sync(self._h))

    def phase_ms(self) -> dict:
        out = {}
        for i, name in enumerate(PHASE_NAMES):
            v = C.c_double()
            check(lib().sigmo_phase_ms(self._h, i, C.byref(v)))
            out[name] = v.value
        return out

    def reset_phases(self):
        check(lib().sigmo_phase_reset(self._h))

    def kernel_stats(self) -> dict:
        """{kernel: (launches, total ms, total algorithmic HBM bytes)} since the last reset."""
        out = {}
        for i, name in enumerate(KERNEL_NAMES):
            n, ms, by = C.c_uint64(), C.c_double(), C.c_double()
            check(lib().sigmo_kernel_stats(self._h, i, C.byref(n), C.byref(ms), C.byref(by)))
            out[name] = (n.value, ms.value, by.value)
        return out

    def device_info(self):
        buf = C.create_string_buffer(256)
        mem = C.c_uint64()
        check(lib().sigmo_device_info(self._h, buf, 256, C.byref(mem)))
        return buf.value.decode(), mem.value


class DeviceGraph:
    def __init__(self, ctx: Context, host: HostGraphs):
        h = C.c_void_p()
        check(lib().sigmo_graph_upload(ctx.handle, host.handle, C.byref(h)))
        self._h, self.ctx = h, ctx
        g, n, e = C.c_uint32(), C.c_uint64(), C.c_uint64()
        check(lib().sigmo_graph_info(h, C.byref(g), C.byref(n), C.byref(e)))
        self.num_graphs, self.total_nodes, self.total_edges = g.value, n.value, e.value

    def free(self):
        if getattr(self, "_h", None):
            lib().sigmo_graph_free(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.free()

    @property
    def handle(self):
        return self._h

    def alloc_size(self):
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().sigmo_graph_alloc_size(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value


class DeviceAMGraph:
    """DeviceBatchedAMGraph (graph.hpp:114-148), made by createDeviceAMGraph (graph.hpp:345-392)
    from a host batch of graphs of at most 64 nodes."""

    def __init__(self, ctx: Context, host: HostGraphs):
        h = C.c_void_p()
        check(lib().sigmo_am_graph_upload(ctx.handle, host.handle, C.byref(h)))
        self._h, self.ctx = h, ctx
        g, n, w = C.c_uint32(), C.c_uint64(), C.c_uint64()
        check(lib().sigmo_am_graph_info(h, C.byref(g), C.byref(n), C.byref(w)))
        self.num_graphs, self.total_nodes, self.adjacency_words = g.value, n.value, w.value

    def free(self):
        if getattr(self, "_h", None):
            lib().sigmo_am_graph_free(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.free()

    @property
    def handle(self):
        return self._h

    def download(self) -> dict:
        """The device arrays: graph_offsets and num_nodes (cumulative) per graph, adjacency words,
        node labels."""
        go = np.zeros(self.num_graphs, np.uint32)
        nn = np.zeros(self.num_graphs, np.uint32)
        adj = np.zeros(self.adjacency_words, np.uint64)
        nl = np.zeros(self.total_nodes, np.uint8)
        check(lib().sigmo_am_graph_download(self.ctx.handle, self._h, _ptr(go), _ptr(nn), _ptr(adj), _ptr(nl)))
        return {"graph_offsets": go, "num_nodes": nn, "adjacency": adj, "node_labels": nl}


class Signature:
    def __init__(self, ctx: Context, data_nodes: int, query_nodes: int):
        h = C.c_void_p()
        check(lib().sigmo_signatures_create(ctx.handle, data_nodes, query_nodes, C.byref(h)))
        self._h, self.ctx = h, ctx
        self.data_nodes, self.query_nodes = data_nodes, query_nodes

    def free(self):
        if getattr(self, "_h", None):
            lib().sigmo_signatures_free(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.free()

    @property
    def handle(self):
        return self._h

    def generate_data(self, g: DeviceGraph):
        check(lib().sigmo_signatures_generate(self.ctx.handle, self._h, g.handle, SCOPE_DATA))

    def generate_query(self, g: DeviceGraph):
        check(lib().sigmo_signatures_generate(self.ctx.handle, self._h, g.handle, SCOPE_QUERY))

    def refine_data(self, g: DeviceGraph, view_size: int = 1):
        check(lib().sigmo_signatures_refine(self.ctx.handle, self._h, g.handle, SCOPE_DATA, view_size))

    def refine_query(self, g: DeviceGraph, view_size: int = 1):
        check(lib().sigmo_signatures_refine(self.ctx.handle, self._h, g.handle, SCOPE_QUERY, view_size))

    def refine_view_based(self, g: DeviceGraph, scope: int, view_size: int = 1):
        """refineCSRSignatures<Algorithm::ViewBased> on the data or query scope."""
        check(lib().sigmo_signatures_refine_view_based(self.ctx.handle, self._h, g.handle, scope, view_size))

    def generate_am(self, g: DeviceAMGraph, scope: int = SCOPE_QUERY):
        """generateAMSignatures (signature.hpp:100-120)."""
        check(lib().sigmo_signatures_generate_am(self.ctx.handle, self._h, g.handle, scope))

    def refine_am(self, g: DeviceAMGraph, view_size: int = 1, scope: int = SCOPE_QUERY):
        """refineAMSignatures<PowerGraph> (signature.hpp:250-296)."""
        check(lib().sigmo_signatures_refine_am(self.ctx.handle, self._h, g.handle, scope, view_size))

    def refine_am_view_based(self, g: DeviceAMGraph, view_size: int = 1, scope: int = SCOPE_QUERY):
        """refineAMSignatures<ViewBased> (signature.hpp:214-248)."""
        check(lib().sigmo_signatures_refine_am_view_based(self.ctx.handle, self._h, g.handle, scope, view_size))

    def download(self, scope: int) -> np.ndarray:
        n = self.data_nodes if scope == SCOPE_DATA else self.query_nodes
        out = np.zeros(n, np.uint64)
        check(lib().sigmo_signatures_download(self.ctx.handle, self._h, scope, _ptr(out), n))
        return out


class Candidates:
    def __init__(self, ctx: Context, source_nodes: int, target_nodes: int):
        h = C.c_void_p()
        check(lib().sigmo_candidates_create(ctx.handle, source_nodes, target_nodes, C.byref(h)))
        self._h, self.ctx = h, ctx
        self.source_nodes, self.target_nodes = source_nodes, target_nodes

    def free(self):
        if getattr(self, "_h", None):
            lib().sigmo_candidates_free(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.free()

    @property
    def handle(self):
        return self._h

    @property
    def stride32(self) -> int:
        return (self.target_nodes + 31) // 32

    def download(self) -> np.ndarray:
        """Reference layout: uint32 words, rows of ceil(target/32) words."""
        out = np.zeros(self.source_nodes * self.stride32, np.uint32)
        check(lib().sigmo_candidates_download(self.ctx.handle, self._h, _ptr(out), out.size))
        return out

    def counts(self) -> np.ndarray:
        out = np.zeros(self.source_nodes, np.uint32)
        check(lib().sigmo_candidates_counts(self.ctx.handle, self._h, _ptr(out), out.size))
        return out


def filter_candidates(ctx: Context, query: DeviceGraph, data: DeviceGraph, sig: Signature, cand: Candidates):
    check(lib().sigmo_candidates_filter(ctx.handle, cand.handle, query.handle, data.handle, sig.handle))


def filter_candidates_data(ctx: Context, query: DeviceGraph, data: DeviceGraph, cand: Candidates):
    """filterCandidates<CandidatesDomain::Data>: cand = Candidates(ctx, data nodes, query nodes)."""
    check(lib().sigmo_candidates_filter_data(ctx.handle, cand.handle, query.handle, data.handle))


def refine_candidates(ctx: Context, query: DeviceGraph, data: DeviceGraph, sig: Signature, cand: Candidates):
    check(lib().sigmo_candidates_refine(ctx.handle, cand.handle, query.handle, data.handle, sig.handle))


class GMCR:
    def __init__(self, ctx: Context):
        h = C.c_void_p()
        check(lib().sigmo_gmcr_create(ctx.handle, C.byref(h)))
        self._h, self.ctx = h, ctx
        self.num_data_graphs = 0

    def free(self):
        if getattr(self, "_h", None):
            lib().sigmo_gmcr_free(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.free()

    @property
    def handle(self):
        return self._h

    def generate(self, query: DeviceGraph, data: DeviceGraph, cand: Candidates):
        check(lib().sigmo_gmcr_generate(self.ctx.handle, self._h, query.handle, data.handle, cand.handle))
        self.num_data_graphs = data.num_graphs

    @property
    def total(self) -> int:
        v = C.c_uint64()
        check(lib().sigmo_gmcr_total(self._h, C.byref(v)))
        return v.value

    def download(self):
        off = np.zeros(self.num_data_graphs + 1, np.uint32)
        idx = np.zeros(self.total, np.uint32)
        check(lib().sigmo_gmcr_download(self.ctx.handle, self._h, _ptr(off), off.size,
                                        _ptr(idx) if idx.size else None, idx.size))
        return off, idx


def join_candidates(ctx: Context, query: DeviceGraph, data: DeviceGraph, cand: Candidates, gmcr: GMCR,
                    find_first: bool = True, per_pair: bool = True):
    """Returns (num_matches, per-pair int32 counts in GMCR order or None)."""
    total = C.c_uint64(0)
    counts = np.zeros(gmcr.total, np.int32) if per_pair else None
    check(lib().sigmo_join(ctx.handle, query.handle, data.handle, cand.handle, gmcr.handle, int(find_first),
                           C.byref(total), _ptr(counts) if per_pair and counts.size else None))
    return total.value, counts


JOIN_NOALLOC_MAX_PAIRS = 1 << 21


def join_candidates_noalloc(ctx: Context, query: DeviceGraph, data: DeviceGraph, cand: Candidates, gmcr: GMCR,
                            find_first: bool = True, per_pair: bool = True):
    """sigmo_join_noalloc: the join with its counters and per-pair counts in the library's static
    device scratch (at most JOIN_NOALLOC_MAX_PAIRS GMCR pairs); returns what join_candidates does."""
    total = C.c_uint64(0)
    counts = np.zeros(gmcr.total, np.int32) if per_pair else None
    check(lib().sigmo_join_noalloc(ctx.handle, query.handle, data.handle, cand.handle, gmcr.handle, int(find_first),
                                   C.byref(total), _ptr(counts) if per_pair and counts.size else None))
    return total.value, counts


def join_wildcard(ctx: Context, query: DeviceGraph, data: DeviceGraph, cand: Candidates, gmcr: GMCR, find_first: bool = True):
    """joinWildcardCandidates: (num_matches, per-pair int32 counts in GMCR order)."""
    total = C.c_uint64(0)
    counts = np.zeros(gmcr.total, np.int32)
    check(lib().sigmo_join_wildcard(ctx.handle, query.handle, data.handle, cand.handle, gmcr.handle, int(find_first), C.byref(total),
                                    _ptr(counts) if counts.size else None))
    return total.value, counts


def pipeline_run(ctx: Context, query: DeviceGraph, data: DeviceGraph, refinement_steps: int, find_first: bool = True) -> dict:
    r = PipelineResult()
    check(lib().sigmo_pipeline_run(ctx.handle, query.handle, data.handle, refinement_steps, int(find_first), C.byref(r)))
    return _result_dict(r)


def _result_dict(r: PipelineResult) -> dict:
    return {
        "num_matches": r.num_matches,
        "gmcr_pairs": r.gmcr_pairs,
        "matched_pairs": r.matched_pairs,
        "phase_ms": dict(zip(PHASE_NAMES, list(r.phase_ms))),
        "total_ms": r.total_ms,
        "dfs_steps": r.dfs_steps,
        "join_wave_steps": r.join_wave_steps,
    }


class Pipeline:
    """Plain per-phase pipeline with buffers allocated once (sigmo_pipeline_*)."""

    def __init__(self, ctx: Context, query: DeviceGraph, data: DeviceGraph):
        h = C.c_void_p()
        check(lib().sigmo_pipeline_create(ctx.handle, query.handle, data.handle, C.byref(h)))
        self._h, self.ctx, self.query, self.data = h, ctx, query, data
        self.last = None

    def free(self):
        if getattr(self, "_h", None):
            lib().sigmo_pipeline_free(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.free()

    def execute(self, refinement_steps: int, find_first: bool = True) -> dict:
        r = PipelineResult()
        check(lib().sigmo_pipeline_execute(self._h, refinement_steps, int(find_first), C.byref(r)))
        self.last = _result_dict(r)
        return self.last

    def capture(self, refinement_steps: int, find_first: bool = True) -> None:
        """Record one whole execute as a hipGraph (after an execute that sizes the buffers)."""
        check(lib().sigmo_pipeline_capture(self._h, refinement_steps, int(find_first)))

    def replay(self) -> dict:
        r = PipelineResult()
        check(lib().sigmo_pipeline_replay(self._h, C.byref(r)))
        self.last = _result_dict(r)
        return self.last

    def pairs(self):
        """(data_graph_offsets, query_graph_indices, counts) of the last execute."""
        P = self.last["gmcr_pairs"]
        off = np.zeros(self.data.num_graphs + 1, np.uint32)
        idx = np.zeros(P, np.uint32)
        cnt = np.zeros(P, np.int32)
        check(lib().sigmo_pipeline_pairs(self._h, _ptr(off), off.size, _ptr(idx) if P else None,
                                         _ptr(cnt) if P else None, P))
        return off, idx, cnt


class Database:
    """Resident, query-independent data side of the batched form (sigmo_db_*)."""

    def __init__(self, ctx: Context, data: DeviceGraph, refinement_steps: int):
        h = C.c_void_p()
        check(lib().sigmo_db_prepare(ctx.handle, data.handle, refinement_steps, C.byref(h)))
        self._h, self.ctx, self.data, self.refinement_steps = h, ctx, data, refinement_steps

    def free(self):
        if getattr(self, "_h", None):
            lib().sigmo_db_free(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.free()

    @property
    def handle(self):
        return self._h

    @property
    def device_bytes(self) -> int:
        v = C.c_uint64()
        check(lib().sigmo_db_alloc_size(self._h, C.byref(v)))
        return v.value

    def refresh(self) -> None:
        check(lib().sigmo_db_refresh(self._h))


class Batch:
    """Batched-form runs of query graphs against a Database (sigmo_batch_*)."""

    def __init__(self, ctx: Context, db: Database):
        h = C.c_void_p()
        check(lib().sigmo_batch_create(ctx.handle, db.handle, C.byref(h)))
        self._h, self.ctx, self.db = h, ctx, db
        self.last = None

    def free(self):
        if getattr(self, "_h", None):
            lib().sigmo_batch_free(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.free()

    def _store(self, r: BatchResult) -> dict:
        self.last = {"num_matches": r.num_matches, "matched_pairs": r.matched_pairs, "gmcr_pairs": r.gmcr_pairs,
                     "dfs_steps": r.dfs_steps, "filter_wave_steps": r.filter_wave_steps,
                     "filter_lane_steps": r.filter_lane_steps, "filter_scans": r.filter_scans,
                     "join_wave_steps": r.join_wave_steps, "one_sync": r.one_sync, "arena_repeats": r.arena_repeats}
        return self.last

    def run(self, query: DeviceGraph, find_first: bool = True) -> dict:
        r = BatchResult()
        check(lib().sigmo_batch_run(self._h, query.handle, int(find_first), C.byref(r)))
        return self._store(r)

    def capture(self, query: DeviceGraph, find_first: bool = True, refresh_db: bool = False) -> None:
        """Record one whole run (sized by a plain run first) as a hipGraph."""
        self._captured_query = query  # the graph holds its device addresses
        check(lib().sigmo_batch_capture(self._h, query.handle, int(find_first), int(refresh_db)))

    def replay(self) -> dict:
        r = BatchResult()
        check(lib().sigmo_batch_replay(self._h, C.byref(r)))
        return self._store(r)

    @property
    def tuning(self) -> dict:
        t = BatchTuning()
        check(lib().sigmo_batch_get_tuning(self._h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in BatchTuning._fields_}

    @tuning.setter
    def tuning(self, values: dict) -> None:
        t = BatchTuning(**values)
        check(lib().sigmo_batch_set_tuning(self._h, C.byref(t)))

    def autotune(self, query: DeviceGraph, find_first: bool = True, repeats: int = 2):
        """Pick launch shapes by timing runs; returns (tuning dict, best ms)."""
        t = BatchTuning()
        ms = C.c_double()
        check(lib().sigmo_batch_autotune(self._h, query.handle, int(find_first), repeats, C.byref(t), C.byref(ms)))
        return {name: getattr(t, name) for name, _ in BatchTuning._fields_}, ms.value

    def request_stats(self, request_graph_offsets) -> list[dict]:
        """Per-request matched pairs, embeddings and order-independent hash of the last run;
        request r owns the query graphs [offsets[r], offsets[r+1])."""
        rgo = np.ascontiguousarray(request_graph_offsets, np.uint32)
        n = rgo.size - 1
        out = (RequestStats * max(n, 1))()
        check(lib().sigmo_batch_request_stats(self._h, _ptr(rgo), n, out))
        return [{"matched_pairs": out[i].matched_pairs, "embeddings": out[i].embeddings, "hash": out[i].hash} for i in range(n)]

    def workspace_bytes(self, num_query_graphs: int, query_nodes: int) -> int:
        v = C.c_uint64()
        check(lib().sigmo_batch_workspace_bytes(self._h, num_query_graphs, query_nodes, C.byref(v)))
        return v.value

    def pairs(self):
        """(data_graph_offsets, query_graph_indices, counts) of the last run, GMCR order."""
        P = self.last["gmcr_pairs"]
        off = np.zeros(self.db.data.num_graphs + 1, np.uint32)
        idx = np.zeros(P, np.uint32)
        cnt = np.zeros(P, np.int32)
        check(lib().sigmo_batch_pairs(self._h, _ptr(off), off.size, _ptr(idx) if P else None, _ptr(cnt) if P else None, P))
        return off, idx, cnt


# ---- serving ------------------------------------------------------------------------------------

def _serve_options(max_batch_requests=0, max_batch_graphs=0, memory_budget=0, bytes_per_graph=0, bytes_per_batch=0,
                   deadline_ms=0.0, find_first=True) -> ServeOptions:
    return ServeOptions(max_batch_requests, max_batch_graphs, memory_budget, bytes_per_graph, bytes_per_batch, deadline_ms,
                        int(find_first))


def _serve_out(res, summ) -> dict:
    n = len(res)
    cols = {name: np.array([getattr(res[i], name) for i in range(n)]) for name, _ in ServeResult._fields_}
    cols["hash"] = np.array([res[i].hash for i in range(n)], np.uint64)
    cols["summary"] = {name: getattr(summ, name) for name, _ in ServeSummary._fields_}
    return cols


def serve(workers: list, queries: HostGraphs, request_graph_offsets, arrival_s, **options) -> dict:
    """Continuous-batching serving loop (sigmo_serve) over batch objects of one database, one per
    worker; returns per-request columns (status, latency_ms, matched_pairs, embeddings, hash, ...)
    and a 'summary' dict."""
    rgo = np.ascontiguousarray(request_graph_offsets, np.uint32)
    arr = np.ascontiguousarray(arrival_s, np.float64)
    n = arr.size
    if rgo.size != n + 1:
        raise ValueError("request_graph_offsets needs one entry more than arrival_s")
    handles = (C.c_void_p * len(workers))(*[w._h for w in workers])
    res = (ServeResult * max(n, 1))()
    summ = ServeSummary()
    opt = _serve_options(**options)
    check(lib().sigmo_serve(handles, len(workers), queries.handle, _ptr(rgo), _ptr(arr), n, C.byref(opt), res, C.byref(summ)))
    return _serve_out(res[:n], summ)


def _executor_shim(executor):
    """A SERVE_EXECUTOR body around ``executor(worker, requests) -> (stats, gmcr_pairs)``."""

    def _run(_user, worker, reqs, count, out, gmcr):
        try:
            stats, gp = executor(int(worker), [reqs[i] for i in range(count)])
            for i, (m, e, h) in enumerate(stats):
                out[i].matched_pairs, out[i].embeddings, out[i].hash = m, e, h
            gmcr[0] = gp
            return OK
        except Exception:  # noqa: BLE001 -- reported as a failed batch
            return ERR_INVALID

    return _run


def serve_schedule(num_workers: int, request_graph_offsets, arrival_s, executor, **options) -> dict:
    """The scheduler of the serving loop with a Python executor
    ``executor(worker, requests) -> (list of (matched_pairs, embeddings, hash), gmcr_pairs)``."""
    rgo = np.ascontiguousarray(request_graph_offsets, np.uint32)
    arr = np.ascontiguousarray(arrival_s, np.float64)
    n = arr.size
    if rgo.size != n + 1:
        raise ValueError("request_graph_offsets needs one entry more than arrival_s")

    cb = SERVE_EXECUTOR(_executor_shim(executor))
    res = (ServeResult * max(n, 1))()
    summ = ServeSummary()
    opt = _serve_options(**options)
    check(lib().sigmo_serve_schedule(num_workers, _ptr(rgo), _ptr(arr), n, C.byref(opt), cb, None, res, C.byref(summ)))
    return _serve_out(res[:n], summ)


class Server:
    """Online serving (sigmo_server_*): requests submitted while the loop runs, from any thread.

    ``submit(HostGraphs) -> id``; ``poll(id, timeout_ms)`` returns the request's result as a dict
    (``status`` is SERVE_PENDING until it settles); ``close()`` serves what is queued and returns
    the summary.  Worker batches must stay alive until close."""

    def __init__(self, workers: list | None = None, *, num_workers: int = 0, executor=None, **options):
        """Device server over ``workers`` (batch objects of one database), or, with ``executor``
        (as for serve_schedule) and ``num_workers``, the scheduler alone; its requests are
        submitted as query-graph counts."""
        self._h = None
        self.workers = list(workers or [])
        opt = _serve_options(**options)
        h = C.c_void_p()
        if executor is None:
            handles = (C.c_void_p * len(self.workers))(*[w._h for w in self.workers])
            check(lib().sigmo_server_open(handles, len(self.workers), C.byref(opt), C.byref(h)))
        else:
            self._cb = SERVE_EXECUTOR(_executor_shim(executor))  # kept alive while the server runs
            check(lib().sigmo_server_open_schedule(num_workers, C.byref(opt), self._cb, None, C.byref(h)))
        self._h = h

    def submit(self, queries) -> int:
        """Queue a request: HostGraphs for a device server, a query-graph count otherwise."""
        rid = C.c_uint64()
        if isinstance(queries, HostGraphs):
            check(lib().sigmo_server_submit(self._h, queries.handle, C.byref(rid)))
        else:
            check(lib().sigmo_server_submit_count(self._h, int(queries), C.byref(rid)))
        return rid.value

    def poll(self, request_id: int, timeout_ms: float = -1.0) -> dict:
        r = ServeResult()
        check(lib().sigmo_server_poll(self._h, request_id, float(timeout_ms), C.byref(r)))
        return {name: getattr(r, name) for name, _ in ServeResult._fields_}

    def close(self) -> dict:
        if not self._h:
            return {}
        summ = ServeSummary()
        h, self._h = self._h, None
        check(lib().sigmo_server_close(h, C.byref(summ)))
        return {name: getattr(summ, name) for name, _ in ServeSummary._fields_}

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().sigmo_server_close(self._h, None)
            except Exception:  # noqa: BLE001 -- interpreter shutdown
                pass
            self._h = None


# ---- streaming ----------------------------------------------------------------------------------

def stream_run(ctx: Context, query: DeviceGraph, data: HostGraphs, chunk_molecules: int = 0, form: str = "batched",
               refinement_steps: int = 6, find_first: bool = True, overlap: bool = True, collect: bool = False,
               digest: bool = False, workers: int = 0) -> dict:
    """Run a host-resident molecule shard in chunks (sigmo_stream_run: one pass of a session).  With
    collect, also returns the whole shard's GMCR (offsets over all molecules, indices) and per-pair
    counts, assembled from the chunks.  With digest, returns instead, per chunk in order, (first
    molecule, molecules, pairs, a 128-bit xxh3 digest of the chunk's offsets, indices and counts)
    under 'chunk_digests': equal chunk plans give equal digests exactly when the results are equal."""
    opt = StreamOptions(chunk_molecules, refinement_steps, int(find_first), FORM_PLAIN if form == "plain" else FORM_BATCHED,
                        int(overlap), workers)
    return _stream_call(lambda cb, r: lib().sigmo_stream_run(ctx.handle, query.handle, data.handle, C.byref(opt), cb, None,
                                                             C.byref(r)), collect, digest)


class Stream:
    """A streaming session (sigmo_stream_open / pass / close): device buffers are allocated by the
    first pass and kept for the later ones.  ``pass_()`` takes the same collect / digest options
    as :func:`stream_run`."""

    def __init__(self, ctx: Context, query: DeviceGraph, data: HostGraphs, chunk_molecules: int = 0, form: str = "batched",
                 refinement_steps: int = 6, find_first: bool = True, overlap: bool = True, workers: int = 0):
        opt = StreamOptions(chunk_molecules, refinement_steps, int(find_first), FORM_PLAIN if form == "plain" else FORM_BATCHED,
                            int(overlap), workers)
        h = C.c_void_p()
        check(lib().sigmo_stream_open(ctx.handle, query.handle, data.handle, C.byref(opt), C.byref(h)))
        self._h, self.ctx, self.query, self.data = h, ctx, query, data

    def pass_(self, collect: bool = False, digest: bool = False) -> dict:
        return _stream_call(lambda cb, r: lib().sigmo_stream_pass(self._h, cb, None, C.byref(r)), collect, digest)

    def close(self):
        if self._h:
            check(lib().sigmo_stream_close(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass


def _stream_call(call, collect: bool, digest: bool) -> dict:
    parts = []

    def _chunk(_user, chunk, first, n, off, idx, cnt, P):
        try:
            o = np.ctypeslib.as_array(off, (n + 1,))
            i = np.ctypeslib.as_array(idx, (P,)) if P else np.zeros(0, np.uint32)
            c = np.ctypeslib.as_array(cnt, (P,)) if P else np.zeros(0, np.int32)
            if digest:
                import xxhash

                h = xxhash.xxh3_128()
                for a in (o, i, c):
                    h.update(memoryview(np.ascontiguousarray(a)).cast("B"))
                parts.append((first, n, P, h.hexdigest()))
            else:
                parts.append((first, o.copy(), i.copy(), c.copy()))
            return 0
        except Exception:  # noqa: BLE001 -- reported as a failed chunk
            return 1

    cb = STREAM_CHUNK_FN(_chunk) if (collect or digest) else STREAM_CHUNK_FN()
    r = StreamResult()
    check(call(cb, r))
    out = {name: getattr(r, name) for name, _ in StreamResult._fields_}
    if digest:
        out["chunk_digests"] = sorted(parts)
    elif collect:
        offs, idxs, cnts, base = [np.zeros(1, np.uint32)], [], [], 0
        for first, off, idx, cnt in sorted(parts, key=lambda p: p[0]):
            offs.append(off[1:] + np.uint32(base))
            idxs.append(idx)
            cnts.append(cnt)
            base += int(off[-1])
        out["offsets"] = np.concatenate(offs).astype(np.uint32)
        out["indices"] = np.concatenate(idxs).astype(np.uint32) if idxs else np.zeros(0, np.uint32)
        out["counts"] = np.concatenate(cnts).astype(np.int32) if cnts else np.zeros(0, np.int32)
    return out
