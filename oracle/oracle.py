"""Python face of the CPU oracle (``oracle/liboracle.so``, built from ``oracle/oracle.cpp``).

TEST INFRASTRUCTURE ONLY: used by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker / CPU reference.  Product code never
imports it.

Graph arguments are dicts of numpy arrays with the keys produced by
``sigmo_amd.capi.HostGraphs.arrays()`` (graph_offsets, row_offsets, column_indices,
node_labels, edge_labels) -- the batched CSR of library/include/graph.hpp:255-314.
Candidate bitsets use the reference layout (uint32 words, row stride ceil(D/32)).

Pinning status per function (what checks it beyond this restatement) is listed in
``tests/test_oracle_pinning.py``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("SIGMO_ORACLE_LIB", Path(__file__).resolve().parent / "liboracle.so"))
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m sigmo_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        vp, u8, u32, u64, i32, i64 = C.c_void_p, C.c_uint8, C.c_uint32, C.c_uint64, C.c_int, C.c_int64
        sigs = {
            "oracle_set_threads": ([i32], None),
            "oracle_get_threads": ([], i32),
            "oracle_sig_set": ([u64, u8, u8], u64),
            "oracle_sig_inc": ([u64, u8, u8], u64),
            "oracle_sig_get": ([u64, u8], u8),
            "oracle_cand_insert": ([vp, u64, u32, u32], None),
            "oracle_cand_remove": ([vp, u64, u32, u32], None),
            "oracle_cand_count": ([vp, u64, u32], u32),
            "oracle_cand_count_range": ([vp, u64, u32, u32, u32], u32),
            "oracle_cand_at": ([vp, u64, u32, u32, u32, u32], u32),
            "oracle_generate_signatures": ([u32, vp, vp, vp, vp, vp], None),
            "oracle_refine_signatures": ([u32, vp, vp, vp, vp, u32, vp, vp, i32], None),
            "oracle_filter": ([u64, vp, u64, vp, vp], None),
            "oracle_refine_candidates": ([u64, vp, u64, vp, vp, i32], None),
            "oracle_candidate_counts": ([u64, u64, vp, vp], None),
            "oracle_gmcr_offsets": ([u32, vp, u32, vp, u64, vp, vp], i64),
            "oracle_gmcr_indices": ([u32, vp, u32, vp, u64, vp, vp, vp], None),
            "oracle_join": ([u32, vp, vp, vp, u32, vp, vp, vp, u64, vp, vp, vp, i32, vp, i32], i64),
            "oracle_refine_view_based": ([u32, vp, vp, vp, vp, u64, vp, i32], None),
            "oracle_filter_data": ([u64, vp, u64, vp, vp], None),
            "oracle_join_wildcard": ([u32, vp, vp, vp, vp, u32, vp, vp, vp, vp, u64, vp, vp, vp, i32, u32, vp, i32], i64),
            "oracle_am_words": ([u32, vp], u64),
            "oracle_am_build": ([u32, vp, vp, vp, vp, vp, vp], None),
            "oracle_am_generate": ([u32, vp, vp, vp, vp, u64, vp], None),
            "oracle_am_refine_pg": ([u32, vp, vp, vp, vp, u64, u64, vp, i32], None),
            "oracle_am_refine_vb": ([u32, vp, vp, vp, vp, u64, u64, vp], None),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        threads = os.environ.get("SIGMO_ORACLE_THREADS")
        if threads:
            L.oracle_set_threads(int(threads))
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a.size else None


def use_library(path) -> None:
    """Load the oracle from another build of oracle.cpp (e.g. :func:`build_native`) from now on."""
    global LIB_PATH, _lib
    LIB_PATH = Path(path)
    _lib = None


def build_native(dest_dir) -> Path:
    """Compile oracle.cpp with ``g++ -O3 -march=native -fopenmp`` for the host it runs on (the CPU
    baseline's build); returns the shared library's path.  Raises CalledProcessError on failure."""
    import subprocess

    out = Path(dest_dir) / "liboracle_native.so"
    src = Path(__file__).resolve().parent / "oracle.cpp"
    subprocess.run(["g++", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-std=c++17", str(src), "-o", str(out)],
                   check=True, capture_output=True, text=True)
    return out


def set_threads(n: int) -> None:
    lib().oracle_set_threads(n)


def threads() -> int:
    return lib().oracle_get_threads()


def _G(g: dict) -> int:
    return len(g["graph_offsets"]) - 1


def _N(g: dict) -> int:
    return int(g["graph_offsets"][-1])


# ---- scalar known-answer hooks ------------------------------------------------------------------

def sig_set(sig: int, label: int, count: int) -> int:
    return lib().oracle_sig_set(sig, label, count)


def sig_inc(sig: int, label: int, add: int = 1) -> int:
    return lib().oracle_sig_inc(sig, label, add)


def sig_get(sig: int, label: int) -> int:
    return lib().oracle_sig_get(sig, label)


class RefCandidates:
    """Reference-layout bitset (candidates.hpp:16-162) over a numpy uint32 buffer."""

    def __init__(self, source_nodes: int, target_nodes: int, words: np.ndarray | None = None):
        self.Q, self.D = source_nodes, target_nodes
        self.stride = (target_nodes + 31) // 32
        self.words = np.zeros(self.Q * self.stride, np.uint32) if words is None else words

    def insert(self, s, c):
        lib().oracle_cand_insert(_p(self.words), self.stride, s, c)

    def remove(self, s, c):
        lib().oracle_cand_remove(_p(self.words), self.stride, s, c)

    def count(self, s, gs=None, ge=None):
        if gs is None:
            return lib().oracle_cand_count(_p(self.words), self.stride, s)
        return lib().oracle_cand_count_range(_p(self.words), self.stride, s, gs, ge)

    def at(self, s, idx, gs, ge):
        return lib().oracle_cand_at(_p(self.words), self.stride, s, idx, gs, ge)


# ---- phases ---------------------------------------------------------------------------------------

def generate_signatures(g: dict, sigs: np.ndarray | None = None) -> np.ndarray:
    out = np.zeros(_N(g), np.uint64) if sigs is None else sigs
    lib().oracle_generate_signatures(_G(g), _p(g["graph_offsets"]), _p(g["row_offsets"]), _p(g["column_indices"]),
                                     _p(g["node_labels"]), _p(out))
    return out


def refine_signatures(g: dict, view_size: int, sigs: np.ndarray, reach: np.ndarray, mutation: int = 0) -> None:
    lib().oracle_refine_signatures(_G(g), _p(g["graph_offsets"]), _p(g["row_offsets"]), _p(g["column_indices"]),
                                   _p(g["node_labels"]), view_size, _p(sigs), _p(reach), mutation)


def filter_candidates(q: dict, d: dict, words: np.ndarray | None = None) -> np.ndarray:
    Q, D = _N(q), _N(d)
    out = np.zeros(Q * ((D + 31) // 32), np.uint32) if words is None else words
    lib().oracle_filter(Q, _p(q["node_labels"]), D, _p(d["node_labels"]), _p(out))
    return out


def refine_candidates(qsig: np.ndarray, dsig: np.ndarray, words: np.ndarray, mutation: int = 0) -> None:
    lib().oracle_refine_candidates(len(qsig), _p(qsig), len(dsig), _p(dsig), _p(words), mutation)


def candidate_counts(Q: int, D: int, words: np.ndarray) -> np.ndarray:
    out = np.zeros(Q, np.uint32)
    lib().oracle_candidate_counts(Q, D, _p(words), _p(out))
    return out


def gmcr(q: dict, d: dict, words: np.ndarray):
    Gq, Gd, D = _G(q), _G(d), _N(d)
    offsets = np.zeros(Gd + 1, np.uint32)
    total = lib().oracle_gmcr_offsets(Gq, _p(q["graph_offsets"]), Gd, _p(d["graph_offsets"]), D, _p(words), _p(offsets))
    if total < 0:
        raise ValueError("query graphs with fewer than 2 nodes are rejected")
    indices = np.zeros(total, np.uint32)
    lib().oracle_gmcr_indices(Gq, _p(q["graph_offsets"]), Gd, _p(d["graph_offsets"]), D, _p(words), _p(offsets), _p(indices))
    return offsets, indices


def join(q: dict, d: dict, words: np.ndarray, offsets: np.ndarray, indices: np.ndarray, find_first: bool = True,
         mutation: int = 0):
    counts = np.zeros(len(indices), np.int32)
    total = lib().oracle_join(_G(q), _p(q["graph_offsets"]), _p(q["row_offsets"]), _p(q["column_indices"]),
                              _G(d), _p(d["graph_offsets"]), _p(d["row_offsets"]), _p(d["column_indices"]), _N(d),
                              _p(words), _p(offsets), _p(indices), int(find_first), _p(counts), mutation)
    if total < 0:
        raise ValueError("join failed: query too large or per-pair count overflow")
    return int(total), counts


def refine_view_based(g: dict, view_size: int, sigs: np.ndarray, mutation: int = 0) -> None:
    """refineCSRSignatures<ViewBased> (library/include/signature.hpp:298-334), in place."""
    lib().oracle_refine_view_based(_G(g), _p(g["graph_offsets"]), _p(g["row_offsets"]), _p(g["column_indices"]),
                                   _p(g["node_labels"]), view_size, _p(sigs), mutation)


def filter_candidates_data(q: dict, d: dict) -> np.ndarray:
    """filterCandidates<CandidatesDomain::Data>: reference-layout words, rows = data nodes."""
    Q, D = _N(q), _N(d)
    words = np.zeros(D * ((Q + 31) // 32), np.uint32)
    lib().oracle_filter_data(Q, _p(q["node_labels"]), D, _p(d["node_labels"]), _p(words))
    return words


def join_wildcard(q: dict, d: dict, words: np.ndarray, offsets: np.ndarray, indices: np.ndarray, find_first: bool = True,
                  sg_size: int = 64, mutation: int = 0):
    """joinWildcardCandidates (library/include/isomorphism.hpp:428-556): matching order by
    candidate count, edge labels compared, one DFS per first-node candidate."""
    counts = np.zeros(len(indices), np.int32)
    total = lib().oracle_join_wildcard(_G(q), _p(q["graph_offsets"]), _p(q["row_offsets"]), _p(q["column_indices"]),
                                       _p(q["edge_labels"]), _G(d), _p(d["graph_offsets"]), _p(d["row_offsets"]),
                                       _p(d["column_indices"]), _p(d["edge_labels"]), _N(d), _p(words), _p(offsets), _p(indices),
                                       int(find_first), sg_size, _p(counts), mutation)
    if total < 0:
        raise ValueError("wildcard join failed: query too large or per-pair count overflow")
    return int(total), counts


def am_build(g: dict) -> dict:
    """toAMGraph + createDeviceAMGraph (library/include/graph.hpp:208-214, 345-392) of a CSR batch:
    per graph an n x n bit matrix in uint64 words (row width W = (uint16)sqrt(float(64*words)),
    words = (n*n+63)/64), cumulative node counts and per-graph word offsets."""
    G = _G(g)
    words = lib().oracle_am_words(G, _p(g["graph_offsets"]))
    if G and not words:
        raise ValueError("adjacency-matrix graphs hold at most 64 nodes")
    out = {"graph_offsets": np.zeros(G, np.uint32), "num_nodes": np.zeros(G, np.uint32), "adjacency": np.zeros(max(words, 1), np.uint64),
           "node_labels": np.ascontiguousarray(g["node_labels"], np.uint8), "G": G, "N": _N(g)}
    lib().oracle_am_build(G, _p(g["graph_offsets"]), _p(g["row_offsets"]), _p(g["column_indices"]), _p(out["graph_offsets"]),
                          _p(out["num_nodes"]), _p(out["adjacency"]))
    out["adjacency"] = out["adjacency"][:words]
    return out


def _am_args(am):
    adj = am["adjacency"] if len(am["adjacency"]) else np.zeros(1, np.uint64)
    return am["G"], _p(am["num_nodes"]), _p(am["graph_offsets"]), _p(adj), _p(am["node_labels"]), am["N"]


def am_generate(am: dict, sigs: np.ndarray | None = None) -> np.ndarray:
    """generateAMSignatures (library/include/signature.hpp:100-120), onto sigs."""
    out = np.zeros(am["N"], np.uint64) if sigs is None else sigs
    lib().oracle_am_generate(*_am_args(am), _p(out))
    return out


def am_refine_power_graph(am: dict, view_size: int, sigs: np.ndarray, mutation: int = 0) -> None:
    """refineAMSignatures<PowerGraph> (library/include/signature.hpp:250-296), in place."""
    lib().oracle_am_refine_pg(*_am_args(am), view_size, _p(sigs), mutation)


def am_refine_view_based(am: dict, view_size: int, sigs: np.ndarray) -> None:
    """refineAMSignatures<ViewBased> (library/include/signature.hpp:214-248), in place."""
    lib().oracle_am_refine_vb(*_am_args(am), view_size, _p(sigs))


def pipeline(q: dict, d: dict, refinement_steps: int, find_first: bool = True, *, keep_levels: bool = False,
             join_mutation: int = 0, modes=None) -> dict:
    """The driver order of library/src/sigmo.cpp:102-159.

    Returns candidates after refinement, the GMCR, per-pair counts for each requested mode
    (``modes`` defaults to just ``find_first``) and, with keep_levels, the signatures of
    every level.
    """
    Q, D = _N(q), _N(d)
    dsig = np.zeros(D, np.uint64)
    qsig = np.zeros(Q, np.uint64)
    dreach = np.zeros(D, np.uint64)
    qreach = np.zeros(Q, np.uint64)
    generate_signatures(d, dsig)
    generate_signatures(q, qsig)
    levels = [(dsig.copy(), qsig.copy())] if keep_levels else None
    words = filter_candidates(q, d)
    filtered = words.copy() if keep_levels else None
    for r in range(1, refinement_steps + 1):
        refine_signatures(d, r, dsig, dreach)
        refine_signatures(q, r, qsig, qreach)
        refine_candidates(qsig, dsig, words)
        if keep_levels:
            levels.append((dsig.copy(), qsig.copy()))
    offsets, indices = gmcr(q, d, words)
    res = {"words": words, "offsets": offsets, "indices": indices, "levels": levels, "filtered": filtered}
    for mode in (modes if modes is not None else [find_first]):
        total, counts = join(q, d, words, offsets, indices, find_first=mode, mutation=join_mutation)
        key = "first" if mode else "all"
        res[f"total_{key}"] = total
        res[f"counts_{key}"] = counts
    return res


def pair_table(offsets: np.ndarray, indices: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """(query graph, data graph, count) rows in GMCR order."""
    dg = np.repeat(np.arange(len(offsets) - 1, dtype=np.uint32), np.diff(offsets.astype(np.int64)))
    return np.stack([indices.astype(np.int64), dg.astype(np.int64), counts.astype(np.int64)], axis=1)


# ---- one rank's share of a .dat file (pure Python) --------------------------------------------


def dat_shard_lines(data: bytes, rank: int, nprocs: int, max_lines: int) -> list[bytes]:
    """loadFileLinesMPI (library/src/sigmo_mpi.cpp:18-80) over the file's bytes, step by step:
    the rank's range starts at rank * (size // nprocs) (:32-33) and ends 2048 * max_lines bytes
    later (:34); the buffer has 1024 more bytes and is zero-filled, so what lies past the end of
    the file reads as NUL (:37-44); ranks > 0 drop through their first newline (:49-54), ranks <
    nprocs - 1 cut at their last newline (:57-62); std::getline splitting with at most max_lines
    lines (:65-70); a line stays if it holds exactly one 'n', one 'e' and one 'l' (:72-77).  The
    NUL run is kept at its full length."""
    size = len(data)
    start = rank * (size // nprocs)
    buf = bytearray(2048 * max_lines + 1024)
    chunk = data[start:start + len(buf)]
    buf[: len(chunk)] = chunk
    s = bytes(buf)
    if rank != 0:
        i = s.find(b"\n")
        if i >= 0:
            s = s[i + 1:]
    if rank != nprocs - 1:
        i = s.rfind(b"\n")
        if i >= 0:
            s = s[:i]
    lines = []
    pos = 0
    while pos < len(s) and len(lines) < max_lines:  # std::getline: a final "\n" yields no empty line
        i = s.find(b"\n", pos)
        if i < 0:
            i = len(s)
        lines.append(s[pos:i])
        pos = i + 1
    return [ln for ln in lines if ln.count(b"n") == 1 and ln.count(b"e") == 1 and ln.count(b"l") == 1]
