"""One rank's share of a .dat file, as the reference's multi-node driver reads it
(loadFileLinesMPI, library/src/sigmo_mpi.cpp:18-80, then io.hpp:37-44).

The oracle's restatement (oracle.dat_shard_lines) is pinned by hand-derived answers on files of
40-byte lines, where every byte range can be worked out on paper; the library
(sigmo_dat_shard_lines, sigmo_host_graphs_load_dat_shard, io::loadFileLinesShard) is compared
with it on random files, rank counts and line limits; two gloo ranks sum their shards' counts as
the reference's MPI_Reduce does (sigmo_mpi.cpp:136-137)."""
from __future__ import annotations

import re
import shutil
import socket
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from sigmo_amd import build as B
from sigmo_amd import capi, shard
from tests._util import GOLDEN


def _line(label: int, width: int = 40) -> bytes:
    """A one-node graph line padded with spaces to `width` bytes including its newline."""
    s = f"n#1 l#1 0 {label} e#0".encode()
    return s + b" " * (width - 1 - len(s)) + b"\n"


L = [_line(i) for i in range(4)]
FILE4 = b"".join(L)  # 160 bytes, line i at [40 i, 40 i + 40)


def _content(i: int) -> bytes:
    return L[i][:-1]


def _collapse(lines):
    return [re.sub(b"\0+", b"\0", x) for x in lines]


# ---- the restatement against hand-derived answers ----------------------------------------------


def test_two_ranks_four_lines():
    # nprocs 2: chunk 80, so rank 1 starts exactly on line 2 and still drops it (the first
    # newline it finds ends line 2).  Rank 0 reads the whole file (its range is 9216 bytes),
    # cuts at its last newline and keeps lines 0-3; rank 1 keeps line 3 and a NUL line that the
    # filter removes.
    assert O.dat_shard_lines(FILE4, 0, 2, 4) == [_content(i) for i in range(4)]
    assert O.dat_shard_lines(FILE4, 1, 2, 4) == [_content(3)]
    # one line each: lines 1 and 2 are read by no rank
    assert O.dat_shard_lines(FILE4, 0, 2, 1) == [_content(0)]
    assert O.dat_shard_lines(FILE4, 1, 2, 1) == [_content(3)]


def test_three_ranks_four_lines():
    # chunk 53: rank 1 starts inside line 1 (40..79) and rank 2 inside line 2 (80..119)
    assert O.dat_shard_lines(FILE4, 0, 3, 2) == [_content(0), _content(1)]
    assert O.dat_shard_lines(FILE4, 1, 3, 2) == [_content(2), _content(3)]
    assert O.dat_shard_lines(FILE4, 2, 3, 2) == [_content(3)]


def test_line_limit_counts_lines_before_the_filter():
    junk = b"# xyz" + b" " * 34 + b"\n"            # no 'n', 'e' or 'l'
    twice = b"n#1 l#1 0 0 e#0 e" + b" " * 22 + b"\n"  # two 'e'
    data = L[0] + junk + twice + L[1] + L[2]
    assert O.dat_shard_lines(data, 0, 1, 2) == [_content(0)]
    assert O.dat_shard_lines(data, 0, 1, 3) == [_content(0)]
    assert O.dat_shard_lines(data, 0, 1, 4) == [_content(0), _content(1)]
    assert O.dat_shard_lines(data, 0, 1, 0) == []


def test_last_line_without_newline_keeps_the_nul_run():
    data = L[0] + b"n#2 l#1 0 1 1 1 e#1 0 1"
    got = O.dat_shard_lines(data, 0, 1, 5)
    assert got[0] == _content(0)
    assert got[1] == b"n#2 l#1 0 1 1 1 e#1 0 1" + b"\0" * (2048 * 5 + 1024 - len(data))
    assert len(got) == 2


def test_ranks_past_the_end_read_nothing():
    tiny = L[0]
    assert O.dat_shard_lines(tiny, 0, 1, 3) == [_content(0)]
    for r in range(1, 4):  # chunk 10: ranks 1-3 start inside the only line, drop it, keep NULs
        assert O.dat_shard_lines(tiny, r, 4, 3) == []


# ---- the library against the restatement ------------------------------------------------------


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("shard")
    rng = np.random.default_rng(5)
    out = []
    mol = capi.HostGraphs.molecules(11, 60)
    base = d / "mol.dat"
    mol.save_dat(base)
    lines = base.read_bytes().split(b"\n")[:-1]
    for i in range(6):
        pick = list(lines)
        for _ in range(int(rng.integers(0, 6))):  # junk and near-miss lines at random places
            junk = rng.choice([b"", b"# comment", b"n#1 l#1 0 0 e#0 e", b"   ", b"n#1 l#1 0 0"])
            pick.insert(int(rng.integers(0, len(pick) + 1)), bytes(junk))
        data = b"\n".join(pick) + (b"\n" if i % 2 == 0 else b"")
        p = d / f"f{i}.dat"
        p.write_bytes(data)
        out.append(p)
    for name, data in (("four.dat", FILE4), ("tail.dat", L[0] + b"n#2 l#1 0 1 1 1 e#1 0 1"), ("empty.dat", b"")):
        p = d / name
        p.write_bytes(data)
        out.append(p)
    return out


def test_library_lines_equal_the_restatement(files):
    rng = np.random.default_rng(6)
    for p in files:
        data = p.read_bytes()
        for nprocs in (1, 2, 3, 5):
            for max_lines in sorted({0, 1, 2, int(rng.integers(1, 70)), 200}):
                for rank in range(nprocs):
                    want = _collapse(O.dat_shard_lines(data, rank, nprocs, max_lines))
                    got = capi.dat_shard_lines(p, rank, nprocs, max_lines)
                    assert got == want, (p.name, rank, nprocs, max_lines)


def test_library_graphs_equal_the_parsed_lines(files):
    for p in files[:6]:
        data = p.read_bytes()
        for nprocs in (1, 2, 3):
            for rank in range(nprocs):
                lines = O.dat_shard_lines(data, rank, nprocs, 25)
                g = capi.HostGraphs.load_dat_shard(p, rank, nprocs, 25)
                want = capi.HostGraphs.parse_dat(b"\n".join(lines).decode() + "\n") if lines else None
                if want is None:
                    assert g.num_graphs == 0
                    continue
                for k, v in want.arrays().items():
                    np.testing.assert_array_equal(g.arrays()[k], v, err_msg=f"{p.name} {rank}/{nprocs} {k}")


def test_one_rank_reads_the_whole_reference_fixture():
    path = GOLDEN / "test_data.dat"
    n = path.read_bytes().count(b"\n")
    whole = capi.HostGraphs.load_dat(path)
    g = capi.HostGraphs.load_dat_shard(path, 0, 1, n)
    assert g.num_graphs == whole.num_graphs
    for k, v in whole.arrays().items():
        np.testing.assert_array_equal(g.arrays()[k], v, err_msg=k)


def test_nul_run_after_a_space_is_a_token(tmp_path):
    # The last line of a file without a final newline takes the NUL run as one more edge token
    # when it ends in a space: "Invalid graph format", as in the reference's parser.
    p = tmp_path / "space.dat"
    p.write_bytes(L[0] + _content(1))
    with pytest.raises(capi.SigmoError, match="Invalid graph format"):
        capi.HostGraphs.load_dat_shard(p, 0, 1, 4)
    # rank 0 of 2 is not the last rank: it cuts at its last newline, dropping the unterminated line
    assert capi.HostGraphs.load_dat_shard(p, 0, 2, 4).num_graphs == 1


def test_shard_errors(tmp_path):
    with pytest.raises(capi.SigmoError):
        capi.dat_shard_lines(tmp_path / "missing.dat", 0, 1, 4)
    p = tmp_path / "f.dat"
    p.write_bytes(FILE4)
    for rank, nprocs in ((1, 1), (0, 0), (3, 2)):
        with pytest.raises(capi.SigmoError):
            capi.HostGraphs.load_dat_shard(p, rank, nprocs, 4)


@pytest.fixture(scope="module")
def shard_cli(tmp_path_factory):
    if not B.LIB.exists() or shutil.which(B.CXX) is None:
        pytest.skip("library or host compiler missing")
    exe = tmp_path_factory.mktemp("shardcpp") / "shard_lines"
    subprocess.run([B.CXX, "-std=c++17", "-O1", "-I", str(B.ROOT / "include"), str(B.ROOT / "tests" / "cpp" / "shard_lines.cpp"),
                    "-L", str(B.PKG), "-lsigmo_hip", f"-Wl,-rpath,{B.PKG}", "-o", str(exe)], check=True)
    return exe


def test_cpp_mirror_lines_and_graphs(shard_cli, files):
    for p in files:
        data = p.read_bytes()
        for nprocs in (1, 3):
            for rank in range(nprocs):
                out = subprocess.run([str(shard_cli), str(p), "30", str(rank), str(nprocs)], check=True, capture_output=True).stdout
                got = out.split(b"\n")[:-1]
                want = _collapse(O.dat_shard_lines(data, rank, nprocs, 30))
                assert got[:-1] == [x.replace(b"\0", b"<NUL>") for x in want], (p.name, rank, nprocs)
                try:
                    n = capi.HostGraphs.load_dat_shard(p, rank, nprocs, 30).num_graphs
                except capi.SigmoError:
                    n = -1
                assert got[-1] == f"graphs {n}".encode()


# ---- two ranks over gloo ---------------------------------------------------------------------


def _free_port() -> int:
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def _rank_main(rank, world, port, path, max_lines, out_dir):
    import torch.distributed as dist

    dist.init_process_group("gloo", init_method=f"tcp://127.0.0.1:{port}", rank=rank, world_size=world)
    try:
        g = capi.HostGraphs.load_dat_shard(path, rank, world, max_lines)
        totals = shard.reduce_totals([g.num_graphs, g.total_nodes, g.total_edges])
        (out_dir / f"rank{rank}.txt").write_text(" ".join(str(x) for x in totals))
    finally:
        dist.destroy_process_group()


def test_gloo_ranks_sum_their_shards(files, tmp_path):
    import torch.multiprocessing as mp

    path = files[0]
    data = path.read_bytes()
    world, max_lines = 2, 40
    mp.spawn(_rank_main, args=(world, _free_port(), path, max_lines, tmp_path), nprocs=world, join=True)
    want = [0, 0, 0]
    for r in range(world):
        g = capi.HostGraphs.parse_dat(b"\n".join(O.dat_shard_lines(data, r, world, max_lines)).decode() + "\n")
        want = [a + b for a, b in zip(want, (g.num_graphs, g.total_nodes, g.total_edges))]
    for r in range(world):
        assert [int(x) for x in (tmp_path / f"rank{r}.txt").read_text().split()] == want
