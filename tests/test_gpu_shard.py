"""Ranks of the multi-node driver (sigmo_mpi.cpp:81-240) run one after another on one GPU
(sigmo_amd.shard.run_rank without a process group): each rank's matches equal the oracle's
pipeline over the lines the oracle's restatement of loadFileLinesMPI gives that rank, in both
join modes."""
from __future__ import annotations

import pytest

from oracle import oracle as O
from sigmo_amd import capi, shard

pytestmark = pytest.mark.gpu

R = 2


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("shard")
    capi.HostGraphs.molecules(21, 1500).save_dat(d / "molecules.dat")
    capi.HostGraphs.queries(21, 40, 3, 10, perturb_frac=0.2).save_dat(d / "queries.dat")
    return d / "queries.dat", d / "molecules.dat"


@pytest.mark.parametrize("nprocs,max_lines", [(1, 1500), (3, 500), (4, 300)])
@pytest.mark.parametrize("find_all", [False, True])
def test_ranks_equal_the_oracle_on_their_shards(ctx, files, nprocs, max_lines, find_all):
    query, data_file = files
    qa = capi.HostGraphs.load_dat(query).arrays()
    data = data_file.read_bytes()
    for rank in range(nprocs):
        print(f"rank {rank}/{nprocs} max_lines={max_lines} find_all={find_all}", flush=True)
        got = shard.run_rank(ctx, query, data_file, max_lines, refinement_steps=R, find_all=find_all, rank=rank, nprocs=nprocs)
        lines = O.dat_shard_lines(data, rank, nprocs, max_lines)
        sh = capi.HostGraphs.parse_dat(b"\n".join(lines).decode() + "\n")
        assert got["data_graphs"] == sh.num_graphs == got["total_data_graphs"]
        want = O.pipeline(qa, sh.arrays(), R, modes=[not find_all])[f"total_{'all' if find_all else 'first'}"] if lines else 0
        assert got["num_matches"] == want == got["total_matches"]
        assert len(got["rank_ms"]) == 1
    if nprocs == 1:
        whole = capi.HostGraphs.load_dat(data_file)
        assert got["data_graphs"] == whole.num_graphs
