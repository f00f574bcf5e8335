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


def _rank_main(rank, world, port, query, data_file, max_lines, out_dir):
    import json

    import torch.distributed as dist

    dist.init_process_group("gloo", init_method=f"tcp://127.0.0.1:{port}", rank=rank, world_size=world)
    try:
        ctx = capi.Context(0)
        got = shard.run_rank(ctx, query, data_file, max_lines, refinement_steps=R)
        got.pop("result")
        (out_dir / f"rank{rank}.json").write_text(json.dumps(got))
    finally:
        dist.destroy_process_group()


@pytest.mark.timeout(600)  # the first import of torch on a fresh machine can take minutes
def test_two_gloo_ranks_on_one_gpu(files, tmp_path):
    # Two processes share the GPU, each a rank of a gloo group: their sums and gathered times
    # are the same on both and equal the oracle's per-shard results added up.
    import json
    import socket

    import torch.multiprocessing as mp

    query, data_file = files
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        port = s.getsockname()[1]
    world, max_lines = 2, 800
    print("two gloo ranks", flush=True)
    mp.spawn(_rank_main, args=(world, port, query, data_file, max_lines, tmp_path), nprocs=world, join=True)
    qa = capi.HostGraphs.load_dat(query).arrays()
    data = data_file.read_bytes()
    graphs = matches = 0
    out = [json.loads((tmp_path / f"rank{r}.json").read_text()) for r in range(world)]
    for r in range(world):
        lines = O.dat_shard_lines(data, r, world, max_lines)
        sh = capi.HostGraphs.parse_dat(b"\n".join(lines).decode() + "\n")
        m = O.pipeline(qa, sh.arrays(), R)["total_first"]
        assert (out[r]["rank"], out[r]["data_graphs"], out[r]["num_matches"]) == (r, sh.num_graphs, m)
        graphs += sh.num_graphs
        matches += m
    for r in range(world):
        assert (out[r]["total_data_graphs"], out[r]["total_matches"], out[r]["average_matches"]) == (graphs, matches, matches // world)
        assert out[r]["rank_ms"] == out[0]["rank_ms"] and len(out[r]["rank_ms"]) == world
