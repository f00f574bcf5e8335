"""Write the golden fixtures of tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py [--config1] [--config2]

* config0.json -- config 0 (2,000 molecules seed 0, 20 queries k in U[4,12], R = 6):
  per-level signature digests, candidate digests, the full GMCR and per-pair counts for
  find-first and find-all.
* config1.json -- config 1 (100,000 molecules seed 1, 100 queries k in U[4,20] with 20%
  perturbed, R = 6): input digests, candidate/GMCR/count digests and match totals.
* config2_subset.json -- config 2's exact 1,000-query set (seed 2) against the first 20,000
  molecules of seed 2: input digests, GMCR/count digests and match totals.

The oracle itself is pinned by tests/test_oracle_pinning.py; these fixtures freeze its
output so the GPU tests can compare against it without re-running the CPU path.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from sigmo_amd import capi  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests._util import digest  # noqa: E402

HERE = Path(__file__).resolve().parent


def config0() -> dict:
    qh = capi.HostGraphs.queries(0, 20, 4, 12, capi.SIZE_NORMAL)
    dh = capi.HostGraphs.molecules(0, 2000)
    qa, da = qh.arrays(), dh.arrays()
    r = O.pipeline(qa, da, 6, keep_levels=True, modes=[True, False])
    return {
        "workload": "config 0: 2000 molecules (seed 0, Normal(24,8) sizes), 20 queries k~U[4,12], R=6",
        "query_digest": digest(*qa.values()),
        "data_digest": digest(*da.values()),
        "level_digests": [[digest(d), digest(q)] for d, q in r["levels"]],
        "filtered_digest": digest(r["filtered"]),
        "words_digest": digest(r["words"]),
        "offsets": r["offsets"].tolist(),
        "indices": r["indices"].tolist(),
        "counts_first": r["counts_first"].tolist(),
        "counts_all": r["counts_all"].tolist(),
        "matches_first": r["total_first"],
        "matches_all": r["total_all"],
    }


def config1() -> dict:
    qh = capi.HostGraphs.queries(1, 100, 4, 20, capi.SIZE_NORMAL, perturb_frac=0.2)
    dh = capi.HostGraphs.molecules(1, 100_000)
    qa, da = qh.arrays(), dh.arrays()
    t = time.time()
    r = O.pipeline(qa, da, 6, modes=[True, False])
    Q, D = len(qa["node_labels"]), len(da["node_labels"])
    return {
        "workload": "config 1: 100000 molecules (seed 1), 100 queries k~U[4,20], 20% perturbed, R=6",
        "oracle_seconds": round(time.time() - t, 1),
        "oracle_threads": O.threads(),
        "query_digest": digest(*qa.values()),
        "data_digest": digest(*da.values()),
        "row_counts_digest": digest(O.candidate_counts(Q, D, r["words"])),
        "words_digest": digest(r["words"]),
        "gmcr_digest": digest(r["offsets"], r["indices"]),
        "gmcr_pairs": int(len(r["indices"])),
        "counts_first_digest": digest(r["counts_first"]),
        "counts_all_digest": digest(r["counts_all"]),
        "matches_first": r["total_first"],
        "matches_all": r["total_all"],
    }


CONFIG2_SUBSET_MOLECULES = 20_000


def config2_subset() -> dict:
    """Config 2's exact query set (1,000 queries, seed 2, k~U[4,32], 30% with a ring) against the
    first 20,000 molecules of its seed-2 LogNormal molecule stream (a prefix of the 1,000,000)."""
    qh = capi.HostGraphs.queries(2, 1000, 4, 32, capi.SIZE_LOGNORMAL, ring_frac=0.3)
    dh = capi.HostGraphs.molecules(2, CONFIG2_SUBSET_MOLECULES, capi.SIZE_LOGNORMAL)
    qa, da = qh.arrays(), dh.arrays()
    t = time.time()
    r = O.pipeline(qa, da, 6, modes=[True, False])
    return {
        "workload": f"config 2 subset: config 2's 1000 queries (seed 2, k~U[4,32], 30% with a ring) x the first "
                    f"{CONFIG2_SUBSET_MOLECULES} molecules of seed 2 (LogNormal(ln 24,0.45) sizes), R=6",
        "molecules": CONFIG2_SUBSET_MOLECULES,
        "oracle_seconds": round(time.time() - t, 1),
        "oracle_threads": O.threads(),
        "query_digest": digest(*qa.values()),
        "data_digest": digest(*da.values()),
        "gmcr_digest": digest(r["offsets"], r["indices"]),
        "gmcr_pairs": int(len(r["indices"])),
        "counts_first_digest": digest(r["counts_first"]),
        "counts_all_digest": digest(r["counts_all"]),
        "matches_first": r["total_first"],
        "matches_all": r["total_all"],
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config1", action="store_true", help="also write config1.json (minutes of CPU)")
    ap.add_argument("--config2", action="store_true", help="also write config2_subset.json (minutes of CPU)")
    args = ap.parse_args()
    (HERE / "config0.json").write_text(json.dumps(config0()) + "\n")
    print("wrote config0.json")
    if args.config1:
        c1 = config1()
        (HERE / "config1.json").write_text(json.dumps(c1, indent=1) + "\n")
        print("wrote config1.json", c1["oracle_seconds"], "s")
    if args.config2:
        c2 = config2_subset()
        (HERE / "config2_subset.json").write_text(json.dumps(c2, indent=1) + "\n")
        print("wrote config2_subset.json", c2["oracle_seconds"], "s")


if __name__ == "__main__":
    main()
