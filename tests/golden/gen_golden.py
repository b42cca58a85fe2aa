#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/.

reference_vectors.json -- the known answers the reference's own tests assert
    (inputs and expected outputs copied as data, with their file:line).
pyref_vectors.json -- values of the independent big-integer restatement in
    tests/pyref.py on seeded inputs (splitmix64, ntHash, canonical k-mers and
    folds, HLL index/rank, CMS cells).

Run: python tests/golden/gen_golden.py
"""
from __future__ import annotations

import json
import random
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from tests import pyref as P  # noqa: E402

REFERENCE = {
    "rev_comp_word": {  # src/common/bitseq.rs:434-446
        "cases": [{"word": 0b000110, "bases": 3, "expect": 0b011011},
                  {"word": 0b00011011, "bases": 4, "expect": 0b00011011}],
    },
    "reverse_complement_ascii": {  # src/common/bitseq.rs:420-431
        "cases": [["ACGTACGT", "ACGTACGT"], ["AAAACCCGGT", "ACCGGGTTTT"]],
    },
    "cms_basic": {  # src/common/sketch.rs:304-319
        "width": 1024, "key": 42, "after0": 0, "adds1": 5, "after1": 5, "adds2": 20, "after2": 15,
    },
    "cms_concurrent": {  # src/common/sketch.rs:321-342
        "width": 4096, "key": 12345, "threads": 4, "per_thread": 3, "expect": 12,
    },
    "count_fastx_exact": {  # src/ingest/estimator.rs:463-486
        "pairs": 10, "r1_len": 12, "r2_len": 12,
        "expect": {"exact_read_pairs": 10, "total_bases_r1": 120, "total_bases_r2": 120, "total_bases": 240,
                   "min_read_len": 12, "max_read_len": 12, "median_read_len": 12, "mean_read_len": 12.0},
    },
    "dynamic_kmers": {  # src/ingest/estimator.rs:528-546
        "median150": {"first": 31, "last": 127, "lo": 31, "hi": 127},
        "median50": {"first": 21},
    },
    "rolling_gate": {  # src/common/kmer.rs:580-651: LCG bases, k = 21/55/99/127
        "seed": 0xCAFEBABEDEADBEEF, "mul": 6364136223846793005, "add": 1442695040888963407, "length": 300,
        "ks": [21, 55, 99, 127],
    },
    "nthash_rolling": {  # src/common/kmer.rs:653-664
        "seq": "ACGTACGTACGTTTGCAACCGGTTACGT", "k": 15,
    },
    "kmer_basic": {  # src/common/kmer.rs:543-577
        "u64": "ACGTACGTACGTACGTACGTACGT",
        "u128": "ACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGT",
        "u256": "ACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGT",
    },
    "hll_gates": {  # src/common/sketch.rs:261-407
        "single_100k": {"n": 100000, "max_err": 0.03},
        "merge": {"split": 50000, "n": 100000, "max_err": 0.03},
        "raw_min": {"n": 200000, "fwd_xor": 0x123456789ABCDEF0, "rev_xor": 0xFEDCBA9876543210,
                    "raw_err_below": -0.50, "fixed_max_err": 0.025},
        "multi_scale": {"ns": [50000, 200000, 1000000, 2500000], "fwd_xor": 0xABCDEF0123456789,
                        "rev_xor": 0x9876543210FEDCBA, "max_err": 0.025},
    },
    "circular_repeat": {  # src/graph/build.rs:604-687
        "k": 31, "repeat_len": 200, "flank_len": 4800, "seeds": [0x1111222233334444, 0x5555666677778888,
                                                               0x9999AAAABBBBCCCC],
        "read_seed": 0xFEEDBEEF, "depth": 40, "read_len": 150, "target_solid_cov": 5, "expect_unitigs": 3,
    },
}


def pyref_vectors() -> dict:
    rnd = random.Random(20261203)
    out = {}
    xs = [0, 1, 2, 0xFFFFFFFFFFFFFFFF, 0x8000000000000000] + [rnd.getrandbits(64) for _ in range(32)]
    out["splitmix64"] = [[x, P.splitmix64(x)] for x in xs]
    seqs = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(130, 160))) for _ in range(4)]
    out["nthash"] = []
    for k in (15, 21, 31, 63, 64, 65, 127):
        s = seqs[0][:k]
        out["nthash"].append({"seq": s, "fwd": P.nthash_fwd(s), "rev": P.nthash_rev(s), "can": P.nthash_canonical(s)})
    out["canonical"] = []
    for k in (1, 5, 21, 31, 32, 33, 55, 63, 64, 65, 77, 96, 99, 127):
        s = seqs[k % 4]
        wins = []
        for i in range(0, len(s) - k + 1, 7):
            x, rc = P.canonical(s[i:i + k])
            wins.append([i, P.ref_words(x, k), int(rc), P.fold(x)])
        out["canonical"].append({"k": k, "seq": s, "windows": wins})
    hs = [0, 1, (1 << 50) - 1, 1 << 50, 0xFFFFFFFFFFFFFFFF] + [rnd.getrandbits(64) for _ in range(40)]
    out["hll_index_rho"] = [[h, p, *P.hll_index_rho(h, p)] for h in hs for p in (10, 14, 16)]
    keys = [rnd.getrandbits(64) for _ in range(16)]
    out["cms_cells"] = [[key, w, P.cms_cells(key, w)] for key in keys for w in (1024, 1 << 22)]
    return out


def main() -> None:
    (HERE / "reference_vectors.json").write_text(json.dumps(REFERENCE, indent=1) + "\n")
    (HERE / "pyref_vectors.json").write_text(json.dumps(pyref_vectors()) + "\n")
    print("wrote", HERE / "reference_vectors.json", HERE / "pyref_vectors.json")


if __name__ == "__main__":
    main()
