"""Writes the committed fixtures under tests/golden/ (run in the build
container, where the reference tree is mounted; nothing at test time reads it):

  ref_query_1000.fasta.gz / ref_target_1000.fasta.gz
      the first 1000 pairs of the reference's own example data
      (test_prog/query_batch.fasta.gz, test_prog/target_batch.fasta.gz), verbatim;
  ref_expected_1000.npz
      the CPU oracle's outputs on those pairs for the modes listed in MODES
      (scores (1,4,6,1), N_CODE 0x4E, N_PENALTY undefined; KSW seed score 30).
"""
import gzip
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

REF = "/root/reference/test_prog"
N = 1000

# name -> oracle kwargs (algo, start, second, head, tail, k_band)
MODES = {
    "local": dict(algo=3, start_pos=0),
    "local_start_second": dict(algo=3, start_pos=1, second_best=True),
    "local_tb_second": dict(algo=3, start_pos=2, second_best=True),
    "global": dict(algo=1, start_pos=0),
    "global_tb": dict(algo=1, start_pos=2),
    "semi_tt_start": dict(algo=2, start_pos=1, head=2, tail=2),
    "semi_bb_second": dict(algo=2, start_pos=0, second_best=True, head=3, tail=3),
    "banded32": dict(algo=5, start_pos=0, k_band=32),
    "ksw": dict(algo=6, start_pos=0),
}


def read_pairs(qpath, tpath, n):
    with gzip.open(qpath, "rt") as fq, gzip.open(tpath, "rt") as ft:
        ql, tl = fq.read().split("\n"), ft.read().split("\n")
    pairs, i = [], 0
    while len(pairs) < n and i + 1 < len(ql):
        pairs.append((ql[i], ql[i + 1], tl[i], tl[i + 1]))
        i += 2
    return pairs


def load_fixture_batch():
    import gasal2_amd as g
    pairs = read_pairs(os.path.join(HERE, "ref_query_1000.fasta.gz"), os.path.join(HERE, "ref_target_1000.fasta.gz"),
                       N)
    ops = {">": 0, "<": 1, "/": 2, "+": 3}
    qs = [p[1].encode() for p in pairs]
    ts = [p[3].encode() for p in pairs]
    qop = [ops.get(p[0][0], 0) for p in pairs]
    top = [ops.get(p[2][0], 0) for p in pairs]
    return g.make_batch(qs, ts, qop=qop, top=top, seeds=[30] * len(qs)), pairs


def main():
    import pyoracle as O
    pairs = read_pairs(os.path.join(REF, "query_batch.fasta.gz"), os.path.join(REF, "target_batch.fasta.gz"), N)
    with gzip.open(os.path.join(HERE, "ref_query_1000.fasta.gz"), "wt") as f:
        for hq, sq, _, _ in pairs:
            f.write(f"{hq}\n{sq}\n")
    with gzip.open(os.path.join(HERE, "ref_target_1000.fasta.gz"), "wt") as f:
        for _, _, ht, st in pairs:
            f.write(f"{ht}\n{st}\n")
    b, _ = load_fixture_batch()
    out = {}
    for name, kw in MODES.items():
        r = O.align(b, **kw)
        for k in ("score", "qend", "tend", "qstart", "tstart", "s_score", "s_qend", "s_tend", "n_ops"):
            out[f"{name}.{k}"] = r[k]
        if "cigars" in r:
            flat = np.frombuffer(b"".join(r["cigars"]), dtype=np.uint8)
            out[f"{name}.cigar_flat"] = flat
    np.savez_compressed(os.path.join(HERE, "ref_expected_1000.npz"), **out)
    print("wrote fixtures for", len(pairs), "pairs")


if __name__ == "__main__":
    main()
