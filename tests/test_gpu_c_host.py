"""The C99 host program (gasal2_amd/csrc/c_host.c: public headers compiled by
gcc, Parameters built through gasal2_parameters_new/set, one batch through
gasal_host_batch_fill, gasal_aln_async and gasal_is_aln_async_done) against
the oracle, unpacked and with isPacked=true."""
import os
import subprocess

import numpy as np
import pytest

import gasal2_amd as g
import pyoracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C_HOST = os.path.join(ROOT, "gasal2_amd", "c_host")


def write_fasta(path, seqs):
    with open(path, "wb") as f:
        for i, s in enumerate(seqs):
            f.write(b">%d\n" % i + s + b"\n")


def run_c_host(tmp_path, b, algo, start, second, packed, n_penalty):
    qf, tf, of = tmp_path / "q.fa", tmp_path / "t.fa", tmp_path / "ops.txt"
    write_fasta(qf, [b.query(i) for i in range(b.n)])
    write_fasta(tf, [b.target(i) for i in range(b.n)])
    of.write_text("".join(f"{int(x)}\n" for x in b.qop))
    cmd = [C_HOST] + (["-p"] if packed else []) + ["-n", str(n_penalty), "-r", str(of),
                                                  str(int(algo)), str(int(start)), str(int(second)), str(qf), str(tf)]
    print("c_host:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [ln.split() for ln in r.stdout.strip().splitlines()]
    assert len(rows) == b.n
    return rows


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("algo,start,second", [(g.LOCAL, g.WITH_TB, True), (g.GLOBAL, g.WITH_START, False),
                                               (g.SEMI_GLOBAL, g.WITH_START, False)])
def test_c_host_matches_oracle(tmp_path, algo, start, second, packed):
    b = g.synth_batch(2, 1400 + int(algo), 400)
    rows = run_c_host(tmp_path, b, algo, start, second, packed, 1)
    ref = O.align(b, algo, start_pos=start, second_best=second, is_rc=True, n_penalty=1)
    cols = {"score": 1, "qend": 2, "tend": 3, "qstart": 4, "tstart": 5, "s_score": 6, "s_qend": 7, "s_tend": 8}
    from parity import fields
    for f in fields(algo, start, second):
        got = np.array([int(r[cols[f]]) for r in rows])
        assert np.array_equal(got, ref[f]), f
    if start == g.WITH_TB:
        for k, r in enumerate(rows):
            assert int(r[9]) == ref["n_ops"][k]
            slot = 8 * ((int(b.qlen[k]) + 7) // 8)
            nb = min(int(ref["n_ops"][k]), slot)
            assert bytes.fromhex(r[10] if len(r) > 10 else "") == ref["cigars"][k][:nb], k
