"""The multi-storage driver (gasal2_pairs_align): every pair's outputs equal
the oracle's, whatever the sub-batch size and storage count, with and
without packed inputs, from a Batch's arrays and from FASTA files read by the
library, CIGARs included."""
import numpy as np
import pytest

import gasal2_amd as g
import pyoracle as O
from parity import fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batch_pairs,storages", [(1000, 1), (333, 3), (64, 2), (5000, 3)])
@pytest.mark.parametrize("packed", [False, True])
def test_stream_matches_oracle(batch_pairs, storages, packed):
    b = g.synth_batch(2, 2900 + batch_pairs, 1500)
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE, isReverseComplement=True, isPacked=packed)
    ps = g.PairSet.from_batch(b, pack=packed)
    cap = 8 * ((int(b.qlen.max()) + 7) // 8)
    got, st = ps.align(p, batch_pairs=batch_pairs, storages=storages, n_penalty=1, cigar_cap=cap)
    assert st["sub_batches"] == (b.n + batch_pairs - 1) // batch_pairs
    ref = O.align(b, g.LOCAL, start_pos=g.WITH_TB, second_best=True, is_rc=True, n_penalty=1)
    for f in fields(g.LOCAL, g.WITH_TB, True):
        assert np.array_equal(got[f], ref[f]), f
    assert np.array_equal(got["n_ops"], ref["n_ops"])
    for k in range(b.n):
        slot = 8 * ((int(b.qlen[k]) + 7) // 8)
        nb = min(int(ref["n_ops"][k]), slot)
        assert bytes(got["cigar_slots"][k, :nb]) == ref["cigars"][k][:nb], k


@pytest.mark.parametrize("algo,start", [(g.GLOBAL, g.WITH_START), (g.SEMI_GLOBAL, g.WITH_START)])
def test_stream_other_modes_from_fasta(tmp_path, algo, start):
    b = g.synth_batch(2, 2950 + algo, 800)
    for name, data, offs, lens, ops in (("q.fa", b.q, b.qoff, b.qlen, b.qop), ("t.fa", b.t, b.toff, b.tlen, b.top)):
        with open(tmp_path / name, "wb") as f:
            for i in range(b.n):
                f.write(b"></+"[ops[i]:ops[i] + 1] + b"p%d\n" % i + bytes(data[offs[i]:offs[i] + lens[i]]) + b"\n")
    ps = g.PairSet.read(str(tmp_path / "q.fa"), str(tmp_path / "t.fa"), pack=True)
    try:
        p = g.Parameters(algo=algo, start_pos=start, isReverseComplement=True, isPacked=True)
        got, _ = ps.align(p, batch_pairs=300, storages=3, n_penalty=1)
    finally:
        ps.close()
    ref = O.align(b, algo, start_pos=start, is_rc=True, n_penalty=1)
    for f in fields(algo, start, False):
        assert np.array_equal(got[f], ref[f]), f
