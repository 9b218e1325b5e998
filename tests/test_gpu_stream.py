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


@pytest.mark.parametrize("order", [0, 1, 2])
def test_stream_orders_give_the_same_results(order):
    b = g.synth_batch(2, 2990 + order, 2000)
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE, isReverseComplement=True)
    ps = g.PairSet.from_batch(b)
    st = g.Stream(p, 500, 3, int(b.qlen.max()), int(b.tlen.max()))
    try:
        st.set_order(order)
        got, stats = ps.align(p, n_penalty=1, stream=st)
        again, _ = ps.align(p, n_penalty=1, stream=st)      # the storages reused
    finally:
        st.close()
    assert stats["sub_batches"] == 4
    ref = O.align(b, g.LOCAL, start_pos=g.WITH_TB, second_best=True, is_rc=True, n_penalty=1, with_cigar=False)
    for f in fields(g.LOCAL, g.WITH_TB, True):
        assert np.array_equal(got[f], ref[f]), f
        assert np.array_equal(again[f], ref[f]), f


def test_stream_sam_from_fasta(tmp_path):
    """FASTA in -> stream -> SAM out; every record equals the one built from
    the oracle's outputs (gasal2_sam_record), long CIGARs included."""
    b = g.synth_batch(2, 2999, 1200)
    for name, data, offs, lens, ops in (("q.fa", b.q, b.qoff, b.qlen, b.qop), ("t.fa", b.t, b.toff, b.tlen, b.top)):
        with open(tmp_path / name, "wb") as f:
            for i in range(b.n):
                f.write(b"></+"[ops[i]:ops[i] + 1] + b"r%d desc\n" % i + bytes(data[offs[i]:offs[i] + lens[i]]) + b"\n")
    ps = g.PairSet.read(str(tmp_path / "q.fa"), str(tmp_path / "t.fa"), names=True)
    try:
        p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE, isReverseComplement=True)
        cap = int(b.qlen.max() + b.tlen.max()) + 8
        got, _ = ps.align(p, batch_pairs=500, storages=2, n_penalty=1, cigar_cap=cap)
        ps.write_sam(tmp_path / "out.sam", got, threads=4)
    finally:
        ps.close()
    ref = O.align(b, g.LOCAL, start_pos=g.WITH_TB, second_best=True, is_rc=True, n_penalty=1)
    lines = (tmp_path / "out.sam").read_text().splitlines(keepends=True)
    body = [l for l in lines if not l.startswith("@")]
    assert len(body) == b.n and sum(l.startswith("@SQ") for l in lines) == b.n
    for k in range(b.n):
        q = bytes(b.q[b.qoff[k]:b.qoff[k] + b.qlen[k]])
        t = bytes(b.t[b.toff[k]:b.toff[k] + b.tlen[k]])
        want = g.sam_record(q, t, int(ref["score"][k]), int(ref["qend"][k]), int(ref["tend"][k]), ref["cigars"][k],
                            qname=f"r{k}", rname=f"r{k}", query_op=int(b.qop[k]), target_op=int(b.top[k]),
                            second_score=int(ref["s_score"][k]), rc_enabled=True)
        assert body[k] == want, k
