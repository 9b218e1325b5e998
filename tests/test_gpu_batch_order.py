"""Batch order across storages (gasal2_set_batch_order): several storages
submitted back to back, collected in reverse order, with the order on and
off, give every pair's oracle outputs; a destroyed storage's events are
never waited on by a later batch."""
import numpy as np
import pytest

import gasal2_amd as g
import pyoracle as O
from parity import fields

pytestmark = pytest.mark.gpu


def run_storages(subs, p, order):
    L = g.lib()
    L.gasal2_set_batch_order(order)
    al = [g.Aligner(p, int(max(s.qlen.max() for s in subs)), int(max(s.tlen.max() for s in subs)),
                    max(s.n for s in subs), n_penalty=1) for _ in subs]
    try:
        filled = [a.fill(sb, per_pair=False) for a, sb in zip(al, subs)]
        for a, (qb, tb, m) in zip(al, filled):
            g.gasal_aln_async(a.s, qb, tb, m, a.params)
        out = [None] * len(al)
        for i in reversed(range(len(al))):   # collected newest first
            while True:
                r = g.gasal_is_aln_async_done(al[i].s)
                if r == 0:
                    break
                assert r == -1, r
            out[i] = al[i]._collect(filled[i][2], filled[i][0])
        return out
    finally:
        for a in al:
            a.close()
        L.gasal2_set_batch_order(0)


@pytest.mark.parametrize("order", [2, 1, 0])
@pytest.mark.parametrize("algo,start,second", [(g.LOCAL, g.WITH_TB, True), (g.GLOBAL, g.WITH_START, False)])
def test_storages_in_order(order, algo, start, second):
    b = g.synth_batch(2, 1500 + order, 1200)
    cuts = [0, 100, 500, 520, 1200]
    subs = [b.subset(np.arange(cuts[i], cuts[i + 1])) for i in range(4)]
    p = g.Parameters(algo=algo, start_pos=start, secondBest=g.TRUE if second else g.FALSE, isReverseComplement=True)
    out = run_storages(subs, p, order)
    for sb, r in zip(subs, out):
        ref = O.align(sb, algo, start_pos=start, second_best=second, is_rc=True, n_penalty=1)
        for f in fields(algo, start, second):
            assert np.array_equal(r[f], ref[f]), f
        if start == g.WITH_TB:
            assert np.array_equal(r["n_ops"], ref["n_ops"])


def test_destroyed_storage_not_waited_on():
    """A storage is destroyed right after its batch; the next storage's batch
    must not wait on the freed events."""
    b = g.synth_batch(2, 1510, 300)
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, isReverseComplement=True)
    for _ in range(3):
        a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, n_penalty=1)
        r = a.run(b)
        a.close()
    ref = O.align(b, g.LOCAL, start_pos=g.WITH_TB, is_rc=True, n_penalty=1)
    assert np.array_equal(r["score"], ref["score"])
