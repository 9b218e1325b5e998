"""Graph mode: the captured device program, replayed several times with eager
work in between, gives the eager results on every replay (and the oracle's);
a change of batch shape triggers a new capture."""
import ctypes as C

import numpy as np
import pytest

import gasal2_amd as g
from parity import assert_parity, fields, run_both

pytestmark = pytest.mark.gpu


def stats(a):
    c, l = C.c_uint64(), C.c_uint64()
    g.lib().gasal2_graph_stats(C.byref(a.s), C.byref(c), C.byref(l))
    return c.value, l.value


def same(x, y, keys):
    for k in keys:
        assert np.array_equal(x[k], y[k]), k


MODES = [
    dict(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE, isReverseComplement=True),
    dict(algo=g.LOCAL, start_pos=g.WITHOUT_START, isReverseComplement=True),
    dict(algo=g.SEMI_GLOBAL, start_pos=g.WITH_START, semiglobal_skipping_head=g.TARGET, semiglobal_skipping_tail=g.TARGET),
    dict(algo=g.GLOBAL, start_pos=g.WITH_TB),
]


@pytest.mark.parametrize("mode", range(len(MODES)))
def test_replays_match_eager(mode):
    kw = MODES[mode]
    b = g.synth_batch(2, 1600 + mode, 1500)
    other = g.synth_batch(1, 1700 + mode, 900)
    p = g.Parameters(**kw)
    keys = fields(p.algo, p.start_pos, bool(p.secondBest))
    L = g.lib()
    try:
        L.gasal2_set_graph_mode(0)
        ea = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, n_penalty=1)
        eager = ea.run(b)
        eager_cg = [ea.cigar(k) for k in range(b.n)] if p.start_pos == g.WITH_TB else None
        ga = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, n_penalty=1)
        eo = g.Aligner(p, int(other.qlen.max()), int(other.tlen.max()), other.n, n_penalty=1)
        for rep in range(4):
            L.gasal2_set_graph_mode(1)
            got = ga.run(b)
            same(got, eager, keys)
            if eager_cg is not None:
                assert [ga.cigar(k) for k in range(b.n)] == eager_cg, f"replay {rep}: CIGARs differ"
            L.gasal2_set_graph_mode(0)
            eo.run(other)            # eager work on another storage between replays
            ea.sync()
        assert stats(ga) == (1, 4)
        assert stats(ea) == (0, 0)
        for x in (ea, ga, eo):
            x.close()
    finally:
        L.gasal2_set_graph_mode(1)
    # the eager result itself against the oracle
    got, ref = run_both(b, p.algo, p.start_pos, bool(p.secondBest), head=p.semiglobal_skipping_head,
                        tail=p.semiglobal_skipping_tail, is_rc=bool(p.isReverseComplement), n_penalty=1)
    assert_parity(got, ref, p.algo, p.start_pos, bool(p.secondBest), b)


def test_shape_change_recaptures():
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE)
    b1 = g.synth_batch(1, 1800, 700)
    b2 = g.synth_batch(2, 1801, 500)
    a = g.Aligner(p, 600, 600, 1000, n_penalty=1)
    try:
        r1 = a.run(b1)
        r2 = a.run(b2)
        r1b = a.run(b1)
        r1c = a.run(b1)
        assert stats(a) == (3, 4)
        keys = fields(g.LOCAL, g.WITH_TB, True)
        same(r1, r1b, keys)
        same(r1, r1c, keys)
        for b, r in ((b1, r1), (b2, r2)):
            _, ref = run_both(b, g.LOCAL, g.WITH_TB, True, n_penalty=1, cigars=False)
            for k in keys:
                assert np.array_equal(r[k], ref[k]), k
    finally:
        a.close()
