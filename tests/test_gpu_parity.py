"""GPU parity: the HIP library (through its C-ABI) against the CPU oracle,
bit-exact on every output each mode defines, on seeded synthetic batches."""
import numpy as np
import pytest

import gasal2_amd as g
from parity import assert_parity, run_both

pytestmark = pytest.mark.gpu

LOCAL_MODES = [(s, b) for s in (g.WITHOUT_START, g.WITH_START, g.WITH_TB) for b in (False, True)]


@pytest.mark.parametrize("start,second", LOCAL_MODES)
@pytest.mark.parametrize("cfg,n_penalty", [(1, None), (2, 1)])
def test_local(cfg, n_penalty, start, second):
    b = g.synth_batch(cfg, 100 + cfg, 700)
    got, ref = run_both(b, g.LOCAL, start, second, is_rc=(cfg == 2), n_penalty=n_penalty)
    assert_parity(got, ref, g.LOCAL, start, second, b)


@pytest.mark.parametrize("start", [g.WITHOUT_START, g.WITH_START, g.WITH_TB])
@pytest.mark.parametrize("cfg,n_penalty", [(1, None), (2, 1)])
def test_global(cfg, n_penalty, start):
    b = g.synth_batch(cfg, 200 + cfg, 500)
    got, ref = run_both(b, g.GLOBAL, start, is_rc=(cfg == 2), n_penalty=n_penalty)
    assert_parity(got, ref, g.GLOBAL, start, False, b)


SEMI = [(h, t) for h in (g.NONE, g.QUERY, g.TARGET, g.BOTH) for t in (g.NONE, g.QUERY, g.TARGET, g.BOTH)]


@pytest.mark.parametrize("head,tail", SEMI)
@pytest.mark.parametrize("start,second", [(g.WITHOUT_START, True), (g.WITH_START, False), (g.WITH_TB, False)])
def test_semi_global(head, tail, start, second):
    b = g.synth_batch(2, 300 + head * 4 + tail, 300)
    got, ref = run_both(b, g.SEMI_GLOBAL, start, second, head=head, tail=tail, n_penalty=1)
    assert_parity(got, ref, g.SEMI_GLOBAL, start, second, b)


@pytest.mark.parametrize("k_band", [8, 32, 64])
def test_banded(k_band):
    b = g.synth_batch(4, 400 + k_band, 600)
    got, ref = run_both(b, g.BANDED, k_band=k_band)
    assert_parity(got, ref, g.BANDED, g.WITHOUT_START, False, b)


@pytest.mark.parametrize("scores", [(1, 4, 6, 1), (2, 9, 5, 2)])
def test_banded_per_cell_scores(scores):
    """BANDED without the score tables: a nibble-class collision ('I' with 'A')
    or a mismatch past the 4-bit range takes the per-cell substitution rule."""
    b0 = g.synth_batch(4, 470, 300)
    qs = [b0.query(i) for i in range(b0.n)]
    ts = [b0.target(i) for i in range(b0.n)]
    qs[7] = qs[7][:30] + b"IIA" + qs[7][33:]
    b = g.make_batch(qs, ts)
    got, ref = run_both(b, g.BANDED, k_band=32, scores=scores, n_penalty=1)
    assert_parity(got, ref, g.BANDED, g.WITHOUT_START, False, b)


@pytest.mark.parametrize("second", [False, True])
def test_ksw(second):
    b = g.synth_batch(4, 500, 600)
    got, ref = run_both(b, g.KSW, second=second)
    assert_parity(got, ref, g.KSW, g.WITHOUT_START, second, b)


@pytest.mark.parametrize("h0", [40, 70000])
def test_ksw_long_queries(h0):
    """Queries over 1024 bases: the kernel reads the query words from global
    memory instead of LDS; h0 = 70000 also takes the int32 row (values past
    16 bits).  Short pairs in the same batch, and pairs whose targets end
    early, leave lanes of a block outside their ranges."""
    rng = np.random.default_rng(4300 + h0)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    qs, ts = [], []
    for lq, lt in [(1100, 1150), (1500, 700), (300, 320), (1030, 1030), (64, 2000), (1700, 90)] * 3:
        q = acgt[rng.integers(0, 4, lq)]
        t = q.copy()[:lt] if lt <= lq else np.concatenate([q, acgt[rng.integers(0, 4, lt - lq)]])
        flip = rng.random(len(t)) < 0.04
        t[flip] = acgt[rng.integers(0, 4, int(flip.sum()))]
        qs.append(q.tobytes())
        ts.append(t.tobytes())
    b = g.make_batch(qs, ts, seeds=np.full(len(qs), h0, np.uint32))
    got, ref = run_both(b, g.KSW)
    assert_parity(got, ref, g.KSW, g.WITHOUT_START, False, b)


@pytest.mark.parametrize("second", [False, True])
@pytest.mark.parametrize("ckpt", [0, 1])
def test_local_tb_checkpoint_switch(ckpt, second):
    """LOCAL WITH_TB with the checkpointed traceback off (direction words) and on."""
    g.lib().gasal2_set_tb_checkpoint(ckpt)
    try:
        b = g.synth_batch(2, 930 + ckpt, 600)
        got, ref = run_both(b, g.LOCAL, g.WITH_TB, second, is_rc=True, n_penalty=1)
        assert_parity(got, ref, g.LOCAL, g.WITH_TB, second, b)
    finally:
        g.lib().gasal2_set_tb_checkpoint(1)   # the default


@pytest.mark.parametrize("start,second", [(g.WITH_TB, True), (g.WITH_TB, False), (g.WITHOUT_START, False),
                                          (g.WITH_START, True)])
def test_local_panels_off(start, second):
    """The fast LOCAL kernels reading each pair's own words (panels off)."""
    g.lib().gasal2_set_panels(0)
    try:
        b = g.synth_batch(2, 940 + start, 500)
        got, ref = run_both(b, g.LOCAL, start, second, is_rc=True, n_penalty=1)
        assert_parity(got, ref, g.LOCAL, start, second, b)
    finally:
        g.lib().gasal2_set_panels(1)


@pytest.mark.parametrize("w,zdrop", [(50, 100), (8, 0), (-1, 20), (-1, 5), (0, 0), (-1, 0)])
def test_ksw_extension(w, zdrop):
    """KSW with bwa-style band and z-drop (the reference runs w < 0, z-drop 0)."""
    import pyoracle as O
    b0 = g.synth_batch(4, 510, 500)
    qs = [b0.query(i) for i in range(b0.n)] + [b"ACGTTGCAAC" * 3 + b"AAAAA" + b"GATTACAGGC" * 10]
    ts = [b0.target(i) for i in range(b0.n)] + [b"ACGTTGCAAC" * 3 + b"CCCCC" + b"GATTACAGGC" * 10]
    sd = list(b0.seeds) + [20]
    b = g.make_batch(qs, ts, seeds=sd)
    g.set_ksw_extension(w, zdrop)
    try:
        p = g.Parameters(algo=g.KSW, start_pos=g.WITHOUT_START)
        a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n)
        try:
            got = a.run(b)
        finally:
            a.close()
    finally:
        g.set_ksw_extension()
    ref = O.align(b, g.KSW, ksw_w=w, ksw_zdrop=zdrop)
    assert_parity(got, ref, g.KSW, g.WITHOUT_START, False, b)


@pytest.mark.parametrize("start,second", [(g.WITH_TB, True), (g.WITH_START, False), (g.WITHOUT_START, True)])
def test_local_reference_order_kernel(start, second):
    """The reference-order kernel (used when the fast path does not apply)
    forced by dispatch policy 1."""
    g.lib().gasal2_set_dispatch(1)
    try:
        b = g.synth_batch(2, 900, 500)
        got, ref = run_both(b, g.LOCAL, start, second, is_rc=True, n_penalty=1)
        assert_parity(got, ref, g.LOCAL, start, second, b)
    finally:
        g.lib().gasal2_set_dispatch(0)


def test_local_nibble_class_collision_falls_back():
    """'I' (nibble 9) shares a class with 'A' (nibble 1): the census routes the
    batch to the reference-order kernel; results stay exact."""
    b0 = g.synth_batch(1, 901, 300)
    qs = [b0.query(i) for i in range(b0.n)]
    ts = [b0.target(i) for i in range(b0.n)]
    qs[5] = qs[5][:20] + b"IIA" + qs[5][23:]
    b = g.make_batch(qs, ts)
    got, ref = run_both(b, g.LOCAL, g.WITH_TB, True)
    assert_parity(got, ref, g.LOCAL, g.WITH_TB, True, b)


@pytest.mark.parametrize("scores", [(1, 9, 6, 1), (8, 4, 6, 1), (2, 3, 5, 2)])
def test_local_other_scores(scores):
    b = g.synth_batch(2, 902, 400)
    got, ref = run_both(b, g.LOCAL, g.WITH_TB, True, is_rc=True, scores=scores, n_penalty=1)
    assert_parity(got, ref, g.LOCAL, g.WITH_TB, True, b)


@pytest.mark.parametrize("n_code", [0x4E, 0x0E])
def test_reverse_complement_both_sides(n_code):
    """Random per-pair ops on query and target; N_CODE 0x0E makes the reverse
    shift by the padding count (src/kernels/pack_rc_seqs.h:112-165)."""
    import numpy as np
    b0 = g.synth_batch(1, 903, 400)
    rng = np.random.default_rng(5)
    qs = [b0.query(i) for i in range(b0.n)]
    ts = [b0.target(i) for i in range(b0.n)]
    b = g.make_batch(qs, ts, qop=rng.integers(0, 4, b0.n), top=rng.integers(0, 4, b0.n), n_code=n_code)
    got, ref = run_both(b, g.LOCAL, g.WITH_TB, True, is_rc=True, n_code=n_code)
    assert_parity(got, ref, g.LOCAL, g.WITH_TB, True, b)


def test_local_tiny_edge_cases():
    qs = [b"A", b"ACGT", b"ACGTACGT", b"ACGTACGTA", b"N", b"NNNNNNNNN", b"TTTT", b"ACGTNACGT" * 5]
    ts = [b"A", b"ACGTTT", b"ACGTACGT", b"TACGTACGT", b"A", b"ACGTACGTAC", b"GGGG", b"ACGTACGT" * 6]
    b = g.make_batch(qs, ts)
    for start in (g.WITHOUT_START, g.WITH_START, g.WITH_TB):
        got, ref = run_both(b, g.LOCAL, start, True, n_penalty=1)
        assert_parity(got, ref, g.LOCAL, start, True, b)


@pytest.mark.parametrize("algo,second", [(g.LOCAL, True), (g.LOCAL, False), (g.GLOBAL, False)])
def test_per_group_tile_regions(algo, second):
    """WITH_TB with per-group direction-tile regions (forced by a zero budget)
    on a batch of uneven lengths, with and without the one-wave path."""
    b = g.synth_batch(3, 1010, 600)
    L = g.lib()
    L.gasal2_set_tile_budget(0)
    try:
        got, ref = run_both(b, algo, g.WITH_TB, second, n_penalty=1)
        assert_parity(got, ref, algo, g.WITH_TB, second, b)
    finally:
        L.gasal2_set_tile_budget(64 << 30)


def test_last_paths_report_the_kernels_taken():
    """gasal2_last_paths: checkpoints for LOCAL WITH_TB, packed 16-bit for the
    score modes of the full-feature shapes, neither for WITH_TB off.  Lane
    kernels only: a batch this small would otherwise run one wave per pair."""
    b = g.synth_batch(2, 990, 300)
    g.lib().gasal2_set_dispatch(2)
    try:
        _last_paths_cases(b)
    finally:
        g.lib().gasal2_set_dispatch(0)


def _last_paths_cases(b):
    cases = [(g.LOCAL, g.WITH_TB, True, 4), (g.LOCAL, g.WITHOUT_START, False, 1), (g.GLOBAL, g.WITH_START, False, 1),
             (g.SEMI_GLOBAL, g.WITH_START, False, 1)]
    for algo, start, second, bit in cases:
        p = g.Parameters(algo=algo, start_pos=start, secondBest=g.TRUE if second else g.FALSE,
                         semiglobal_skipping_head=g.TARGET, semiglobal_skipping_tail=g.TARGET)
        a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, n_penalty=1)
        try:
            a.run(b)
            assert a.last_paths() & bit, (algo, start, a.last_paths())
        finally:
            a.close()
    g.lib().gasal2_set_tb_checkpoint(0)
    try:
        p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB)
        a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, n_penalty=1)
        try:
            a.run(b)
            assert not a.last_paths() & 4
        finally:
            a.close()
    finally:
        g.lib().gasal2_set_tb_checkpoint(1)   # the default


def _many_shapes_batch(n, seed, n_code=0x4E):
    """n pairs, query and target lengths 1..400 drawn independently (so a block
    of 8192 pairs holds more distinct (query, target) word shapes than the
    sort's 2048-slot hash table), each target a 5%-mutated copy of the query
    with random flanks or a mutated piece of it; the first 16 pairs are long
    (2000 x 2600 bp); random reverse / complement ops on both sides."""
    import numpy as np
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)

    def rand(m):
        return acgt[rng.integers(0, 4, m)]

    def mutate(s):
        s = s.copy()
        m = rng.random(len(s)) < 0.05
        s[m] = rand(int(m.sum()))
        return s

    qs, ts = [], []
    for k in range(n):
        lq, lt = (2000, 2600) if k < 16 else (int(rng.integers(1, 401)), int(rng.integers(1, 401)))
        q = rand(lq)
        if lt >= lq:
            a = int(rng.integers(0, lt - lq + 1))
            t = np.concatenate([rand(a), mutate(q), rand(lt - lq - a)])
        else:
            a = int(rng.integers(0, lq - lt + 1))
            t = mutate(q[a:a + lt])
        qs.append(q.tobytes())
        ts.append(t.tobytes())
    return g.make_batch(qs, ts, qop=rng.integers(0, 4, n), top=rng.integers(0, 4, n), n_code=n_code)


@pytest.mark.parametrize("start,second,n_code", [(g.WITHOUT_START, False, 0x4E), (g.WITH_TB, True, 0x4E),
                                                 (g.WITH_TB, False, 0x0E)])
def test_pair_order_many_shapes(start, second, n_code):
    """The shape sort over two blocks of pairs, one of them overflowing its
    hash table into global counters, with long one-wave pairs and multi-round
    reverse / complement (N_CODE 0x0E: the reference's shifting reverse):
    every pair is aligned once, into its own results."""
    b = _many_shapes_batch(9000, 77, n_code)
    got, ref = run_both(b, g.LOCAL, start, second, is_rc=True, n_code=n_code)
    assert_parity(got, ref, g.LOCAL, start, second, b)


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("second", [False, True])
def test_local_tb_fused_walk(fused, second):
    """Checkpointed LOCAL WITH_TB with the traceback walked inside the DP kernel
    (default) or by the separate traceback kernel: identical CIGARs and starts,
    also on the many-shapes batch with one-wave pairs beside the lane path."""
    L = g.lib()
    L.gasal2_set_tb_fused(fused)
    try:
        b = g.synth_batch(2, 1212, 700)
        got, ref = run_both(b, g.LOCAL, g.WITH_TB, second, is_rc=True, n_penalty=1)
        assert_parity(got, ref, g.LOCAL, g.WITH_TB, second, b)
        b = _many_shapes_batch(2000, 78)
        got, ref = run_both(b, g.LOCAL, g.WITH_TB, second, is_rc=True)
        assert_parity(got, ref, g.LOCAL, g.WITH_TB, second, b)
    finally:
        L.gasal2_set_tb_fused(0)


def test_short_read_set_whole():
    """The short-read workload itself, every pair: seed 1, 10,000 pairs of
    150 bp queries against 50 + q + 50 targets, LOCAL WITHOUT_START (scores,
    ends) and LOCAL WITH_TB + second best (CIGARs included)."""
    b = g.synth_batch(1, 1, 10_000)
    got, ref = run_both(b, g.LOCAL, g.WITHOUT_START, False)
    assert_parity(got, ref, g.LOCAL, g.WITHOUT_START, False, b)
    got, ref = run_both(b, g.LOCAL, g.WITH_TB, True)
    assert_parity(got, ref, g.LOCAL, g.WITH_TB, True, b)
