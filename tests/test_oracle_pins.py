"""Pins the CPU oracle against what the mathematics fixes (CPU only).

The reference ships no golden outputs (its only fixture is input FASTA), so
each oracle function is checked against an independent formulation: closed
forms for pack / reverse-complement, textbook automaton or Gotoh DPs and an
exhaustive enumerator for the DP scores, definitional re-computations of the
order-dependent outputs (end tie-break, second best, start positions), and
known-answer tests for the documented quirks.  A planted mistake in the
oracle (dropped term, wrong sign or index) fails at least one of these.
"""
import random

import numpy as np
import pytest

import gasal2_amd as g
import pyoracle as O
from refmath import (BOTH_, NONE_, QUERY_, TARGET_, automaton_global, automaton_local, banded_reference,
                     enumerate_local, gotoh_local, gotoh_semiglobal_target_target, key_order_first,
                     ksw_unpruned_reference, local_traceback_reference, second_best_definition,
                     semiglobal_reference)

ACGT = b"ACGT"


def rnd_seq(rng, n, alphabet=ACGT):
    return bytes(rng.choice(alphabet) for _ in range(n))


def mutate(rng, s, p=0.15):
    out = bytearray()
    for c in s:
        u = rng.random()
        if u < p / 3:
            continue
        if u < 2 * p / 3:
            out += bytes([rng.choice(ACGT)])
        if u < p:
            out += bytes([rng.choice(ACGT)])
            continue
        out.append(c)
    return bytes(out) or b"A"


# ---------------------------------------------------------------- pack -----
def test_pack_all_bytes_closed_form():
    data = np.arange(256, dtype=np.uint8)
    data = np.concatenate([data, np.random.default_rng(0).integers(0, 256, 4096, dtype=np.uint8)])
    words = O.pack(data)
    # word w holds the low nibbles of bytes 8w..8w+7, first byte most significant
    for w in range(len(words)):
        hexs = "".join("%x" % (b & 15) for b in data[8 * w:8 * w + 8])
        assert int(words[w]) == int(hexs, 16)


# ------------------------------------------------- reverse / complement -----
COMP = {1: 4, 4: 1, 3: 7, 7: 3}


def words_of(seq_nibbles, pad_nib):
    n = len(seq_nibbles)
    padded = list(seq_nibbles) + [pad_nib] * ((-n) % 8)
    return [int("".join("%x" % x for x in padded[8 * k:8 * k + 8]), 16) for k in range(len(padded) // 8)]


def nibbles_of(words):
    return [int(c, 16) for c in "".join("%08x" % int(w) for w in words)]


@pytest.mark.parametrize("length", list(range(1, 41)))
@pytest.mark.parametrize("op", [1, 2, 3])
def test_rc_n_code_0x4e_reverses_padded_words(length, op):
    # N_CODE 0x4E never equals a nibble, so the reverse flips the whole padded
    # word sequence: padding nibbles (0xE) move to the front.
    rng = random.Random(length * 7 + op)
    nibs = [rng.choice([1, 3, 4, 7, 0xE]) for _ in range(length)]
    w = words_of(nibs, 0xE)
    got = nibbles_of(O.rc_one(w, length, op, 0x4E))
    exp = nibbles_of(w)
    if op & 1:
        exp = exp[::-1]
    if op & 2:
        exp = [COMP.get(x, x) for x in exp]
    assert got == exp


@pytest.mark.parametrize("length", list(range(1, 41)))
def test_rc_small_n_code_keeps_padding_at_end(length):
    # With a nibble-sized N_CODE the reverse shifts by the padding count, so the
    # real bases are reversed in place and the padding stays at the end (valid
    # when no real base in the last word equals the padding nibble).
    rng = random.Random(length)
    nibs = [rng.choice([1, 3, 4, 7]) for _ in range(length)]
    w = words_of(nibs, 0xE)
    got = nibbles_of(O.rc_one(w, length, 1, 0xE))
    pad = (-length) % 8
    assert got == nibs[::-1] + [0xE] * pad


# ---------------------------------------------------------- LOCAL score -----
def oracle_one(q, t, algo, **kw):
    b = g.make_batch([q], [t])
    return O.align(b, algo, **kw)


@pytest.mark.parametrize("seed", range(40))
def test_local_score_vs_enumeration_tiny(seed):
    rng = random.Random(seed)
    q = rnd_seq(rng, rng.randint(1, 4))
    t = rnd_seq(rng, rng.randint(1, 4))
    r = oracle_one(q, t, g.LOCAL)
    assert r["score"][0] == enumerate_local(q, t)


@pytest.mark.parametrize("seed", range(60))
def test_local_score_and_ends_vs_automaton(seed):
    rng = random.Random(1000 + seed)
    q = rnd_seq(rng, rng.randint(1, 40), b"ACGTN" if seed % 3 == 0 else ACGT)
    t = mutate(rng, q) if seed % 2 else rnd_seq(rng, rng.randint(1, 40))
    nval = -1 if seed % 4 == 1 else 0
    r = oracle_one(q, t, g.LOCAL, n_penalty=(1 if nval == -1 else None))
    # the reference computes the padded matrix: pad with N as the host batch does
    qp = q + b"N" * ((-len(q)) % 8)
    tp = t + b"N" * ((-len(t)) % 8)
    best, H = automaton_local(qp, tp, nval=nval)
    assert r["score"][0] == best
    # end = first maximal cell in key order (strip, column, row in strip)
    if best > 0:
        rr, cc = key_order_first(H, len(qp), len(tp), len(qp), lambda h: h == best)
        assert (r["tend"][0], r["qend"][0]) == (rr, cc)
    else:
        assert (r["tend"][0], r["qend"][0]) == (0, 0)


@pytest.mark.parametrize("seed", range(40))
def test_local_second_best_vs_definition(seed):
    rng = random.Random(2000 + seed)
    q = rnd_seq(rng, rng.randint(2, 30))
    t = mutate(rng, q, 0.3) if seed % 2 else rnd_seq(rng, rng.randint(2, 30))
    r = oracle_one(q, t, g.LOCAL, second_best=True)
    qp = q + b"N" * ((-len(q)) % 8)
    tp = t + b"N" * ((-len(t)) % 8)
    _, H = automaton_local(qp, tp)
    cells, rows = [], []
    for s in range((len(tp) + 7) // 8):
        for c in range(len(qp)):
            for r8 in range(8):
                rr = 8 * s + r8
                cells.append(H[rr][c])
                rows.append(rr)
    v, idx = second_best_definition(cells)
    assert r["s_score"][0] == v
    assert r["s_tend"][0] == (rows[idx] if idx is not None else 0)
    # second query end compares against the best score: last padded column
    assert r["s_qend"][0] == (len(qp) - 1 if r["score"][0] > 0 else 0)


def test_local_symmetric_under_swap():
    rng = random.Random(5)
    for _ in range(30):
        q = rnd_seq(rng, rng.randint(1, 30))
        t = mutate(rng, q)
        a = oracle_one(q, t, g.LOCAL)["score"][0]
        b = oracle_one(t, q, g.LOCAL)["score"][0]
        assert a == b


def test_local_identical_sequences():
    for n in (1, 7, 8, 9, 64, 150):
        s = bytes(random.Random(n).choice(ACGT) for _ in range(n))
        r = oracle_one(s, s, g.LOCAL, start_pos=g.WITH_TB)
        assert r["score"][0] == n
        assert (r["qend"][0], r["tend"][0]) == (n - 1, n - 1)
        assert (r["qstart"][0], r["tstart"][0]) == (0, 0)
        assert g.cigar_to_string(r["cigars"][0]) == f"{n}M"


# ---------------------------------------------------- LOCAL WITH_START -----
@pytest.mark.parametrize("seed", range(40))
def test_local_start_positions(seed):
    rng = random.Random(3000 + seed)
    core = rnd_seq(rng, rng.randint(8, 40))
    q = rnd_seq(rng, rng.randint(0, 12)) + core + rnd_seq(rng, rng.randint(0, 12))
    t = rnd_seq(rng, rng.randint(0, 20)) + mutate(rng, core, 0.1) + rnd_seq(rng, rng.randint(0, 20))
    r = oracle_one(q, t, g.LOCAL, start_pos=g.WITH_START)
    score, qe, te = int(r["score"][0]), int(r["qend"][0]), int(r["tend"][0])
    qs, ts_rep = int(r["qstart"][0]), int(r["tstart"][0])
    if score == 0:
        assert (qs, ts_rep) == (0, 0)
        return
    # the target start is reported as true + 2*(7 - (true & 7)): invert it
    cands = [x for x in range(0, te + 1) if x + 2 * (7 - (x & 7)) == ts_rep]
    assert len(cands) == 1
    ts = cands[0]
    # an alignment of that score fits exactly in [ts..te] x [qs..qe]
    qp = q + b"N" * ((-len(q)) % 8)
    tp = t + b"N" * ((-len(t)) % 8)
    sub, _ = automaton_local(qp[qs:qe + 1], tp[ts:te + 1])
    assert sub == score


# ------------------------------------------------------------- GLOBAL -----
def test_global_kat_leading_query_gap_quirk():
    # one leading query base is free: global("AC", "C") = +1 (textbook -6)
    assert oracle_one(b"AC", b"C", g.GLOBAL)["score"][0] == 1


@pytest.mark.parametrize("seed", range(60))
def test_global_vs_automaton(seed):
    rng = random.Random(4000 + seed)
    q = rnd_seq(rng, rng.randint(1, 40))
    t = mutate(rng, q, 0.2) if seed % 2 else rnd_seq(rng, rng.randint(1, 40))
    r = oracle_one(q, t, g.GLOBAL)
    assert r["score"][0] == automaton_global(q, t)


# -------------------------------------------------------- SEMI_GLOBAL -----
@pytest.mark.parametrize("seed", range(40))
def test_semiglobal_target_target_vs_gotoh(seed):
    rng = random.Random(5000 + seed)
    q = rnd_seq(rng, rng.randint(1, 30))
    t = rnd_seq(rng, rng.randint(0, 15)) + mutate(rng, q, 0.1) + rnd_seq(rng, rng.randint(0, 15))
    r = oracle_one(q, t, g.SEMI_GLOBAL, head=g.TARGET, tail=g.TARGET)
    assert r["score"][0] == gotoh_semiglobal_target_target(q, t)


# --------------------------------------------------------------- BANDED -----
@pytest.mark.parametrize("seed", range(40))
def test_banded_full_band_is_gotoh_local(seed):
    rng = random.Random(6000 + seed)
    q = rnd_seq(rng, rng.randint(1, 40))
    t = mutate(rng, q, 0.2) if seed % 2 else rnd_seq(rng, rng.randint(1, 40))
    qp = q + b"N" * ((-len(q)) % 8)
    tp = t + b"N" * ((-len(t)) % 8)
    r = oracle_one(q, t, g.BANDED, k_band=8 * (len(qp) // 8 + len(tp) // 8 + 1))
    assert r["score"][0] == gotoh_local(qp, tp)


@pytest.mark.parametrize("seed", range(30))
def test_banded_never_exceeds_full_band(seed):
    rng = random.Random(6500 + seed)
    q = rnd_seq(rng, rng.randint(10, 80))
    t = mutate(rng, q, 0.2)
    full = oracle_one(q, t, g.BANDED, k_band=1024)["score"][0]
    for kb in (8, 16, 32):
        assert oracle_one(q, t, g.BANDED, k_band=kb)["score"][0] <= full


# ------------------------------------------------------------------ KSW -----
@pytest.mark.parametrize("n,h0", [(8, 10), (9, 30), (50, 20), (150, 60)])
def test_ksw_identical_extension(n, h0):
    s = bytes(random.Random(n).choice(ACGT) for _ in range(n))
    b = g.make_batch([s], [s], seeds=[h0])
    r = O.align(b, g.KSW)
    # extending an exact match to both ends: global score h0 + n reaches the max
    assert (r["score"][0], r["qend"][0], r["tend"][0]) == (h0 + n, n, n)


def test_ksw_random_never_below_seed():
    rng = random.Random(9)
    for _ in range(40):
        q = rnd_seq(rng, rng.randint(5, 60))
        t = rnd_seq(rng, rng.randint(5, 60))
        h0 = rng.randint(10, 60)
        r = O.align(g.make_batch([q], [t], seeds=[h0]), g.KSW)
        assert r["score"][0] >= h0 - 5


# KSW extension (band w, z-drop).  bwa's ksw_extend2, which these settings
# follow, is not in the image: parity of the extension is unpinned beyond the
# properties below (no-op settings, the band constraint on the end cell,
# z-drop only removing rows, and two constructed cases each setting changes).
def _ksw(q, t, h0, w=-1, zdrop=0, end_bonus=5):
    r = O.align(g.make_batch([q], [t], seeds=[h0]), g.KSW, ksw_w=w, ksw_zdrop=zdrop, ksw_end_bonus=end_bonus)
    return int(r["score"][0]), int(r["qend"][0]), int(r["tend"][0])


@pytest.mark.parametrize("n,h0", [(20, 10), (45, 30), (120, 25)])
def test_ksw_end_bonus_closed_form(n, h0):
    # query == target except the last base: the best cell is (n-2, n-2) with
    # h0 + n - 1; reaching the query's end costs the mismatch, h0 + n - 1 - 4.
    # The to-end score wins exactly when it exceeds best - end_bonus, i.e. when
    # end_bonus > 4 (the reference's PEN_CLIP5 = 5 takes the end).
    rng = random.Random(n)
    s = rnd_seq(rng, n)
    q, t = s, s[:-1] + (b"A" if s[-1:] != b"A" else b"C")
    best, to_end = h0 + n - 1, h0 + n - 1 - 4
    assert _ksw(q, t, h0) == (to_end, n, n)
    assert _ksw(q, t, h0, end_bonus=5) == (to_end, n, n)
    for eb in (4, 3, 0):
        assert _ksw(q, t, h0, end_bonus=eb) == (best, n - 1, n - 1)
    assert _ksw(q, t, h0, end_bonus=100) == (to_end, n, n)
    # with a band, the end bonus widens bwa's clamp of w: w = 0 is never clamped
    assert _ksw(q, t, h0, w=0, end_bonus=eb)[0] == best


def test_ksw_extension_noop_settings_equal_reference():
    rng = random.Random(31)
    for _ in range(60):
        q = rnd_seq(rng, rng.randint(5, 120))
        t = mutate(rng, q, 0.15) if rng.random() < 0.7 else rnd_seq(rng, rng.randint(5, 120))
        h0 = rng.randint(10, 60)
        ref = _ksw(q, t, h0)
        assert _ksw(q, t, h0, w=10 ** 6) == ref
        assert _ksw(q, t, h0, zdrop=10 ** 6) == ref


def test_ksw_extension_band_and_zdrop_properties():
    rng = random.Random(32)
    for _ in range(80):
        q = rnd_seq(rng, rng.randint(20, 150))
        t = mutate(rng, q, 0.2)
        h0 = rng.randint(10, 60)
        full = _ksw(q, t, h0)
        for w in (0, 3, 12):
            s, qe, te = _ksw(q, t, h0, w=w)
            # a best-cell end lies in the band (a query-end result, qe == Lq,
            # takes the row's last computed H: when Lq is a multiple of 8 the
            # reference's loop nesting reads that as column Lq on every row)
            if qe != len(q):
                assert abs(qe - te) <= w, (w, qe, te)
        # z-drop only removes rows, so its best cell is at most the full run's;
        # the full run may report its query-end score instead, which lies
        # within 4 of its best (PEN_CLIP5)
        for zd in (1, 10, 40):
            assert _ksw(q, t, h0, zdrop=zd)[0] <= full[0] + 4


def test_ksw_extension_band_cuts_off_a_long_indel():
    rng = random.Random(33)
    a, b, z = rnd_seq(rng, 100), rnd_seq(rng, 100), rnd_seq(rng, 20)
    q, t = a + b, a + z + b
    full = _ksw(q, t, 20)
    assert full == (20 + 200 - 6 - 20, 200, 220)   # 200 matches, one 20-base deletion
    assert _ksw(q, t, 20, w=25) == full
    assert _ksw(q, t, 20, w=5)[0] < full[0]


def test_ksw_extension_zdrop_stops_after_a_divergent_stretch():
    # 30 matches, 5 mismatches, 100 matches.  z-drop 5 (below the gap-open
    # penalty 6) fires two rows into the mismatches: the best cell stays at
    # the end of the first block.  Without z-drop the extension runs through.
    rng = random.Random(34)
    a, b = rnd_seq(rng, 30), rnd_seq(rng, 100)
    q, t = a + b"AAAAA" + b, a + b"CCCCC" + b
    full = _ksw(q, t, 20)
    assert full[0] >= 20 + 130 - 20
    assert _ksw(q, t, 20, zdrop=5) == (50, 30, 30)
    assert _ksw(q, t, 20, zdrop=100) == full


# --------------------------------------------------------------- get_tb -----
def apply_cigar_score(q, t, qs, ts, cg, sc=(1, 4, 6, 1)):
    m, mm, go, ge = sc
    ops = []
    for b in reversed(cg):
        ops += [b & 3] * (b >> 2)
    i, j, s, prev = ts, qs, 0, None
    for op in ops:
        if op in (0, 1):
            s += m if (q[j] & 15) == (t[i] & 15) else -mm
            i += 1
            j += 1
        elif op == 2:
            s -= (go + ge) if prev != 2 else ge
            i += 1
        else:
            s -= (go + ge) if prev != 3 else ge
            j += 1
        prev = op
    return s, i - 1, j - 1


@pytest.mark.parametrize("seed", range(60))
def test_local_cigar_replays_to_score(seed):
    rng = random.Random(7000 + seed)
    core = rnd_seq(rng, rng.randint(10, 60))
    q = rnd_seq(rng, rng.randint(0, 10)) + core + rnd_seq(rng, rng.randint(0, 10))
    t = rnd_seq(rng, rng.randint(0, 10)) + mutate(rng, core, 0.08) + rnd_seq(rng, rng.randint(0, 10))
    r = oracle_one(q, t, g.LOCAL, start_pos=g.WITH_TB)
    cg = r["cigars"][0]
    assert len(cg) == r["n_ops"][0]
    assert all((b >> 2) <= 63 for b in cg)
    score = int(r["score"][0])
    qs, ts = int(r["qstart"][0]), int(r["tstart"][0])
    ops = []
    for b in reversed(cg):
        ops += [b & 3] * (b >> 2)
    # the walk replays from the reported start to the reported end, to the
    # score, on all 60 pairs (19 of them with gaps); the reference's
    # accounting can move the stop only where it labels a mismatch a match on
    # leaving a gap (src/kernels/get_tb.h:74-82)
    replay, ie, je = apply_cigar_score(q + b"N" * 8, t + b"N" * 8, qs, ts, cg)
    assert replay == score
    assert (ie, je) == (int(r["tend"][0]), int(r["qend"][0]))


def test_cigar_gap_kat():
    # t has a 3-base insertion relative to q: query ACGTACGTAC GGGTACCATG
    q = b"ACGTACGTACGGGTACCATGCA"
    t = q[:10] + b"TTT" + q[10:]
    r = oracle_one(q, t, g.LOCAL, start_pos=g.WITH_TB)
    assert r["score"][0] == len(q) - (6 + 3)
    assert g.cigar_to_string(r["cigars"][0]) == "10M3D12M"
    assert (r["qstart"][0], r["tstart"][0]) == (0, 0)


def test_global_cigar_covers_both_sequences():
    rng = random.Random(11)
    for _ in range(40):
        q = rnd_seq(rng, rng.randint(1, 40))
        t = mutate(rng, q, 0.2)
        r = oracle_one(q, t, g.GLOBAL, start_pos=g.WITH_TB)
        ops = []
        for b in reversed(r["cigars"][0]):
            ops += [b & 3] * (b >> 2)
        consumed_t = sum(1 for o in ops if o in (0, 1, 2))
        consumed_q = sum(1 for o in ops if o in (0, 1, 3))
        # the walk starts one past the last cell (Lt, Lq) and ends at (-1, -1)
        assert consumed_t == len(t) + 1 and consumed_q == len(q) + 1


# ---------------------------------------------------- planted mistakes -----
def test_pins_reject_a_planted_mistake():
    """Sanity of the pins themselves: a local DP that opens gaps from H instead
    of M (a plausible mistake) disagrees with the automaton on some input."""
    rng = random.Random(3)
    sc = (2, 20, 1, 1)   # mismatch > 2*(open+extend): an insertion next to a deletion beats a mismatch
    diffs = 0
    for _ in range(200):
        q = rnd_seq(rng, rng.randint(3, 20))
        t = mutate(rng, q, 0.3)
        qp = q + b"N" * ((-len(q)) % 8)
        tp = t + b"N" * ((-len(t)) % 8)
        if gotoh_local(qp, tp, sc) != automaton_local(qp, tp, sc)[0]:
            diffs += 1
    assert diffs > 0


@pytest.mark.parametrize("seed", range(40))
def test_local_gap_opening_from_m_under_cheap_gaps(seed):
    """With mismatch > 2*(open+extend) the gap-opening source is visible in the
    score: the oracle must follow the no-I/D-adjacency automaton."""
    sc = (2, 20, 1, 1)
    rng = random.Random(8000 + seed)
    q = rnd_seq(rng, rng.randint(3, 30))
    t = mutate(rng, q, 0.3)
    r = O.align(g.make_batch([q], [t]), g.LOCAL, scores=sc)
    qp = q + b"N" * ((-len(q)) % 8)
    tp = t + b"N" * ((-len(t)) % 8)
    assert r["score"][0] == automaton_local(qp, tp, sc)[0]
    rg = O.align(g.make_batch([q], [t]), g.GLOBAL, scores=sc)
    assert rg["score"][0] == automaton_global(q, t, sc)


# ------------------------------------------ full-matrix pins (every variant) -----
SCORE_SETS = [(1, 4, 6, 1), (2, 20, 1, 1), (3, 2, 5, 2), (1500, 4, 6, 1)]


def _semi_case(rng, kind):
    al = b"ACGTN" if kind == 5 else ACGT
    if kind == 0:
        return rnd_seq(rng, rng.randint(1, 30), al), rnd_seq(rng, rng.randint(1, 35))
    if kind == 1:   # query inside the target
        q = rnd_seq(rng, rng.randint(1, 30), al)
        return q, rnd_seq(rng, rng.randint(0, 12)) + mutate(rng, q, 0.1) + rnd_seq(rng, rng.randint(0, 12))
    if kind == 2:   # short query, long target (columns past the query's words)
        q = rnd_seq(rng, rng.randint(1, 9), al)
        return q, rnd_seq(rng, rng.randint(0, 20)) + mutate(rng, q, 0.1) + rnd_seq(rng, rng.randint(10, 25))
    if kind == 3:   # target ends with the query
        q = rnd_seq(rng, rng.randint(1, 9), al)
        return q, rnd_seq(rng, rng.randint(10, 25)) + q
    t = rnd_seq(rng, rng.randint(1, 9))   # long query, short target
    return rnd_seq(rng, rng.randint(0, 12), al) + mutate(rng, t, 0.1) + rnd_seq(rng, rng.randint(8, 25), al), t


@pytest.mark.parametrize("head", [NONE_, QUERY_, TARGET_, BOTH_])
@pytest.mark.parametrize("tail", [NONE_, QUERY_, TARGET_, BOTH_])
def test_semiglobal_every_head_tail_vs_full_matrix(head, tail):
    """Score, both ends, second best (score and ends) and WITH_START starts of
    every head/tail variant against the full-matrix formulation with the
    reference's boundary rows and end bookkeeping (refmath.semiglobal_reference),
    over four score sets (one large enough that the 16-bit row between strips
    wraps), N bases with and without N_PENALTY, and shapes that reach the
    columns past the query's padded words."""
    for seed in range(36):
        rng = random.Random(seed * 16 + head * 4 + tail)
        q, t = _semi_case(rng, seed % 6)
        sc = SCORE_SETS[seed % 4]
        npen = 1 if seed % 2 == 0 else None
        r = oracle_one(q, t, g.SEMI_GLOBAL, start_pos=g.WITH_START, second_best=True, head=head, tail=tail,
                       n_penalty=npen, scores=sc)
        ref = semiglobal_reference(q, t, head, tail, sc=sc, nval=(-1 if npen else 0), with_start=True)
        got = {k: int(r[k][0]) for k in ref}
        assert got == ref, (seed, q, t, sc)


@pytest.mark.parametrize("head", [NONE_, QUERY_, TARGET_, BOTH_])
def test_semiglobal_columns_past_the_query_words_kat(head):
    # A one-base query against a 13-base target: the second best of a QUERY
    # tail can come from a column the DP never computed (the row's initial
    # value), which only exists for targets longer than the query's words.
    q, t = b"A", b"C" * 12 + b"A"
    r = oracle_one(q, t, g.SEMI_GLOBAL, second_best=True, head=head, tail=QUERY_, scores=(1, 20, 1, 1))
    ref = semiglobal_reference(q, t, head, QUERY_, sc=(1, 20, 1, 1))
    assert {k: int(r[k][0]) for k in ("score", "qend", "tend", "s_score", "s_qend", "s_tend")} == \
        {k: ref[k] for k in ("score", "qend", "tend", "s_score", "s_qend", "s_tend")}


@pytest.mark.parametrize("seed", range(150))
def test_banded_narrow_bands_vs_masked_matrix(seed):
    """Tile-banded LOCAL at bands 0..64 against Smith-Waterman-Gotoh over the
    reference's per-strip word ranges (refmath.banded_reference)."""
    rng = random.Random(9000 + seed)
    q = rnd_seq(rng, rng.randint(1, 60), b"ACGTN" if seed % 5 == 0 else ACGT)
    t = (rnd_seq(rng, rng.randint(0, 20)) + mutate(rng, q, 0.15) + rnd_seq(rng, rng.randint(0, 20))
         if seed % 3 else rnd_seq(rng, rng.randint(1, 60)))
    if seed % 7 == 0:
        q, t = t, q
    sc = SCORE_SETS[seed % 3]
    npen = 1 if seed % 2 else None
    kb = rng.choice([0, 8, 16, 24, 32, 48, 64])
    r = O.align(g.make_batch([q], [t]), g.BANDED, k_band=kb, scores=sc, n_penalty=npen)
    got = (int(r["score"][0]), int(r["qend"][0]), int(r["tend"][0]))
    assert got == banded_reference(q, t, kb, sc=sc, nval=(-1 if npen else 0))


@pytest.mark.parametrize("seed", range(300))
def test_ksw_reference_form_vs_full_matrix(seed):
    """Reference-form KSW (no band, z-drop 0) against the full-matrix seed
    extension, on inputs where the seed score keeps every cell positive so
    that row pruning and the zero-row exit provably never act
    (refmath.ksw_unpruned_reference); includes divergent query ends, where
    the best cell (not the query end) is reported."""
    rng = random.Random(9500 + seed)
    q = rnd_seq(rng, rng.randint(1, 40), b"ACGTN" if seed % 5 == 0 else ACGT)
    t = mutate(rng, q, 0.2) if seed % 3 else rnd_seq(rng, rng.randint(1, 40))
    if seed % 4 == 3:
        core = rnd_seq(rng, rng.randint(3, 20))
        q, t = core + rnd_seq(rng, rng.randint(5, 20)), core + rnd_seq(rng, rng.randint(5, 20))
    t = t[:40]
    sc = SCORE_SETS[seed % 3]
    npen = 1 if seed % 2 else None
    h0 = 1000 + rng.randint(0, 50)
    r = O.align(g.make_batch([q], [t], seeds=[h0]), g.KSW, scores=sc, n_penalty=npen)
    got = (int(r["score"][0]), int(r["qend"][0]), int(r["tend"][0]))
    assert got == ksw_unpruned_reference(q, t, h0, sc=sc, nval=(-1 if npen else 0))


@pytest.mark.parametrize("seed", range(300))
def test_local_traceback_vs_automaton_walk(seed):
    """LOCAL WITH_TB: score, ends, starts and every CIGAR byte against a
    traceback over the automaton's M / D / I matrices
    (refmath.local_traceback_reference) -- no direction nibbles, tiles or
    strips -- over four score sets, with and without N bases and N_PENALTY."""
    rng = random.Random(7700 + seed)
    core = rnd_seq(rng, rng.randint(5, 50), b"ACGTN" if seed % 5 == 0 else ACGT)
    q = rnd_seq(rng, rng.randint(0, 10)) + core + rnd_seq(rng, rng.randint(0, 10))
    t = rnd_seq(rng, rng.randint(0, 10)) + mutate(rng, core, 0.12) + rnd_seq(rng, rng.randint(0, 10))
    sc = [(1, 4, 6, 1), (2, 20, 1, 1), (3, 2, 5, 2), (2, 3, 2, 1)][seed % 4]
    npen = 1 if seed % 2 else None
    r = O.align(g.make_batch([q], [t]), g.LOCAL, start_pos=g.WITH_TB, scores=sc, n_penalty=npen)
    ref = local_traceback_reference(q, t, sc=sc, nval=(-1 if npen else 0))
    got = dict(score=int(r["score"][0]), qend=int(r["qend"][0]), tend=int(r["tend"][0]),
               qstart=int(r["qstart"][0]), tstart=int(r["tstart"][0]), cigar=r["cigars"][0])
    assert got == ref
    assert int(r["n_ops"][0]) == len(ref["cigar"])
