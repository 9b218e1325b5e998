"""Packed 16-bit LOCAL and GLOBAL (two strips per lane in register halves): bit-exact
against the oracle and against the int32 kernels, including odd strip
counts, ties, other scorings and the fall back to int32 when a score could
exceed int16."""
import numpy as np
import pytest

import gasal2_amd as g
from parity import assert_parity, fields, run_both

pytestmark = pytest.mark.gpu


def both_ways(b, **kw):
    L = g.lib()
    try:
        L.gasal2_set_pk16(0)
        g32, _ = run_both(b, g.LOCAL, g.WITHOUT_START, False, **kw)
        L.gasal2_set_pk16(1)
        g16, ref = run_both(b, g.LOCAL, g.WITHOUT_START, False, **kw)
    finally:
        L.gasal2_set_pk16(1)
    assert_parity(g16, ref, g.LOCAL, g.WITHOUT_START, False, b)
    for k in ("score", "qend", "tend"):
        assert np.array_equal(g16[k], g32[k]), k


@pytest.mark.parametrize("cfg,n_penalty,is_rc", [(1, None, False), (2, 1, True), (3, None, False)])
def test_pk16_synthetic(cfg, n_penalty, is_rc):
    b = g.synth_batch(cfg, 2500 + cfg, 700)
    both_ways(b, n_penalty=n_penalty, is_rc=is_rc)


@pytest.mark.parametrize("scores", [(2, 3, 5, 2), (7, 1, 1, 1), (1, 8, 7, 1)])
def test_pk16_other_scores(scores):
    b = g.synth_batch(2, 2510, 500)
    both_ways(b, scores=scores, n_penalty=1)


def test_pk16_odd_strips_ties_and_tiny():
    rng = np.random.default_rng(2520)
    unit = b"ACGTTGCA" * 5
    qs = [b"A", b"ACGT", b"ACGTACGTA", unit * 3, unit, b"AC" * 70, b"N" * 9]
    ts = [b"A", b"TACGT", b"ACGTACGT" * 3, unit * 2 + b"G" + unit, unit * 5, b"AC" * 90 + b"T", b"NACGTNACGTN"]
    for _ in range(40):   # target word counts both odd and even
        lq, lt = int(rng.integers(1, 200)), int(rng.integers(1, 200))
        qs.append(bytes(rng.choice(list(b"ACGT"), lq).tolist()))
        ts.append(bytes(rng.choice(list(b"ACGT"), lt).tolist()))
    b = g.make_batch(qs, ts)
    both_ways(b, n_penalty=1)


def test_pk16_falls_back_when_scores_could_exceed_int16():
    """Match 7 over 5000 identical bases: 35000 > int16, the int32 kernels run."""
    rng = np.random.default_rng(2530)
    q = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 5000)].tobytes()
    b = g.make_batch([q, q[:300]], [q, q[:300]])
    got, ref = run_both(b, g.LOCAL, g.WITHOUT_START, False, scores=(7, 2, 3, 1))
    assert_parity(got, ref, g.LOCAL, g.WITHOUT_START, False, b)


# ---------------------------------------------------------------- GLOBAL ----
def both_ways_global(b, start=g.WITH_START, **kw):
    L = g.lib()
    try:
        L.gasal2_set_pk16(0)
        g32, _ = run_both(b, g.GLOBAL, start, False, **kw)
        L.gasal2_set_pk16(1)
        g16, ref = run_both(b, g.GLOBAL, start, False, **kw)
    finally:
        L.gasal2_set_pk16(1)
    assert_parity(g16, ref, g.GLOBAL, start, False, b)
    assert np.array_equal(g16["score"], g32["score"])


@pytest.mark.parametrize("start", [g.WITHOUT_START, g.WITH_START])
@pytest.mark.parametrize("cfg,n_penalty,is_rc", [(1, None, False), (2, 1, True), (3, None, False)])
def test_pk16_global_synthetic(cfg, n_penalty, is_rc, start):
    b = g.synth_batch(cfg, 2600 + cfg, 700)
    both_ways_global(b, start, n_penalty=n_penalty, is_rc=is_rc)


@pytest.mark.parametrize("scores", [(2, 3, 5, 2), (7, 1, 1, 1), (1, 8, 7, 1)])
def test_pk16_global_other_scores(scores):
    b = g.synth_batch(2, 2610, 500)
    both_ways_global(b, scores=scores, n_penalty=1)


def test_pk16_global_odd_strips_and_tiny():
    """Odd and even strip counts, one-base pairs, lengths off multiples of 8
    (the score row of the last strip in either register half), unrelated
    pairs (deeply negative scores), N bases under both N rules."""
    rng = np.random.default_rng(2620)
    qs = [b"A", b"ACGT", b"ACGTACGTA", b"N" * 9, b"ACGTN" * 30]
    ts = [b"A", b"TACGT", b"ACGTACGT" * 3, b"NACGTNACGTN", b"ACGT" * 2]
    for _ in range(60):
        lq, lt = int(rng.integers(1, 220)), int(rng.integers(1, 220))
        qs.append(bytes(rng.choice(list(b"ACGT"), lq).tolist()))
        ts.append(bytes(rng.choice(list(b"ACGT"), lt).tolist()))
    b = g.make_batch(qs, ts)
    both_ways_global(b)
    both_ways_global(b, n_penalty=1)


def test_pk16_global_falls_back_when_values_could_exceed_int16():
    """Unrelated pairs of 5000 bases: scores far below -16000 are possible, the
    int32 kernel runs."""
    rng = np.random.default_rng(2630)
    q = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 5000)].tobytes()
    t = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 4800)].tobytes()
    b = g.make_batch([q, q[:300]], [t, t[:280]])
    got, ref = run_both(b, g.GLOBAL, g.WITH_START, False, scores=(1, 4, 6, 1))
    assert_parity(got, ref, g.GLOBAL, g.WITH_START, False, b)


# ----------------------------------------------------------- SEMI_GLOBAL ----
SEMI_PK = [(h, t) for h in (g.NONE, g.QUERY, g.TARGET, g.BOTH) for t in (g.NONE, g.TARGET)]


def both_ways_semi(b, head, tail, start, second, **kw):
    L = g.lib()
    try:
        L.gasal2_set_pk16(0)
        g32, _ = run_both(b, g.SEMI_GLOBAL, start, second, head=head, tail=tail, **kw)
        L.gasal2_set_pk16(1)
        g16, ref = run_both(b, g.SEMI_GLOBAL, start, second, head=head, tail=tail, **kw)
    finally:
        L.gasal2_set_pk16(1)
    assert_parity(g16, ref, g.SEMI_GLOBAL, start, second, b)
    for k in fields(g.SEMI_GLOBAL, start, second):
        assert np.array_equal(g16[k], g32[k]), k


@pytest.mark.parametrize("head,tail", SEMI_PK)
@pytest.mark.parametrize("start,second", [(g.WITHOUT_START, True), (g.WITH_START, False)])
def test_pk16_semi_global(head, tail, start, second):
    b = g.synth_batch(2, 2700 + 4 * head + tail, 400)
    both_ways_semi(b, head, tail, start, second, n_penalty=1, is_rc=True)


def test_pk16_semi_global_odd_strips_tiny_and_scores():
    rng = np.random.default_rng(2720)
    qs = [b"A", b"ACGT", b"ACGTACGTA", b"N" * 9, b"ACGTN" * 30]
    ts = [b"A", b"TACGT", b"ACGTACGT" * 3, b"NACGTNACGTN", b"ACGT" * 2]
    for _ in range(60):
        lq, lt = int(rng.integers(1, 220)), int(rng.integers(1, 220))
        qs.append(bytes(rng.choice(list(b"ACGT"), lq).tolist()))
        ts.append(bytes(rng.choice(list(b"ACGT"), lt).tolist()))
    b = g.make_batch(qs, ts)
    for head, tail in ((g.TARGET, g.TARGET), (g.NONE, g.TARGET), (g.QUERY, g.NONE)):
        both_ways_semi(b, head, tail, g.WITH_START, False, n_penalty=1)
        both_ways_semi(b, head, tail, g.WITHOUT_START, True, scores=(2, 3, 5, 2))
