"""Packed 16-bit LOCAL (two strips per lane in register halves): bit-exact
against the oracle and against the int32 kernels, including odd strip
counts, ties, other scorings and the fall back to int32 when a score could
exceed int16."""
import numpy as np
import pytest

import gasal2_amd as g
from parity import assert_parity, run_both

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
