"""GPU parity of the one-wave-per-pair LOCAL path (score and ends) and of the
per-pair switch between it and the one-lane-per-pair kernels, against the CPU
oracle, bit-exact.  Shapes are chosen to cross the path's internal boundaries:
64-strip passes (512 target bases) and 1024-column blocks of the query."""
import numpy as np
import pytest

import gasal2_amd as g
from parity import assert_parity, run_both

pytestmark = pytest.mark.gpu



class policy:
    """Dispatch policy / switch thresholds for one block, restored after."""

    def __init__(self, dispatch=0, wf=None):
        self.dispatch, self.wf = dispatch, wf

    def __enter__(self):
        g.lib().gasal2_set_dispatch(self.dispatch)
        if self.wf:
            g.lib().gasal2_set_wavefront(*self.wf)

    def __exit__(self, *exc):
        g.lib().gasal2_set_dispatch(0)
        g.lib().gasal2_set_wavefront(0, 0, 0)   # the per-mode defaults


def related_pairs(shapes, seed, sub=0.03, indel=0.01, alphabet=b"ACGT"):
    """Query of length lq; target = a mutated copy of it cut or extended to
    length lt with random flanks, so the best local path runs across many
    strips and column blocks."""
    rng = np.random.default_rng(seed)
    al = np.frombuffer(alphabet, np.uint8)
    qs, ts = [], []
    for lq, lt in shapes:
        q = al[rng.integers(0, len(al), lq)]
        t = []
        for b in q:
            r = rng.random()
            if r < sub:
                t.append(al[rng.integers(0, len(al))])
            elif r < sub + indel / 2:
                continue
            elif r < sub + indel:
                t.extend([b, al[rng.integers(0, len(al))]])
            else:
                t.append(b)
        t = np.array(t, np.uint8)
        if len(t) >= lt:
            o = int(rng.integers(0, len(t) - lt + 1))
            t = t[o:o + lt]
        else:
            extra = lt - len(t)
            left = int(rng.integers(0, extra + 1))
            t = np.concatenate([al[rng.integers(0, len(al), left)], t, al[rng.integers(0, len(al), extra - left)]])
        qs.append(q.tobytes())
        ts.append(t.tobytes())
    return qs, ts


def check(b, **kw):
    got, ref = run_both(b, g.LOCAL, g.WITHOUT_START, False, **kw)
    assert_parity(got, ref, g.LOCAL, g.WITHOUT_START, False, b)


@pytest.mark.parametrize("cfg,n_penalty,is_rc", [(1, None, False), (2, 1, True)])
def test_every_pair_one_wave(cfg, n_penalty, is_rc):
    """Short pairs (one pass, idle lanes below the target's last strip)."""
    b = g.synth_batch(cfg, 1100 + cfg, 600)
    with policy(dispatch=3):
        check(b, is_rc=is_rc, n_penalty=n_penalty)


SHAPES = [(1017, 1500), (1024, 1024), (1025, 1031), (1031, 504), (2048, 511), (2049, 512),
          (1500, 513), (3000, 1025), (700, 4100), (4100, 700), (1, 3000), (3000, 1), (9, 2600), (2600, 9)]


@pytest.mark.parametrize("scores", [(1, 4, 6, 1), (2, 3, 5, 2), (3, 1, 2, 1)])
def test_pass_and_block_boundaries(scores):
    qs, ts = related_pairs(SHAPES, 1200 + scores[0])
    b = g.make_batch(qs, ts)
    with policy(dispatch=3):
        check(b, scores=scores, n_penalty=1)


def test_uneven_batch_mixed_switch():
    """Log-normal lengths: long pairs one wave each, the rest one lane each, in
    one launch; outputs land at the right pairs."""
    b = g.synth_batch(3, 1300, 400)
    with policy(dispatch=0, wf=(16, 64 * 16 * 16, 64)):
        check(b)


def test_uneven_batch_default_thresholds():
    """Default policy: a batch of a few pairs per CU runs every pair one wave
    (run_both also runs the lane kernels alone and compares)."""
    b = g.synth_batch(3, 1301, 300)
    check(b, n_penalty=1)


def test_int16_boundary_truncation_long_identical():
    """Match 7 over 5000 identical bases exceeds int16 in the boundary rows;
    the position encoding of the lane kernel does not fit either, so short
    pairs go to the reference-order kernel while long ones run one wave each."""
    rng = np.random.default_rng(1400)
    qs, ts = [], []
    for lq in (5000, 4700, 300, 200, 150):
        q = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, lq)].tobytes()
        qs.append(q)
        ts.append(q)
    b = g.make_batch(qs, ts)
    with policy(dispatch=0, wf=(16, 64 * 8 * 16, 2048)):
        check(b, scores=(7, 2, 3, 1))


def test_tied_maxima_across_strips_and_blocks():
    """Repeats give equal maxima at many cells; the first in (strip, column,
    row) order must be reported."""
    unit = b"ACGTTGCA" * 40
    qs = [unit * 4, unit + b"GG" + unit, (b"AC" * 600)]
    ts = [unit * 3 + b"T" * 100 + unit, b"CCCC" + unit * 5, (b"AC" * 700)]
    b = g.make_batch(qs, ts)
    with policy(dispatch=3):
        check(b)


def test_census_collision_every_pair_one_wave():
    """Nibble classes collide ('I' vs 'A'): every pair runs one wave per pair
    with the substitution rule per cell instead of score tables."""
    qs, ts = related_pairs([(1500, 1600), (900, 800), (2100, 2000)], 1500)
    qs[1] = qs[1][:50] + b"IIA" + qs[1][53:]
    b = g.make_batch(qs, ts)
    with policy(dispatch=3):
        check(b)


def test_autotune_choices_identical_results():
    """Every switch setting the tuner tries gives the same outputs (the tuner
    raises otherwise); the tuned table is then used by the library."""
    from gasal2_amd import autotune
    b = g.synth_batch(3, 1600, 1500)
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITHOUT_START)
    cands = [(16, 64 * 16 * 16, 64), (48, 1 << 20, 2048), (1, 0, 256), (1 << 20, 1 << 30, 1)]
    r = autotune.tune(b, p, candidates=cands, repeats=1)
    assert len(r["timings"]) == len(cands)
    autotune.clear()
    g.lib().gasal2_tune_set(g.LOCAL, g.WITHOUT_START, 0, 0, 1 << 31, *r["best"])
    try:
        check(b)
    finally:
        autotune.clear()


def check_tb(b, ckpt=2, **kw):
    """ckpt 2: checkpoints and tile recompute on both paths; 1: lane path only
    (default); 0: direction words."""
    g.lib().gasal2_set_tb_checkpoint(ckpt)
    try:
        got, ref = run_both(b, g.LOCAL, g.WITH_TB, False, **kw)
    finally:
        g.lib().gasal2_set_tb_checkpoint(1)   # the default
    assert_parity(got, ref, g.LOCAL, g.WITH_TB, False, b)


@pytest.mark.parametrize("ckpt", [2, 1, 0])
@pytest.mark.parametrize("cfg,n_penalty,is_rc", [(1, None, False), (2, 1, True)])
def test_tb_every_pair_one_wave(cfg, n_penalty, is_rc, ckpt):
    """WITH_TB: per-pair tiles (checkpoints or direction words), traceback through them."""
    b = g.synth_batch(cfg, 1700 + cfg, 500)
    with policy(dispatch=3):
        check_tb(b, ckpt=ckpt, is_rc=is_rc, n_penalty=n_penalty)


@pytest.mark.parametrize("ckpt", [2, 1, 0])
@pytest.mark.parametrize("scores", [(1, 4, 6, 1), (2, 3, 5, 2)])
def test_tb_pass_and_block_boundaries(scores, ckpt):
    qs, ts = related_pairs(SHAPES, 1800 + scores[0])
    b = g.make_batch(qs, ts)
    with policy(dispatch=3):
        check_tb(b, ckpt=ckpt, scores=scores, n_penalty=1)


def test_tb_mixed_switch_and_lane_groups():
    """Long pairs one wave each, the rest one lane each: the traceback finds
    both tile layouts (lane groups count after the long pairs)."""
    b = g.synth_batch(3, 1900, 700)
    with policy(dispatch=0, wf=(16, 64 * 16 * 16, 64)):
        check_tb(b, n_penalty=1)


def test_census_collision_runs_one_wave_without_tables():
    """Nibble classes collide: the one-wave path scores per cell instead of
    from tables, in both modes."""
    qs, ts = related_pairs([(1500, 1600), (900, 800), (2100, 2000), (300, 280)], 1950)
    qs[1] = qs[1][:50] + b"IIA" + qs[1][53:]
    b = g.make_batch(qs, ts)
    with policy(dispatch=0, wf=(16, 64 * 16 * 16, 64)):
        check(b)
        check_tb(b)
        check_tb(b, ckpt=1)
        check_tb(b, ckpt=0)


def test_small_batch_runs_one_wave_per_pair():
    """Auto policy: a batch of at most 128 pairs per CU (64 for WITH_TB) runs
    every pair one wave (bit 1 of gasal2_last_paths); the lane-only policy and
    thresholds set by the caller turn the rule off."""
    b = g.synth_batch(1, 1350, 2000)
    for start in (g.WITHOUT_START, g.WITH_TB):
        p = g.Parameters(algo=g.LOCAL, start_pos=start)
        a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n)
        try:
            a.run(b)
            assert a.last_paths() & 2, start
            with policy(dispatch=2):
                a.run(b)
                assert not a.last_paths() & 2, start
            with policy(dispatch=0, wf=(48, 1 << 20, 2048)):
                a.run(b)
                assert not a.last_paths() & 2, start
        finally:
            a.close()
