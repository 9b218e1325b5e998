"""LOCAL with two pairs per lane in packed 16-bit halves (k_local_pk2): every
mode bit-exact against the oracle and against the int32 kernels -- second
best, WITH_START, the checkpointed traceback (CIGARs included) -- with lane
pairs of different shapes, lanes whose second pair is missing, per-group tile
regions, other scores and the fall back to int32 when a rank could exceed 16
bits."""
import numpy as np
import pytest

import gasal2_amd as g
from parity import assert_parity, fields, run_both

pytestmark = pytest.mark.gpu

MODES = [(s, b) for s in (g.WITHOUT_START, g.WITH_START, g.WITH_TB) for b in (False, True)]


def three_ways(b, start, second, packed_expected=True, **kw):
    """pk16 policy 0 (int32), 1 (default) and 2 (two pairs per lane everywhere):
    every output identical, and equal to the oracle."""
    L = g.lib()
    out = {}
    try:
        for mode in (0, 1, 2):
            L.gasal2_set_pk16(mode)
            out[mode], ref = run_both(b, g.LOCAL, start, second, **kw)
    finally:
        L.gasal2_set_pk16(1)
    assert_parity(out[2], ref, g.LOCAL, start, second, b)
    for mode in (0, 1):
        for f in fields(g.LOCAL, start, second) + (["n_ops"] if start == g.WITH_TB else []):
            assert np.array_equal(out[mode][f], out[2][f]), (mode, f)
        if start == g.WITH_TB:
            assert out[mode]["cigars"] == out[2]["cigars"], mode


def paths_of(b, start, second, mode=2, scores=(1, 4, 6, 1)):
    L = g.lib()
    p = g.Parameters(algo=g.LOCAL, start_pos=start, secondBest=g.TRUE if second else g.FALSE)
    a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, scores=scores)
    try:
        L.gasal2_set_pk16(mode)
        L.gasal2_set_dispatch(2)   # lane kernels only: small batches would otherwise run one wave per pair
        a.run(b)
        return a.last_paths()
    finally:
        L.gasal2_set_dispatch(0)
        L.gasal2_set_pk16(1)
        a.close()


@pytest.mark.parametrize("start,second", MODES)
@pytest.mark.parametrize("cfg,n_penalty,is_rc", [(1, None, False), (2, 1, True)])
def test_pk2_synthetic(cfg, n_penalty, is_rc, start, second):
    b = g.synth_batch(cfg, 2700 + cfg, 900)
    three_ways(b, start, second, n_penalty=n_penalty, is_rc=is_rc)


@pytest.mark.parametrize("start,second", [(g.WITH_TB, True), (g.WITHOUT_START, True), (g.WITH_TB, False)])
def test_pk2_kernel_ran(start, second):
    b = g.synth_batch(2, 2710, 300)
    assert paths_of(b, start, second) & 1, "the packed two-pairs-per-lane kernel did not run"


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 128, 129, 64 * 5 + 17])
def test_pk2_lane_pairs_missing(n):
    """Lane groups beyond the batch: the high half of a lane runs with no pair."""
    b = g.synth_batch(2, 2720 + n, n)
    three_ways(b, g.WITH_TB, True, n_penalty=1, is_rc=True)


def test_pk2_mixed_shapes_in_a_lane():
    """Lengths 1..400 uniformly, so the two pairs of a lane often differ in
    query and target words (the smaller one's cells past its matrix must not
    reach its results or checkpoints); ties, N runs and tiny pairs included."""
    rng = np.random.default_rng(2730)
    unit = b"ACGTTGCA" * 5
    qs = [b"A", b"ACGT", b"ACGTACGTA", unit * 3, unit, b"AC" * 70, b"N" * 9, b"ACGTN" * 20]
    ts = [b"A", b"TACGT", b"ACGTACGT" * 3, unit * 2 + b"G" + unit, unit * 5, b"AC" * 90 + b"T", b"NACGTNACGTN",
          b"ACGTN" * 33]
    for _ in range(600):
        lq, lt = int(rng.integers(1, 400)), int(rng.integers(1, 400))
        qs.append(bytes(rng.choice(list(b"ACGT"), lq).tolist()))
        t = bytes(rng.choice(list(b"ACGT"), lt).tolist())
        if rng.random() < 0.5 and lt > 20:   # embed part of the query so scores are real
            k = min(lq, lt - 10)
            o = int(rng.integers(0, lt - k + 1))
            t = t[:o] + qs[-1][:k] + t[o + k:]
        ts.append(t)
    b = g.make_batch(qs, ts)
    for start, second in MODES:
        three_ways(b, start, second, n_penalty=1)


def test_pk2_per_group_tile_regions():
    """Tile budget 1 byte: every lane group gets its own region sized for its
    own shape; pair B's checkpoints must stay inside group 2w+1's region."""
    L = g.lib()
    b = g.synth_batch(3, 2740, 700)
    try:
        L.gasal2_set_tile_budget(1)
        three_ways(b, g.WITH_TB, True)
        three_ways(b, g.WITH_TB, False)
    finally:
        L.gasal2_set_tile_budget(64 << 30)


@pytest.mark.parametrize("scores", [(2, 3, 5, 2), (7, 1, 1, 1), (1, 8, 7, 1)])
def test_pk2_other_scores(scores):
    b = g.synth_batch(2, 2750, 500)
    three_ways(b, g.WITH_TB, True, scores=scores, n_penalty=1, is_rc=True)


def test_pk2_falls_back_when_ranks_could_exceed_16_bits():
    """Match 7 over 700 identical bases: h up to 4900 > 4095, so the ranked
    cells would not fit 16 bits; the int32 kernels run, results exact."""
    rng = np.random.default_rng(2760)
    q = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 700)].tobytes()
    b = g.make_batch([q, q[:300], q[100:400]], [q, q[:300], q])
    got, ref = run_both(b, g.LOCAL, g.WITH_TB, True, scores=(7, 2, 3, 1))
    assert_parity(got, ref, g.LOCAL, g.WITH_TB, True, b)
    assert not (paths_of(b, g.WITH_TB, True, mode=2, scores=(7, 2, 3, 1)) & 1)
