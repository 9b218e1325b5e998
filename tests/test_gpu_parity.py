"""GPU parity: the HIP library (through its C-ABI) against the CPU oracle,
bit-exact on every output each mode defines, on seeded synthetic batches."""
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


@pytest.mark.parametrize("second", [False, True])
def test_ksw(second):
    b = g.synth_batch(4, 500, 600)
    got, ref = run_both(b, g.KSW, second=second)
    assert_parity(got, ref, g.KSW, g.WITHOUT_START, second, b)


def test_local_tiny_edge_cases():
    qs = [b"A", b"ACGT", b"ACGTACGT", b"ACGTACGTA", b"N", b"NNNNNNNNN", b"TTTT", b"ACGTNACGT" * 5]
    ts = [b"A", b"ACGTTT", b"ACGTACGT", b"TACGTACGT", b"A", b"ACGTACGTAC", b"GGGG", b"ACGTACGT" * 6]
    b = g.make_batch(qs, ts)
    for start in (g.WITHOUT_START, g.WITH_START, g.WITH_TB):
        got, ref = run_both(b, g.LOCAL, start, True, n_penalty=1)
        assert_parity(got, ref, g.LOCAL, start, True, b)
