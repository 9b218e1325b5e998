"""gasal2_aln_noalloc (the entry point that allocates nothing: its scratch
is one file-scope device array) against gasal2_run_resident, the entry point
whose device program it runs, through the C-ABI: two calls in a row, the
second with a smaller batch, every output equal; pairs that break the
declared maxima are counted, never read past their buffers; sizes beyond the
scratch are rejected."""
import ctypes as C

import numpy as np
import pytest

import gasal2_amd as g

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

def noalloc_fn():
    return g.lib().gasal2_aln_noalloc   # argtypes bound by gasal2_amd (include/gasal2_ext.h)


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a).astype(dtype)).cuda()


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def run_noalloc(b, p, packed=False, max_q=None, max_t=None):
    n = b.n
    q = b.packed_words(True) if packed else b.q
    t = b.packed_words(False) if packed else b.t
    dq, dt = dev(q, np.uint8), dev(t, np.uint8)
    dqo, dto, dql, dtl = (dev(x, np.int32) for x in (b.qoff, b.toff, b.qlen, b.tlen))
    dqop, dtop = dev(b.qop, np.uint8), dev(b.top, np.uint8)
    outs = {k: torch.full((n,), -7, dtype=torch.int32, device="cuda")
            for k in ("score", "qend", "tend", "qstart", "tstart", "s_score", "s_qend", "s_tend")}
    bad = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = noalloc_fn()(ptr(dq), ptr(dt), len(b.q), len(b.t), ptr(dqo), ptr(dto), ptr(dql), ptr(dtl), ptr(dqop),
                      ptr(dtop), n, int(max_q or b.qlen.max()), int(max_t or b.tlen.max()),
                      *[ptr(outs[k]) for k in ("score", "qend", "tend", "qstart", "tstart", "s_score", "s_qend",
                                                "s_tend")],
                      ptr(bad), p.native(), None)
    torch.cuda.synchronize()
    return rc, {k: v.cpu().numpy() for k, v in outs.items()}, int(bad.item())


def run_resident(b, p):
    a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, n_penalty=1)
    try:
        a.load(b)
        a.run_resident()
        return a.download()
    finally:
        a.close()


def keys_of(p):
    k = ["score"] if p.algo == g.GLOBAL else ["score", "qend", "tend"]
    if (p.algo == g.LOCAL and p.start_pos == g.WITH_START) or (p.algo == g.SEMI_GLOBAL and p.start_pos == g.WITH_START):
        k += ["qstart", "tstart"]
    if p.secondBest and p.algo in (g.LOCAL, g.SEMI_GLOBAL):
        k += ["s_score", "s_qend", "s_tend"]
    return k


MODES = [(g.LOCAL, g.WITH_START, True), (g.LOCAL, g.WITHOUT_START, False), (g.GLOBAL, g.WITH_START, False),
         (g.SEMI_GLOBAL, g.WITH_START, False), (g.SEMI_GLOBAL, g.WITHOUT_START, True)]


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("algo,start,second", MODES)
def test_noalloc_matches_run_resident_twice(algo, start, second, packed):
    g.set_n_penalty(1)
    big = g.synth_batch(2, 1600 + algo, 3000)
    small = g.synth_batch(2, 1700 + algo, 700)
    for b in (big, small):   # two calls in a row, the second smaller
        p = g.Parameters(algo=algo, start_pos=start, secondBest=g.TRUE if second else g.FALSE,
                         isReverseComplement=True, isPacked=packed)
        want = run_resident(b, p)
        rc, got, bad = run_noalloc(b, p, packed)
        assert rc == 0 and bad == 0
        for k in keys_of(p):
            assert np.array_equal(got[k], want[k]), k


def test_noalloc_rejects_oversized_batches():
    b = g.synth_batch(2, 1800, 200)
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITHOUT_START)
    # declared maxima so large that the boundary rows exceed the scratch
    rc, _, _ = run_noalloc(b, p, max_q=4_000_000)
    assert rc == -1
    p2 = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB)
    assert run_noalloc(b, p2)[0] == -1


def test_noalloc_counts_pairs_beyond_declared_maxima():
    """Pairs longer than the declared maximum length are counted and replaced;
    every other pair's outputs are unaffected."""
    g.set_n_penalty(1)
    b = g.synth_batch(2, 1810, 500)
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITHOUT_START, isReverseComplement=True)
    want = run_resident(b, p)
    cap = int(np.sort(b.qlen)[-20])   # about 20 pairs longer than this
    rc, got, bad = run_noalloc(b, p, max_q=cap)
    assert rc == 0
    long_ = b.qlen > cap
    assert bad == int(long_.sum()) > 0
    for k in ("score", "qend", "tend"):
        assert np.array_equal(got[k][~long_], want[k][~long_]), k
