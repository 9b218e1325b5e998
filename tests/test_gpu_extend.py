"""Two-sided seed extension (gasal2_extend_seeds) and KSW batches through the
stream driver.  Each side is checked against the oracle's KSW run on the
one-sided problem built here from the pair and its seed (reversed prefixes
from the seed's score, then suffixes from the left side's score), in the
reference form and with band, z-drop and end bonuses; an exact match
extends over both whole sequences."""
import random

import numpy as np
import pytest

import gasal2_amd as g
import pyoracle as O

pytestmark = pytest.mark.gpu

ACGT = b"ACGT"


def rnd(rng, n):
    return bytes(rng.choice(ACGT) for _ in range(n))


def mutate(rng, s, p):
    out = bytearray()
    for c in s:
        u = rng.random()
        if u < p / 3:
            continue
        out.append(rng.choice(ACGT) if u < 2 * p / 3 else c)
        if u > 1 - p / 6:
            out += rnd(rng, rng.randint(1, 3))
    return bytes(out)


def seeded_pairs(seed, n, p=0.06):
    rng = random.Random(seed)
    qs, ts, seeds = [], [], []
    for k in range(n):
        lq, rq = rnd(rng, rng.randint(0, 120)), rnd(rng, rng.randint(0, 120))
        core = rnd(rng, rng.randint(12, 40))
        lt = rnd(rng, rng.randint(0, 20)) + mutate(rng, lq, p)
        rt = mutate(rng, rq, p) + rnd(rng, rng.randint(0, 20))
        if k % 7 == 0:
            lq, lt = b"", lt        # seed at the query's start: no left extension
        if k % 11 == 0:
            rq = b""                # seed at the query's end: no right extension
        qs.append(lq + core + rq)
        ts.append(lt + core + rt)
        seeds.append((len(lq), len(lt), len(core), len(core)))
    return qs, ts, seeds


def oracle_side(qs, ts, h0s, w, zdrop, eb):
    if not qs:
        return []
    r = O.align(g.make_batch(qs, ts, seeds=h0s), g.KSW, ksw_w=w, ksw_zdrop=zdrop, ksw_end_bonus=eb)
    return list(zip(r["score"].tolist(), r["qend"].tolist(), r["tend"].tolist()))


def reference_extend(qs, ts, seeds, w, zdrop, eb):
    n = len(qs)
    left = [(sd[3], 0, 0) for sd in seeds]
    idx = [k for k in range(n) if seeds[k][0] and seeds[k][1]]
    got = oracle_side([qs[k][:seeds[k][0]][::-1] for k in idx], [ts[k][:seeds[k][1]][::-1] for k in idx],
                      [seeds[k][3] for k in idx], w, zdrop, eb[0])
    for k, v in zip(idx, got):
        left[k] = v
    right = [(left[k][0], 0, 0) for k in range(n)]
    idx = [k for k in range(n) if seeds[k][0] + seeds[k][2] < len(qs[k]) and seeds[k][1] + seeds[k][2] < len(ts[k])]
    got = oracle_side([qs[k][seeds[k][0] + seeds[k][2]:] for k in idx], [ts[k][seeds[k][1] + seeds[k][2]:] for k in idx],
                      [max(left[k][0], 0) for k in idx], w, zdrop, eb[1])
    for k, v in zip(idx, got):
        right[k] = v
    out = {"score": [], "query_beg": [], "query_end": [], "target_beg": [], "target_end": [], "left_score": []}
    for k, (qb, tb, ln, _) in enumerate(seeds):
        out["score"].append(right[k][0])
        out["left_score"].append(left[k][0])
        out["query_beg"].append(qb - left[k][1])
        out["target_beg"].append(tb - left[k][2])
        out["query_end"].append(qb + ln + right[k][1])
        out["target_end"].append(tb + ln + right[k][2])
    return {k: np.array(v, np.int32) for k, v in out.items()}


def run_extend(qs, ts, seeds, w, zdrop, eb, batch_pairs=700, storages=2, packed=False):
    b = g.make_batch(qs, ts)
    ps = g.PairSet.from_batch(b, pack=packed)
    p = g.Parameters(algo=g.KSW, isPacked=packed)
    st = g.Stream(p, batch_pairs, storages, int(b.qlen.max()), int(b.tlen.max()))
    try:
        a = np.array(seeds, np.int64)
        return ps.extend(st, a[:, 0], a[:, 1], a[:, 2], a[:, 3], w=w, zdrop=zdrop, end_bonus=eb)
    finally:
        st.close()


@pytest.mark.parametrize("w,zdrop,eb", [(-1, 0, (5, 5)), (50, 100, (5, 5)), (20, 30, (0, 12)), (-1, 0, (3, 100))])
@pytest.mark.parametrize("packed", [False, True])
def test_extend_matches_one_sided_oracle(w, zdrop, eb, packed):
    qs, ts, seeds = seeded_pairs(400 + w + zdrop, 2000)
    got, stats = run_extend(qs, ts, seeds, w, zdrop, eb, packed=packed)
    ref = reference_extend(qs, ts, seeds, w, zdrop, eb)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(got["right_score"], got["score"])
    assert stats["sub_batches"] >= 2


def test_exact_match_extends_over_both_sequences():
    rng = random.Random(3)
    qs, ts, seeds = [], [], []
    for _ in range(300):
        s = rnd(rng, rng.randint(30, 200))
        a = rng.randint(0, len(s) - 10)
        ln = rng.randint(5, len(s) - a)
        qs.append(s)
        ts.append(s)
        seeds.append((a, a, ln, ln))
    got, _ = run_extend(qs, ts, seeds, -1, 0, (5, 5), batch_pairs=128, storages=3)
    L = np.array([len(s) for s in qs], np.int32)
    assert np.array_equal(got["score"], L)
    assert np.array_equal(got["query_beg"], np.zeros_like(L)) and np.array_equal(got["target_beg"], np.zeros_like(L))
    assert np.array_equal(got["query_end"], L) and np.array_equal(got["target_end"], L)


def test_ksw_through_the_stream_matches_oracle():
    b = g.synth_batch(4, 4100, 3000)
    p = g.Parameters(algo=g.KSW)
    ps = g.PairSet.from_batch(b)
    got, st = ps.align(p, batch_pairs=1000, storages=3)
    ref = O.align(b, g.KSW)
    for k in ("score", "qend", "tend"):
        assert np.array_equal(got[k], ref[k]), k
    assert st["sub_batches"] == 3
