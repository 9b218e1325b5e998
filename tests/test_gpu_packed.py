"""isPacked=true: the caller hands over 4-bit packed words (the reference
skips its pack kernel, src/gasal_align.cu:190, and aliases the packed buffers
with the unpacked ones, src/ctors.cpp:64-67).  Byte counts, offsets and the
CIGAR layout stay in bases.  Every output, CIGAR bytes in the reference layout
included, must equal the oracle's on the bases and the isPacked=false run."""
import numpy as np
import pytest

import gasal2_amd as g
from parity import assert_parity, fields, run_both

pytestmark = pytest.mark.gpu

# the full-feature batch's modes: LOCAL WITH_TB + second best, GLOBAL
# WITH_START, SEMI_GLOBAL skipping target head and tail (with and without
# starts); reverse-complemented queries on half the pairs, 0.5% N, N_PENALTY 1
MODES = [(g.LOCAL, g.WITH_TB, True), (g.GLOBAL, g.WITH_START, False),
         (g.SEMI_GLOBAL, g.WITHOUT_START, True), (g.SEMI_GLOBAL, g.WITH_START, False),
         (g.LOCAL, g.WITHOUT_START, False), (g.LOCAL, g.WITH_START, True)]


@pytest.mark.parametrize("per_pair", [False, True])
@pytest.mark.parametrize("algo,start,second", MODES)
def test_packed_inputs_match_oracle(algo, start, second, per_pair):
    b = g.synth_batch(2, 1200 + 10 * algo + start, 900 if per_pair else 4000)
    got, ref = run_both(b, algo, start, second, is_rc=True, n_penalty=1, packed=True, per_pair=per_pair)
    if per_pair:   # offsets in the host batch differ from the batch's own: compare the CIGARs per pair only
        assert_parity(got, ref, algo, start, second, None)
    else:
        assert_parity(got, ref, algo, start, second, b)
    unp, _ = run_both(b, algo, start, second, is_rc=True, n_penalty=1, packed=False)
    for f in fields(algo, start, second):
        assert np.array_equal(got[f], unp[f]), f
    if start == g.WITH_TB:
        assert got["cigars"] == unp["cigars"]


def test_packed_local_tb_reference_layout_per_pair():
    """Per-sequence packed fills: the CIGAR sits at each pair's own (base)
    offset in host_res->cigar."""
    b = g.synth_batch(2, 1299, 600)
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE, isReverseComplement=True, isPacked=True)
    a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, n_penalty=1)
    try:
        got = a.run(b, per_pair=True)
        offs = np.ctypeslib.as_array(a.s.host_query_batch_offsets, shape=(b.n,)).copy()
        cig = [a.cigar(k) for k in range(b.n)]
    finally:
        a.close()
    ql = got["cigar_qlayout"]
    for k in range(b.n):
        slot = 8 * ((int(b.qlen[k]) + 7) // 8)
        nb = min(int(got["n_ops"][k]), slot)
        assert bytes(ql[offs[k]:offs[k] + nb]) == cig[k][:nb], k


@pytest.mark.parametrize("packed", [False, True])
def test_resident_reruns_keep_inputs(packed):
    """gasal2_run_resident three times on one upload: with isPacked the
    reverse/complement step rewrites the caller's words in place, so reruns
    must not apply it again; every run equals the oracle."""
    import pyoracle as O
    b = g.synth_batch(2, 1310, 700)
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE, isReverseComplement=True, isPacked=packed)
    a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n, n_penalty=1)
    try:
        a.load(b)
        outs = []
        for _ in range(3):
            a.run_resident()
            outs.append(a.download())
    finally:
        a.close()
    ref = O.align(b, g.LOCAL, start_pos=g.WITH_TB, second_best=True, is_rc=True, n_penalty=1)
    for r in outs:
        for f in fields(g.LOCAL, g.WITH_TB, True):
            assert np.array_equal(r[f], ref[f]), f
        assert np.array_equal(r["n_ops"], ref["n_ops"])
