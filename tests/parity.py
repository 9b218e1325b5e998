"""Shared helpers: run one batch through the HIP library (via the C-ABI) and
through the CPU oracle, and compare every output the mode defines."""
import numpy as np

import gasal2_amd as g
import pyoracle as O

# outputs each (algo, start) defines (reference src/res.cpp:8-73, src/gasal_align.cu:266-304)
def fields(algo, start, second):
    f = ["score"]
    if algo != g.GLOBAL:
        f += ["qend", "tend"]
    # starts are written by the LOCAL reverse pass or traceback and by the
    # SEMI_GLOBAL reverse pass; other modes leave them unwritten in the reference
    if (algo == g.LOCAL and start in (g.WITH_START, g.WITH_TB)) or (algo == g.SEMI_GLOBAL and start == g.WITH_START):
        f += ["qstart", "tstart"]
    if second and algo in (g.LOCAL, g.SEMI_GLOBAL):
        f += ["s_score", "s_qend", "s_tend"]
    return f


def run_both(batch, algo, start=g.WITHOUT_START, second=False, head=g.TARGET, tail=g.TARGET, k_band=0,
             is_rc=False, scores=(1, 4, 6, 1), n_penalty=None, n_code=0x4E, cigars=True, packed=False,
             per_pair=None):
    """packed: the host supplies 4-bit packed words (isPacked=true); the
    oracle always reads the bases."""
    p = g.Parameters(algo=algo, start_pos=start, secondBest=g.TRUE if second else g.FALSE,
                     semiglobal_skipping_head=head, semiglobal_skipping_tail=tail, k_band=k_band,
                     isReverseComplement=is_rc, sa=scores[0], sb=scores[1], gapo=scores[2], gape=scores[3],
                     isPacked=packed)
    a = g.Aligner(p, int(batch.qlen.max()), int(batch.tlen.max()), batch.n, scores=scores, n_code=n_code,
                  n_penalty=n_penalty)
    L = g.lib()
    try:
        got = a.run(batch, per_pair=per_pair)
        if start == g.WITH_TB and cigars:
            got["cigars"] = [a.cigar(k) for k in range(batch.n)]
        # The auto policy runs every pair of a small LOCAL batch one wave each;
        # the lane kernels then run too, with the one-wave path switched off,
        # and must give the same outputs.
        if algo == g.LOCAL and not second and start != g.WITH_START and L.gasal2_get_dispatch() == 0:
            L.gasal2_set_dispatch(2)
            try:
                lane = a.run(batch, per_pair=per_pair)
                if start == g.WITH_TB and cigars:
                    lane["cigars"] = [a.cigar(k) for k in range(batch.n)]
            finally:
                L.gasal2_set_dispatch(0)
            keys = fields(algo, start, second) + (["n_ops"] + (["cigars"] if cigars else []) if start == g.WITH_TB else [])
            assert_same(got, lane, keys, "lane path")
    finally:
        a.close()
    ref = O.align(batch, algo, start_pos=start, second_best=second, head=head, tail=tail, k_band=k_band,
                  is_rc=is_rc, scores=scores, n_code=n_code, n_penalty=n_penalty)
    return got, ref


def assert_same(got, other, keys, what):
    """The outputs `keys` of two GPU runs identical."""
    for f in keys:
        x, y = got[f], other[f]
        if isinstance(x, np.ndarray):
            bad = np.nonzero(x != y)[0]
            assert len(bad) == 0, f"{what}: {f} differs at {len(bad)} pairs, first {bad[0]}"
        else:
            assert x == y, f"{what}: {f} differs"


def assert_parity(got, ref, algo, start, second, batch=None):
    for f in fields(algo, start, second):
        a, b = got[f], ref[f]
        assert a is not None, f"{f} missing from GPU results"
        bad = np.nonzero(a != b)[0]
        assert len(bad) == 0, (f"{f}: {len(bad)} mismatches, first pair {bad[0]}: gpu {a[bad[0]]} oracle {b[bad[0]]}"
                               + (f" q={batch.query(bad[0])!r} t={batch.target(bad[0])!r}" if batch is not None
                                  and batch.qlen[bad[0]] < 64 else ""))
    if start == g.WITH_TB:
        bad = np.nonzero(got["n_ops"] != ref["n_ops"])[0]
        assert len(bad) == 0, f"n_ops: {len(bad)} mismatches, first {bad[0]}: {got['n_ops'][bad[0]]} vs {ref['n_ops'][bad[0]]}"
        if "cigars" in got:
            for k, (x, y) in enumerate(zip(got["cigars"], ref["cigars"])):
                assert x == y, f"CIGAR of pair {k}: gpu {g.cigar_to_string(x)} oracle {g.cigar_to_string(y)}"
            # reference layout: bytes at the pair's query offset, clipped to its slot
            if batch is not None:
                ql = got["cigar_qlayout"]
                for k in range(batch.n):
                    slot = 8 * ((int(batch.qlen[k]) + 7) // 8)
                    nb = min(int(ref["n_ops"][k]), slot)
                    o = int(batch.qoff[k])
                    assert bytes(ql[o:o + nb]) == ref["cigars"][k][:nb], f"query-layout CIGAR of pair {k}"
