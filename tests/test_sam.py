"""CIGAR text and SAM records (include/gasal2_sam.h), on the CPU.

Known answers pin the text forms and the sequence the DP saw (reverse /
complement closed forms).  SAM records built from the oracle's alignments are
read back by an independent SAM reader: field syntax, SEQ length equal to the
CIGAR's query length, the span ending at the reported ends, NM recounted from
the bases, and the record's alignment rescored with the affine model -- never
above the reported optimum, and equal to it whenever the reference's
traceback accounts the walk correctly.  The writer's file is compared line by
line with single records.
"""
import random
import re

import ctypes as C
import numpy as np
import pytest

import gasal2_amd as g
import pyoracle as O

ACGT = b"ACGT"
COMP = bytes.maketrans(b"ACGTN", b"TGCAN")


def rnd(rng, n):
    return bytes(rng.choice(ACGT) for _ in range(n))


def mutate(rng, s, p=0.1):
    out = bytearray()
    for c in s:
        u = rng.random()
        if u < p / 3:
            continue                      # deletion
        out.append(rng.choice(ACGT) if u < 2 * p / 3 else c)
        if u > 1 - p / 3:
            out += rnd(rng, rng.randint(1, 3))   # insertion
    return bytes(out)


def cig_bytes(forward):
    """Reverse-order CIGAR bytes of a forward op list [(op, count), ...]."""
    return bytes((op | (n << 2)) for op, n in reversed(forward))


# ------------------------------------------------------------ CIGAR text ---
@pytest.mark.parametrize("forward,harness,sam", [
    ([(0, 10), (3, 2), (0, 5)], "10M2I5M", "10M2I5M"),
    ([(0, 3), (1, 1), (0, 4)], "3M1X4M", "8M"),
    ([(0, 63), (0, 2)], "65M", "65M"),
    ([(0, 5), (2, 3), (0, 5)], "5M3D5M", "5M3D5M"),
    ([(1, 2), (0, 1), (2, 1), (3, 1), (0, 7)], "2X1M1D1I7M", "3M1D1I7M"),
    ([], "", ""),
])
def test_cigar_text_known_answers(forward, harness, sam):
    cg = cig_bytes(forward)
    assert g.cigar_text(cg, g.CIGAR_HARNESS) == harness
    assert g.cigar_text(cg, g.CIGAR_SAM) == sam


def test_cigar_text_matches_harness_printer_and_reports_length():
    rng = random.Random(5)
    for _ in range(200):
        cg = bytes(rng.randrange(4) | (rng.randint(1, 63) << 2) for _ in range(rng.randint(1, 30)))
        assert g.cigar_text(cg, g.CIGAR_HARNESS) == g.cigar_to_string(cg)
    L = g.lib()
    cg = cig_bytes([(0, 12), (3, 1), (0, 40)])
    buf = C.create_string_buffer(4)
    assert L.gasal2_cigar_text(cg, len(cg), 1, buf, 4) == len("12M1I40M")   # snprintf-style length
    assert buf.value == b"12M"
    with pytest.raises(ValueError):
        g.cigar_text(cg, 7)


# -------------------------------------------------- the sequence as aligned ---
def seq_field(line):
    return line.rstrip("\n").split("\t")[9]


@pytest.mark.parametrize("length", [8, 16, 13, 1, 7, 31])
def test_seq_is_the_query_as_aligned(length):
    rng = random.Random(length)
    q = rnd(rng, length)
    pad = (-length) % 8
    t = b"ACGT" * 10
    # no reverse/complement step: the bytes as given
    for op in range(4):
        line = g.sam_record(q, t, 0, query_op=op, rc_enabled=False)
        assert seq_field(line) == q.decode()
    # with the step (N_CODE 'N' never counts as padding, so the whole padded
    # word sequence is reversed and the padding leads)
    rev = (b"N" * pad + q[::-1])[:length]
    cases = {1: rev, 2: q.translate(COMP), 3: rev.translate(COMP)}
    for op, want in cases.items():
        line = g.sam_record(q, t, 0, query_op=op, rc_enabled=True)
        assert seq_field(line) == want.decode(), op
        assert (int(line.split("\t")[1]) & 16) == (16 if op == 3 else 0)


def test_unmapped_and_names():
    line = g.sam_record(b"ACGTAC", b"ACGT", 0, qname="read 1 extra", rname="chr1 x").rstrip("\n").split("\t")
    assert line[:9] == ["read", "4", "*", "0", "0", "*", "*", "0", "0"]
    assert line[10] == "*" and "AS:i:0" in line
    line = g.sam_record(b"ACGT", b"ACGT", 4, qname=None, rname="", query_end=3, target_end=3,
                        cigar=cig_bytes([(0, 4)])).split("\t")
    assert line[0] == "*" and line[2] == "*"
    line = g.sam_record(b"ACGT", b"ACGT", 4, query_end=3, target_end=3, cigar=cig_bytes([(0, 4)]),
                        flags=g.SAM_NO_SEQ, second_score=2).rstrip("\n").split("\t")
    assert line[9] == "*" and "XS:i:2" in line


# ------------------------------------------------------ an independent reader ---
CIGAR_RE = re.compile(r"(\d+)([MIDNSHP=X])")


def read_record(line):
    f = line.rstrip("\n").split("\t")
    assert len(f) >= 11
    rec = dict(qname=f[0], flag=int(f[1]), rname=f[2], pos=int(f[3]), mapq=int(f[4]), cigar=f[5],
               seq=f[9], qual=f[10], tags={})
    assert f[6] == "*" and f[7] == "0" and f[8] == "0"
    for t in f[11:]:
        k, ty, v = t.split(":", 2)
        rec["tags"][k] = int(v) if ty == "i" else v
    if rec["cigar"] != "*":
        ops = CIGAR_RE.findall(rec["cigar"])
        assert "".join(a + b for a, b in ops) == rec["cigar"]
        rec["ops"] = [(int(a), b) for a, b in ops]
    return rec


def rescore(rec, t, sc=(1, 4, 6, 1)):
    """Affine score of the record's alignment; (score, query end, target end, NM)."""
    m, mm, go, ge = sc
    q = rec["seq"].encode()
    qi, ti, s, nm = 0, rec["pos"] - 1, 0, 0
    for i, (n, op) in enumerate(rec["ops"]):
        if op == "S":
            assert i in (0, len(rec["ops"]) - 1)
            if i == 0:
                qi += n
        elif op in "M=X":
            for _ in range(n):
                same = q[qi] == t[ti]
                assert op == "M" or (op == "=") == same
                s += m if same else -mm
                nm += not same
                qi += 1
                ti += 1
        elif op == "I":
            s -= go + ge * n
            qi += n
            nm += n
        elif op == "D":
            s -= go + ge * n
            ti += n
            nm += n
    return s, qi, ti, nm


def oracle_pairs(seed, n):
    rng = random.Random(seed)
    qs, ts = [], []
    for _ in range(n):
        core = rnd(rng, rng.randint(12, 70))
        qs.append(rnd(rng, rng.randint(0, 12)) + core + rnd(rng, rng.randint(0, 12)))
        ts.append(rnd(rng, rng.randint(0, 12)) + mutate(rng, core) + rnd(rng, rng.randint(0, 12)))
    b = g.make_batch(qs, ts)
    return qs, ts, b


@pytest.mark.parametrize("eqx", [False, True])
def test_local_records_read_back(eqx):
    qs, ts, b = oracle_pairs(11, 300)
    r = O.align(b, g.LOCAL, start_pos=g.WITH_TB, second_best=True)
    equal = 0
    for k in range(b.n):
        cg = r["cigars"][k]
        line = g.sam_record(qs[k], ts[k], int(r["score"][k]), int(r["qend"][k]), int(r["tend"][k]), cg,
                            qname=f"q{k}", rname=f"t{k}", second_score=int(r["s_score"][k]),
                            flags=g.SAM_EQX if eqx else 0)
        rec = read_record(line)
        assert rec["tags"]["AS"] == r["score"][k] and rec["tags"]["XS"] == r["s_score"][k]
        if r["score"][k] <= 0:
            assert rec["flag"] & 4
            continue
        assert rec["flag"] == 0 and rec["rname"] == f"t{k}" and rec["mapq"] == 255
        assert rec["seq"] == qs[k].decode()
        qlen = sum(n for n, op in rec["ops"] if op in "MIS=X")
        assert qlen == len(qs[k])
        s, qe, te, nm = rescore(rec, ts[k])
        lead = rec["ops"][0][0] if rec["ops"][0][1] == "S" else 0
        trail = rec["ops"][-1][0] if rec["ops"][-1][1] == "S" else 0
        assert qe - 1 == r["qend"][k] and te - 1 == r["tend"][k]          # anchored at the reported ends
        assert trail == len(qs[k]) - 1 - r["qend"][k]
        assert rec["tags"]["NM"] == nm
        assert s <= r["score"][k]                                         # a local alignment never beats the optimum
        if s == r["score"][k]:                                            # a correctly accounted walk: its starts
            equal += 1
            assert (lead, rec["pos"] - 1) == (r["qstart"][k], r["tstart"][k])
    assert equal >= 0.9 * int((r["score"] > 0).sum())


def test_global_records_anchor_at_the_start():
    qs, ts, b = oracle_pairs(12, 120)
    r = O.align(b, g.GLOBAL, start_pos=g.WITH_TB)
    for k in range(b.n):
        line = g.sam_record(qs[k], ts[k], int(r["score"][k]), -1, -1, r["cigars"][k])
        rec = read_record(line)
        if rec["flag"] & 4:
            continue
        assert rec["pos"] == 1 and rec["ops"][0][1] != "S"
        assert sum(n for n, op in rec["ops"] if op in "MIS=X") == len(qs[k])
        _, qe, te, _ = rescore(rec, ts[k])
        assert qe <= len(qs[k]) and te <= len(ts[k])


def results_struct(r, n, cap, keep):
    R = g.gasal2_results_t()
    def p(a, ct):
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(ct))
    R.score = p(r["score"].astype(np.int32), C.c_int32)
    R.query_end = p(r["qend"].astype(np.int32), C.c_int32)
    R.target_end = p(r["tend"].astype(np.int32), C.c_int32)
    R.second_score = p(r["s_score"].astype(np.int32), C.c_int32)
    R.n_cigar_ops = p(r["n_ops"].astype(np.uint32), C.c_uint32)
    slots = np.zeros((n, cap), np.uint8)
    for k, cg in enumerate(r["cigars"]):
        slots[k, :min(len(cg), cap)] = np.frombuffer(cg, np.uint8)[:cap]
    R.cigar = p(slots.reshape(-1), C.c_uint8)
    R.cigar_cap = cap
    return R


def test_writer_matches_single_records(tmp_path):
    rng = random.Random(13)
    qs = [rnd(rng, rng.randint(8, 40)) for _ in range(40000)]
    ts = [mutate(rng, q) + rnd(rng, rng.randint(0, 6)) for q in qs]
    ops = np.array([rng.choice((0, 3)) for _ in qs], np.uint8)
    b = g.make_batch(qs, ts)
    b.qop[:] = ops
    r = O.align(b, g.LOCAL, start_pos=g.WITH_TB, second_best=True, is_rc=True)
    cap = 24                      # some CIGARs do not fit: their records carry '*'
    keep = []
    R = results_struct(r, b.n, cap, keep)
    ps = g.PairSet.from_batch(b)
    path = tmp_path / "out.sam"
    assert g.lib().gasal2_sam_write(str(path).encode(), C.byref(ps.p), C.byref(R), 1, g.SAM_HEADER, 3) == 0
    lines = path.read_text().splitlines(keepends=True)
    head = [l for l in lines if l.startswith("@")]
    body = lines[len(head):]
    assert head[0].startswith("@HD\tVN:") and head[-1].startswith("@PG\t")
    sq = head[1:-1]
    assert len(sq) == b.n and sq[7] == f"@SQ\tSN:t7\tLN:{len(ts[7])}\n"
    assert len(body) == b.n
    clipped = 0
    for k in range(0, b.n, 97):
        cg = r["cigars"][k]
        fits = len(cg) <= cap
        clipped += not fits
        want = g.sam_record(qs[k], ts[k], int(r["score"][k]), int(r["qend"][k]), int(r["tend"][k]),
                            cg if fits else b"", qname=f"q{k}", rname=f"t{k}", query_op=int(ops[k]),
                            second_score=int(r["s_score"][k]), rc_enabled=True)
        assert body[k] == want, k
    assert any(read_record(l)["flag"] & 16 for l in body[:200])


@pytest.mark.parametrize("n_code", [0x0E, 0x04])
def test_seq_follows_the_reverse_with_a_nibble_n_code(n_code):
    """With an N_CODE that a nibble can equal, the reverse shifts by the
    count of such nibbles in the last word (the reference's nbr_N): SEQ must
    be the oracle's reverse/complement of the padded words, letter by letter."""
    rng = random.Random(n_code)
    for length in list(range(1, 25)) + [40, 63]:
        q = rnd(rng, length)
        pad = (-length) % 8
        nib = [c & 15 for c in q] + [n_code & 15] * pad
        words = [sum(nib[8 * w + i] << (28 - 4 * i) for i in range(8)) for w in range(len(nib) // 8)]
        for op in (1, 2, 3):
            got = O.rc_one(words, length, op, n_code)
            want = "".join("ACGT"["\x01\x03\x07\x04".index(chr((int(got[i // 8]) >> (28 - 4 * (i % 8))) & 15))]
                           if ((int(got[i // 8]) >> (28 - 4 * (i % 8))) & 15) in (1, 3, 4, 7) else "N"
                           for i in range(length))
            line = g.sam_record(q, b"ACGT", 0, query_op=op, rc_enabled=True, n_code=n_code)
            assert seq_field(line) == want, (length, op)
