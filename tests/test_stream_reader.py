"""The host input path's reader (gasal2_pairs_read, include/gasal2_stream.h),
CPU only: FASTA with wrapped lines and the reference harness's operation
markers ('>' '<' '/' '+', test_prog/test_prog.cpp:79-137), FASTQ, CRLF line
ends, N_CODE padding to 8 bases, packed words, names, and the errors it
reports."""
import numpy as np
import pytest

import gasal2_amd as g


def write(path, text):
    path.write_bytes(text.encode() if isinstance(text, str) else text)
    return str(path)


def test_fasta_markers_wrapping_and_padding(tmp_path):
    q = write(tmp_path / "q.fa", ">a\nACGT\nAC\n<b\nGGGGGGGGG\n/c\n\n+d\nA\n")
    t = write(tmp_path / "t.fa", ">a\nTT\n>b\nC\n>c\nACGTACGT\n>d\nN\n")
    ps = g.PairSet.read(q, t, names=True, pack=True)
    try:
        b = ps.to_batch()
        assert ps.n == 4
        assert [b.query(i) for i in range(4)] == [b"ACGTAC", b"GGGGGGGGG", b"", b"A"]
        assert [b.target(i) for i in range(4)] == [b"TT", b"C", b"ACGTACGT", b"N"]
        assert list(b.qop) == [0, 1, 2, 3] and list(b.top) == [0, 0, 0, 0]
        assert list(b.qoff) == [0, 8, 24, 24] and list(b.qlen) == [6, 9, 0, 1]
        assert b.q[6:8].tobytes() == b"NN" and len(b.q) == 32
        assert ps.names() == (["a", "b", "c", "d"], ["a", "b", "c", "d"])
        words = np.ctypeslib.as_array(ps.p.query_words, shape=(len(b.q) // 8,)).copy()
        assert bytes(words.view(np.uint8)) == bytes(g.pack_bases(b.q))
    finally:
        ps.close()


def test_fastq_and_crlf(tmp_path):
    q = write(tmp_path / "q.fq", b"@r1 x\r\nACGTACGTAC\r\n+\r\nIIIIIIIIII\r\n@r2\r\nGG\r\n+r2\r\nII\r\n")
    t = write(tmp_path / "t.fq", "@r1\nTTTT\n+\n!!!!\n@r2\nCCCCCCCCCCCCCCCCC\n+\n" + "!" * 17 + "\n")
    ps = g.PairSet.read(q, t, names=True)
    try:
        b = ps.to_batch()
        assert [b.query(i) for i in range(2)] == [b"ACGTACGTAC", b"GG"]
        assert [b.target(i) for i in range(2)] == [b"TTTT", b"C" * 17]
        assert ps.names()[0] == ["r1 x", "r2"]
    finally:
        ps.close()


def test_matches_a_batch_written_to_fasta(tmp_path):
    b = g.synth_batch(2, 31, 500)
    for name, data, offs, lens, ops in (("q.fa", b.q, b.qoff, b.qlen, b.qop), ("t.fa", b.t, b.toff, b.tlen, b.top)):
        with open(tmp_path / name, "wb") as f:
            for i in range(b.n):
                f.write(b"></+"[ops[i]:ops[i] + 1] + b"%d\n" % i)
                s = bytes(data[offs[i]:offs[i] + lens[i]])
                for j in range(0, len(s), 60):
                    f.write(s[j:j + 60] + b"\n")
    ps = g.PairSet.read(str(tmp_path / "q.fa"), str(tmp_path / "t.fa"))
    try:
        r = ps.to_batch()
        for k in ("q", "t", "qoff", "toff", "qlen", "tlen", "qop", "top"):
            assert np.array_equal(getattr(r, k), getattr(b, k)), k
    finally:
        ps.close()


@pytest.mark.parametrize("q,t", [(">a\nAC\n", ">a\nAC\n>b\nGG\n"), ("ACGT\n", ">a\nA\n"), ("@r\nAC\n", "@r\nA\n+\nI\n")])
def test_errors(tmp_path, q, t):
    with pytest.raises(RuntimeError):
        g.PairSet.read(write(tmp_path / "q", q), write(tmp_path / "t", t))


def test_missing_file(tmp_path):
    with pytest.raises(RuntimeError):
        g.PairSet.read(str(tmp_path / "nope.fa"), str(tmp_path / "nope2.fa"))
