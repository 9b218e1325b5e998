"""Golden fixtures: the reference's own example pairs (first 1000 of
test_prog/*.fasta.gz) with committed expected outputs.

CPU: the oracle still reproduces the committed outputs (guards the checker).
GPU: the HIP library and the reference-compatible test program reproduce them.
"""
import os
import subprocess

import numpy as np
import pytest

import gasal2_amd as g
import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")

import sys
sys.path.insert(0, GOLD)
from make_fixtures import MODES, load_fixture_batch  # noqa: E402

FIELDS = ("score", "qend", "tend", "qstart", "tstart", "s_score", "s_qend", "s_tend", "n_ops")


@pytest.fixture(scope="module")
def fixture():
    b, pairs = load_fixture_batch()
    exp = np.load(os.path.join(GOLD, "ref_expected_1000.npz"), allow_pickle=False)
    return b, pairs, exp


def split_cigars(flat, n_ops):
    out, o = [], 0
    for k in n_ops:
        out.append(bytes(flat[o:o + int(k)]))
        o += int(k)
    return out


def test_fixture_shape(fixture):
    b, pairs, exp = fixture
    assert b.n == 1000
    assert set(np.unique(b.qlen)) == {150}          # reference data: 150 bp queries
    assert b.tlen.min() >= 150


@pytest.mark.parametrize("mode", sorted(MODES))
def test_oracle_reproduces_golden(fixture, mode):
    b, _, exp = fixture
    r = O.align(b, **MODES[mode])
    for k in FIELDS:
        np.testing.assert_array_equal(r[k], exp[f"{mode}.{k}"], err_msg=f"{mode}.{k}")
    if f"{mode}.cigar_flat" in exp:
        assert r["cigars"] == split_cigars(exp[f"{mode}.cigar_flat"], exp[f"{mode}.n_ops"])


def _gpu_fields(mode):
    kw = MODES[mode]
    from parity import fields
    return fields(kw["algo"], kw.get("start_pos", 0), kw.get("second_best", False))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", sorted(MODES))
def test_gpu_reproduces_golden(fixture, mode):
    b, _, exp = fixture
    kw = MODES[mode]
    p = g.Parameters(algo=kw["algo"], start_pos=kw.get("start_pos", 0),
                     secondBest=g.TRUE if kw.get("second_best") else g.FALSE,
                     semiglobal_skipping_head=kw.get("head", 2), semiglobal_skipping_tail=kw.get("tail", 2),
                     k_band=kw.get("k_band", 0))
    a = g.Aligner(p, int(b.qlen.max()), int(b.tlen.max()), b.n)
    try:
        got = a.run(b)
        for k in _gpu_fields(mode):
            np.testing.assert_array_equal(got[k], exp[f"{mode}.{k}"], err_msg=f"{mode}.{k}")
        if f"{mode}.cigar_flat" in exp:
            np.testing.assert_array_equal(got["n_ops"], exp[f"{mode}.n_ops"])
            cg = split_cigars(exp[f"{mode}.cigar_flat"], exp[f"{mode}.n_ops"])
            assert [a.cigar(k) for k in range(b.n)] == cg
    finally:
        a.close()


def expected_lines(mode, b, pairs, exp, args):
    """The reference program's result lines (test_prog/test_prog.cpp:353-429)
    built from the golden outputs."""
    kw = MODES[mode]
    algo, start = kw["algo"], kw.get("start_pos", 0)
    head = kw.get("head", 2)
    cg = split_cigars(exp[f"{mode}.cigar_flat"], exp[f"{mode}.n_ops"]) if f"{mode}.cigar_flat" in exp else None
    lines = []
    for i in range(b.n):
        s = f"query_name={pairs[i][0][1:]}\ttarget_name={pairs[i][2][1:]}\tscore={exp[mode + '.score'][i]}"
        if start in (1, 2) and ((algo == 2 and head != 0) or algo > 2):
            s += f"\tquery_batch_start={exp[mode + '.qstart'][i]}\ttarget_batch_start={exp[mode + '.tstart'][i]}"
        if algo != 1:
            s += f"\tquery_batch_end={exp[mode + '.qend'][i]}\ttarget_batch_end={exp[mode + '.tend'][i]}"
        if kw.get("second_best"):
            s += (f"\t2nd_score={exp[mode + '.s_score'][i]}\t2nd_query_batch_end={exp[mode + '.s_qend'][i]}"
                  f"\t2nd_target_batch_end={exp[mode + '.s_tend'][i]}")
        if start == 2:
            s += "\tCIGAR=" + g.cigar_to_string(cg[i])
        lines.append(s)
    return lines


PROG_MODES = {
    "local_tb_second": ["-y", "local", "-t", "--second-best"],
    "local_start_second": ["-y", "local", "-s", "--second-best"],
    "global": ["-y", "global"],
    "semi_tt_start": ["-y", "semi_global", "-s", "-x", "TARGET", "TARGET"],
}


@pytest.mark.gpu
@pytest.mark.parametrize("mode", sorted(PROG_MODES))
def test_test_prog_output_matches(fixture, mode, tmp_path):
    b, pairs, exp = fixture
    prog = os.path.join(os.path.dirname(HERE), "gasal2_amd", "test_prog")
    assert os.path.exists(prog), "test_prog not built"
    qf, tf = tmp_path / "q.fa", tmp_path / "t.fa"
    qf.write_text("".join(f"{p[0]}\n{p[1]}\n" for p in pairs))
    tf.write_text("".join(f"{p[2]}\n{p[3]}\n" for p in pairs))
    cmd = [prog, "-p", "-n", "2"] + PROG_MODES[mode] + [str(qf), str(tf)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = sorted(r.stdout.strip().split("\n"))
    want = sorted(expected_lines(mode, b, pairs, exp, PROG_MODES[mode]))
    assert len(got) == len(want)
    assert got == want
