"""Growth paths of a storage, through the reference's own call sequence:
host pages appended by gasal_host_batch_fill / gasal_host_batch_add when a
batch outgrows them (src/host_batch.cpp:79-153,162-223), per-pair host arrays
resized by gasal_host_alns_resize (src/interfaces.cpp:26-78), and device
buffers grown by gasal_aln_async (src/gasal_align.cu:70-145).  A storage made
for 8 pairs of 16 bases takes hundreds of pairs of 100-400 bases; the pages'
chain, contents and sizes are read back, and every batch's outputs equal the
oracle's."""
import ctypes as C

import numpy as np
import pytest

import gasal2_amd as g
import pyoracle as O
from parity import fields

pytestmark = pytest.mark.gpu


def pages(head):
    out = []
    p = head
    while p:
        h = p.contents
        out.append((h.offset, h.page_size, h.data_size, bytes(h.data[:h.data_size])))
        p = h.next
    return out


def make_storage(p, max_len=16, max_n=8):
    L = g.lib()
    v = L.gasal_init_gpu_storage_v(1)
    L.gasal_init_streams(C.byref(v), max_len, max_len, max_n, p.native())
    return v


def fill(s, b, use_add=False):
    L = g.lib()
    qi = ti = 0
    for k in range(b.n):
        s.host_query_batch_offsets[k] = qi
        s.host_target_batch_offsets[k] = ti
        q = bytes(b.q[b.qoff[k]:b.qoff[k] + b.qlen[k]])
        t = bytes(b.t[b.toff[k]:b.toff[k] + b.tlen[k]])
        if use_add:
            # the bases in two chunks and one at a time, then N padding to 8
            h = len(q) // 2
            qi = L.gasal_host_batch_add(C.byref(s), qi, q[:h], h, g.QUERY)
            for c in q[h:]:
                qi = L.gasal_host_batch_addbase(C.byref(s), qi, bytes([c]), g.QUERY)
            pad = (-len(q)) % 8
            qi = L.gasal_host_batch_add(C.byref(s), qi, b"N" * pad, pad, g.QUERY) if pad else qi
            ti = L.gasal_host_batch_add(C.byref(s), ti, t, len(t), g.TARGET)
            pad = (-len(t)) % 8
            ti = L.gasal_host_batch_add(C.byref(s), ti, b"N" * pad, pad, g.TARGET) if pad else ti
        else:
            qi = L.gasal_host_batch_fill(C.byref(s), qi, q, len(q), g.QUERY)
            ti = L.gasal_host_batch_fill(C.byref(s), ti, t, len(t), g.TARGET)
        s.host_query_batch_lens[k] = len(q)
        s.host_target_batch_lens[k] = len(t)
    C.memmove(s.host_query_op, b.qop.ctypes.data, b.n)
    C.memmove(s.host_target_op, b.top.ctypes.data, b.n)
    s.current_n_alns = b.n
    return qi, ti


def run(s, qb, tb, n, p):
    L = g.lib()
    L.gasal_aln_async(C.byref(s), qb, tb, n, p.native())
    while True:
        r = L.gasal_is_aln_async_done(C.byref(s))
        if r == 0:
            break
        assert r == -1, r
    h, h2 = s.host_res.contents, s.host_res_second.contents
    arr = lambda ptr: np.ctypeslib.as_array(ptr, shape=(n,)).copy()
    return {"score": arr(h.aln_score), "qend": arr(h.query_batch_end), "tend": arr(h.target_batch_end),
            "qstart": arr(h.query_batch_start), "tstart": arr(h.target_batch_start), "n_ops": arr(h.n_cigar_ops),
            "s_score": arr(h2.aln_score), "s_qend": arr(h2.query_batch_end), "s_tend": arr(h2.target_batch_end)}


def check(got, b):
    ref = O.align(b, g.LOCAL, start_pos=g.WITH_TB, second_best=True, is_rc=True, n_penalty=1, with_cigar=False)
    for f in fields(g.LOCAL, g.WITH_TB, True) + ["n_ops"]:
        assert np.array_equal(got[f], ref[f]), f


@pytest.mark.parametrize("use_add", [False, True])
def test_storage_grows_host_pages_arrays_and_device_buffers(use_add):
    L = g.lib()
    g.set_n_penalty(1)
    g.gasal_copy_subst_scores(g.gasal_subst_scores(1, 4, 6, 1))
    p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE, isReverseComplement=True)
    v = make_storage(p)
    s = v.a[0]
    try:
        assert s.host_max_n_alns == 8 and s.host_max_query_batch_bytes == 128
        b = g.synth_batch(2, 3100 + use_add, 300)
        L.gasal_host_alns_resize(C.byref(s), b.n, p.native())          # 8 -> 300 pairs
        assert s.host_max_n_alns == b.n
        qb, tb = fill(s, b, use_add)
        assert (qb, tb) == (len(b.q), len(b.t))
        # the host pages: a chain covering the batch, contents at their offsets
        for head, data in ((s.extensible_host_unpacked_query_batch, b.q), (s.extensible_host_unpacked_target_batch, b.t)):
            pg = pages(head)
            assert len(pg) > 3
            assert b"".join(d for _, _, _, d in pg) == bytes(data)
            for (o0, _, d0, _), (o1, _, _, _) in zip(pg, pg[1:]):
                assert o1 == o0 + d0                                    # each page starts where the last one's data ends
            if not use_add:                                             # fill doubles the page (src/host_batch.cpp:107-123)
                assert all(p1[1] == 2 * p0[1] for p0, p1 in zip(pg, pg[1:]))
        assert s.host_max_query_batch_bytes >= qb and s.host_max_target_batch_bytes >= tb
        gq0, gt0, gn0 = s.gpu_max_query_batch_bytes, s.gpu_max_target_batch_bytes, s.gpu_max_n_alns
        got = run(s, qb, tb, b.n, p)
        assert s.gpu_max_query_batch_bytes >= qb > gq0 and s.gpu_max_target_batch_bytes >= tb > gt0
        assert s.gpu_max_n_alns >= b.n > gn0
        check(got, b)
        # a smaller batch on the grown storage: the pages are reused, nothing shrinks
        grown = (s.gpu_max_query_batch_bytes, s.gpu_max_target_batch_bytes, s.gpu_max_n_alns)
        b2 = g.synth_batch(2, 3200 + use_add, 120)
        qb, tb = fill(s, b2, use_add)
        got = run(s, qb, tb, b2.n, p)
        assert (s.gpu_max_query_batch_bytes, s.gpu_max_target_batch_bytes, s.gpu_max_n_alns) == grown
        check(got, b2)
    finally:
        L.gasal_destroy_streams(C.byref(v), p.native())
        L.gasal_destroy_gpu_storage_v(C.byref(v))
