"""The C-ABI library (CPU checks only, no compute): it loads, it exports every
function include/*.h declares, and the host-side API logic that needs no GPU
behaves like the reference's (validation messages, enum values, defaults)."""
import ctypes as C
import os
import re

import pytest

import gasal2_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b((?:gasal|gasal2)_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    L = g.lib()
    decl = declared_functions()
    assert len(decl) >= 40
    missing = [n for n in sorted(decl) if not hasattr(L, n)]
    assert not missing, missing
    # and the Python mirror binds every one of them
    assert decl <= set(g.exported_symbols()) | {"gasal_kernel_launcher"}


def test_enum_values_match_reference():
    assert [int(x) for x in g.comp_start] == [0, 1, 2]
    assert int(g.LOCAL) == 3 and int(g.KSW) == 6 and int(g.BANDED) == 5 and int(g.SEMI_GLOBAL) == 2
    assert [int(x) for x in g.data_source] == [0, 1, 2, 3]


def test_parameters_defaults_roundtrip():
    p = g.Parameters()
    assert (p.sa, p.sb, p.gapo, p.gape) == (1, 4, 6, 1)
    assert p.semiglobal_skipping_head == g.TARGET and p.semiglobal_skipping_tail == g.TARGET
    assert p.algo == g.UNKNOWN and p.start_pos == g.WITHOUT_START and p.k_band == 0
    p.algo = g.LOCAL
    p.secondBest = g.TRUE
    ptr = p.native()
    pod = g.gasal2_params_t()
    g.lib().gasal2_parameters_get(ptr, C.byref(pod))
    assert pod.algo == 3 and pod.secondBest == 1 and pod.sa == 1 and pod.semiglobal_skipping_head == 2


def test_struct_layout_is_c_layout():
    # 20 pointers, 6 result pointers, 7 uint32 sizes, uint64, pointer, stream, 2 ints, impl
    s = g.gasal_gpu_storage_t
    assert s.packed_tb_matrix_size.offset % 8 == 0
    assert s.current_n_alns.offset == s.gpu_max_query_batch_bytes.offset + 6 * 4
    assert C.sizeof(g.gasal_res_t) == 7 * 8


@pytest.mark.parametrize("args,msg", [
    ((8, 8, 0), "actual_n_alns <= 0"),
    ((0, 8, 1), "actual_query_batch_bytes <= 0"),
    ((8, 0, 1), "actual_target_batch_bytes <= 0"),
    ((8, 8, 1 << 30), "actual_n_alns(1073741824) > host_max_n_alns(16)"),
    ((12, 8, 1), "actual_query_batch_bytes=12 is not a multiple of 8"),
    ((8, 20, 1), "actual_target_batch_bytes=20 is not a multiple of 8"),
    ((1 << 20, 8, 1), "actual_query_batch_bytes(1048576) > host_max_query_batch_bytes(4096)"),
])
def test_aln_async_validation_messages(args, msg):
    """src/gasal_align.cu:32-67: misuse is reported before any device work."""
    L = g.lib()
    L.gasal2_set_error_mode(1)
    try:
        s = g.gasal_gpu_storage_t()
        s.host_max_query_batch_bytes = 4096
        s.host_max_target_batch_bytes = 4096
        s.host_max_n_alns = 16
        before = L.gasal2_error_count()
        p = g.Parameters(algo=g.LOCAL)
        qb, tb, n = args
        L.gasal_aln_async(C.byref(s), qb, tb, n, p.native())
        assert L.gasal2_error_count() == before + 1
        assert msg in L.gasal2_last_error().decode()
        assert s.is_free == 0  # untouched
    finally:
        L.gasal2_set_error_mode(0)


def test_is_done_on_idle_storage_returns_minus_two():
    s = g.gasal_gpu_storage_t()
    s.is_free = 1
    assert g.lib().gasal_is_aln_async_done(C.byref(s)) == -2


def test_n_code_and_penalty_switches():
    L = g.lib()
    g.set_n_code(0x4E)
    assert L.gasal2_get_n_code() == 0x4E
    g.set_n_code(0x4E)
    g.set_n_penalty(None)


def test_cigar_text_format():
    # reverse-order bytes op | count << 2 -> merged forward text (test_prog.cpp:382-428)
    cg = bytes([(5 << 2) | 0, (3 << 2) | 2, (63 << 2) | 0, (2 << 2) | 0])
    assert g.cigar_to_string(cg) == "65M3D5M"


def test_failed_growth_keeps_storage_and_launches_nothing():
    """Error mode 1: a device allocation that fails while growing the storage
    (here: no GPU in this process) ends gasal_aln_async before any copy or
    launch, and the storage keeps its old buffers and recorded sizes."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("needs a process where device allocation fails")
    L = g.lib()
    L.gasal2_set_error_mode(1)
    try:
        s = g.gasal_gpu_storage_t()
        s.host_max_query_batch_bytes = s.host_max_target_batch_bytes = 4096
        s.gpu_max_query_batch_bytes = s.gpu_max_target_batch_bytes = 8
        s.host_max_n_alns = s.gpu_max_n_alns = 16
        s.is_free = 1
        p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB)
        before = L.gasal2_error_count()
        L.gasal_aln_async(C.byref(s), 64, 64, 1, p.native())
        assert L.gasal2_error_count() > before
        assert "CUDA error" in L.gasal2_last_error().decode()
        assert s.gpu_max_query_batch_bytes == 8 and s.gpu_max_target_batch_bytes == 8
        assert not s.unpacked_query_batch and not s.packed_query_batch
        assert s.is_free == 1   # nothing enqueued
    finally:
        L.gasal2_set_error_mode(0)


def _page(nbytes, data_size):
    buf = (C.c_uint8 * nbytes)()
    pg = g.host_batch_t()
    pg.data = C.cast(buf, C.POINTER(C.c_uint8))
    pg.page_size = nbytes
    pg.data_size = data_size
    pg.offset = 0
    pg.next = None
    return pg, buf


@pytest.mark.parametrize("packed,qb,page_bytes,msg", [
    (True, 64, 64, "isPacked: the query host batch holds 64 bytes of packed words (128 bases) but "
                   "actual_query_batch_bytes is 64"),
    (False, 64, 72, "the query host batch holds 72 bytes, more than actual_query_batch_bytes(64)"),
])
def test_batch_bytes_checked_against_host_pages(packed, qb, page_bytes, msg):
    """With isPacked the byte counts count bases: a caller passing the packed
    byte count is told so before anything reaches the device."""
    L = g.lib()
    L.gasal2_set_error_mode(1)
    try:
        s = g.gasal_gpu_storage_t()
        s.host_max_query_batch_bytes = s.host_max_target_batch_bytes = 4096
        s.host_max_n_alns = 16
        s.is_free = 1
        qp, qbuf = _page(256, page_bytes)
        tp, tbuf = _page(256, 8)
        s.extensible_host_unpacked_query_batch = C.pointer(qp)
        s.extensible_host_unpacked_target_batch = C.pointer(tp)
        p = g.Parameters(algo=g.LOCAL, isPacked=packed)
        before = L.gasal2_error_count()
        L.gasal_aln_async(C.byref(s), qb, 64, 1, p.native())
        assert L.gasal2_error_count() == before + 1
        assert msg in L.gasal2_last_error().decode()
        assert s.is_free == 1
    finally:
        L.gasal2_set_error_mode(0)


@pytest.mark.parametrize("off,ln,msg", [(4, 8, "pair 0: query offset 4, length 8"),
                                        (56, 9, "pair 0: query offset 56, length 9")])
def test_pair_outside_batch_rejected(off, ln, msg):
    """Every pair's sequence must lie inside the batch at a multiple of 8
    bases; otherwise the call fails before any device work."""
    L = g.lib()
    L.gasal2_set_error_mode(1)
    try:
        s = g.gasal_gpu_storage_t()
        s.host_max_query_batch_bytes = s.host_max_target_batch_bytes = 4096
        s.host_max_n_alns = 16
        s.is_free = 1
        arrs = [(C.c_uint32 * 1)(v) for v in (off, 0, ln, 8)]
        s.host_query_batch_offsets = C.cast(arrs[0], C.POINTER(C.c_uint32))
        s.host_target_batch_offsets = C.cast(arrs[1], C.POINTER(C.c_uint32))
        s.host_query_batch_lens = C.cast(arrs[2], C.POINTER(C.c_uint32))
        s.host_target_batch_lens = C.cast(arrs[3], C.POINTER(C.c_uint32))
        p = g.Parameters(algo=g.LOCAL)
        before = L.gasal2_error_count()
        L.gasal_aln_async(C.byref(s), 64, 64, 1, p.native())
        assert L.gasal2_error_count() == before + 1
        assert msg in L.gasal2_last_error().decode()
    finally:
        L.gasal2_set_error_mode(0)


def test_host_packing_matches_the_pack_kernel_oracle():
    """pack_bases (what an isPacked=true caller hands over) is the oracle's
    pack of the same bytes (src/kernels/pack_rc_seqs.h:24-31), for bytes and
    arrays, every byte value."""
    import pyoracle as O
    import numpy as np
    data = np.arange(256, dtype=np.uint8).repeat(3)[:760]
    want = O.pack(data).view(np.uint8)
    assert bytes(g.pack_bases(data)) == bytes(want)
    assert bytes(g.pack_bases(data.tobytes())) == bytes(want)
    with pytest.raises(ValueError):
        g.pack_bases(b"ACGT")
