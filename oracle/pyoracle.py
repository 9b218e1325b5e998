"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / the reported CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class Scores(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("match", "mismatch", "gap_open", "gap_extend", "n_code", "has_n_penalty",
                                         "n_penalty")]


class Params(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("algo", "start_pos", "second_best", "head", "tail", "k_band", "is_packed",
                                         "is_reverse_complement", "ksw_w", "ksw_zdrop", "ksw_end_bonus")]


class Out(C.Structure):
    _fields_ = [("aln_score", C.c_void_p), ("query_batch_end", C.c_void_p), ("target_batch_end", C.c_void_p),
                ("query_batch_start", C.c_void_p), ("target_batch_start", C.c_void_p),
                ("second_score", C.c_void_p), ("second_query_end", C.c_void_p), ("second_target_end", C.c_void_p),
                ("n_cigar_ops", C.c_void_p), ("cigar", C.c_void_p), ("cigar_offsets", C.c_void_p)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_cigar_capacity.restype = C.c_uint32
        L.oracle_cigar_capacity.argtypes = [C.c_uint32, C.c_uint32]
        L.oracle_pack.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.oracle_rc_one.argtypes = [C.c_void_p, C.c_uint32, C.c_uint8, C.c_int32]
        L.oracle_align_batch.restype = C.c_int
        L.oracle_align_batch.argtypes = [C.POINTER(Scores), C.POINTER(Params), C.c_void_p, C.c_uint32, C.c_void_p,
                                         C.c_uint32] + [C.c_void_p] * 7 + [C.c_uint32, C.POINTER(Out), C.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def pack(data):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.zeros(len(data) // 8, np.uint32)
    lib().oracle_pack(_p(data), len(data), _p(out))
    return out


def rc_one(words, length, op, n_code=0x4E):
    w = np.array(words, dtype=np.uint32)
    lib().oracle_rc_one(_p(w), length, op, n_code)
    return w


def align(batch, algo, start_pos=0, second_best=False, head=2, tail=2, k_band=0, is_packed=False,
          is_rc=False, scores=(1, 4, 6, 1), n_code=0x4E, n_penalty=None, n_threads=0, with_cigar=True,
          ksw_w=-1, ksw_zdrop=0, ksw_end_bonus=5):
    """Runs the oracle on a gasal2_amd.Batch; returns a dict of numpy arrays.
    CIGARs (WITH_TB) are returned as a list of bytes, one per pair."""
    L = lib()
    n = batch.n
    sc = Scores(scores[0], scores[1], scores[2], scores[3], n_code, 0 if n_penalty is None else 1,
                0 if n_penalty is None else int(n_penalty))
    pr = Params(int(algo), int(start_pos), int(bool(second_best)), int(head), int(tail), int(k_band),
                int(bool(is_packed)), int(bool(is_rc)), int(ksw_w), int(ksw_zdrop), int(ksw_end_bonus))
    res = {k: np.zeros(n, np.int32) for k in ("score", "qend", "tend", "qstart", "tstart", "s_score", "s_qend",
                                               "s_tend")}
    res["n_ops"] = np.zeros(n, np.uint32)
    cig = cig_off = None
    if with_cigar and int(start_pos) == 2:
        caps = np.array([L.oracle_cigar_capacity(int(a), int(b)) for a, b in zip(batch.qlen, batch.tlen)],
                        dtype=np.uint64)
        cig_off = np.zeros(n, np.uint64)
        if n > 1:
            cig_off[1:] = np.cumsum(caps)[:-1]
        cig = np.zeros(int(caps.sum()) + 16, np.uint8)
    out = Out(_p(res["score"]), _p(res["qend"]), _p(res["tend"]), _p(res["qstart"]), _p(res["tstart"]),
              _p(res["s_score"]), _p(res["s_qend"]), _p(res["s_tend"]), _p(res["n_ops"]),
              _p(cig) if cig is not None else None, _p(cig_off) if cig_off is not None else None)
    rc = L.oracle_align_batch(C.byref(sc), C.byref(pr), _p(batch.q), len(batch.q), _p(batch.t), len(batch.t),
                              _p(batch.qoff), _p(batch.toff), _p(batch.qlen), _p(batch.tlen), _p(batch.qop),
                              _p(batch.top), _p(batch.seeds), n, C.byref(out), n_threads)
    if rc != 0:
        raise ValueError(f"oracle rejected the batch (code {rc})")
    if cig is not None:
        res["cigars"] = [bytes(cig[int(o):int(o) + int(k)]) for o, k in zip(cig_off, res["n_ops"])]
    return res
