"""gasal2_amd -- MI355X-native batched pairwise alignment with GASAL2's API.

Python mirror of the reference's C/C++ interface (nahmedraja/GASAL2
src/gasal.h, src/args_parser.h, src/gasal_align.h, src/ctors.h,
src/host_batch.h, src/interfaces.h) over the C-ABI of ``libgasal2_amd.so``:
the same function names, argument meanings, enum values and struct fields.
The device work is hand-written gfx950 HIP in that library; nothing here
computes alignments.  If the library is missing, every entry point raises.
"""
import ctypes as C
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GASAL2_AMD_LIB selects another build of the library (e.g. one compiled with
# different kernel options); the in-tree build is the default
LIB_PATH = os.environ.get("GASAL2_AMD_LIB") or os.path.join(_HERE, "libgasal2_amd.so")
SYNTH_PATH = os.path.join(_HERE, "libgasal2_synth.so")


# ----------------------------------------------------------------- enums ---
class comp_start(enum.IntEnum):          # src/gasal.h:37-41
    WITHOUT_START = 0
    WITH_START = 1
    WITH_TB = 2


class Bool(enum.IntEnum):                # src/gasal.h:45-48
    FALSE = 0
    TRUE = 1


class data_source(enum.IntEnum):         # src/gasal.h:50-55
    NONE = 0
    QUERY = 1
    TARGET = 2
    BOTH = 3


class algo_type(enum.IntEnum):           # src/gasal.h:57-65
    UNKNOWN = 0
    GLOBAL = 1
    SEMI_GLOBAL = 2
    LOCAL = 3
    MICROLOCAL = 4
    BANDED = 5
    KSW = 6


class operation_on_seq(enum.IntEnum):    # src/gasal.h:67-72
    FORWARD_NATURAL = 0
    REVERSE_NATURAL = 1
    FORWARD_COMPLEMENT = 2
    REVERSE_COMPLEMENT = 3


globals().update({m.name: m for e in (comp_start, Bool, data_source, algo_type, operation_on_seq) for m in e})

N_CODE_DEFAULT = 0x4E
CIGAR_HARNESS, CIGAR_SAM = 0, 1          # include/gasal2_sam.h
SAM_EQX, SAM_NO_SEQ, SAM_HEADER = 1, 2, 4


# --------------------------------------------------------------- structs ---
class host_batch_t(C.Structure):          # src/gasal.h:75-83
    pass


host_batch_t._fields_ = [("data", C.POINTER(C.c_uint8)), ("page_size", C.c_uint32), ("data_size", C.c_uint32),
                         ("offset", C.c_uint32), ("is_locked", C.c_int), ("next", C.POINTER(host_batch_t))]


class gasal_res_t(C.Structure):           # src/gasal.h:86-95
    _fields_ = [("aln_score", C.POINTER(C.c_int32)), ("query_batch_end", C.POINTER(C.c_int32)),
                ("target_batch_end", C.POINTER(C.c_int32)), ("query_batch_start", C.POINTER(C.c_int32)),
                ("target_batch_start", C.POINTER(C.c_int32)), ("cigar", C.POINTER(C.c_uint8)),
                ("n_cigar_ops", C.POINTER(C.c_uint32))]


_P8, _P32 = C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)


class gasal_gpu_storage_t(C.Structure):   # src/gasal.h:98-150 (+ impl)
    _fields_ = [("unpacked_query_batch", _P8), ("unpacked_target_batch", _P8),
                ("packed_query_batch", _P32), ("packed_target_batch", _P32),
                ("query_batch_offsets", _P32), ("target_batch_offsets", _P32),
                ("query_batch_lens", _P32), ("target_batch_lens", _P32),
                ("host_seed_scores", _P32), ("seed_scores", _P32),
                ("extensible_host_unpacked_query_batch", C.POINTER(host_batch_t)),
                ("extensible_host_unpacked_target_batch", C.POINTER(host_batch_t)),
                ("host_query_op", _P8), ("host_target_op", _P8), ("query_op", _P8), ("target_op", _P8),
                ("host_query_batch_offsets", _P32), ("host_target_batch_offsets", _P32),
                ("host_query_batch_lens", _P32), ("host_target_batch_lens", _P32),
                ("host_res", C.POINTER(gasal_res_t)), ("device_cpy", C.POINTER(gasal_res_t)),
                ("device_res", C.POINTER(gasal_res_t)), ("host_res_second", C.POINTER(gasal_res_t)),
                ("device_res_second", C.POINTER(gasal_res_t)), ("device_cpy_second", C.POINTER(gasal_res_t)),
                ("gpu_max_query_batch_bytes", C.c_uint32), ("gpu_max_target_batch_bytes", C.c_uint32),
                ("host_max_query_batch_bytes", C.c_uint32), ("host_max_target_batch_bytes", C.c_uint32),
                ("gpu_max_n_alns", C.c_uint32), ("host_max_n_alns", C.c_uint32), ("current_n_alns", C.c_uint32),
                ("packed_tb_matrix_size", C.c_uint64), ("packed_tb_matrices", C.c_void_p),
                ("str", C.c_void_p), ("is_free", C.c_int), ("id", C.c_int), ("impl", C.c_void_p)]


class gasal_gpu_storage_v(C.Structure):   # src/gasal.h:153-156
    _fields_ = [("n", C.c_int), ("a", C.POINTER(gasal_gpu_storage_t))]


class gasal_subst_scores(C.Structure):    # src/gasal.h:160-165
    _fields_ = [("match", C.c_int32), ("mismatch", C.c_int32), ("gap_open", C.c_int32), ("gap_extend", C.c_int32)]


class gasal2_pairs_t(C.Structure):        # include/gasal2_stream.h
    _fields_ = [("query", _P8), ("target", _P8), ("query_bytes", C.c_uint64), ("target_bytes", C.c_uint64),
                ("query_words", _P32), ("target_words", _P32),
                ("query_offsets", _P32), ("target_offsets", _P32), ("query_lens", _P32), ("target_lens", _P32),
                ("query_ops", _P8), ("target_ops", _P8),
                ("query_names", C.POINTER(C.c_char_p)), ("target_names", C.POINTER(C.c_char_p)), ("n", C.c_uint32),
                ("seed_scores", _P32)]


class gasal2_seed_t(C.Structure):         # include/gasal2_stream.h
    _fields_ = [("query_beg", C.c_uint32), ("target_beg", C.c_uint32), ("len", C.c_uint32), ("score", C.c_int32)]


class gasal2_extend_opts_t(C.Structure):  # include/gasal2_stream.h
    _fields_ = [(k, C.c_int32) for k in ("w", "zdrop", "end_bonus_left", "end_bonus_right")]


class gasal2_extend_out_t(C.Structure):   # include/gasal2_stream.h
    _fields_ = [(k, C.POINTER(C.c_int32)) for k in ("score", "query_beg", "query_end", "target_beg", "target_end",
                                                      "left_score", "right_score")]


class gasal2_results_t(C.Structure):      # include/gasal2_stream.h
    _fields_ = [(k, C.POINTER(C.c_int32)) for k in ("score", "query_end", "target_end", "query_start", "target_start",
                                                      "second_score", "second_query_end", "second_target_end")] + \
               [("n_cigar_ops", _P32), ("cigar", _P8), ("cigar_cap", C.c_uint32)]


class gasal2_stream_stats_t(C.Structure):  # include/gasal2_stream.h
    _fields_ = [("seconds", C.c_double), ("seconds_from_fill", C.c_double), ("sub_batches", C.c_uint32),
                ("cells", C.c_uint64), ("fill_seconds", C.c_double), ("launch_seconds", C.c_double),
                ("wait_seconds", C.c_double), ("copy_seconds", C.c_double)]


class gasal2_sam_in_t(C.Structure):      # include/gasal2_sam.h
    _fields_ = [("qname", C.c_char_p), ("rname", C.c_char_p), ("query", _P8), ("target", _P8),
                ("query_len", C.c_uint32), ("target_len", C.c_uint32), ("query_op", C.c_uint8), ("target_op", C.c_uint8),
                ("score", C.c_int32), ("query_end", C.c_int32), ("target_end", C.c_int32),
                ("has_second", C.c_int32), ("second_score", C.c_int32), ("cigar", _P8), ("n_cigar_ops", C.c_uint32)]


class gasal2_params_t(C.Structure):       # include/gasal2_ext.h
    _fields_ = [(n, C.c_int32) for n in ("sa", "sb", "gapo", "gape", "start_pos", "print_out", "n_threads", "k_band",
                                         "secondBest", "isPacked", "isReverseComplement",
                                         "semiglobal_skipping_head", "semiglobal_skipping_tail", "algo")]


# ------------------------------------------------------------ the library ---
class NativeLibraryMissing(RuntimeError):
    pass


_lib = None

# name -> (restype, argtypes); every symbol include/*.h declares.
_SIGS = {
    "gasal_copy_subst_scores": (None, [C.POINTER(gasal_subst_scores)]),
    "gasal_aln_async": (None, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "gasal_is_aln_async_done": (C.c_int, [C.POINTER(gasal_gpu_storage_t)]),
    "gasal_init_gpu_storage_v": (gasal_gpu_storage_v, [C.c_int]),
    "gasal_init_streams": (None, [C.POINTER(gasal_gpu_storage_v), C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "gasal_gpu_mem_alloc": (None, [C.POINTER(gasal_gpu_storage_t), C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "gasal_gpu_mem_free": (None, [C.POINTER(gasal_gpu_storage_t), C.c_void_p]),
    "gasal_destroy_streams": (None, [C.POINTER(gasal_gpu_storage_v), C.c_void_p]),
    "gasal_destroy_gpu_storage_v": (None, [C.POINTER(gasal_gpu_storage_v)]),
    "gasal_host_batch_new": (C.POINTER(host_batch_t), [C.c_uint32, C.c_uint32]),
    "gasal_host_batch_destroy": (None, [C.POINTER(host_batch_t)]),
    "gasal_host_batch_getlast": (C.POINTER(host_batch_t), [C.POINTER(host_batch_t)]),
    "gasal_host_batch_reset": (None, [C.POINTER(gasal_gpu_storage_t)]),
    "gasal_host_batch_fill": (C.c_uint32, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_char_p, C.c_uint32, C.c_int]),
    "gasal_host_batch_add": (C.c_uint32, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_char_p, C.c_uint32, C.c_int]),
    "gasal_host_batch_addbase": (C.c_uint32, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_char, C.c_int]),
    "gasal_host_batch_print": (None, [C.POINTER(host_batch_t)]),
    "gasal_host_batch_printall": (None, [C.POINTER(host_batch_t)]),
    "gasal_host_alns_resize": (None, [C.POINTER(gasal_gpu_storage_t), C.c_int, C.c_void_p]),
    "gasal_op_fill": (None, [C.POINTER(gasal_gpu_storage_t), C.POINTER(C.c_uint8), C.c_uint32, C.c_int]),
    "gasal_set_device": (None, [C.c_int, C.c_bool]),
    "gasal_res_new_host": (C.POINTER(gasal_res_t), [C.c_uint32, C.c_void_p]),
    "gasal_res_new_device": (C.POINTER(gasal_res_t), [C.POINTER(gasal_res_t)]),
    "gasal_res_new_device_cpy": (C.POINTER(gasal_res_t), [C.c_uint32, C.c_void_p]),
    "gasal_res_destroy_host": (None, [C.POINTER(gasal_res_t)]),
    "gasal_res_destroy_device": (None, [C.POINTER(gasal_res_t), C.POINTER(gasal_res_t)]),
    "gasal2_fail": (None, [C.c_int]),
    "gasal2_parameters_new": (C.c_void_p, []),
    "gasal2_parameters_free": (None, [C.c_void_p]),
    "gasal2_parameters_get": (None, [C.c_void_p, C.POINTER(gasal2_params_t)]),
    "gasal2_parameters_set": (None, [C.c_void_p, C.POINTER(gasal2_params_t)]),
    "gasal2_set_n_code": (None, [C.c_int32]),
    "gasal2_get_n_code": (C.c_int32, []),
    "gasal2_set_n_penalty": (None, [C.c_int32, C.c_int32]),
    "gasal2_set_error_mode": (None, [C.c_int32]),
    "gasal2_error_count": (C.c_int32, []),
    "gasal2_last_error": (C.c_char_p, []),
    "gasal2_upload": (None, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "gasal2_run_resident": (None, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "gasal2_download": (None, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_uint32, C.c_void_p]),
    "gasal2_synchronize": (None, [C.POINTER(gasal_gpu_storage_t)]),
    "gasal2_enable_timing": (None, [C.POINTER(gasal_gpu_storage_t), C.c_int32]),
    "gasal2_stage_ms": (C.c_float, [C.POINTER(gasal_gpu_storage_t), C.c_int32]),
    "gasal2_get_cigar": (C.c_uint32, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.POINTER(C.c_uint8), C.c_uint32]),
    "gasal2_set_dispatch": (None, [C.c_int32]),
    "gasal2_get_dispatch": (C.c_int32, []),
    "gasal2_set_wavefront": (None, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "gasal2_set_graph_mode": (None, [C.c_int32]),
    "gasal2_set_batch_order": (None, [C.c_int32]),
    "gasal2_pairs_read": (C.POINTER(gasal2_pairs_t), [C.c_char_p, C.c_char_p, C.c_uint32]),
    "gasal2_pairs_free": (None, [C.POINTER(gasal2_pairs_t)]),
    "gasal2_pairs_align": (C.c_int32, [C.POINTER(gasal2_pairs_t), C.POINTER(gasal2_results_t), C.c_void_p, C.c_uint32,
                                       C.c_int32, C.POINTER(gasal2_stream_stats_t)]),
    "gasal2_stream_new": (C.c_void_p, [C.c_void_p, C.c_uint32, C.c_int32, C.c_uint32, C.c_uint32]),
    "gasal2_stream_free": (None, [C.c_void_p]),
    "gasal2_stream_set_order": (None, [C.c_void_p, C.c_int32]),
    "gasal2_extend_seeds": (C.c_int32, [C.c_void_p, C.POINTER(gasal2_pairs_t), C.POINTER(gasal2_seed_t),
                                        C.POINTER(gasal2_extend_opts_t), C.POINTER(gasal2_extend_out_t),
                                        C.POINTER(gasal2_stream_stats_t)]),
    "gasal2_order_new": (C.c_void_p, [C.c_int32]),
    "gasal2_order_free": (None, [C.c_void_p]),
    "gasal2_storage_set_order": (None, [C.c_void_p, C.c_void_p]),
    "gasal2_stream_align": (C.c_int32, [C.c_void_p, C.POINTER(gasal2_pairs_t), C.POINTER(gasal2_results_t),
                                        C.POINTER(gasal2_stream_stats_t)]),
    "gasal2_cigar_text": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_int32, C.c_char_p, C.c_uint32]),
    "gasal2_sam_record": (C.c_int32, [C.POINTER(gasal2_sam_in_t), C.c_int32, C.c_int32, C.c_uint32, C.c_char_p,
                                      C.c_uint32]),
    "gasal2_sam_write": (C.c_int32, [C.c_char_p, C.POINTER(gasal2_pairs_t), C.POINTER(gasal2_results_t), C.c_int32,
                                     C.c_uint32, C.c_int32]),
    "gasal2_aln_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32] + [C.c_void_p] * 6 +
                           [C.c_uint32] * 3 + [C.c_void_p] * 9 + [C.c_void_p, C.c_void_p]),
    "gasal2_set_tile_budget": (None, [C.c_uint64]),
    "gasal2_set_pk16": (None, [C.c_int32]),
    "gasal2_set_ksw_extension": (None, [C.c_int32, C.c_int32]),
    "gasal2_set_ksw_end_bonus": (None, [C.c_int32]),
    "gasal2_storage_set_ksw": (None, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "gasal2_set_tb_checkpoint": (None, [C.c_int32]),
    "gasal2_set_tb_fused": (None, [C.c_int32]),
    "gasal2_set_panels": (None, [C.c_int32]),
    "gasal2_last_paths": (C.c_int32, [C.POINTER(gasal_gpu_storage_t)]),
    "gasal2_tune_set": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "gasal2_tune_clear": (None, []),
    "gasal2_tune_count": (C.c_int32, []),
    "gasal2_prep_resident": (C.c_int32, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "gasal2_result_views": (None, [C.POINTER(gasal_gpu_storage_t)] + [C.POINTER(C.c_void_p)] * 5 + [C.POINTER(C.c_uint32)]),
    "gasal2_device_views": (None, [C.POINTER(gasal_gpu_storage_t)] + [C.POINTER(C.c_void_p)] * 6),
    "gasal2_run_banded": (C.c_int32, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int32]),
    "gasal2_run_subset": (C.c_int32, [C.POINTER(gasal_gpu_storage_t), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_uint64]),
    # learned band (include/gasal2_band.h): device pointers as void*
    "gasal2_lb_features": (C.c_int32, [C.c_void_p] * 6 + [C.c_uint32] + [C.c_void_p] * 4),
    "gasal2_lb_dense_fwd": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                        C.c_void_p, C.c_int32, C.c_void_p]),
    "gasal2_lb_dense_bwd": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_uint32, C.c_uint32,
                                        C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "gasal2_lb_mse": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "gasal2_lb_adam": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float,
                                   C.c_float, C.c_float, C.c_int32, C.c_void_p, C.c_void_p]),
    "gasal2_lb_to_bf16": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "gasal2_lb_band_from_net": (C.c_int32, [C.c_void_p] * 4 + [C.c_uint32, C.c_float, C.c_float, C.c_float]
                                + [C.c_void_p] * 4),
    "gasal2_lb_band_peak": (C.c_int32, [C.c_void_p] * 5 + [C.c_uint32, C.c_float, C.c_float, C.c_float]
                            + [C.c_void_p] * 4),
    "gasal2_lb_path_diag": (C.c_int32, [C.c_void_p, C.c_uint32] + [C.c_void_p] * 3 + [C.c_uint32] + [C.c_void_p] * 3),
    "gasal2_get_graph_mode": (C.c_int32, []),
    "gasal2_graph_stats": (None, [C.POINTER(gasal_gpu_storage_t), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "gasal2_device_info": (C.c_int32, [C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p, C.c_int32]),
}


def lib():
    """The loaded native library; raises NativeLibraryMissing if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(f"{LIB_PATH} not built: run `python -m gasal2_amd.build`")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_info():
    n_cu, clk = C.c_int32(0), C.c_int32(0)
    name = C.create_string_buffer(256)
    if lib().gasal2_device_info(C.byref(n_cu), C.byref(clk), name, 256) != 0:
        raise RuntimeError("no HIP device")
    return {"n_cu": n_cu.value, "clock_mhz": clk.value / 1000.0, "name": name.value.decode(errors="replace")}


def exported_symbols():
    return sorted(_SIGS)


# ------------------------------------------------------------ Parameters ---
class Parameters:
    """class Parameters (src/args_parser.h:24-68) with the reference defaults
    (src/args_parser.cpp:12-31).  Holds a native Parameters object; field
    changes are pushed to it before every call that takes it."""

    _FIELDS = [f for f, _ in gasal2_params_t._fields_]

    def __init__(self, **kw):
        self.sa, self.sb, self.gapo, self.gape = 1, 4, 6, 1
        self.start_pos = comp_start.WITHOUT_START
        self.print_out, self.n_threads, self.k_band = 0, 1, 0
        self.secondBest = Bool.FALSE
        self.isPacked = False
        self.isReverseComplement = False
        self.semiglobal_skipping_head = data_source.TARGET
        self.semiglobal_skipping_tail = data_source.TARGET
        self.algo = algo_type.UNKNOWN
        for k, v in kw.items():
            if k not in self._FIELDS:
                raise AttributeError(k)
            setattr(self, k, v)
        self._native = None

    def native(self):
        L = lib()
        if self._native is None:
            self._native = L.gasal2_parameters_new()
        pod = gasal2_params_t(**{f: int(getattr(self, f)) for f in self._FIELDS})
        L.gasal2_parameters_set(self._native, C.byref(pod))
        return self._native

    def __del__(self):
        if getattr(self, "_native", None) is not None and _lib is not None:
            _lib.gasal2_parameters_free(self._native)
            self._native = None


# ----------------------------------------------------- reference-named API ---
def gasal_copy_subst_scores(subst):
    lib().gasal_copy_subst_scores(C.byref(subst))


def gasal_init_gpu_storage_v(n_streams):
    return lib().gasal_init_gpu_storage_v(n_streams)


def gasal_init_streams(v, max_query_len, max_target_len, max_n_alns, params):
    lib().gasal_init_streams(C.byref(v), max_query_len, max_target_len, max_n_alns, params.native())


def gasal_destroy_streams(v, params):
    lib().gasal_destroy_streams(C.byref(v), params.native())


def gasal_destroy_gpu_storage_v(v):
    lib().gasal_destroy_gpu_storage_v(C.byref(v))


def gasal_host_batch_fill(storage, idx, data, src):
    return lib().gasal_host_batch_fill(C.byref(storage), idx, bytes(data), len(data), int(src))


def gasal_op_fill(storage, ops, src):
    arr = (C.c_uint8 * len(ops))(*[int(o) for o in ops])
    lib().gasal_op_fill(C.byref(storage), arr, len(ops), int(src))


def gasal_host_alns_resize(storage, new_max, params):
    lib().gasal_host_alns_resize(C.byref(storage), new_max, params.native())


def gasal_aln_async(storage, q_bytes, t_bytes, n_alns, params):
    lib().gasal_aln_async(C.byref(storage), q_bytes, t_bytes, n_alns, params.native())


def gasal_is_aln_async_done(storage):
    return lib().gasal_is_aln_async_done(C.byref(storage))


def gasal_set_device(gpu_select=0, is_printing_prop=True):
    lib().gasal_set_device(gpu_select, is_printing_prop)


def set_n_code(n_code):
    lib().gasal2_set_n_code(n_code)


def set_ksw_extension(w=-1, zdrop=0, end_bonus=5):
    """KSW seed-extension settings: band half-width w (< 0: none), z-drop
    (0: off) and end bonus.  The defaults reproduce the reference's KSW
    kernel: no band, z-drop 0, PEN_CLIP5 = 5
    (src/kernels/ksw_kernel_template.h:20,27-29)."""
    lib().gasal2_set_ksw_extension(int(w), int(zdrop))
    lib().gasal2_set_ksw_end_bonus(int(end_bonus))


def set_n_penalty(penalty):
    """penalty None == N_PENALTY undefined (src/gasal_kernels.h:47-55)."""
    lib().gasal2_set_n_penalty(0 if penalty is None else 1, 0 if penalty is None else int(penalty))


# ----------------------------------------------------------------- batches ---
class Batch:
    """Host-side pair batch laid out as a GASAL2 host batch: each sequence
    padded with N_CODE to a multiple of 8 (src/host_batch.cpp:100-102)."""

    def __init__(self, q, t, qoff, toff, qlen, tlen, qop=None, top=None, seeds=None):
        n = len(qlen)
        self.q = np.ascontiguousarray(q, dtype=np.uint8)
        self.t = np.ascontiguousarray(t, dtype=np.uint8)
        self.qoff = np.ascontiguousarray(qoff, dtype=np.uint32)
        self.toff = np.ascontiguousarray(toff, dtype=np.uint32)
        self.qlen = np.ascontiguousarray(qlen, dtype=np.uint32)
        self.tlen = np.ascontiguousarray(tlen, dtype=np.uint32)
        self.qop = np.zeros(n, np.uint8) if qop is None else np.ascontiguousarray(qop, dtype=np.uint8)
        self.top = np.zeros(n, np.uint8) if top is None else np.ascontiguousarray(top, dtype=np.uint8)
        self.seeds = np.zeros(n, np.uint32) if seeds is None else np.ascontiguousarray(seeds, dtype=np.uint32)

    @property
    def n(self):
        return len(self.qlen)

    def packed_words(self, query=True):
        """The query (or target) batch as 4-bit packed words (pack_bases),
        computed once per batch: what an isPacked=true caller holds."""
        key = "_pq" if query else "_pt"
        if getattr(self, key, None) is None:
            setattr(self, key, pack_bases(self.q if query else self.t))
        return getattr(self, key)

    def subset(self, idx):
        """A new batch of the selected pairs, re-laid-out contiguously."""
        qs = [self.query(i) for i in idx]
        ts = [self.target(i) for i in idx]
        return make_batch(qs, ts, qop=self.qop[idx], top=self.top[idx], seeds=self.seeds[idx])

    def query(self, i):
        return bytes(self.q[self.qoff[i]:self.qoff[i] + self.qlen[i]])

    def target(self, i):
        return bytes(self.t[self.toff[i]:self.toff[i] + self.tlen[i]])


def pack_bases(buf):
    """Host form of the 4-bit packed words an isPacked=true caller supplies:
    nibble = byte & 0xF, 8 bases per uint32, first base in bits 31:28
    (src/kernels/pack_rc_seqs.h:24-31).  `buf` holds whole 8-base words
    (sequences padded with N_CODE); returns their bytes as stored in memory."""
    b = np.frombuffer(buf, np.uint8) if isinstance(buf, (bytes, bytearray)) else np.ascontiguousarray(buf, np.uint8)
    if len(b) % 8:
        raise ValueError("packing needs whole 8-base words")
    nib = (b.reshape(-1, 8) & 0xF).astype(np.uint32)
    w = np.zeros(len(nib), np.uint32)
    for i in range(8):
        w |= nib[:, i] << np.uint32(28 - 4 * i)
    return w.view(np.uint8)


def make_batch(queries, targets, qop=None, top=None, seeds=None, n_code=N_CODE_DEFAULT):
    def lay(seqs):
        offs, lens, chunks, o = [], [], [], 0
        for s in seqs:
            s = bytes(s)
            pad = (-len(s)) % 8
            offs.append(o)
            lens.append(len(s))
            chunks.append(s + bytes([n_code]) * pad)
            o += len(s) + pad
        buf = np.frombuffer(b"".join(chunks), dtype=np.uint8).copy() if chunks else np.zeros(0, np.uint8)
        return buf, offs, lens

    q, qoff, qlen = lay(queries)
    t, toff, tlen = lay(targets)
    return Batch(q, t, qoff, toff, qlen, tlen, qop, top, seeds)


_synth = None


def synth_batch(cfg, seed, n, n_code=N_CODE_DEFAULT):
    """Seeded synthetic workload `cfg` (1..5, see csrc/synth.cpp)."""
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise NativeLibraryMissing(f"{SYNTH_PATH} not built: run `python -m gasal2_amd.build`")
        S = C.CDLL(SYNTH_PATH)
        S.synth_generate.restype = C.c_void_p
        S.synth_generate.argtypes = [C.c_int, C.c_uint64, C.c_uint32]
        S.synth_query_bytes.restype = C.c_uint64
        S.synth_query_bytes.argtypes = [C.c_void_p]
        S.synth_target_bytes.restype = C.c_uint64
        S.synth_target_bytes.argtypes = [C.c_void_p]
        S.synth_export.argtypes = [C.c_void_p, C.c_uint8] + [C.c_void_p] * 9
        S.synth_free.argtypes = [C.c_void_p]
        _synth = S
    h = _synth.synth_generate(cfg, seed, n)
    try:
        qb, tb = _synth.synth_query_bytes(h), _synth.synth_target_bytes(h)
        q, t = np.empty(qb, np.uint8), np.empty(tb, np.uint8)
        qoff, toff, qlen, tlen, seeds = (np.empty(n, np.uint32) for _ in range(5))
        qop, top = np.empty(n, np.uint8), np.empty(n, np.uint8)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        _synth.synth_export(h, n_code, ptr(q), ptr(t), ptr(qoff), ptr(toff), ptr(qlen), ptr(tlen), ptr(qop),
                            ptr(top), ptr(seeds))
    finally:
        _synth.synth_free(h)
    return Batch(q, t, qoff, toff, qlen, tlen, qop, top, seeds)


# ------------------------------------------------------------ one-shot run ---
class Results(dict):
    pass


def _arr(ptr, n, dtype):
    """A copy of n elements of a host result array (one memcpy)."""
    if not ptr:
        return None
    addr = C.cast(ptr, C.c_void_p).value
    raw = (C.c_uint8 * (n * C.sizeof(ptr._type_))).from_address(addr)
    return np.frombuffer(raw, dtype=ptr._type_).astype(dtype, copy=True)


class Aligner:
    """Owns one storage (one HIP stream) sized for a batch and runs batches
    through the reference call sequence: gasal_host_batch_fill per sequence (or
    one bulk page copy), gasal_op_fill, gasal_aln_async, poll
    gasal_is_aln_async_done (src/gasal_align.cu:29-326)."""

    def __init__(self, params, max_query_len, max_target_len, max_n_alns, scores=(1, 4, 6, 1),
                 n_code=N_CODE_DEFAULT, n_penalty=None):
        L = lib()
        self.params = params
        self.n_code = n_code
        set_n_code(n_code)
        set_n_penalty(n_penalty)
        gasal_copy_subst_scores(gasal_subst_scores(*scores))
        self.v = gasal_init_gpu_storage_v(1)
        # packed per-sequence fills round each sequence to 16 bases (two words)
        slack = 15 if params.isPacked else 7
        gasal_init_streams(self.v, max_query_len + slack, max_target_len + slack, max_n_alns, params)
        self.s = self.v.a[0]
        self._L = L

    def close(self):
        if self.v is not None:
            gasal_destroy_streams(self.v, self.params)
            gasal_destroy_gpu_storage_v(self.v)
            self.v = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fill(self, batch, per_pair):
        s = self.s
        n = batch.n
        if n > s.host_max_n_alns:
            gasal_host_alns_resize(s, n, self.params)
        # a new batch starts on empty pages (gasal_is_aln_async_done resets them
        # after gasal_aln_async; the resident path never calls it)
        self._L.gasal_host_batch_reset(C.byref(s))
        if self.params.isPacked:
            return self._fill_packed(batch, per_pair)
        if per_pair:
            qi = ti = 0
            for i in range(n):
                s.host_query_batch_offsets[i] = qi
                s.host_target_batch_offsets[i] = ti
                qi = gasal_host_batch_fill(s, qi, batch.query(i), data_source.QUERY)
                ti = gasal_host_batch_fill(s, ti, batch.target(i), data_source.TARGET)
                s.host_query_batch_lens[i] = int(batch.qlen[i])
                s.host_target_batch_lens[i] = int(batch.tlen[i])
            return qi, ti
        # bulk: the batch is already padded; one copy into the first page
        for src, buf, offs, lens, dst_off, dst_len in (
                (data_source.QUERY, batch.q, batch.qoff, batch.qlen, s.host_query_batch_offsets, s.host_query_batch_lens),
                (data_source.TARGET, batch.t, batch.toff, batch.tlen, s.host_target_batch_offsets, s.host_target_batch_lens)):
            page = (s.extensible_host_unpacked_query_batch if src == data_source.QUERY
                    else s.extensible_host_unpacked_target_batch).contents
            if page.page_size < len(buf):
                raise ValueError("batch larger than the storage's first host page; use per_pair=True")
            C.memmove(page.data, buf.ctypes.data, len(buf))
            page.data_size = len(buf)
            C.memmove(dst_off, offs.ctypes.data, 4 * n)
            C.memmove(dst_len, lens.ctypes.data, 4 * n)
        return len(batch.q), len(batch.t)

    def _fill_packed(self, batch, per_pair):
        """isPacked=true: the host pages carry packed words (4 bytes per 8
        bases); offsets, lengths and the returned byte counts are in bases."""
        s = self.s
        n = batch.n
        nc = self.n_code
        if per_pair:
            # each sequence packed on its own; gasal_host_batch_fill pads the
            # packed bytes to a multiple of 8 (a word of N_CODE bytes after
            # odd-word sequences), so every offset stays a multiple of 8 bases
            qi = ti = 0
            for i in range(n):
                s.host_query_batch_offsets[i] = 2 * qi
                s.host_target_batch_offsets[i] = 2 * ti
                for src, seq in ((data_source.QUERY, batch.query(i)), (data_source.TARGET, batch.target(i))):
                    words = pack_bases(seq + bytes([nc]) * ((-len(seq)) % 8)).tobytes()
                    if src == data_source.QUERY:
                        qi = gasal_host_batch_fill(s, qi, words, src)
                    else:
                        ti = gasal_host_batch_fill(s, ti, words, src)
                s.host_query_batch_lens[i] = int(batch.qlen[i])
                s.host_target_batch_lens[i] = int(batch.tlen[i])
            return 2 * qi, 2 * ti
        for src, buf, offs, lens, dst_off, dst_len in (
                (data_source.QUERY, batch.q, batch.qoff, batch.qlen, s.host_query_batch_offsets, s.host_query_batch_lens),
                (data_source.TARGET, batch.t, batch.toff, batch.tlen, s.host_target_batch_offsets, s.host_target_batch_lens)):
            page = (s.extensible_host_unpacked_query_batch if src == data_source.QUERY
                    else s.extensible_host_unpacked_target_batch).contents
            words = batch.packed_words(src == data_source.QUERY)
            if page.page_size < len(words):
                raise ValueError("batch larger than the storage's first host page; use per_pair=True")
            C.memmove(page.data, words.ctypes.data, len(words))
            page.data_size = len(words)
            C.memmove(dst_off, offs.ctypes.data, 4 * n)
            C.memmove(dst_len, lens.ctypes.data, 4 * n)
        return len(batch.q), len(batch.t)

    def fill(self, batch, per_pair=None):
        """Host half of a submission: fills the storage's host batch, lengths,
        offsets and per-pair ops (what a caller does with
        gasal_host_batch_fill before gasal_aln_async)."""
        s, p = self.s, self.params
        n = batch.n
        if per_pair is None:
            per_pair = n <= 2000
        qb, tb = self._fill(batch, per_pair)
        s.current_n_alns = n
        if p.isReverseComplement:
            C.memmove(s.host_query_op, batch.qop.ctypes.data, n)
            C.memmove(s.host_target_op, batch.top.ctypes.data, n)
        if p.algo == algo_type.KSW:
            C.memmove(s.host_seed_scores, batch.seeds.ctypes.data, 4 * n)
        return qb, tb, n

    def align_filled(self, qb, tb, n, timing=False, poll_s=0.0005):
        """gasal_aln_async on a filled host batch, then gasal_is_aln_async_done
        until the results are on the host."""
        s, p, L = self.s, self.params, self._L
        if timing:
            L.gasal2_enable_timing(C.byref(s), 1)
        gasal_aln_async(s, qb, tb, n, p)
        import time
        while True:
            r = gasal_is_aln_async_done(s)
            if r == 0:
                break
            if r != -1:
                raise RuntimeError(f"gasal_is_aln_async_done returned {r}")
            if poll_s:
                time.sleep(poll_s)

    def run(self, batch, per_pair=None, timing=False):
        qb, tb, n = self.fill(batch, per_pair)
        self.align_filled(qb, tb, n, timing)
        return self._collect(n, qb)

    # ---- resident-input form (gasal2_upload / gasal2_run_resident / gasal2_download)
    def load(self, batch, per_pair=None):
        """Fills the host batch and copies it to device memory once."""
        s, p, L = self.s, self.params, self._L
        n = batch.n
        qb, tb = self._fill(batch, per_pair if per_pair is not None else n <= 2000)
        s.current_n_alns = n
        if p.isReverseComplement:
            C.memmove(s.host_query_op, batch.qop.ctypes.data, n)
            C.memmove(s.host_target_op, batch.top.ctypes.data, n)
        if p.algo == algo_type.KSW:
            C.memmove(s.host_seed_scores, batch.seeds.ctypes.data, 4 * n)
        L.gasal2_upload(C.byref(s), qb, tb, n, p.native())
        self._loaded = (qb, tb, n)
        return self._loaded

    def run_resident(self, timing=False):
        qb, tb, n = self._loaded
        if timing:
            self._L.gasal2_enable_timing(C.byref(self.s), 1)
        self._L.gasal2_run_resident(C.byref(self.s), qb, tb, n, self.params.native())

    def sync(self):
        self._L.gasal2_synchronize(C.byref(self.s))

    def download(self):
        qb, tb, n = self._loaded
        self._L.gasal2_download(C.byref(self.s), qb, n, self.params.native())
        return self._collect(n, qb)

    # ---- learned-band path (gasal2_prep_resident / gasal2_run_banded / gasal2_run_subset)
    def prep(self):
        """Packs and reverse-complements the loaded batch on the device."""
        qb, tb, n = self._loaded
        r = self._L.gasal2_prep_resident(C.byref(self.s), qb, tb, n, self.params.native())
        if r != 0:
            raise RuntimeError(f"gasal2_prep_resident returned {r}")
        self.sync()

    def device_views(self):
        """Device pointers (ints) of packed query/target words, offsets, lengths."""
        ptrs = [C.c_void_p() for _ in range(6)]
        self._L.gasal2_device_views(C.byref(self.s), *[C.byref(p) for p in ptrs])
        return [p.value for p in ptrs]

    def result_views(self):
        """Device pointers (ints) of the last run's score, qend, tend, n_ops,
        CIGAR slots, and the slot size."""
        ptrs = [C.c_void_p() for _ in range(5)]
        cap = C.c_uint32()
        self._L.gasal2_result_views(C.byref(self.s), *[C.byref(p) for p in ptrs], C.byref(cap))
        return [p.value for p in ptrs] + [cap.value]

    def run_banded(self, dlo, dhi, check_widen=128):
        """Banded LOCAL inside per-pair bands; returns the per-pair flags."""
        qb, tb, n = self._loaded
        lo = np.ascontiguousarray(dlo, dtype=np.int32)
        hi = np.ascontiguousarray(dhi, dtype=np.int32)
        flags = np.zeros(n, np.uint8)
        r = self._L.gasal2_run_banded(C.byref(self.s), qb, tb, n, self.params.native(), lo.ctypes.data,
                                      hi.ctypes.data, flags.ctypes.data, int(check_widen))
        if r < 0:
            raise RuntimeError(f"gasal2_run_banded returned {r}")
        return flags

    def run_subset(self, idx, max_tb_bytes=0):
        """Full DP over pairs idx of the last run, results replacing theirs."""
        qb, tb, n = self._loaded
        ix = np.ascontiguousarray(idx, dtype=np.uint32)
        if len(ix) == 0:
            return
        r = self._L.gasal2_run_subset(C.byref(self.s), qb, tb, n, self.params.native(), ix.ctypes.data, len(ix),
                                      int(max_tb_bytes))
        if r != 0:
            raise RuntimeError(f"gasal2_run_subset returned {r}")

    def _collect(self, n, qb):
        s, p = self.s, self.params
        h = s.host_res.contents
        out = Results()
        out["score"] = _arr(h.aln_score, n, np.int32)
        out["qend"] = _arr(h.query_batch_end, n, np.int32)
        out["tend"] = _arr(h.target_batch_end, n, np.int32)
        out["qstart"] = _arr(h.query_batch_start, n, np.int32)
        out["tstart"] = _arr(h.target_batch_start, n, np.int32)
        if p.secondBest and s.host_res_second:
            h2 = s.host_res_second.contents
            out["s_score"] = _arr(h2.aln_score, n, np.int32)
            out["s_qend"] = _arr(h2.query_batch_end, n, np.int32)
            out["s_tend"] = _arr(h2.target_batch_end, n, np.int32)
        if p.start_pos == comp_start.WITH_TB:
            out["n_ops"] = _arr(h.n_cigar_ops, n, np.uint32)
            out["cigar_qlayout"] = _arr(h.cigar, qb, np.uint8)
        return out

    def cigar(self, k, cap=1 << 20):
        buf = (C.c_uint8 * cap)()
        n = self._L.gasal2_get_cigar(C.byref(self.s), k, buf, cap)
        return bytes(buf[:min(n, cap)])

    def stage_ms(self, stage):
        return self._L.gasal2_stage_ms(C.byref(self.s), stage)

    def last_paths(self):
        """gasal2_last_paths: bit 0 packed 16-bit lane kernel, bit 1 one-wave
        blocks, bit 2 checkpointed traceback."""
        return int(self._L.gasal2_last_paths(C.byref(self.s)))


class PairSet:
    """Pairs for the multi-storage driver (gasal2_pairs_align,
    include/gasal2_stream.h): read from paired FASTA/FASTQ files by the
    library's reader (gasal2_pairs_read) or wrapping a Batch's arrays."""

    def __init__(self, ptr=None, batch=None, pack=False):
        self._ptr = ptr          # owned by the library (gasal2_pairs_read)
        self._keep = []
        if ptr is not None:
            self.p = ptr.contents
        else:
            P = gasal2_pairs_t()
            keep = self._keep
            def arr(a, dt, ct):
                a = np.ascontiguousarray(a, dtype=dt)
                keep.append(a)
                return a.ctypes.data_as(C.POINTER(ct))
            P.query, P.target = arr(batch.q, np.uint8, C.c_uint8), arr(batch.t, np.uint8, C.c_uint8)
            P.query_bytes, P.target_bytes = len(batch.q), len(batch.t)
            if pack:
                P.query_words = arr(batch.packed_words(True).view(np.uint32), np.uint32, C.c_uint32)
                P.target_words = arr(batch.packed_words(False).view(np.uint32), np.uint32, C.c_uint32)
            P.query_offsets, P.target_offsets = arr(batch.qoff, np.uint32, C.c_uint32), arr(batch.toff, np.uint32, C.c_uint32)
            P.query_lens, P.target_lens = arr(batch.qlen, np.uint32, C.c_uint32), arr(batch.tlen, np.uint32, C.c_uint32)
            P.query_ops, P.target_ops = arr(batch.qop, np.uint8, C.c_uint8), arr(batch.top, np.uint8, C.c_uint8)
            P.seed_scores = arr(batch.seeds, np.uint32, C.c_uint32)
            P.n = batch.n
            self.p = P

    @classmethod
    def read(cls, query_path, target_path, names=False, pack=False):
        flags = (1 if names else 0) | (2 if pack else 0)
        ptr = lib().gasal2_pairs_read(os.fsencode(query_path), os.fsencode(target_path), flags)
        if not ptr:
            raise RuntimeError(f"gasal2_pairs_read failed on {query_path}, {target_path}")
        return cls(ptr=ptr)

    @classmethod
    def from_batch(cls, batch, pack=False):
        return cls(batch=batch, pack=pack)

    @property
    def n(self):
        return int(self.p.n)

    def names(self):
        if not self.p.query_names:
            return None, None
        return ([self.p.query_names[k].decode() for k in range(self.n)],
                [self.p.target_names[k].decode() for k in range(self.n)])

    def to_batch(self):
        """The pairs as a Batch (host arrays copied)."""
        n = self.n
        a = lambda ptr, cnt, dt: np.ctypeslib.as_array(ptr, shape=(cnt,)).astype(dt, copy=True) if cnt else np.zeros(0, dt)
        return Batch(a(self.p.query, self.p.query_bytes, np.uint8), a(self.p.target, self.p.target_bytes, np.uint8),
                     a(self.p.query_offsets, n, np.uint32), a(self.p.target_offsets, n, np.uint32),
                     a(self.p.query_lens, n, np.uint32), a(self.p.target_lens, n, np.uint32),
                     a(self.p.query_ops, n, np.uint8), a(self.p.target_ops, n, np.uint8))

    def align(self, params, batch_pairs=250_000, storages=3, scores=(1, 4, 6, 1), n_code=N_CODE_DEFAULT,
              n_penalty=None, cigar_cap=0, stream=None):
        """gasal2_pairs_align (or gasal2_stream_align on `stream`, a Stream
        whose storages stay allocated across calls); returns (Results, stats)."""
        L = lib()
        set_n_code(n_code)
        set_n_penalty(n_penalty)
        gasal_copy_subst_scores(gasal_subst_scores(*scores))
        n = self.n
        out = Results()
        R = gasal2_results_t()
        keep = []
        def slot(name, dt, ct, field):
            a = np.zeros(n, dt)
            keep.append(a)
            out[name] = a
            setattr(R, field, a.ctypes.data_as(C.POINTER(ct)))
        slot("score", np.int32, C.c_int32, "score")
        if params.algo != algo_type.GLOBAL:
            slot("qend", np.int32, C.c_int32, "query_end")
            slot("tend", np.int32, C.c_int32, "target_end")
            if params.start_pos in (comp_start.WITH_START, comp_start.WITH_TB):
                slot("qstart", np.int32, C.c_int32, "query_start")
                slot("tstart", np.int32, C.c_int32, "target_start")
        if params.secondBest:
            slot("s_score", np.int32, C.c_int32, "second_score")
            slot("s_qend", np.int32, C.c_int32, "second_query_end")
            slot("s_tend", np.int32, C.c_int32, "second_target_end")
        if params.start_pos == comp_start.WITH_TB:
            slot("n_ops", np.uint32, C.c_uint32, "n_cigar_ops")
            if cigar_cap:
                cg = np.zeros(n * cigar_cap, np.uint8)
                keep.append(cg)
                R.cigar = cg.ctypes.data_as(C.POINTER(C.c_uint8))
                R.cigar_cap = cigar_cap
                out["cigar_slots"] = cg.reshape(n, cigar_cap)
        st = gasal2_stream_stats_t()
        if stream is not None:
            r = L.gasal2_stream_align(stream.h, C.byref(self.p), C.byref(R), C.byref(st))
        else:
            r = L.gasal2_pairs_align(C.byref(self.p), C.byref(R), params.native(), batch_pairs, storages, C.byref(st))
        if r != 0:
            raise RuntimeError("gasal2_stream_align failed")
        out.native = (R, keep, bool(params.isReverseComplement))
        return out, {k: getattr(st, k) for k, _ in gasal2_stream_stats_t._fields_}

    def write_sam(self, path, results, flags=SAM_HEADER, threads=0):
        """gasal2_sam_write: the SAM records of `results` (what align()
        returned, with cigar_cap > 0 for CIGARs) into `path`."""
        R, _, rc = results.native
        if lib().gasal2_sam_write(os.fsencode(str(path)), C.byref(self.p), C.byref(R), int(rc), flags, threads) != 0:
            raise RuntimeError(f"gasal2_sam_write failed on {path}")

    def extend(self, stream, query_beg, target_beg, length, score, w=-1, zdrop=0, end_bonus=(5, 5), scores=(1, 4, 6, 1),
               n_code=N_CODE_DEFAULT, n_penalty=None):
        """gasal2_extend_seeds: two-sided extension of one seed per pair on
        `stream` (a Stream made with algo = KSW); returns (dict, stats)."""
        set_n_code(n_code)
        set_n_penalty(n_penalty)
        gasal_copy_subst_scores(gasal_subst_scores(*scores))
        n = self.n
        seeds = (gasal2_seed_t * max(n, 1))()
        for k, (a, b, l, sc) in enumerate(zip(query_beg, target_beg, length, score)):
            seeds[k] = gasal2_seed_t(int(a), int(b), int(l), int(sc))
        out, keep = {}, []
        O = gasal2_extend_out_t()
        for name in ("score", "query_beg", "query_end", "target_beg", "target_end", "left_score", "right_score"):
            a = np.zeros(n, np.int32)
            out[name] = a
            keep.append(a)
            setattr(O, name, a.ctypes.data_as(C.POINTER(C.c_int32)))
        opts = gasal2_extend_opts_t(int(w), int(zdrop), int(end_bonus[0]), int(end_bonus[1]))
        st = gasal2_stream_stats_t()
        if lib().gasal2_extend_seeds(stream.h, C.byref(self.p), seeds, C.byref(opts), C.byref(O), C.byref(st)) != 0:
            raise RuntimeError("gasal2_extend_seeds failed")
        return out, {k: getattr(st, k) for k, _ in gasal2_stream_stats_t._fields_}

    def close(self):
        if self._ptr is not None:
            lib().gasal2_pairs_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    """gasal2_stream_new: storages (one HIP stream and host thread each) kept
    across PairSet.align(..., stream=this) calls, so their workspace is
    allocated once."""

    def __init__(self, params, batch_pairs, storages, max_query_len, max_target_len):
        self.params = params
        self.h = lib().gasal2_stream_new(params.native(), batch_pairs, storages, max_query_len, max_target_len)
        if not self.h:
            raise RuntimeError("gasal2_stream_new failed")

    def set_order(self, mode):
        """gasal2_stream_set_order: 2 (default) sub-batches' H2D and device
        programs in submission order, 1 H2D only, 0 independent streams."""
        lib().gasal2_stream_set_order(self.h, mode)

    def close(self):
        if self.h:
            lib().gasal2_stream_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cigar_text(cg, style=CIGAR_SAM):
    """gasal2_cigar_text: forward CIGAR text of the reverse-order CIGAR bytes."""
    cg = bytes(cg)
    n = lib().gasal2_cigar_text(cg, len(cg), style, None, 0)
    if n < 0:
        raise ValueError(f"unknown CIGAR style {style}")
    buf = C.create_string_buffer(n + 1)
    lib().gasal2_cigar_text(cg, len(cg), style, buf, n + 1)
    return buf.value.decode()


def sam_record(query, target, score, query_end=-1, target_end=-1, cigar=b"", qname="q", rname="t", query_op=0,
               target_op=0, second_score=None, rc_enabled=False, n_code=N_CODE_DEFAULT, flags=0):
    """gasal2_sam_record: one SAM line (with its newline) for one pair."""
    q, t, cg = bytes(query), bytes(target), bytes(cigar)
    I = gasal2_sam_in_t()
    I.qname, I.rname = (qname.encode() if qname is not None else None), (rname.encode() if rname is not None else None)
    I.query = C.cast(C.c_char_p(q), _P8)
    I.target = C.cast(C.c_char_p(t), _P8)
    I.query_len, I.target_len = len(q), len(t)
    I.query_op, I.target_op = query_op, target_op
    I.score, I.query_end, I.target_end = score, query_end, target_end
    I.has_second, I.second_score = (0, 0) if second_score is None else (1, second_score)
    I.cigar = C.cast(C.c_char_p(cg), _P8) if cg else None
    I.n_cigar_ops = len(cg)
    L = lib()
    n = L.gasal2_sam_record(C.byref(I), int(rc_enabled), n_code, flags, None, 0)
    if n < 0:
        raise ValueError("gasal2_sam_record rejected its input")
    buf = C.create_string_buffer(n + 1)
    L.gasal2_sam_record(C.byref(I), int(rc_enabled), n_code, flags, buf, n + 1)
    return buf.value.decode()


def cigar_to_string(cg):
    """Forward CIGAR text with equal adjacent runs merged, as the reference
    harness prints it (test_prog/test_prog.cpp:382-428)."""
    if not cg:
        return ""
    ops = "MXDI"
    out = []
    last = cg[-1] & 3
    count = cg[-1] >> 2
    for b in reversed(cg[:-1]):
        op = b & 3
        if op == last:
            count += b >> 2
        else:
            out.append(f"{count}{ops[last]}")
            count = b >> 2
        last = op
    out.append(f"{count}{ops[last]}")
    return "".join(out)
