"""ctypes mirror of include/gicops_amd.h and include/gicops_index.h.

Only plain C types cross the boundary; numpy arrays are passed by pointer.
Loading fails loudly (OSError with the path) when a library has not been
built: there is no fallback implementation behind these bindings.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIBDIR = Path(__file__).resolve().parent / "lib"

# ---- status codes (slmerr.h) ------------------------------------------------
GX_SUCCESS = 0
GX_ERR_BASE = -3000
STATUS_NAMES = {
    0: "SLM_SUCCESS",
    GX_ERR_BASE - 1: "ERR_INVLD_MEMORY",
    GX_ERR_BASE - 2: "ERR_INVLD_PARAM",
    GX_ERR_BASE - 3: "ERR_INVLD_INDEX",
    GX_ERR_BASE - 4: "ERR_INVLD_SIZE",
    GX_ERR_BASE - 5: "ERR_SLM_ALREADY_INIT",
    GX_ERR_BASE - 6: "ERR_FILE_NOT_FOUND",
    GX_ERR_BASE - 7: "ERR_BAD_MEM_ALLOC",
    GX_ERR_BASE - 8: "ERR_INVLD_PTR",
    GX_ERR_BASE - 9: "ERR_INVLD_SORT",
    GX_ERR_BASE - 10: "ERR_INVLD_MOD",
    GX_ERR_BASE - 12: "ERR_NOT_ENOUGH_DATA",
    GX_ERR_BASE - 13: "ERR_FILE_ERROR",
    GX_ERR_BASE - 50: "GX_ERR_DEVICE",
    GX_ERR_BASE - 51: "GX_ERR_NOT_READY",
    GX_ERR_BASE - 52: "GX_ERR_NO_DEVICE",
}
GX_ERR_INVLD_PARAM = GX_ERR_BASE - 2
GX_ERR_INVLD_INDEX = GX_ERR_BASE - 3
GX_ERR_INVLD_SIZE = GX_ERR_BASE - 4
GX_ERR_INVLD_PTR = GX_ERR_BASE - 8
GX_ERR_DEVICE = GX_ERR_BASE - 50
GX_ERR_NOT_READY = GX_ERR_BASE - 51
GX_ERR_NO_DEVICE = GX_ERR_BASE - 52


class GxError(RuntimeError):
    def __init__(self, func: str, code: int):
        super().__init__(f"{func} failed with status {code} ({STATUS_NAMES.get(code, 'unknown')})")
        self.code = code


def check(func: str, code: int) -> None:
    if code != GX_SUCCESS:
        err = GxError(func, code)
        if code in (GX_ERR_DEVICE, GX_ERR_INVLD_PARAM, GX_ERR_NOT_READY, GX_ERR_INVLD_INDEX) and _HIP is not None:
            buf = C.create_string_buffer(512)
            _HIP.gx_last_error(buf, 512)
            err = GxError(f"{func} [{buf.value.decode(errors='replace')}]", code)
        raise err


# ---- structs ----------------------------------------------------------------
class gx_bucket(C.Structure):
    _fields_ = [
        ("peplen", C.c_int32), ("n_peps", C.c_uint32), ("n_entries", C.c_uint32),
        ("speclen", C.c_uint32), ("n_ions", C.c_uint64), ("nbins", C.c_uint32),
        ("seqs", C.c_void_p), ("mass", C.c_void_p), ("seqid", C.c_void_p),
        ("modnum", C.c_void_p), ("sites", C.c_void_p), ("bA", C.c_void_p), ("iA", C.c_void_p),
    ]


class gx_query_batch(C.Structure):
    _fields_ = [
        ("n_spectra", C.c_uint32), ("n_peaks", C.c_uint64), ("idx", C.c_void_p),
        ("moz", C.c_void_p), ("intensity", C.c_void_p), ("precurse", C.c_void_p), ("dm", C.c_void_p),
    ]


class gx_search_params(C.Structure):
    _fields_ = [
        ("dF", C.c_int32), ("scale", C.c_int32), ("max_mass", C.c_int32), ("min_shp", C.c_int32),
        ("min_cpsm", C.c_int32), ("evalue_mode", C.c_int32), ("dM", C.c_double),
        ("expect_max", C.c_double), ("surrogate_conf", C.c_float), ("engine", C.c_int32),
        ("config", C.c_int32), ("use_graph", C.c_int32), ("split", C.c_int32), ("whole_bin", C.c_int32),
        ("wide_stream", C.c_int32),
    ]


class gx_search_stats(C.Structure):
    _fields_ = [
        ("ions", C.c_uint64), ("bins", C.c_uint64), ("window", C.c_uint64), ("candidates", C.c_uint64),
        ("spectra", C.c_uint64), ("peaks", C.c_uint64), ("fallbacks", C.c_uint64),
        ("ions_scanned", C.c_uint64), ("errors", C.c_uint64), ("items", C.c_uint64), ("chunks", C.c_uint32), ("config", C.c_int32),
        ("kernel_ms", C.c_double), ("evalue_ms", C.c_double), ("plan_ms", C.c_double), ("total_ms", C.c_double),
    ]


GX_TUNE_MAX = 16


class gx_tune_result(C.Structure):
    _fields_ = [
        ("config", C.c_int32), ("split", C.c_int32), ("whole_bin", C.c_int32), ("n_tried", C.c_int32),
        ("tried_config", C.c_int32 * GX_TUNE_MAX), ("tried_split", C.c_int32 * GX_TUNE_MAX),
        ("tried_whole_bin", C.c_int32 * GX_TUNE_MAX), ("tried_ms", C.c_double * GX_TUNE_MAX),
    ]


GX_TRAIN_MAX_EPOCHS = 64
MLP_FEAT, MLP_HIDDEN, MLP_OUT_PAD, MLP_OUT, MLP_TARGET = 128, 256, 16, 3, 4


class gx_train_params(C.Structure):
    _fields_ = [
        ("batch", C.c_int32), ("epochs", C.c_int32), ("lr", C.c_float), ("lr_decay", C.c_float),
        ("conf_weight", C.c_float), ("pad", C.c_int32), ("seed", C.c_uint64),
    ]


class gx_train_report(C.Structure):
    _fields_ = [
        ("samples", C.c_uint64), ("epochs", C.c_int32), ("pad", C.c_int32),
        ("mse_mark", C.c_float * GX_TRAIN_MAX_EPOCHS), ("mse_mark2", C.c_float * GX_TRAIN_MAX_EPOCHS),
        ("bce", C.c_float * GX_TRAIN_MAX_EPOCHS), ("in_window", C.c_float * GX_TRAIN_MAX_EPOCHS),
    ]


class gx_mod_spec(C.Structure):
    _fields_ = [("residues", C.c_char * 8), ("mass", C.c_float), ("max_per_pep", C.c_int32)]


class gx_db_params(C.Structure):
    _fields_ = [
        ("min_len", C.c_int32), ("max_len", C.c_int32), ("maxz", C.c_int32), ("scale", C.c_int32),
        ("min_mass", C.c_float), ("max_mass", C.c_float), ("max_mods_per_pep", C.c_int32),
        ("n_mod_types", C.c_int32), ("mods", gx_mod_spec * 4), ("modmass_all", C.c_int32),
        ("threads", C.c_int32),
    ]


GX_MAX_MOD_TYPES = 15


class gx_pick_params(C.Structure):
    _fields_ = [("base_int", C.c_int32), ("min_int", C.c_int32), ("qalen", C.c_int32)]



class gx_index_build_params(C.Structure):
    _fields_ = [
        ("maxz", C.c_int32), ("scale", C.c_int32), ("min_mass", C.c_float), ("max_mass", C.c_float),
        ("n_mod_types", C.c_int32), ("mod_mass", C.c_uint32 * GX_MAX_MOD_TYPES), ("modmass_all", C.c_int32),
    ]


class gx_partial(C.Structure):
    _fields_ = [("key", C.c_uint64), ("eq", C.c_uint32), ("cpsms", C.c_uint32), ("sharedions", C.c_uint16),
                ("totalions", C.c_uint16), ("reserved", C.c_uint32)]


PARTIAL_DTYPE = np.dtype([("key", np.uint64), ("eq", np.uint32), ("cpsms", np.uint32), ("sharedions", np.uint16),
                          ("totalions", np.uint16), ("reserved", np.uint32)])


class gx_spec_params(C.Structure):
    _fields_ = [
        ("n_spectra", C.c_uint32), ("keep_frac", C.c_float), ("mz_sigma", C.c_float),
        ("peak_mode", C.c_int32), ("noise_min", C.c_int32), ("noise_max", C.c_int32),
        ("peaks_min", C.c_int32), ("peaks_max", C.c_int32), ("power_alpha", C.c_float),
        ("noise_mz_lo", C.c_float), ("noise_mz_hi", C.c_float), ("int_sigma", C.c_float),
        ("charge_min", C.c_int32), ("charge_max", C.c_int32), ("qalen", C.c_int32),
        ("base_int", C.c_int32), ("min_int", C.c_int32), ("n_dm", C.c_int32), ("dm_choices", C.c_float * 4),
    ]


class gx_queries(C.Structure):
    _fields_ = [
        ("n_spectra", C.c_uint32), ("n_peaks", C.c_uint64), ("idx", C.POINTER(C.c_uint32)),
        ("moz", C.POINTER(C.c_int32)), ("intensity", C.POINTER(C.c_int32)),
        ("precurse", C.POINTER(C.c_float)), ("charges", C.POINTER(C.c_int32)),
        ("rtimes", C.POINTER(C.c_float)), ("dm", C.POINTER(C.c_float)),
        ("truth_bucket", C.POINTER(C.c_int32)), ("truth_psid", C.POINTER(C.c_int32)),
    ]


# numpy view of gx_psm (40 bytes, natural alignment)
PSM_DTYPE = np.dtype([
    ("hyperscore", np.float32), ("psid", np.int32), ("idxoffset", np.uint16),
    ("sharedions", np.uint16), ("totalions", np.uint16), ("flags", np.uint16),
    ("cpsms", np.int32), ("hyp", np.int32), ("mu", np.int32), ("beta", np.int32),
    ("evalue", np.float64),
])
assert PSM_DTYPE.itemsize == 40

GX_HIST_BINS = 802
GX_PSM_HAS_HIT, GX_PSM_MODELED, GX_PSM_FIT_OK, GX_PSM_REPORTED, GX_PSM_SURROGATE, GX_PSM_TIE = 1, 2, 4, 8, 16, 32
GX_EVALUE_TAILFIT, GX_EVALUE_LEARNED = 0, 1
GX_ENGINE_BATCHED, GX_ENGINE_SPECTRUM = 0, 1
GX_PEAKS_NOISE_UNIFORM, GX_PEAKS_TOTAL_UNIFORM, GX_PEAKS_TOTAL_POWER = 0, 1, 2


def ptr(a: np.ndarray | None) -> C.c_void_p | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the C ABI must be contiguous"
    return C.c_void_p(a.ctypes.data)


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


def u8ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


# ---- library loading --------------------------------------------------------
_HOST = None
_HIP = None


def _load(name: str) -> C.CDLL:
    path = LIBDIR / name
    if not path.exists():
        raise OSError(f"{path} is not built; run `python -m gicops_amd._build` (or __graft_entry__.build())")
    return C.CDLL(str(path))


def host_lib() -> C.CDLL:
    global _HOST
    if _HOST is None:
        lib = _load("libgicops_host.so")
        vp, i32, u64 = C.c_void_p, C.c_int, C.c_uint64
        lib.gx_db_create.argtypes = [C.POINTER(gx_db_params), C.POINTER(vp)]
        lib.gx_db_add_peptides.argtypes = [vp, C.c_char_p, vp, u64, C.POINTER(u64)]
        lib.gx_db_generate_peptides.argtypes = [vp, u64, u64]
        lib.gx_db_finalize.argtypes = [vp]
        lib.gx_db_finalize_entries.argtypes = [vp]
        lib.gx_db_build_params.argtypes = [vp, C.POINTER(gx_index_build_params)]
        lib.gx_db_num_buckets.argtypes = [vp]
        lib.gx_db_get_bucket.argtypes = [vp, i32, C.POINTER(gx_bucket)]
        lib.gx_db_destroy.argtypes = [vp]
        lib.gx_db_destroy.restype = None
        lib.gx_spectra_generate.argtypes = [vp, C.POINTER(gx_spec_params), u64, C.POINTER(C.POINTER(gx_queries))]
        lib.gx_queries_destroy.argtypes = [C.POINTER(gx_queries)]
        lib.gx_queries_destroy.restype = None
        lib.gx_spectra_generate_raw.argtypes = [vp, C.POINTER(gx_spec_params), u64,
                                                C.POINTER(C.POINTER(gx_queries))]
        lib.gx_ms2_read.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.POINTER(gx_queries))]
        lib.gx_queries_pick.argtypes = [C.POINTER(gx_queries), C.POINTER(gx_pick_params),
                                        C.POINTER(C.POINTER(gx_queries))]
        lib.gx_pbin_write.argtypes = [C.c_char_p, C.POINTER(gx_queries)]
        lib.gx_pbin_read.argtypes = [C.c_char_p, C.POINTER(C.POINTER(gx_queries))]
        _HOST = lib
    return _HOST


def hip_lib() -> C.CDLL:
    global _HIP
    if _HIP is None:
        lib = _load("libgicops_hip.so")
        vp = C.c_void_p
        lib.gx_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        lib.gx_ctx_destroy.argtypes = [vp]
        lib.gx_ctx_destroy.restype = None
        lib.gx_index_upload.argtypes = [vp, C.POINTER(gx_bucket), C.c_int, C.c_int, C.POINTER(vp)]
        lib.gx_index_free.argtypes = [vp]
        lib.gx_index_free.restype = None
        lib.gx_index_build.argtypes = [vp, C.POINTER(gx_bucket), C.c_int, C.POINTER(gx_index_build_params),
                                       C.POINTER(vp)]
        lib.gx_index_download.argtypes = [vp, C.c_int, vp, vp]
        lib.gx_queries_upload.argtypes = [vp, C.POINTER(gx_query_batch), C.POINTER(vp)]
        lib.gx_queries_free.argtypes = [vp]
        lib.gx_queries_free.restype = None
        lib.gx_queries_preprocess.argtypes = [vp, C.POINTER(gx_query_batch), C.POINTER(gx_pick_params),
                                              C.POINTER(vp)]
        lib.gx_queries_shape.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_float)]
        lib.gx_queries_download.argtypes = [vp, vp, vp, vp]
        lib.gx_measure_read_bandwidth.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        lib.gx_partials_export.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int]
        lib.gx_partials_merge.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(vp), C.c_uint32,
                                          C.POINTER(gx_search_params), C.c_int, vp, vp, C.POINTER(C.c_float)]
        lib.gx_search_device.argtypes = [vp, vp, vp, C.POINTER(gx_search_params), C.POINTER(gx_search_stats)]
        lib.gx_results_download.argtypes = [vp, vp, vp, vp]
        lib.gx_search_spectra_noalloc.argtypes = [vp, vp, C.POINTER(gx_query_batch), C.POINTER(gx_search_params),
                                                  vp, vp, C.POINTER(gx_search_stats)]
        lib.gx_search.argtypes = [vp, vp, C.POINTER(gx_query_batch), C.POINTER(gx_search_params), vp, vp,
                                  C.POINTER(gx_search_stats)]
        lib.gx_ctx_device_info.argtypes = [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]
        lib.gx_autotune.argtypes = [vp, vp, vp, C.POINTER(gx_search_params), C.POINTER(gx_tune_result)]
        lib.gx_last_error.argtypes = [C.c_char_p, C.c_size_t]
        fp, u8p = C.POINTER(C.c_float), C.POINTER(C.c_uint8)
        lib.gx_mlp_param_count.argtypes = []
        lib.gx_mlp_create.argtypes = [vp, C.POINTER(vp)]
        lib.gx_mlp_destroy.argtypes = [vp]
        lib.gx_mlp_destroy.restype = None
        lib.gx_mlp_init.argtypes = [vp, C.c_uint64]
        lib.gx_mlp_set_params.argtypes = [vp, fp, C.c_size_t, fp]
        lib.gx_mlp_get_params.argtypes = [vp, fp, C.c_size_t, fp]
        lib.gx_mlp_add_training.argtypes = [vp, vp, C.POINTER(C.c_uint64)]
        lib.gx_mlp_train.argtypes = [vp, C.POINTER(gx_train_params), C.POINTER(gx_train_report)]
        lib.gx_mlp_grad.argtypes = [vp, fp, fp, C.c_uint32, fp, C.c_float, fp, fp, fp]
        lib.gx_mlp_predict.argtypes = [vp, fp, C.c_uint32, fp]
        lib.gx_ctx_set_surrogate.argtypes = [vp, vp]
        lib.gx_features_download.argtypes = [vp, vp, fp, fp, fp, u8p, u8p]
        lib.gx_surrogate_outputs_download.argtypes = [vp, vp, fp]
        _HIP = lib
    return _HIP
