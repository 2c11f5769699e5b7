"""Python face of oracle/liboracle.so - TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
never by the gicops_amd package. It is the checker the HIP path is compared
against, restating the reference CPU search (dslim_query.cpp:749-1242) and
tail fit (expeRT.cpp:396-559).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from gicops_amd import _ffi as F

_LIB = None
ORACLE_DIR = Path(__file__).resolve().parent

COUNTER_DTYPE = np.dtype([("ions", np.uint64), ("bin_ions", np.uint64), ("bins", np.uint32),
                          ("window", np.uint32), ("legal", np.uint32), ("buckets", np.uint32)])


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = ORACLE_DIR / "liboracle.so"
        if not path.exists():
            raise OSError(f"{path} is not built; run `make -C oracle`")
        L = C.CDLL(str(path))
        vp = C.c_void_p
        L.orc_lower_bound.argtypes = [vp, C.c_int64, C.c_uint32]
        L.orc_lower_bound.restype = C.c_int64
        L.orc_upper_bound.argtypes = [vp, C.c_int64, C.c_uint32]
        L.orc_upper_bound.restype = C.c_int64
        L.orc_window.argtypes = [vp, C.c_uint32, C.c_float, C.c_double, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]
        L.orc_lgfact.argtypes = [vp, C.c_int]
        L.orc_lgfact.restype = None
        L.orc_tailfit.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_search.argtypes = [C.POINTER(F.gx_bucket), C.c_int, C.POINTER(F.gx_query_batch),
                                 C.POINTER(F.gx_search_params), vp, vp, vp, C.c_int]
        L.orc_scorecard.argtypes = [C.POINTER(F.gx_bucket), vp, vp, C.c_uint32, C.POINTER(F.gx_search_params),
                                    C.c_int32, C.c_int32, vp, vp, vp, vp]
        L.orc_pickpeaks.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32]
        L.orc_pickpeaks.restype = C.c_int64
        _LIB = L
    return _LIB


def pickpeaks_one(intensity, mz, base_int: int = 100_000, min_int: int = 1000, qalen: int = 100):
    """One raw spectrum -> (sorted intensities, sorted m/z, kept count); the
    kept peaks are the last `kept` entries."""
    intn = np.array(intensity, dtype=np.int32)
    m = np.array(mz, dtype=np.int32)
    k = lib().orc_pickpeaks(intn.ctypes.data if intn.size else None, m.ctypes.data if m.size else None,
                            len(intn), base_int, min_int, qalen)
    return intn, m, int(k)


def pickpeaks(batch, base_int: int = 100_000, min_int: int = 1000, qalen: int = 100):
    """Peak picking of every spectrum of a raw batch -> (idx, moz, intensity)."""
    idx = [0]
    moz, inten = [], []
    for s in range(batch.n_spectra):
        a, b = int(batch.idx[s]), int(batch.idx[s + 1])
        i, m, k = pickpeaks_one(batch.intensity[a:b], batch.moz[a:b], base_int, min_int, qalen)
        inten.append(i[len(i) - k:])
        moz.append(m[len(m) - k:])
        idx.append(idx[-1] + k)
    cat = (lambda xs: np.concatenate(xs).astype(np.int32) if xs else np.zeros(0, np.int32))
    return np.array(idx, dtype=np.uint32), cat(moz), cat(inten)


def lower_bound(a: np.ndarray, v: int) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return lib().orc_lower_bound(F.ptr(a), len(a), v)


def upper_bound(a: np.ndarray, v: int) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return lib().orc_upper_bound(F.ptr(a), len(a), v)


def window(mass: np.ndarray, precmass: float, dM: float):
    mass = np.ascontiguousarray(mass, dtype=np.float32)
    lo, hi = C.c_int32(), C.c_int32()
    ok = lib().orc_window(F.ptr(mass), len(mass), precmass, dM, C.byref(lo), C.byref(hi))
    return bool(ok), lo.value, hi.value


def lgfact(n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.float64)
    lib().orc_lgfact(F.ptr(out), n)
    return out


def tailfit(hist: np.ndarray, hyp: int, cpsms: int):
    """-> (ok, slope, intercept, mu, beta)"""
    h = np.ascontiguousarray(hist, dtype=np.uint32)
    assert h.shape == (F.GX_HIST_BINS,)
    sl, ic, mu, be = C.c_double(), C.c_double(), C.c_int32(), C.c_int32()
    ok = lib().orc_tailfit(F.ptr(h), hyp, cpsms, C.byref(sl), C.byref(ic), C.byref(mu), C.byref(be))
    return bool(ok), sl.value, ic.value, mu.value, be.value


def default_params(dM: float = 10.0, **kw) -> F.gx_search_params:
    p = F.gx_search_params()
    p.dF, p.scale, p.max_mass, p.min_shp, p.min_cpsm = 2, 100, 5000, 4, 4
    p.evalue_mode = F.GX_EVALUE_TAILFIT
    p.dM, p.expect_max, p.surrogate_conf = dM, 20.0, 0.5
    p.engine, p.config, p.use_graph, p.split = 0, -1, 0, 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def search(index, batch, params: F.gx_search_params, threads: int = 0, want_hist: bool = False,
           want_counters: bool = False):
    """Run the CPU restatement over a batch -> (psms, hist | None, counters | None)."""
    n = batch.n_spectra
    psm = np.zeros(n, dtype=F.PSM_DTYPE)
    hist = np.zeros((n, F.GX_HIST_BINS), dtype=np.uint32) if want_hist else None
    cnt = np.zeros(n, dtype=COUNTER_DTYPE) if want_counters else None
    q = batch.as_ffi()
    st = lib().orc_search(index.raw_array, len(index.buckets), C.byref(q), C.byref(params), F.ptr(psm),
                          F.ptr(hist), F.ptr(cnt), threads)
    F.check("orc_search", st)
    return psm, hist, cnt


def scorecard(bucket, moz, inten, params, minlimit: int, maxlimit: int):
    w = maxlimit - minlimit + 1
    bc = np.zeros(w, np.uint16)
    yc = np.zeros(w, np.uint16)
    ibc = np.zeros(w, np.uint32)
    iyc = np.zeros(w, np.uint32)
    moz = np.ascontiguousarray(moz, np.int32)
    inten = np.ascontiguousarray(inten, np.int32)
    st = lib().orc_scorecard(C.byref(bucket.raw), F.ptr(moz), F.ptr(inten), len(moz), C.byref(params),
                             minlimit, maxlimit, F.ptr(bc), F.ptr(yc), F.ptr(ibc), F.ptr(iyc))
    F.check("orc_scorecard", st)
    return bc, yc, ibc, iyc
