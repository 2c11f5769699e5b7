"""gx_search_spectra_noalloc: the per-spectrum search with its work buffers in
the library's static device array. Its results must equal the existing
per-spectrum entry point's (gx_search with GX_ENGINE_SPECTRUM) byte for byte,
on two calls in a row (the second on a smaller batch, so the array holds the
first call's contents), and sizes beyond its limits must come back as errors
before any search runs."""
import ctypes as C

import numpy as np
import pytest
# torch provides the caller's device buffers; it is imported before the
# library's HIP runtime is loaded (by the fixtures), the order in which both
# see the GPU
import torch  # noqa: F401

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import SearchParams

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _dev(torch, a):
    """A device copy of a host array (torch allocates it, not the library)."""
    t = torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to("cuda")
    return t


def _device_batch(torch, batch):
    bufs = {k: _dev(torch, getattr(batch, k)) for k in ("idx", "moz", "intensity", "precurse", "dm")}
    q = F.gx_query_batch()
    q.n_spectra = batch.n_spectra
    q.n_peaks = batch.n_peaks
    q.idx = C.c_void_p(bufs["idx"].data_ptr())
    q.moz = C.c_void_p(bufs["moz"].data_ptr())
    q.intensity = C.c_void_p(bufs["intensity"].data_ptr())
    q.precurse = C.c_void_p(bufs["precurse"].data_ptr())
    q.dm = C.c_void_p(bufs["dm"].data_ptr())
    return q, bufs


def _noalloc(torch, searcher, batch, params):
    n = batch.n_spectra
    q, bufs = _device_batch(torch, batch)
    psm_d = torch.zeros(n * F.PSM_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    hist_d = torch.zeros(n * F.GX_HIST_BINS, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    st = F.gx_search_stats()
    p = params.to_ffi()
    rc = searcher._lib.gx_search_spectra_noalloc(searcher._ctx, searcher._didx, C.byref(q), C.byref(p),
                                                 C.c_void_p(psm_d.data_ptr()), C.c_void_p(hist_d.data_ptr()),
                                                 C.byref(st))
    F.check("gx_search_spectra_noalloc", rc)
    psm = np.frombuffer(psm_d.cpu().numpy().tobytes(), dtype=F.PSM_DTYPE)
    hist = hist_d.cpu().numpy().view(np.uint32).reshape(n, F.GX_HIST_BINS)
    return psm, hist, st


def test_noalloc_equals_the_spectrum_engine(small_index, searcher_small):
    params = SearchParams(dM=100.0, engine=F.GX_ENGINE_SPECTRUM)
    big = slm.generate_spectra(small_index, slm.SpecConfig(n_spectra=1200, seed=41, dm_choices=(10.0, 100.0, 500.0)))
    small = slm.generate_spectra(small_index, slm.SpecConfig(n_spectra=150, seed=42, dm_choices=(10.0, 500.0)))
    for name, batch in (("first call, 1200 spectra", big), ("second call, 150 spectra", small)):
        say(f"case: no-alloc search, {name}")
        want, want_h, want_st = searcher_small.search(batch, params, want_hist=True)
        got, got_h, st = _noalloc(torch, searcher_small, batch, params)
        assert got.tobytes() == want.tobytes()
        assert np.array_equal(got_h, want_h)
        assert st.ions == want_st["ions"] and st.candidates == want_st["candidates"]
        assert st.spectra == batch.n_spectra and st.errors == 0
        assert int((got["flags"] & F.GX_PSM_HAS_HIT).astype(bool).sum()) > 0.9 * batch.n_spectra


def test_noalloc_rejects_what_it_cannot_hold(searcher_small):
    say("case: no-alloc limits and argument errors")
    lib = searcher_small._lib
    n_peaks = 3000  # one spectrum above the 2048-peak limit: rejected by the batch check, nothing searched
    idx = torch.tensor([0, n_peaks], dtype=torch.int32, device="cuda")
    zeros = torch.zeros(n_peaks, dtype=torch.int32, device="cuda")
    prec = torch.full((1,), 1500.0, dtype=torch.float32, device="cuda")
    out = torch.zeros(F.PSM_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    q = F.gx_query_batch()
    q.n_spectra, q.n_peaks = 1, n_peaks
    q.idx, q.moz, q.intensity = C.c_void_p(idx.data_ptr()), C.c_void_p(zeros.data_ptr()), C.c_void_p(zeros.data_ptr())
    q.precurse, q.dm = C.c_void_p(prec.data_ptr()), None
    p = SearchParams(dM=100.0, engine=F.GX_ENGINE_SPECTRUM).to_ffi()
    args = (searcher_small._ctx, searcher_small._didx, C.byref(q), C.byref(p), C.c_void_p(out.data_ptr()), None, None)
    assert lib.gx_search_spectra_noalloc(*args) == F.GX_ERR_INVLD_SIZE
    # idx not ending at n_peaks
    q.n_peaks = n_peaks - 1
    assert lib.gx_search_spectra_noalloc(*args) == F.GX_ERR_INVLD_SIZE
    # 1000 peaks x 5 bins: above the 4096-slice limit
    idx2 = torch.tensor([0, 1000], dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    q.idx, q.n_peaks = C.c_void_p(idx2.data_ptr()), 1000
    assert lib.gx_search_spectra_noalloc(*args) == F.GX_ERR_INVLD_SIZE
    # the batched engine and the learned e-value are not what this entry point runs
    pb = SearchParams(dM=100.0).to_ffi()
    assert lib.gx_search_spectra_noalloc(searcher_small._ctx, searcher_small._didx, C.byref(q), C.byref(pb),
                                         C.c_void_p(out.data_ptr()), None, None) == F.GX_ERR_INVLD_PARAM
