"""Index-sharded search on one GPU: partial results per shard
(gx_partials_export) merged into PSMs (gx_partials_merge).

Bar: with every bucket's entries split into contiguous psid ranges, the
merged PSMs (top hit, tie flag, shared ions, cpsms, hyp, mu, beta, e-value)
and histograms are byte-identical to the unsharded search, for 1-5 shards,
with the exchange through host memory and through device memory, and the
shards' work counters add up to the unsharded ones.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import Searcher, SearchParams, ShardedSearcher

pytestmark = pytest.mark.gpu

MODS = slm.DBConfig(n_peptides=40_000, seed=13, maxz=2, mods=[("M", 15.9949, 2), ("STY", 79.96633, 1)],
                    max_mods_per_pep=2)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mods_index():
    return slm.SlmIndex(MODS)


def unsharded(index, batch, p):
    s = Searcher(index)
    psm, hist, st = s.search(batch, p, want_hist=True)
    s.close()
    return psm, hist, st


@pytest.mark.parametrize("n_shards", [1, 2, 3, 5])
@pytest.mark.parametrize("which", ["small", "mods"])
def test_sharded_merge_equals_unsharded(small_index, mods_index, which, n_shards):
    say(f"case: {which} index in {n_shards} shards")
    index = small_index if which == "small" else mods_index
    sc = slm.SpecConfig(n_spectra=600, seed=21, charge_min=2, charge_max=3, dm_choices=(10.0, 100.0, 500.0))
    batch = slm.generate_spectra(index, sc)
    p = SearchParams(dM=500.0)
    want, want_h, st = unsharded(index, batch, p)
    sh = ShardedSearcher(index, n_shards)
    got, got_h, sst = sh.search(batch, p, want_hist=True)
    sh.close()
    assert got.tobytes() == want.tobytes()
    assert np.array_equal(got_h, want_h)
    assert sst["ions"] == st["ions"]
    ties = int(np.count_nonzero(want["flags"] & F.GX_PSM_TIE))
    say(f"  {ties} tied top hits, merge {sst['merge_ms']:.3f} ms")


def test_sharded_merge_through_device_memory(small_index):
    import torch

    say("case: partials exchanged in device memory")
    batch = slm.generate_spectra(small_index, slm.SPEC_PRESETS["small"])
    p = SearchParams(dM=100.0)
    want, want_h, _ = unsharded(small_index, batch, p)
    k, n = 3, batch.n_spectra
    sh = ShardedSearcher(small_index, k)
    dq = sh.upload(batch)
    rec = torch.zeros((k, n * F.PARTIAL_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    hist = torch.zeros((k, n * F.GX_HIST_BINS), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    lib, pf = sh._lib, p.to_ffi()
    for i, d in enumerate(sh._didx):
        st = F.gx_search_stats()
        F.check("gx_search_device", lib.gx_search_device(sh._ctx, d, dq._h, C.byref(pf), C.byref(st)))
        F.check("gx_partials_export", lib.gx_partials_export(sh._ctx, d, dq._h, F.ptr(sh.offsets[i]),
                                                             C.c_void_p(rec[i].data_ptr()),
                                                             C.c_void_p(hist[i].data_ptr()), 1))
    psm = np.zeros(n, dtype=F.PSM_DTYPE)
    mh = np.zeros((n, F.GX_HIST_BINS), np.uint32)
    rp = (C.c_void_p * k)(*[rec[i].data_ptr() for i in range(k)])
    hp = (C.c_void_p * k)(*[hist[i].data_ptr() for i in range(k)])
    F.check("gx_partials_merge", lib.gx_partials_merge(sh._ctx, k, rp, hp, n, C.byref(pf), 1, F.ptr(psm),
                                                       F.ptr(mh), None))
    assert psm.tobytes() == want.tobytes() and np.array_equal(mh, want_h)
    dq.free()
    sh.close()


def test_partials_error_codes(small_index):
    say("case: partial-result error codes")
    batch = slm.generate_spectra(small_index, slm.SPEC_PRESETS["tiny"])
    sh = ShardedSearcher(small_index, 2)
    lib = sh._lib
    dq = sh.upload(batch)
    n = batch.n_spectra
    rec = np.zeros(n, F.PARTIAL_DTYPE)
    # nothing searched yet
    assert lib.gx_partials_export(sh._ctx, sh._didx[0], dq._h, None, F.ptr(rec), None, 0) == F.GX_ERR_NOT_READY
    pf = SearchParams(dM=10.0).to_ffi()
    st = F.gx_search_stats()
    F.check("gx_search_device", lib.gx_search_device(sh._ctx, sh._didx[0], dq._h, C.byref(pf), C.byref(st)))
    # searched on shard 0, exported as if from shard 1
    assert lib.gx_partials_export(sh._ctx, sh._didx[1], dq._h, None, F.ptr(rec), None, 0) == F.GX_ERR_INVLD_INDEX
    # the plain engine leaves no partials
    pp = SearchParams(dM=10.0, engine=F.GX_ENGINE_SPECTRUM).to_ffi()
    F.check("gx_search_device", lib.gx_search_device(sh._ctx, sh._didx[0], dq._h, C.byref(pp), C.byref(st)))
    assert lib.gx_partials_export(sh._ctx, sh._didx[0], dq._h, None, F.ptr(rec), None, 0) == F.GX_ERR_NOT_READY
    # global psids beyond 26 bits
    big = np.full(len(small_index.buckets), 1 << 26, np.uint32)
    F.check("gx_search_device", lib.gx_search_device(sh._ctx, sh._didx[0], dq._h, C.byref(pf), C.byref(st)))
    assert lib.gx_partials_export(sh._ctx, sh._didx[0], dq._h, F.ptr(big), F.ptr(rec), None, 0) == F.GX_ERR_INVLD_SIZE
    psm = np.zeros(n, dtype=F.PSM_DTYPE)
    assert lib.gx_partials_merge(sh._ctx, 0, None, None, n, C.byref(pf), 0, F.ptr(psm), None, None) in (
        F.GX_ERR_INVLD_PTR, F.GX_ERR_INVLD_PARAM)
    dq.free()
    sh.close()


def test_graph_replay_follows_the_index(small_index):
    """A batch searched with use_graph on two live shard indexes in turn: each
    replay must equal that shard's own eager result (a graph captured on one
    index is never replayed on another, also after the first is freed and a
    new index may take its address)."""
    say("case: graph replays on two shard indexes")
    batch = slm.generate_spectra(small_index, slm.SPEC_PRESETS["small"])
    sh = ShardedSearcher(small_index, 2)
    lib = sh._lib
    dq = sh.upload(batch)
    n = batch.n_spectra

    def run(d, graph):
        pf = SearchParams(dM=100.0, use_graph=graph).to_ffi()
        st = F.gx_search_stats()
        F.check("gx_search_device", lib.gx_search_device(sh._ctx, d, dq._h, C.byref(pf), C.byref(st)))
        psm = np.zeros(n, dtype=F.PSM_DTYPE)
        F.check("gx_results_download", lib.gx_results_download(sh._ctx, dq._h, F.ptr(psm), None))
        return psm

    eager = [run(d, False) for d in sh._didx]
    assert eager[0].tobytes() != eager[1].tobytes()
    for rnd in range(2):
        for k, d in enumerate(sh._didx):
            for rep in range(2):  # capture (or re-capture) on this index, then replay
                got = run(d, True)
                assert got.tobytes() == eager[k].tobytes(), (rnd, k, rep)
    # free shard 1 and build a new index in its place: the batch's graph was
    # last captured on shard 1, and the new index may reuse its address
    lib.gx_index_free(sh._didx[1])
    bp = small_index.build_params()
    d2 = C.c_void_p()
    F.check("gx_index_build", lib.gx_index_build(sh._ctx, small_index.raw_array, len(small_index.buckets),
                                                 C.byref(bp), C.byref(d2)))
    sh._didx[1] = d2
    whole = run(d2, False)
    assert run(d2, True).tobytes() == whole.tobytes()
    assert run(d2, True).tobytes() == whole.tobytes()
    dq.free()
    sh.close()
