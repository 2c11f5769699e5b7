"""The batched engine at the benchmark's full sizes, on the 2M-peptide open
index (2M tryptic peptides + Met oxidation, ~5.6M entries, ~290M ions):

- the open workload (100k spectra, 50-150 peaks, dM +-500 Da) and the uneven
  workload (200k spectra, 20-2000 peaks, dM drawn from {10, 100, 500} Da per
  spectrum), each searched in one call (planning chunks, one persistent
  launch per chunk) and replayed from its hipGraph;
- a seeded random sample of spectra checked against the CPU oracle to the
  parity bar (bit-exact counts, candidates, histograms and top hits);
- size-independent properties over every spectrum: each histogram sums to the
  spectrum's candidate count, the top hit's bin is occupied, graph replays are
  bit-identical to the eager run, the work counters equal the oracle's on the
  sample, and the e-value is cpsms * 10^(mu/1e6 * hyp + beta/1e6).
"""
import time

import numpy as np
import pytest

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import Searcher, SearchParams, ShardedSearcher
from oracle import pyoracle as O

from test_gpu_parity import compare  # noqa: E402

pytestmark = pytest.mark.gpu

SAMPLE = 2000


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def big_index():
    say("building the 2M-peptide open index")
    t0 = time.perf_counter()
    index = slm.SlmIndex(slm.DB_PRESETS["open2m"])
    say(f"  host build: {(time.perf_counter() - t0) * 1e3:.0f} ms")
    return index


@pytest.fixture(scope="module")
def big_searcher(big_index):
    s = Searcher(big_index)
    yield s
    s.close()


def _invariants(psm, hist):
    assert np.array_equal(hist.sum(axis=1, dtype=np.int64), psm["cpsms"].astype(np.int64))
    hit = (psm["flags"] & F.GX_PSM_HAS_HIT) != 0
    assert np.all(hist[np.nonzero(hit)[0], psm["hyp"][hit]] > 0)
    assert np.all(psm["psid"][~hit] == -1)
    fit = (psm["flags"] & F.GX_PSM_FIT_OK) != 0
    e = psm["cpsms"][fit] * np.power(10.0, psm["mu"][fit] / 1e6 * psm["hyp"][fit] + psm["beta"][fit] / 1e6)
    assert np.allclose(e, psm["evalue"][fit], rtol=1e-12, atol=0)
    assert np.all(np.isinf(psm["evalue"][~fit]))
    rep = (psm["flags"] & F.GX_PSM_REPORTED) != 0
    assert np.all(psm["evalue"][rep] < 20.0) and np.all(~rep | fit)


@pytest.mark.parametrize("workload", ["open2m", "uneven"])
def test_full_size_workload(big_index, big_searcher, workload):
    sc = slm.SPEC_PRESETS[workload]
    say(f"case: full-size {workload}: {sc.n_spectra} spectra")
    batch = slm.generate_spectra(big_index, sc)
    dq = big_searcher.upload(batch)
    p = SearchParams(dM=500.0)
    st = big_searcher.search_device(dq, p)
    psm, hist = big_searcher.download(dq, want_hist=True)
    say(f"  eager: {st['total_ms']:.1f} ms, {st['chunks']} chunks, errors {st['errors']}")
    assert st["spectra"] == batch.n_spectra and st["errors"] == 0 and st["ions_scanned"] == st["ions"]
    _invariants(psm, hist)
    # graph capture, then two replays: identical to the eager results
    g = SearchParams(dM=500.0, use_graph=True)
    big_searcher.search_device(dq, g)
    for _ in range(2):
        big_searcher.search_device(dq, g)
        p2, h2 = big_searcher.download(dq, want_hist=True)
        assert p2.tobytes() == psm.tobytes() and np.array_equal(h2, hist)
    # a seeded sample against the oracle, with the work counters
    rng = np.random.default_rng(2024)
    idx = np.sort(rng.choice(batch.n_spectra, SAMPLE, replace=False))
    sub = batch.subset(idx)
    c, hc, cnt = O.search(big_index, sub, O.default_params(dM=500.0), want_hist=True, want_counters=True)
    compare(psm[idx], c, hist[idx], hc, f"full-size {workload} sample")
    gs, _, gst = big_searcher.search(sub, SearchParams(dM=500.0))
    assert gst["ions"] == int(cnt["ions"].sum()) and gst["window"] == int(cnt["window"].sum())
    assert gs.tobytes() == psm[idx].tobytes()
    dq.free()


def test_full_size_device_build(big_index):
    """The 2M-peptide open index built on the GPU equals the host build."""
    say("case: full-size open2m index built on the GPU")
    t0 = time.perf_counter()
    gs = Searcher(big_index, build_on_device=True)
    say(f"  device build (entries upload, fragments, sort, bin starts, derived arrays): "
        f"{(time.perf_counter() - t0) * 1e3:.0f} ms for {big_index.n_ions} ions")
    for i, b in enumerate(big_index.buckets):
        if b.n_entries == 0:
            continue
        bA, iA = gs.index_arrays(i)
        assert np.array_equal(bA, b.bA), f"bucket {i}: bA differs"
        assert np.array_equal(iA, b.iA), f"bucket {i}: iA differs"
    gs.close()


@pytest.mark.parametrize("workload", ["open2m", "uneven"])
def test_full_size_preprocess(big_index, big_searcher, workload):
    """Raw spectra of the full-size workloads picked on the GPU equal the host
    picking of the same raw spectra (and the generator's picked batch)."""
    sc = slm.SPEC_PRESETS[workload]
    say(f"case: full-size {workload} peak picking: {sc.n_spectra} raw spectra")
    raw = slm.generate_raw_spectra(big_index, sc)
    pick = slm.PickParams(sc.base_int, sc.min_int, sc.qalen)
    t0 = time.perf_counter()
    host = slm.pick_peaks(raw, pick)
    t_host = time.perf_counter() - t0
    t0 = time.perf_counter()
    dq = big_searcher.preprocess(raw, pick)
    t_call = time.perf_counter() - t0
    ns, npk, dev_ms = dq.shape()
    say(f"  {raw.n_peaks} raw peaks -> {npk} kept; GPU {dev_ms:.2f} ms on the device "
        f"({t_call * 1e3:.0f} ms with the upload), host (all cores) {t_host * 1e3:.0f} ms")
    idx, moz, inten = dq.peaks()
    assert np.array_equal(idx, host.idx) and np.array_equal(moz, host.moz) and np.array_equal(inten, host.intensity)
    dq.free()


def test_full_size_sharded(big_index, big_searcher):
    """The open workload over the 2M-peptide index in 4 index shards, merged:
    identical to the unsharded search."""
    say("case: full-size open2m, 4 index shards merged")
    batch = slm.generate_spectra(big_index, slm.SPEC_PRESETS["open2m"])
    p = SearchParams(dM=500.0)
    want, want_h, st = big_searcher.search(batch, p, want_hist=True)
    sh = ShardedSearcher(big_index, 4)
    got, got_h, sst = sh.search(batch, p, want_hist=True)
    sh.close()
    say(f"  unsharded {st['total_ms']:.1f} ms; shards " + ", ".join(f"{x:.1f}" for x in sst["shard_ms"]) +
        f" ms; merge {sst['merge_ms']:.2f} ms")
    assert sst["ions"] == st["ions"]
    assert got.tobytes() == want.tobytes() and np.array_equal(got_h, want_h)
