"""HIP search vs the CPU oracle on the same seeded inputs (through the C ABI).

Bar: shared-peak counts, candidate counts, histograms and top-hit ids are
bit-exact; hyperscores within 1e-5 relative; e-values within 1e-3 relative of
log10.
"""
import sys

import numpy as np
import pytest

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import SearchParams
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

HYP_RTOL = 1e-5
LOGE_RTOL = 1e-3
EXACT = ["psid", "idxoffset", "sharedions", "totalions", "cpsms", "hyp", "flags", "mu", "beta"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare(gpu, cpu, hg=None, hc=None, label=""):
    assert len(gpu) == len(cpu)
    for f in EXACT:
        bad = np.nonzero(gpu[f] != cpu[f])[0]
        assert len(bad) == 0, f"{label}: field {f} differs at {bad[:10]}: gpu={gpu[f][bad[:5]]} cpu={cpu[f][bad[:5]]}"
    h_g, h_c = gpu["hyperscore"].astype(np.float64), cpu["hyperscore"].astype(np.float64)
    rel = np.abs(h_g - h_c) / np.maximum(np.abs(h_c), 1e-30)
    assert np.all(rel <= HYP_RTOL), f"{label}: hyperscore rel err {rel.max()}"
    fin = np.isfinite(cpu["evalue"])
    assert np.array_equal(fin, np.isfinite(gpu["evalue"])), label
    lg, lc = np.log10(gpu["evalue"][fin]), np.log10(cpu["evalue"][fin])
    err = np.abs(lg - lc) / np.maximum(np.abs(lc), 1e-12)
    assert np.all((err <= LOGE_RTOL) | (np.abs(lg - lc) < 1e-9)), f"{label}: log10 e rel err {err.max()}"
    if hg is not None:
        assert np.array_equal(hg, hc), f"{label}: histograms differ"


# every search engine and batched launch configuration
# (wide_stream: the kernel instance with 64-bit ion-stream addresses, which an
# index with a bucket of about 2^30 ions or more takes)
ENGINES = [dict(engine=F.GX_ENGINE_BATCHED, config=0), dict(engine=F.GX_ENGINE_BATCHED, config=1),
           dict(engine=F.GX_ENGINE_BATCHED, config=2), dict(engine=F.GX_ENGINE_BATCHED, config=3),
           dict(engine=F.GX_ENGINE_BATCHED, config=0, wide_stream=1),
           dict(engine=F.GX_ENGINE_BATCHED, config=1, wide_stream=1), dict(engine=F.GX_ENGINE_SPECTRUM)]
ENGINE_IDS = ["batched-c0", "batched-c1", "batched-c2", "batched-c3", "batched-c0-wide", "batched-c1-wide",
              "spectrum"]


def check_stats(st, cnt, n):
    assert st["spectra"] == n
    assert st["errors"] == 0
    assert st["ions_scanned"] == st["ions"], (st["ions_scanned"], st["ions"])
    assert st["ions"] == int(cnt["ions"].sum())
    assert st["bins"] == int(cnt["bins"].sum())
    assert st["window"] == int(cnt["window"].sum())


@pytest.mark.parametrize("eng", ENGINES, ids=ENGINE_IDS)
def test_small_closed_search_parity(small_index, small_batch, searcher_small, eng):
    say(f"case: small closed search, 50k peptides, 1000 spectra, dM 10, {eng}")
    p = SearchParams(dM=10.0, **eng)
    g, hg, st = searcher_small.search(small_batch, p, want_hist=True)
    c, hc, cnt = O.search(small_index, small_batch, O.default_params(dM=10.0), want_hist=True,
                          want_counters=True)
    check_stats(st, cnt, small_batch.n_spectra)
    compare(g, c, hg, hc, "small")
    assert st["candidates"] == int(c["cpsms"].sum())


@pytest.mark.parametrize("eng", ENGINES, ids=ENGINE_IDS)
def test_open_search_parity(open_index, searcher_open, eng):
    say(f"case: open search, 200k peptides + Met-ox, 300 spectra, dM 500, {eng}")
    cfg = slm.SpecConfig(n_spectra=300, seed=21, peak_mode=F.GX_PEAKS_TOTAL_UNIFORM, peaks_min=50,
                         peaks_max=150, charge_min=2, charge_max=3, dm_choices=(500.0,))
    batch = slm.generate_spectra(open_index, cfg)
    g, hg, st = searcher_open.search(batch, SearchParams(dM=500.0, **eng), want_hist=True)
    c, hc, cnt = O.search(open_index, batch, O.default_params(dM=500.0), want_hist=True, want_counters=True)
    check_stats(st, cnt, batch.n_spectra)
    compare(g, c, hg, hc, "open")
    assert (c["flags"] & F.GX_PSM_MODELED).sum() > 250


@pytest.mark.parametrize("eng", ENGINES, ids=ENGINE_IDS)
def test_uneven_batch_parity(open_index, searcher_open, eng):
    """power-law peak counts up to 2000 (slice tables spill to scratch) and
    per-spectrum dM from {10, 100, 500}"""
    say(f"case: uneven batch, power-law peaks 20-2000, dM mix, {eng}")
    cfg = slm.SpecConfig(n_spectra=160, seed=33, peak_mode=F.GX_PEAKS_TOTAL_POWER, peaks_min=20,
                         peaks_max=2000, power_alpha=1.2, charge_min=2, charge_max=3, qalen=2048,
                         dm_choices=(10.0, 100.0, 500.0))
    batch = slm.generate_spectra(open_index, cfg)
    assert (np.diff(batch.idx) > 1024 // 5).any()
    # spectra of more than 512 legal peaks: the batched plan sorts them with its bitonic network
    assert (np.diff(batch.idx) > 600).sum() >= 10
    g, hg, st = searcher_open.search(batch, SearchParams(dM=500.0, **eng), want_hist=True)
    c, hc, cnt = O.search(open_index, batch, O.default_params(dM=500.0), want_hist=True, want_counters=True)
    check_stats(st, cnt, batch.n_spectra)
    compare(g, c, hg, hc, "uneven")


_WINDOW_CASES = {}


@pytest.mark.parametrize("config", [0, 1])
@pytest.mark.parametrize("dM, split, whole_bin", [(5000.0, 1, 32), (5000.0, 4, 1), (0.05, 1, 32), (500.0, 4, 1),
                                                  (500.0, 2, 4096)])
def test_window_extremes_parity(open_index, searcher_open, config, dM, split, whole_bin):
    """Windows that span whole buckets (the plan needs no boundary search at
    either end), windows of a few entries, windows cut into up to 4 tiles per
    count tile (interior tile boundaries), every bin searched (whole-bin
    threshold 1: every slice's ends bracketed from the 32x samples) and every
    bin streamed whole (4096)."""
    say(f"case: window extremes, dM {dM}, split {split}, whole bins <= {whole_bin}, config {config}")
    cfg = slm.SpecConfig(n_spectra=120, seed=57, peak_mode=F.GX_PEAKS_TOTAL_UNIFORM, peaks_min=40,
                         peaks_max=160, charge_min=2, charge_max=3, dm_choices=(dM,))
    batch = slm.generate_spectra(open_index, cfg)
    if dM not in _WINDOW_CASES:
        _WINDOW_CASES[dM] = O.search(open_index, batch, O.default_params(dM=dM), want_hist=True,
                                     want_counters=True)
    c, hc, cnt = _WINDOW_CASES[dM]
    p = SearchParams(dM=dM, engine=F.GX_ENGINE_BATCHED, config=config, split=split, whole_bin=whole_bin)
    g, hg, st = searcher_open.search(batch, p, want_hist=True)
    check_stats(st, cnt, batch.n_spectra)
    compare(g, c, hg, hc, f"window extremes dM {dM}")
    if dM >= 5000.0:  # the case's premise: every window is its whole bucket
        assert st["window"] == batch.n_spectra * sum(b.n_entries for b in open_index.buckets)


Z3_MODS = slm.DBConfig(n_peptides=20_000, seed=31, maxz=3,
                       mods=[("M", 15.9949, 2), ("STY", 79.96633, 1), ("^", 42.0106, 1)], max_mods_per_pep=3)


@pytest.fixture(scope="module")
def z3_index():
    return slm.SlmIndex(Z3_MODS)


@pytest.fixture(scope="module")
def searcher_z3(z3_index):
    from gicops_amd.search import Searcher

    s = Searcher(z3_index)
    yield s
    s.close()


@pytest.mark.parametrize("eng", ENGINES, ids=ENGINE_IDS)
def test_charge3_modified_index_parity(z3_index, searcher_z3, eng):
    """fragment charges 1-3 (speclen 6(L-1)), three variable modification
    types incl. the N-terminal one, dM mix"""
    say(f"case: maxz 3, 3 modification types, dM mix, {eng}")
    cfg = slm.SpecConfig(n_spectra=240, seed=35, peak_mode=F.GX_PEAKS_TOTAL_UNIFORM, peaks_min=50,
                         peaks_max=300, charge_min=2, charge_max=4, qalen=256, dm_choices=(10.0, 100.0, 500.0))
    batch = slm.generate_spectra(z3_index, cfg)
    g, hg, st = searcher_z3.search(batch, SearchParams(dM=500.0, **eng), want_hist=True)
    c, hc, cnt = O.search(z3_index, batch, O.default_params(dM=500.0), want_hist=True, want_counters=True)
    check_stats(st, cnt, batch.n_spectra)
    compare(g, c, hg, hc, "maxz 3")


@pytest.mark.parametrize("eng", ENGINES, ids=ENGINE_IDS)
def test_device_resident_path_matches_one_call(small_batch, searcher_small, eng):
    say(f"case: device-resident upload/search/download vs one-call search {eng}")
    p = SearchParams(dM=10.0, **eng)
    g1, _, _ = searcher_small.search(small_batch, p)
    dq = searcher_small.upload(small_batch)
    searcher_small.search_device(dq, p)
    searcher_small.search_device(dq, p)  # repeatable: state is reset per call
    g2, _ = searcher_small.download(dq)
    dq.free()
    assert g1.tobytes() == g2.tobytes()


def _edge_batch(index):
    b = index.buckets[10]
    m = float(b.mass[len(b.mass) // 2])
    base = slm.generate_spectra(index, slm.SpecConfig(n_spectra=4, seed=99))
    specs = [base.peaks(i) for i in range(4)]
    empty = (np.zeros(0, np.int32), np.zeros(0, np.int32))
    illegal = (np.array([0, 1, 2, 499998, 499999, -5], np.int32), np.full(6, 1000, np.int32))
    dup = (np.repeat(specs[0][0], 3), np.repeat(specs[0][1], 3))  # every peak three times
    rows = [specs[0], empty, illegal, specs[1], dup, specs[2], specs[3]]
    prec = [base.precurse[0], m, m, 1.0e6, base.precurse[0], -50.0, base.precurse[3]]
    idx = np.zeros(len(rows) + 1, np.uint32)
    idx[1:] = np.cumsum([len(r[0]) for r in rows])
    return slm.QueryBatch(idx, np.concatenate([r[0] for r in rows]), np.concatenate([r[1] for r in rows]),
                          np.array(prec, np.float32))


@pytest.mark.parametrize("eng", ENGINES, ids=ENGINE_IDS)
@pytest.mark.parametrize("kw", [dict(), dict(min_shp=1), dict(min_shp=0, min_cpsm=0), dict(dM=-1.0),
                                dict(dM=0.0), dict(min_cpsm=1, expect_max=1e30)])
def test_edge_cases_parity(small_index, searcher_small, kw, eng):
    say(f"case: edge batch {kw} {eng}")
    batch = _edge_batch(small_index)
    pd = dict(dM=10.0)
    pd.update(kw)
    pd_gpu = dict(pd, **eng)
    g, hg, _ = searcher_small.search(batch, SearchParams(**pd_gpu), want_hist=True)
    c, hc, _ = O.search(small_index, batch, O.default_params(**pd), want_hist=True)
    compare(g, c, hg, hc, f"edge {kw}")


@pytest.mark.parametrize("eng", ENGINES, ids=ENGINE_IDS)
def test_empty_batch(searcher_small, eng):
    say(f"case: empty batch {eng}")
    batch = slm.QueryBatch(np.zeros(1, np.uint32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                           np.zeros(0, np.float32))
    g, _, st = searcher_small.search(batch, SearchParams(**eng))
    assert len(g) == 0 and st["spectra"] == 0


def test_invalid_arguments_return_status(searcher_small, small_batch):
    say("case: invalid arguments")
    with pytest.raises(F.GxError) as e:
        searcher_small.search(small_batch, SearchParams(scale=50))  # nbins mismatch with the index
    assert e.value.code == F.GX_ERR_INVLD_INDEX
    with pytest.raises(F.GxError) as e:
        searcher_small.search(small_batch, SearchParams(dF=-1))
    assert e.value.code == F.GX_ERR_INVLD_PARAM
