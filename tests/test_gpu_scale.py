"""The 50M-peptide database: 50M tryptic peptides, Met oxidation (<= 2) and
N-terminal acetylation, <= 2 modifications per peptide (~244M entries,
~12.3G fragment ions). The host builds the mass-sorted entries only; the GPU
builds the fragment-ion arrays (gx_index_build).

Checked: every bucket within the engine's limits (< 2^26 entries, < 2^31
ions); 100k spectra of the uneven mix (20-2000 raw peaks, QALEN 2048, dM
drawn from {10, 100, 500} Da) searched in one call, with the size-independent
properties over every spectrum; a seeded sample of 300 against the CPU oracle
on a host copy of the GPU-built arrays, to the parity bar, with the work
counters; and the learned e-value on the same batch, with a surrogate trained
on held-out spectra of this index: top hits and histograms identical to the
plain search, every e-value the surrogate did not serve bit-identical, every
served one within 1e-3 relative of the plain fit's log10 e, and the fallback
rate reported.
"""
import dataclasses
import sys
import time

import numpy as np
import pytest

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import Searcher, SearchParams
from gicops_amd.surrogate import Surrogate, TrainParams
from oracle import pyoracle as O

from test_gpu_fullsize import _invariants  # noqa: E402
from test_gpu_parity import compare  # noqa: E402

pytestmark = pytest.mark.gpu

SAMPLE = 300


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c4():
    say("building the 50M-peptide entries on the host")
    t0 = time.perf_counter()
    index = slm.SlmIndex(slm.DB_PRESETS["c4"], fragments=False)
    t_host = time.perf_counter() - t0
    t0 = time.perf_counter()
    s = Searcher(index)
    t_dev = time.perf_counter() - t0
    say(f"  {index.n_entries} entries, {index.n_ions} ions; host entries {t_host:.1f} s, "
        f"GPU fragment build {t_dev:.1f} s")
    yield index, s
    s.close()
    index.close()


def test_c4_limits(c4):
    index, s = c4
    assert s.built_on_device and not index.has_fragments
    assert index.n_entries > 200_000_000 and index.n_ions > 10_000_000_000
    assert max(b.n_entries for b in index.buckets) < 2**26
    assert max(b.n_ions for b in index.buckets) < 2**31


def _learned_leg(s, index, sc, dq, psm, hist):
    """The learned e-value on the 50M-peptide batch: a surrogate trained on
    held-out spectra of the same index, judged by the plain fit."""
    sur = Surrogate(s, seed=5)
    rows = 0
    for seed in (4401, 4402, 4403):
        b = slm.generate_spectra(index, dataclasses.replace(sc, n_spectra=20_000, seed=seed))
        tq = s.upload(b)
        s.search_device(tq, SearchParams(dM=500.0))
        rows += sur.add_training(tq)
        tq.free()
    rep = sur.train(TrainParams(batch=1024, epochs=6, lr=2e-3, lr_decay=0.7, conf_weight=0.5, seed=1))
    sur.attach()
    lst = s.search_device(dq, SearchParams(dM=500.0, evalue_mode=F.GX_EVALUE_LEARNED, surrogate_conf=0.5))
    lp, lh = s.download(dq, want_hist=True)
    assert np.array_equal(lh, hist)
    for k in ("hyperscore", "psid", "idxoffset", "sharedions", "totalions", "cpsms", "hyp"):
        assert np.array_equal(lp[k], psm[k]), k
    sur_m = (lp["flags"] & F.GX_PSM_SURROGATE) != 0
    modeled = (psm["flags"] & F.GX_PSM_MODELED) != 0
    rest = ~sur_m
    for k in ("mu", "beta", "flags"):
        assert np.array_equal(lp[k][rest], psm[k][rest]), k
    assert np.array_equal(lp["evalue"][rest].view(np.uint64), psm["evalue"][rest].view(np.uint64))
    assert np.all((psm["flags"][sur_m] & F.GX_PSM_FIT_OK) != 0)
    le_ref, le = np.log10(psm["evalue"][sur_m]), np.log10(lp["evalue"][sur_m])
    rel = np.abs(le - le_ref) / np.maximum(np.abs(le_ref), 1e-3)
    assert np.all(rel <= 1e-3)
    assert lst["fallbacks"] == int(modeled.sum() - sur_m.sum())
    say(f"  learned: {rows} training rows, in window {rep['in_window'][-1]:.3f}; served {int(sur_m.sum())} of "
        f"{int(modeled.sum())} modeled ({sur_m.sum() / max(1, modeled.sum()):.3f}), fallback rate "
        f"{lst['fallbacks'] / max(1, modeled.sum()):.3f}, e-value stage {lst['evalue_ms']:.2f} ms, "
        f"max rel. log10 e error {rel.max() if len(rel) else 0.0:.1e}")
    assert sur_m.sum() > 0.3 * modeled.sum()
    sur.close()


def test_c4_search_sample_parity(c4):
    index, s = c4
    sc = dataclasses.replace(slm.SPEC_PRESETS["c4"], n_spectra=100_000)
    say(f"case: 50M-peptide search, {sc.n_spectra} spectra of the uneven mix")
    batch = slm.generate_spectra(index, sc)
    dq = s.upload(batch)
    p = SearchParams(dM=500.0)
    st = s.search_device(dq, p)
    psm, hist = s.download(dq, want_hist=True)
    say(f"  eager: {st['total_ms']:.0f} ms (plan {st['plan_ms']:.0f}, search {st['kernel_ms']:.0f}), "
        f"{st['chunks']} chunks, {st['ions']} ions, errors {st['errors']}")
    assert st["spectra"] == batch.n_spectra and st["errors"] == 0 and st["ions_scanned"] == st["ions"]
    _invariants(psm, hist)
    _learned_leg(s, index, sc, dq, psm, hist)
    dq.free()
    say("  host copy of the GPU-built arrays for the oracle")
    host = s.host_index()
    rng = np.random.default_rng(44)
    idx = np.sort(rng.choice(batch.n_spectra, SAMPLE, replace=False))
    sub = batch.subset(idx)
    t0 = time.perf_counter()
    c, hc, cnt = O.search(host, sub, O.default_params(dM=500.0), want_hist=True, want_counters=True)
    say(f"  oracle on {SAMPLE} spectra: {time.perf_counter() - t0:.1f} s")
    compare(psm[idx], c, hist[idx], hc, "50M-peptide sample")
    gs, _, gst = s.search(sub, p)
    assert gst["ions"] == int(cnt["ions"].sum()) and gst["window"] == int(cnt["window"].sum())
    assert gs.tobytes() == psm[idx].tobytes()
