"""The 50M-peptide database: 50M tryptic peptides, Met oxidation (<= 2) and
N-terminal acetylation, <= 2 modifications per peptide (~244M entries,
~12.3G fragment ions). The host builds the mass-sorted entries only; the GPU
builds the fragment-ion arrays (gx_index_build).

Checked: every bucket within the engine's limits (< 2^26 entries, < 2^31
ions); 100k spectra of the uneven mix (20-2000 raw peaks, QALEN 2048, dM
drawn from {10, 100, 500} Da) searched in one call, with the size-independent
properties over every spectrum; a seeded sample against the CPU oracle on a
host copy of the GPU-built arrays, to the parity bar, with the work counters.
"""
import dataclasses
import sys
import time

import numpy as np
import pytest

from gicops_amd import slm
from gicops_amd.search import Searcher, SearchParams
from oracle import pyoracle as O

from test_gpu_fullsize import _invariants  # noqa: E402
from test_gpu_parity import compare  # noqa: E402

pytestmark = pytest.mark.gpu

SAMPLE = 48


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
