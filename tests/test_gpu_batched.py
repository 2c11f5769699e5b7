"""Batched form: hipGraph capture/replay, window split, autotuner."""
import sys

import numpy as np
import pytest

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import SearchParams
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

from test_gpu_parity import check_stats, compare  # noqa: E402


def say(msg):
    print(msg, flush=True)


def _uneven(index, n=200, seed=41):
    cfg = slm.SpecConfig(n_spectra=n, seed=seed, peak_mode=F.GX_PEAKS_TOTAL_POWER, peaks_min=20, peaks_max=2000,
                         power_alpha=1.5, charge_min=2, charge_max=3, qalen=2048, dm_choices=(10.0, 100.0, 500.0))
    return slm.generate_spectra(index, cfg)


@pytest.mark.parametrize("split", [2, 4])
def test_window_split_parity(open_index, searcher_open, split):
    say(f"case: window split {split}, uneven batch")
    batch = _uneven(open_index)
    g, hg, st = searcher_open.search(batch, SearchParams(dM=500.0, config=1, split=split), want_hist=True)
    c, hc, cnt = O.search(open_index, batch, O.default_params(dM=500.0), want_hist=True, want_counters=True)
    check_stats(st, cnt, batch.n_spectra)
    compare(g, c, hg, hc, f"split {split}")


@pytest.mark.parametrize("whole_bin", [1, 256, 4096])
def test_whole_bin_threshold_parity(open_index, searcher_open, whole_bin):
    """Bins streamed whole and filtered by id give the same counts as bins
    searched for their tile boundaries."""
    say(f"case: whole-bin threshold {whole_bin}, uneven batch")
    batch = _uneven(open_index, n=150, seed=47)
    g, hg, st = searcher_open.search(batch, SearchParams(dM=500.0, config=1, whole_bin=whole_bin), want_hist=True)
    c, hc, cnt = O.search(open_index, batch, O.default_params(dM=500.0), want_hist=True, want_counters=True)
    check_stats(st, cnt, batch.n_spectra)
    compare(g, c, hg, hc, f"whole_bin {whole_bin}")


def test_graph_replays_match_eager(open_index, searcher_open):
    say("case: hipGraph capture and replays")
    batch = _uneven(open_index, n=150, seed=43)
    dq = searcher_open.upload(batch)
    eager = SearchParams(dM=500.0, config=1)
    searcher_open.search_device(dq, eager)
    ref_psm, ref_hist = searcher_open.download(dq, want_hist=True)
    graph = SearchParams(dM=500.0, config=1, use_graph=True)
    searcher_open.search_device(dq, graph)  # eager run + capture
    other = _uneven(open_index, n=60, seed=44)
    for rep in range(4):
        st = searcher_open.search_device(dq, graph)  # replay
        psm, hist = searcher_open.download(dq, want_hist=True)
        assert psm.tobytes() == ref_psm.tobytes(), f"replay {rep}"
        assert np.array_equal(hist, ref_hist), f"replay {rep}"
        assert st["errors"] == 0 and st["ions_scanned"] == st["ions"]
        # eager work on the same context between replays
        searcher_open.search(other, SearchParams(dM=500.0, config=0))
    c, hc, _ = O.search(open_index, batch, O.default_params(dM=500.0), want_hist=True)
    compare(ref_psm, c, ref_hist, hc, "graph")
    dq.free()


def test_autotune_choice_keeps_parity(open_index, searcher_open):
    say("case: autotuner")
    sample = _uneven(open_index, n=120, seed=45)
    dq = searcher_open.upload(sample)
    r = searcher_open.autotune(dq, SearchParams(dM=500.0))
    dq.free()
    assert r["config"] in (0, 1, 2, 3) and r["split"] in (1, 2) and r["whole_bin"] in (8, 16, 32, 64, 128)
    # 4 configurations x 2 splits, then 4 more thresholds for each of the two best pairs
    assert len(r["tried"]) == 16 and all(ms > 0 for _, _, _, ms in r["tried"])
    first = sorted(r["tried"][:8], key=lambda x: x[3])[:2]
    swept = {(c, sp) for c, sp, _, _ in r["tried"][8:]}
    assert swept == {(c, sp) for c, sp, _, _ in first}
    assert (r["config"], r["split"]) in swept
    batch = _uneven(open_index, n=100, seed=46)
    g, hg, st = searcher_open.search(batch, SearchParams(dM=500.0), want_hist=True)  # tuned defaults
    assert st["config"] == r["config"]
    c, hc, _ = O.search(open_index, batch, O.default_params(dM=500.0), want_hist=True)
    compare(g, c, hg, hc, "autotuned")


def test_measured_read_bandwidth(searcher_small):
    """The bench's measured denominators: sequential and random-segment reads
    report plausible rates (a few TB/s, below the 8 TB/s published peak);
    other segment sizes are rejected."""
    say("case: measured read bandwidth")
    seq = searcher_small.read_bandwidth(0)
    rnd = searcher_small.read_bandwidth(2)
    say(f"  sequential {seq:.0f} GB/s, random 2 x 128 B segments {rnd:.0f} GB/s")
    assert 500.0 < seq < 8000.0 and 200.0 < rnd < 8000.0
    with pytest.raises(F.GxError) as e:
        searcher_small.read_bandwidth(3)
    assert e.value.code == F.GX_ERR_INVLD_PARAM
