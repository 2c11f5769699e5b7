"""Peak picking on the GPU (gx_queries_preprocess) against the oracle.

Bar: the device batch's CSR offsets, m/z and intensities bit-identical to the
oracle's pickpeaks (the reference's quicksort tie order included), and
searches of the GPU-preprocessed batch byte-identical to searches of the
host-picked batch uploaded as is. Cases: tie-heavy random batches with empty,
single-peak, all-equal and descending spectra; QALEN 5, 100 and 2048; the
reference's sample MS2 spectra; the uneven workload's raw spectra (20-2000
peaks); bad parameters.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import Searcher, SearchParams
from oracle import pyoracle as O

from test_pickpeaks import tie_heavy_batch  # noqa: E402

pytestmark = pytest.mark.gpu

SAMPLE = Path(__file__).parent / "golden" / "olsen_s01_extracted.ms2"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def searcher(small_index):
    s = Searcher(small_index)
    yield s
    s.close()


def check_device_batch(dq, idx, moz, inten, label):
    ns, npk, ms = dq.shape()
    assert ns == len(idx) - 1 and npk == int(idx[-1]), label
    gi, gm, gn = dq.peaks()
    assert np.array_equal(gi, idx), f"{label}: offsets differ"
    bad = np.nonzero(gn != inten)[0]
    assert len(bad) == 0, f"{label}: intensities differ at {bad[:8]}"
    bad = np.nonzero(gm != moz)[0]
    assert len(bad) == 0, f"{label}: m/z differ at {bad[:8]}"


@pytest.mark.parametrize("qalen, hi", [(100, 20), (2048, 20), (100, 10**5), (5, 3)])
def test_gpu_pick_matches_oracle(searcher, qalen, hi):
    say(f"case: GPU peak picking, qalen {qalen}, intensities < {hi}")
    raw = tie_heavy_batch(qalen + hi, hi=hi)
    pick = slm.PickParams(100_000, 1000, qalen)
    dq = searcher.preprocess(raw, pick)
    idx, moz, inten = O.pickpeaks(raw, 100_000, 1000, qalen)
    check_device_batch(dq, idx, moz, inten, f"qalen {qalen}")
    dq.free()


def test_gpu_pick_sample_ms2(searcher):
    say("case: GPU peak picking of the sample MS2 spectra")
    raw = slm.read_ms2(SAMPLE)
    dq = searcher.preprocess(raw)
    idx, moz, inten = O.pickpeaks(raw)
    check_device_batch(dq, idx, moz, inten, "sample")
    dq.free()


def test_gpu_pick_then_search_equals_host_pick(small_index, searcher):
    say("case: GPU-preprocessed uneven batch searched like the host-picked one")
    sc = slm.SpecConfig(n_spectra=2000, seed=8, peak_mode=F.GX_PEAKS_TOTAL_POWER, peaks_min=20, peaks_max=2000,
                        power_alpha=2.0, charge_min=2, charge_max=3, qalen=2048, dm_choices=(10.0, 100.0, 500.0))
    raw = slm.generate_raw_spectra(small_index, sc)
    picked = slm.generate_spectra(small_index, sc)
    pick = slm.PickParams(sc.base_int, sc.min_int, sc.qalen)
    dq = searcher.preprocess(raw, pick)
    check_device_batch(dq, picked.idx, picked.moz, picked.intensity, "uneven")
    p = SearchParams(dM=500.0)
    searcher.search_device(dq, p)
    a, ha = searcher.download(dq, want_hist=True)
    b, hb, _ = searcher.search(picked, p, want_hist=True)
    assert a.tobytes() == b.tobytes() and np.array_equal(ha, hb)
    dq.free()


def test_gpu_pick_rejects_bad_parameters(searcher):
    say("case: GPU peak picking error codes")
    raw = tie_heavy_batch(1, n_spectra=10)
    for pick in [slm.PickParams(100_000, 1000, 0), slm.PickParams(100_000, 1000, 4096), slm.PickParams(0, 1000, 100)]:
        with pytest.raises(F.GxError) as e:
            searcher.preprocess(raw, pick)
        assert e.value.code == F.GX_ERR_INVLD_PARAM
    empty = slm.QueryBatch(np.zeros(1, np.uint32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                           np.zeros(0, np.float32))
    dq = searcher.preprocess(empty)
    assert dq.shape()[:2] == (0, 0)
    dq.free()
