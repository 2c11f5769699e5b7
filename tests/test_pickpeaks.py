"""Peak picking (MSQuery::pickpeaks, msquery.cpp:338-423), the MS2 reader and
the .pbin format on the CPU.

The oracle's pickpeaks is pinned by partition orders of tied intensities
traced by hand through the reference's quicksort (keyval.h:49-103), by numpy's
sort where intensities are distinct, and by hand-computed normalisations.
The host library's picking (the CPU path and the synthetic generator) must
equal the oracle on tie-heavy batches; the MS2 reader is checked against an
independent parse of the reference's sample spectra
(samples/sample_data/14Sep18_Olsen_S01_extracted.ms2, kept as a fixture) and
hand-made edge cases; the .pbin writer against the documented byte layout.
"""
import struct
from pathlib import Path

import numpy as np
import pytest

from gicops_amd import slm
from oracle import pyoracle as O

GOLD = Path(__file__).parent / "golden"
SAMPLE = GOLD / "olsen_s01_extracted.ms2"
INT_MIN = -2**31


# ---- the oracle ------------------------------------------------------------

@pytest.mark.parametrize("keys, vals, want_keys, want_vals", [
    # pivot key[1]; one swap of the ends, then the middle with itself
    ([5, 5, 5], [1, 2, 3], [5, 5, 5], [3, 2, 1]),
    # pivot key[0]; the two ends swap
    ([5, 5], [1, 2], [5, 5], [2, 1]),
    # pivot 1 at index 1: swap(0,1) -> [1,3,3,2]; right part [3,3,2], pivot
    # key[2] = 3: swap(1,3) -> [1,2,3,3] with the index-2 "3" before the
    # index-0 one (a stable sort would keep them the other way round)
    ([3, 1, 3, 2], [10, 11, 12, 13], [1, 2, 3, 3], [11, 13, 12, 10]),
    # pivot key[2] = 2: i stops at 0, j at 4: swap, values [4,1,2,3,0]; i
    # stops at 1, j at 3 (the 1): swap, keys [2,1,2,2,2], values [4,3,2,1,0];
    # i = j = 2: self swap, i=3, j=1. Left [0,1] = keys [2,1]: pivot 2, swap
    # -> values [3,4]. Right [3,4] = keys [2,2]: swap -> values [0,1].
    ([2, 2, 2, 1, 2], [0, 1, 2, 3, 4], [1, 2, 2, 2, 2], [3, 4, 2, 0, 1]),
])
def test_oracle_tie_order_hand_traced(keys, vals, want_keys, want_vals):
    k, v, kept = O.pickpeaks_one(keys, vals, base_int=10**6, min_int=-2**31, qalen=10_000)
    # the m/z values travel with their intensities: their order is the sort's
    assert list(v) == want_vals
    assert kept == len(keys)


def test_oracle_distinct_intensities_match_numpy_sort():
    rng = np.random.default_rng(5)
    for n in [1, 2, 3, 17, 100, 101, 257, 2000]:
        inten = rng.choice(10**6, n, replace=False).astype(np.int32) + 1
        mz = rng.integers(0, 500_000, n).astype(np.int32)
        k, m, kept = O.pickpeaks_one(inten, mz, base_int=100_000, min_int=0, qalen=10**6)
        order = np.argsort(inten, kind="stable")
        assert np.array_equal(m, mz[order])
        assert kept == n


def test_oracle_normalisation_by_hand():
    # intensities 1000, 500, 20, 7 -> factor 100000/1000 = 100: 50000, 2000 kept, 700 < 1000 stops
    k, m, kept = O.pickpeaks_one([500, 7, 1000, 20], [1, 2, 3, 4], 100_000, 1000, 100)
    assert kept == 3
    assert list(k[-3:]) == [2000, 50000, 100000] and list(m[-3:]) == [4, 1, 3]
    # qalen 3 of 5: only the top 3 are scaled and kept
    k, m, kept = O.pickpeaks_one([10, 20, 30, 40, 50], [1, 2, 3, 4, 5], 1000, 0, 3)
    assert kept == 3 and list(k[-3:]) == [600, 800, 1000] and list(m[-3:]) == [3, 4, 5]
    # factor 100000/3 = 33333.33..: 2 -> 66666, 1 -> 33333 (truncation)
    k, m, kept = O.pickpeaks_one([1, 3, 2], [7, 8, 9], 100_000, 0, 100)
    assert list(k) == [33333, 66666, 100000] and kept == 3
    # empty and single-peak spectra
    assert O.pickpeaks_one([], [])[2] == 0
    k, m, kept = O.pickpeaks_one([42], [9], 100_000, 1000, 100)
    assert kept == 1 and list(k) == [100000]
    # a zero base peak: 0 * inf is NaN, stored as INT32_MIN (x86), so only the base peak stays
    k, m, kept = O.pickpeaks_one([0, 0, 0], [1, 2, 3], 100_000, 1000, 100)
    assert kept == 1 and k[-1] == 100000 and k[-2] == INT_MIN


# ---- host library vs oracle -------------------------------------------------

def tie_heavy_batch(seed, n_spectra=400, max_len=300, hi=20):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len, n_spectra)
    lens[:6] = [0, 1, 2, 3, max_len, 150]
    idx = np.zeros(n_spectra + 1, np.uint32)
    idx[1:] = np.cumsum(lens)
    n = int(idx[-1])
    inten = rng.integers(1, hi, n).astype(np.int32)
    inten[idx[4]:idx[5]] = 7                     # one all-equal spectrum
    inten[idx[5]:idx[6]] = np.arange(150, 0, -1)  # a descending one
    moz = rng.integers(0, 500_000, n).astype(np.int32)
    prec = rng.uniform(600, 3000, n_spectra).astype(np.float32)
    return slm.QueryBatch(idx, moz, inten, prec)


def assert_same_picked(got, idx, moz, inten, label):
    assert np.array_equal(got.idx, idx), label
    assert np.array_equal(got.moz, moz), label
    assert np.array_equal(got.intensity, inten), label


@pytest.mark.parametrize("qalen, hi", [(100, 20), (2048, 20), (100, 10**5), (5, 3)])
def test_host_pick_matches_oracle(qalen, hi):
    raw = tie_heavy_batch(qalen + hi, hi=hi)
    pp = slm.PickParams(base_int=100_000, min_int=1000, qalen=qalen)
    got = slm.pick_peaks(raw, pp)
    idx, moz, inten = O.pickpeaks(raw, 100_000, 1000, qalen)
    assert_same_picked(got, idx, moz, inten, f"qalen {qalen}, intensities < {hi}")
    assert np.array_equal(got.precurse, raw.precurse)


def test_generator_picks_its_raw_spectra(small_index):
    sc = slm.SpecConfig(n_spectra=300, seed=4, peak_mode=2, peaks_min=20, peaks_max=400, qalen=100)
    picked = slm.generate_spectra(small_index, sc)
    raw = slm.generate_raw_spectra(small_index, sc)
    assert raw.n_spectra == picked.n_spectra and raw.n_peaks >= picked.n_peaks
    assert int(raw.intensity.max()) == 10000  # base peak 100 x YAXISMULTIPLIER
    again = slm.pick_peaks(raw, slm.PickParams(sc.base_int, sc.min_int, sc.qalen))
    assert_same_picked(again, picked.idx, picked.moz, picked.intensity, "generate == pick(generate_raw)")
    idx, moz, inten = O.pickpeaks(raw, sc.base_int, sc.min_int, sc.qalen)
    assert_same_picked(picked, idx, moz, inten, "generate == oracle")
    assert np.array_equal(raw.precurse, picked.precurse) and np.array_equal(raw.truth_psid, picked.truth_psid)


# ---- MS2 reader --------------------------------------------------------------

def parse_ms2_independently(text, scale=100):
    """A straightforward parse of the MS2 text with Python floats."""
    specs = []
    for line in text.splitlines():
        if not line or line[0] in "HD":
            continue
        if line[0] == "S":
            specs.append({"z": 1, "prec": np.float32(0.01), "rt": np.float32(0.0), "mz": [], "in": []})
        elif not specs:
            continue
        elif line[0] == "Z":
            f = line.split()
            specs[-1]["z"] = max(1, int(f[1]))
            specs[-1]["prec"] = np.float32(float(f[2]))
        elif line[0] == "I":
            f = line.split()
            if f[1] == "RTime":
                specs[-1]["rt"] = np.float32(max(0.0, float(f[2])))
        else:
            a, b = line.split(" ")
            specs[-1]["mz"].append(int(float(a) * scale))
            specs[-1]["in"].append(int(float(b) * 100.0))
    return specs


def test_ms2_reader_on_reference_sample():
    raw = slm.read_ms2(SAMPLE, scale=100)
    want = parse_ms2_independently(SAMPLE.read_text())
    assert raw.n_spectra == len(want) == 4
    for s, w in enumerate(want):
        m, i = raw.peaks(s)
        assert list(m) == w["mz"] and list(i) == w["in"]
        assert raw.charges[s] == w["z"] and raw.precurse[s] == w["prec"] and raw.rtimes[s] == w["rt"]
    assert raw.charges[0] == 2 and raw.precurse[0] == np.float32(788.4626)
    assert raw.rtimes[0] == np.float32(28.66084) and raw.peaks(0)[0][0] == 10106 and raw.peaks(0)[1][0] == 24237
    # picked: at most QALEN peaks each, intensities ascending, the base peak 100000
    picked = slm.pick_peaks(raw)
    idx, moz, inten = O.pickpeaks(raw)
    assert_same_picked(picked, idx, moz, inten, "sample")
    for s in range(picked.n_spectra):
        m, i = picked.peaks(s)
        assert 0 < len(i) <= 100 and i[-1] == 100000 and np.all(np.diff(i) >= 0) and np.all(i >= 1000)


def test_ms2_reader_edge_cases(tmp_path):
    text = "\n".join([
        "H\tCreationDate x",
        "100.5 7",             # before the first S: ignored
        "S\t1\t1\t0",
        "Z\t3",                # no precursor field: 0.01
        "I\tRTime\t-4.5",      # negative: 0
        "200.25 10.5",
        "300.5",               # no intensity field: 0.01 -> 1
        "400.75\t12",          # tab inside a peak line: one token
        "S\t2\t2\t0",          # no Z, no RTime
        "D\tsomething",
        "150.015 3.999",
        "",
    ])
    f = tmp_path / "edge.ms2"
    f.write_text(text)
    raw = slm.read_ms2(f, scale=100)
    assert raw.n_spectra == 2
    assert raw.charges[0] == 3 and raw.precurse[0] == np.float32(0.01) and raw.rtimes[0] == 0.0
    assert list(raw.peaks(0)[0]) == [20025, 30050, 40075]
    assert list(raw.peaks(0)[1]) == [1050, 1, 1]
    assert raw.charges[1] == 1 and raw.precurse[1] == np.float32(0.01) and raw.rtimes[1] == 0.0
    assert list(raw.peaks(1)[0]) == [int(150.015 * 100)] and list(raw.peaks(1)[1]) == [399]
    empty = tmp_path / "none.ms2"
    empty.write_text("H\tonly headers\n")
    assert slm.read_ms2(empty).n_spectra == 0


# ---- .pbin -------------------------------------------------------------------

def test_pbin_layout_and_round_trip(tmp_path):
    raw = slm.read_ms2(SAMPLE)
    picked = slm.pick_peaks(raw)
    f = tmp_path / "sample.pbin"
    slm.write_pbin(f, picked)
    blob = f.read_bytes()
    want = b""
    for s in range(picked.n_spectra):
        m, i = picked.peaks(s)
        want += struct.pack("<fifi", picked.precurse[s], int(picked.charges[s]), picked.rtimes[s], len(m))
        want += struct.pack(f"<{len(m)}i", *m.tolist()) + struct.pack(f"<{len(i)}i", *i.tolist())
    assert blob == want
    back = slm.read_pbin(f)
    assert_same_picked(back, picked.idx, picked.moz, picked.intensity, "pbin round trip")
    assert np.array_equal(back.precurse, picked.precurse) and np.array_equal(back.charges, picked.charges)
    assert np.array_equal(back.rtimes, picked.rtimes)
    # a truncated file is an error, not a short batch
    (tmp_path / "cut.pbin").write_bytes(blob[:-3])
    with pytest.raises(Exception):
        slm.read_pbin(tmp_path / "cut.pbin")
