"""Pins of the oracle's building blocks against independent answers.

- bound search: the reference's testapp known-answer vectors
  (source/apps/testapp/testapp.cpp:45,60-81, expected by bisect);
- precursor window: brute force over float32 masses (dslim_query.cpp:1102-1164);
- lgfact: 40-digit decimal logarithms (utils.h:115-131);
- tail fit: closed forms on hand-built survival histograms, one per branch
  of expeRT::ModelTailFit (expeRT.cpp:396-559).
"""
import json
import math
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as O

GOLD = Path(__file__).parent / "golden"


def test_testapp_kat_bounds():
    d = json.loads((GOLD / "testapp_kat.json").read_text())
    arr = np.array(d["array"], dtype=np.uint32)
    # the two targets the reference's testapp prints
    assert O.lower_bound(arr, d["reference_targets"]["lower_bound"]) == 318
    assert O.upper_bound(arr, d["reference_targets"]["upper_bound"]) == 319
    for c in d["cases"]:
        assert O.lower_bound(arr, c["target"]) == c["lower_bound"], c
        assert O.upper_bound(arr, c["target"]) == c["upper_bound"], c


def test_bounds_empty_and_duplicates():
    assert O.lower_bound(np.zeros(0, np.uint32), 5) == 0
    a = np.array([1, 3, 3, 3, 7], np.uint32)
    assert O.lower_bound(a, 3) == 1 and O.upper_bound(a, 3) == 4
    assert O.lower_bound(a, 8) == 5 and O.upper_bound(a, 0) == 0


def _window_brute(mass, p, dM):
    pm1 = np.float32(np.float64(np.float32(p)) - dM)
    pm2 = np.float32(np.float64(np.float32(p)) + dM)
    if dM < 0 or len(mass) == 0:
        return False, None
    if pm1 > mass[-1] or pm2 < mass[0]:
        return False, None
    inside = np.nonzero((mass >= pm1) & (mass <= pm2))[0]
    if len(inside) == 0:
        return "empty", None
    return True, (int(inside[0]), int(inside[-1]))


def test_window_matches_brute_force():
    rng = np.random.default_rng(5)
    for trial in range(300):
        n = int(rng.integers(1, 60))
        mass = np.sort(rng.choice(np.linspace(500, 700, 120), size=n)).astype(np.float32)
        p = float(rng.uniform(480, 720))
        dM = float(rng.choice([0.0, 0.5, 3.0, 10.0, 50.0, -1.0]))
        ok, lo, hi = O.window(mass, p, dM)
        exp, rng_ = _window_brute(mass, p, dM)
        if exp is True:
            assert ok and (lo, hi) == rng_, (mass, p, dM)
        else:
            assert (not ok) or hi < lo, (mass, p, dM, ok, lo, hi)


def test_window_exact_edges():
    mass = np.array([500.0, 510.0, 510.0, 520.0], np.float32)
    ok, lo, hi = O.window(mass, 510.0, 0.0)
    assert ok and (lo, hi) == (1, 2)
    ok, lo, hi = O.window(mass, 505.0, 5.0)  # [500, 510] inclusive on both sides
    assert ok and (lo, hi) == (0, 2)


def test_lgfact_closed_form():
    getcontext().prec = 40
    lg = O.lgfact(400)
    acc = Decimal(0)
    for i in range(400):
        if i >= 2:
            acc += Decimal(i).log10()
        # within a few ulps of the exactly rounded value
        assert abs(lg[i] - float(acc)) <= 4 * math.ulp(max(float(acc), 1.0)), i


def _hist(pairs):
    h = np.zeros(802, np.uint32)
    for b, c in pairs:
        h[b] += c
    return h


def test_tailfit_geometric_tail_has_closed_form_slope():
    # counts v/2, v/4, ... make s(x) = 2^-(i+1): log10 s is exactly linear
    v = 1 << 12
    stt = 30
    pairs = [(stt + i, v >> (i + 1)) for i in range(10)]
    top = stt + 20
    used = sum(c for _, c in pairs)
    pairs.append((top, v - used))
    ok, slope, icept, mu, beta = O.tailfit(_hist(pairs), top, v)
    assert ok
    assert abs(slope - (-math.log10(2))) < 1e-12
    # X starts at stt1: log10 s(X) = -(X - stt + 1) log10 2
    assert abs(icept - (stt - 1) * math.log10(2)) < 1e-10
    assert mu == int(slope * 1e6) and beta == int(icept * 1e6)


def test_tailfit_two_point_replacement_branch():
    # end1 == stt1 -> end1 += 1 pulls the top bin in; s = [0.25, 0] -> 0 replaced by 0.25
    ok, slope, icept, mu, beta = O.tailfit(_hist([(5, 3), (6, 1)]), 6, 4)
    assert ok
    assert slope == 0.0
    assert abs(icept - math.log10(0.25)) < 1e-15
    assert (mu, beta) == (0, int(math.log10(0.25) * 1e6))


def test_tailfit_three_point_branch():
    # n == 3: mark = first s <= s0 + 0.22*hgt, minus one; mark2 = 2
    h = _hist([(10, 4), (11, 2), (12, 1), (20, 1)])
    ok, slope, icept, mu, beta = O.tailfit(h, 20, 8)
    s = np.array([1 - 4 / 8, 1 - 6 / 8, 1 - 7 / 8])
    ls = np.log10(s)
    thr = ls[0] + (ls[-1] - ls[0]) * 0.22
    mark = int(np.nonzero(ls <= thr)[0][0]) - 1
    x = np.arange(10 + mark, 13, dtype=float)
    a, b = np.polyfit(x, ls[mark:], 1)
    assert ok and abs(slope - a) < 1e-12 and abs(icept - b) < 1e-10


def test_tailfit_flat_slice_is_undefined():
    # every candidate in the top bin: slice of zeros, hgt = 0, mark = -1
    ok, *_ = O.tailfit(_hist([(40, 6)]), 40, 6)
    assert not ok


def test_tailfit_hyp_zero_is_undefined():
    ok, *_ = O.tailfit(_hist([(0, 5)]), 0, 5)
    assert not ok


def test_tailfit_general_matches_polyfit_on_marked_region():
    # a decaying histogram; marks recomputed independently from their definition
    rng = np.random.default_rng(3)
    for trial in range(50):
        stt = int(rng.integers(0, 30))
        n = int(rng.integers(5, 40))
        counts = np.maximum(1, (2000 * np.exp(-0.3 * np.arange(n)) * rng.uniform(0.7, 1.3, n)).astype(int))
        top = stt + n + int(rng.integers(1, 5))
        pairs = [(stt + i, int(c)) for i, c in enumerate(counts)] + [(top, 1)]
        vaa = int(counts.sum()) + 1
        ok, slope, icept, mu, beta = O.tailfit(_hist(pairs), top, vaa)
        s = 1 - np.cumsum(counts) / vaa
        ls = np.log10(s)
        hgt = ls[-1] - ls[0]
        mark = int(np.nonzero(ls <= ls[0] + hgt * 0.22)[0][0]) - 1
        mark2 = int(np.nonzero(ls >= ls[0] + hgt * 0.87)[0][-1])
        if mark >= mark2:
            mark = mark2 - 1
        if mark < 0:
            assert not ok
            continue
        x = np.arange(stt + mark, stt + mark2 + 1, dtype=float)
        if len(x) == 1:
            a, b = 0.0, ls[mark]
        else:
            a, b = np.polyfit(x, ls[mark:mark2 + 1], 1)
        assert ok
        assert abs(slope - a) < 1e-9 * max(1, abs(a)) and abs(icept - b) < 1e-8 * max(1, abs(b)), trial
