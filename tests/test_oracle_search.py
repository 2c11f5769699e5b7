"""The oracle's search against brute force, and the index builder's invariants.

Brute force: for every entry in the window, count the legal query peaks within
+-dF bins of each of its fragment bins (taken from the inverted index), i.e.
what the scatter of dslim_query.cpp:853-917 sums, computed a different way.
"""
import numpy as np
import pytest

from gicops_amd import slm
from oracle import pyoracle as O


def invert(bucket):
    """ion id -> bin, from bA/iA"""
    counts = np.diff(bucket.bA.astype(np.int64))
    bins = np.repeat(np.arange(bucket.nbins, dtype=np.int64), counts)
    ion_bin = np.empty(bucket.n_ions, dtype=np.int64)
    ion_bin[bucket.iA] = bins
    return ion_bin


def brute_scorecard(bucket, ion_bin, moz, inten, lo, hi, dF=2, nbins=500000):
    legal = (moz > dF) & (moz < nbins - 1 - dF)
    q = moz[legal].astype(np.int64)
    w = inten[legal].astype(np.int64)
    order = np.argsort(q, kind="stable")
    q, w = q[order], w[order]
    cw = np.concatenate([[0], np.cumsum(w)])
    S = bucket.speclen
    ids = (np.arange(lo, hi + 1)[:, None] * S + np.arange(S)[None, :])
    x = ion_bin[ids]
    a = np.searchsorted(q, x - dF, side="left")
    b = np.searchsorted(q, x + dF, side="right")
    n = b - a
    s = cw[b] - cw[a]
    half = S // 2
    return (n[:, :half].sum(1), n[:, half:].sum(1), s[:, :half].sum(1), s[:, half:].sum(1))


def test_index_invariants(small_index):
    for b in small_index.buckets:
        if b.n_entries == 0:
            continue
        assert b.bA[0] == 0 and b.bA[-1] == b.n_ions
        assert np.all(np.diff(b.bA.astype(np.int64)) >= 0)
        assert np.all(np.diff(b.mass) >= 0)
        # iA is a permutation of all ion ids, ascending within every bin
        assert np.array_equal(np.sort(b.iA), np.arange(b.n_ions, dtype=np.uint32))
        d = np.diff(b.iA.astype(np.int64))
        starts = b.bA[1:-1][np.diff(b.bA[:-1].astype(np.int64)) >= 0]
        mask = np.ones(len(d), bool)
        bounds = b.bA.astype(np.int64)
        inner = bounds[(bounds > 0) & (bounds < b.n_ions)] - 1
        mask[inner] = False
        assert np.all(d[mask] > 0)


def test_fragment_bins_closed_form():
    """b/y bins of hand-made peptides vs exact decimal masses (float rounding
    may move a value sitting on a bin edge by one bin)."""
    from decimal import Decimal as D
    aa = {"A": "71.03712", "C": "160.030654", "D": "115.030", "E": "129.0426", "F": "147.068", "G": "57.02146",
          "H": "137.060", "I": "113.084", "K": "128.094", "L": "113.084", "M": "131.0405", "N": "114.043",
          "P": "97.0527", "Q": "128.05858", "R": "156.1012", "S": "87.032", "T": "101.0476", "V": "99.06841",
          "W": "186.0793", "Y": "163.0633"}
    peps = ["PEPTIDEK", "ACDEFGHIKR", "MMMMMMMK", "WYVTSRQPNK", "GGGGGGGGGR"]
    idx = slm.SlmIndex(slm.DBConfig(min_len=8, max_len=10), peptides=peps)
    proton, h2o = D("1.00727647"), D("18.015")
    agree = total = 0
    for b in idx.buckets:
        if b.n_entries == 0:
            continue
        inv = invert(b)
        for psid in range(b.n_entries):
            seq = b.sequence(psid)
            L = len(seq)
            ms = [D(aa[c]) for c in seq]
            bions = [sum(ms[:i + 1]) + proton for i in range(L - 1)]
            yions = [sum(ms[L - 1 - i:]) + h2o + proton for i in range(L - 1)]
            exp = [int(v * 100) for v in bions + yions]
            got = inv[psid * b.speclen:(psid + 1) * b.speclen]
            assert np.all(np.abs(got - np.array(exp)) <= 1), seq
            agree += int(np.sum(got == np.array(exp)))
            total += len(exp)
            mass = float(sum(ms) + h2o)
            assert abs(float(b.mass[psid]) - mass) < 1e-3
    assert agree >= 0.9 * total


def test_variant_enumeration_order_and_masses():
    idx = slm.SlmIndex(slm.DBConfig(min_len=8, max_len=8, mods=[("M", 15.9949, 2)], max_mods_per_pep=2),
                       peptides=["AMAMAMAK"])
    b = idx.buckets[0]
    # 3 single + 3 double oxidations + the peptide itself
    assert b.n_entries == 7
    sites = sorted(int(s) for s in b.sites)
    assert sites == sorted([0, 2, 8, 32, 2 | 8, 2 | 32, 8 | 32])
    base = float(b.mass[b.sites == 0][0])
    for m, s in zip(b.mass, b.sites):
        n = bin(int(s)).count("1")
        # reference quirk: the precursor mass adds only the first modification
        assert abs(float(m) - base - (15.99 if n else 0.0)) < 2e-3


def test_oracle_scorecard_matches_brute_force(tiny_index):
    batch = slm.generate_spectra(tiny_index, slm.SPEC_PRESETS["tiny"])
    p = O.default_params(dM=10.0)
    checked = 0
    for s in range(batch.n_spectra):
        moz, inten = batch.peaks(s)
        for b in tiny_index.buckets:
            if b.n_entries == 0:
                continue
            ok, lo, hi = O.window(b.mass, float(batch.precurse[s]), 10.0)
            if not ok or hi < lo:
                continue
            bc, yc, ibc, iyc = O.scorecard(b, moz, inten, p, lo, hi)
            inv = invert(b)
            ebc, eyc, eibc, eiyc = brute_scorecard(b, inv, moz, inten, lo, hi)
            assert np.array_equal(bc, ebc) and np.array_equal(yc, eyc)
            assert np.array_equal(ibc, eibc) and np.array_equal(iyc, eiyc)
            checked += 1
    assert checked > 50


def test_oracle_search_finds_source_peptide(small_index, small_batch):
    psm, hist, cnt = O.search(small_index, small_batch, O.default_params(dM=10.0), want_hist=True,
                              want_counters=True)
    hit = (psm["idxoffset"] == small_batch.truth_bucket) & (psm["psid"] == small_batch.truth_psid)
    assert hit.mean() > 0.95
    # histogram totals are the candidate counts
    assert np.array_equal(hist.sum(1), psm["cpsms"])
    # the top hit sits in the histogram's highest non-empty bin
    top = np.array([np.nonzero(h)[0].max() if h.any() else 0 for h in hist])
    has = psm["cpsms"] > 0
    assert np.array_equal(top[has], psm["hyp"][has])


def test_oracle_thread_count_does_not_change_results(small_index, small_batch):
    sub = small_batch.subset(np.arange(200))
    a, ha, _ = O.search(small_index, sub, O.default_params(dM=10.0), threads=1, want_hist=True)
    b, hb, _ = O.search(small_index, sub, O.default_params(dM=10.0), threads=8, want_hist=True)
    assert a.tobytes() == b.tobytes() and np.array_equal(ha, hb)


def test_hyperscore_closed_form(tiny_index):
    """hyperscore = log10(bc!) + log10(yc!) + log10(1+ibc) + log10(1+iyc) - 4
    (dslim_query.cpp:938), recomputed from the oracle's own scorecard."""
    batch = slm.generate_spectra(tiny_index, slm.SPEC_PRESETS["tiny"])
    p = O.default_params(dM=10.0)
    psm, _, _ = O.search(tiny_index, batch, p)
    import math
    for s in np.nonzero(psm["cpsms"] > 0)[0][:20]:
        r = psm[s]
        b = tiny_index.buckets[r["idxoffset"]]
        moz, inten = batch.peaks(int(s))
        bc, yc, ibc, iyc = O.scorecard(b, moz, inten, p, int(r["psid"]), int(r["psid"]))
        h = (math.lgamma(bc[0] + 1) + math.lgamma(yc[0] + 1)) / math.log(10) \
            + math.log10(1 + int(ibc[0])) + math.log10(1 + int(iyc[0])) - 4
        assert abs(min(h, 79.0) - float(r["hyperscore"])) < 1e-5 * max(1.0, h)
        assert int(bc[0] + yc[0]) == int(r["sharedions"])


def test_entries_only_index_and_build_params():
    """gx_db_finalize_entries gives the same mass-sorted entries as the full
    build, without host fragment arrays; gx_db_build_params reports the
    build's charges, scale, mass range and modification masses."""
    cfg = slm.DBConfig(n_peptides=3_000, seed=9, maxz=2, mods=[("M", 15.9949, 2), ("STY", 79.96633, 1)],
                       max_mods_per_pep=2)
    full = slm.SlmIndex(cfg)
    eo = slm.SlmIndex(cfg, fragments=False)
    assert full.has_fragments and not eo.has_fragments
    assert len(full.buckets) == len(eo.buckets)
    for a, b in zip(full.buckets, eo.buckets):
        assert a.n_entries == b.n_entries and a.n_ions == b.n_ions
        for f in ("mass", "seqid", "modnum", "sites", "seqs"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        assert b.bA.size == 0 and b.iA.size == 0
    bp = eo.build_params()
    assert (bp.maxz, bp.scale, bp.min_mass, bp.max_mass) == (2, 100, 500.0, 5000.0)
    assert bp.n_mod_types == 2 and bp.modmass_all == 0
    assert list(bp.mod_mass[:2]) == [int(float(np.float32(15.9949)) * 100.0), int(float(np.float32(79.96633)) * 100.0)]


def test_nterm_modification_variants_and_fragments():
    """The "^" (N-terminus) modification: a site-0 modification of any residue,
    exclusive with the residue's own; it moves every b-ion and no y-ion."""
    mods = [("M", 15.9949, 2), ("^", 42.0106, 1)]
    idx = slm.SlmIndex(slm.DBConfig(min_len=8, max_len=8, mods=mods, max_mods_per_pep=3),
                       peptides=["AMAMAMAK", "MAAAAAAK"])
    b = idx.buckets[0]
    seqs = [b.sequence(k) for k in range(b.n_entries)]
    # AMAMAMAK: 7 oxidation subsets (<= 2 of 3 sites) x acetyl yes/no; MAAAAAAK: none, ox0, ac0
    assert seqs.count("AMAMAMAK") == 14 and seqs.count("MAAAAAAK") == 3
    a_sites = sorted(int(s) for s, q in zip(b.sites, seqs) if q == "AMAMAMAK")
    ox = [0, 2, 8, 32, 2 | 8, 2 | 32, 8 | 32]
    assert a_sites == sorted(ox + [x | 1 for x in ox])
    m_sites = sorted(int(s) for s, q in zip(b.sites, seqs) if q == "MAAAAAAK")
    assert m_sites == [0, 1, 1]
    m_mod = sorted(int(n) for s, n, q in zip(b.sites, b.modnum, seqs) if q == "MAAAAAAK" and s)
    assert m_mod == [1, 2]  # type 0 (oxidation) and type 1 (acetyl) at site 0
    inv = invert(b)
    L, n = 8, 7
    base = next(k for k in range(b.n_entries) if seqs[k] == "AMAMAMAK" and b.sites[k] == 0)
    acet = next(k for k in range(b.n_entries) if seqs[k] == "AMAMAMAK" and b.sites[k] == 1)
    f0 = inv[base * b.speclen:(base + 1) * b.speclen].astype(np.int64)
    f1 = inv[acet * b.speclen:(acet + 1) * b.speclen].astype(np.int64)
    assert np.all(np.isin(f1[:n] - f0[:n], [4201, 4202]))  # b-ions + 42.0106 Da
    assert np.array_equal(f1[n:], f0[n:])                  # y-ions unchanged
    # precursor: the first modification's mass only (reference quirk) -> + 42.0106
    assert abs(float(b.mass[acet]) - float(b.mass[base]) - 42.0106) < 2e-3

