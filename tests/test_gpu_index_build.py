"""GPU construction of the fragment-ion index (gx_index_build) against the
host build of the same entries.

Bar: bA and iA of every bucket bit-identical to the CPU build (which the
oracle tests pin by closed forms and brute force), and searches over the
device-built index byte-identical to searches over the uploaded one.
Covered: unmodified peptides; variable modifications of several types with
maxz 3 and every modification's mass in the precursor; N-terminal acetylation; a narrow precursor
range that drops entries (zeroed fragments in bin 0); an entries-only host
index (no host fragment arrays at all); bad build parameters.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import Searcher, SearchParams

pytestmark = pytest.mark.gpu

MODS3 = slm.DBConfig(n_peptides=30_000, seed=5, maxz=3,
                     mods=[("M", 15.9949, 2), ("STY", 79.96633, 1), ("NQ", 0.984016, 1)],
                     max_mods_per_pep=3, modmass_all=True)
NARROW = slm.DBConfig(n_peptides=20_000, seed=11, maxz=2, min_mass=900.0, max_mass=2500.0,
                      mods=[("M", 15.9949, 1)], max_mods_per_pep=1)
NTERM = slm.DBConfig(n_peptides=20_000, seed=17, mods=[("M", 15.9949, 2), ("^", 42.0106, 1)],
                     max_mods_per_pep=3)
CASES = {"small": slm.DB_PRESETS["small"], "mods3_z3": MODS3, "narrow_mass": NARROW, "nterm_acetyl": NTERM}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_same_index(gs, index, label):
    for i, b in enumerate(index.buckets):
        bA, iA = gs.index_arrays(i)
        if b.n_entries == 0:
            continue
        assert np.array_equal(bA, b.bA), f"{label} bucket {i} (L={b.peplen}): bA differs at " \
                                         f"{np.nonzero(bA != b.bA)[0][:8]}"
        bad = np.nonzero(iA != b.iA)[0]
        assert len(bad) == 0, f"{label} bucket {i} (L={b.peplen}): iA differs at {bad[:8]}"


@pytest.mark.parametrize("case", list(CASES))
def test_device_build_matches_host_build(case):
    cfg = CASES[case]
    say(f"case: index build {case}")
    index = slm.SlmIndex(cfg)
    gs = Searcher(index, build_on_device=True)
    assert gs.built_on_device
    assert_same_index(gs, index, case)
    # searches over both indexes agree byte for byte
    sc = slm.SpecConfig(n_spectra=300, seed=3, charge_min=2, charge_max=3, dm_choices=(10.0, 100.0))
    batch = slm.generate_spectra(index, sc)
    hs = Searcher(index)
    p = SearchParams(dM=100.0, max_mass=int(cfg.max_mass))
    a, ha, sa = gs.search(batch, p, want_hist=True)
    b, hb, sb = hs.search(batch, p, want_hist=True)
    assert a.tobytes() == b.tobytes() and np.array_equal(ha, hb)
    assert sa["ions"] == sb["ions"] and sa["errors"] == 0
    gs.close()
    hs.close()


def test_entries_only_index_builds_on_device(small_index):
    say("case: entries-only index, fragments built on the GPU")
    eo = slm.SlmIndex(slm.DB_PRESETS["small"], fragments=False)
    assert not eo.has_fragments and all(b.iA.size == 0 for b in eo.buckets)
    with pytest.raises(ValueError):
        Searcher(eo, build_on_device=False)
    gs = Searcher(eo)
    assert gs.built_on_device
    for i, b in enumerate(small_index.buckets):
        assert np.array_equal(eo.buckets[i].mass, b.mass)
    assert_same_index(gs, small_index, "entries-only")
    batch = slm.generate_spectra(small_index, slm.SPEC_PRESETS["small"])
    hs = Searcher(small_index)
    a, _, _ = gs.search(batch, SearchParams(dM=10.0))
    b, _, _ = hs.search(batch, SearchParams(dM=10.0))
    assert a.tobytes() == b.tobytes()
    gs.close()
    hs.close()


def test_build_rejects_bad_parameters(tiny_index):
    say("case: index build error codes")
    lib = F.hip_lib()
    ctx = C.c_void_p()
    F.check("gx_ctx_create", lib.gx_ctx_create(0, C.byref(ctx)))
    try:
        d = C.c_void_p()
        bp = tiny_index.build_params()
        nb = len(tiny_index.buckets)
        for field, value in [("maxz", 0), ("scale", 0), ("n_mod_types", 16)]:
            bad = F.gx_index_build_params.from_buffer_copy(bp)
            setattr(bad, field, value)
            assert lib.gx_index_build(ctx, tiny_index.raw_array, nb, C.byref(bad), C.byref(d)) == F.GX_ERR_INVLD_PARAM
        # a mass range that does not give the buckets' bin count
        bad = F.gx_index_build_params.from_buffer_copy(bp)
        bad.max_mass = bp.max_mass + 1.0
        assert lib.gx_index_build(ctx, tiny_index.raw_array, nb, C.byref(bad), C.byref(d)) == F.GX_ERR_INVLD_INDEX
        assert lib.gx_index_build(ctx, tiny_index.raw_array, nb, None, C.byref(d)) == F.GX_ERR_INVLD_PTR
        assert not d.value
        F.check("gx_index_build", lib.gx_index_build(ctx, tiny_index.raw_array, nb, C.byref(bp), C.byref(d)))
        assert lib.gx_index_download(d, nb, None, None) == F.GX_ERR_INVLD_INDEX
        lib.gx_index_free(d)
    finally:
        lib.gx_ctx_destroy(ctx)
