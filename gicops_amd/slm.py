"""Host-side SLM index and query batches (Python face of libgicops_host.so).

`SlmIndex` owns one database: per-length buckets in the reference's layout
(`Index` / `spmat_t`, slm_dsts.h:298-346). `QueryBatch` is the reference's
`Queries<int>` CSR batch (slm_dsts.h:520-658). Both are plain numpy views
over memory the native library owns or copied out of it.

Workload presets (`DB_PRESETS`, `SPEC_PRESETS`) name the synthetic workloads:
  small   - 50k tryptic peptides, no variable mods, 1,000 spectra, dM +-10 Da
  open2m  - 2M peptides + Met oxidation (<=2), 100k spectra, dM +-500 Da
  uneven  - the open2m index, 200k spectra with power-law peak counts over
            20-2000 and dM drawn from {10, 100, 500} per spectrum
  c4      - 50M peptides + Met oxidation (<=2) + N-terminal acetylation,
            <=2 modifications per peptide (~244M entries, ~12.3G ions; the
            fragment arrays are built on the GPU), 1M spectra of the uneven mix
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _ffi as F


@dataclass
class DBConfig:
    n_peptides: int = 50_000
    seed: int = 0
    min_len: int = 7
    max_len: int = 40
    maxz: int = 1
    scale: int = 100
    min_mass: float = 500.0
    max_mass: float = 5000.0
    mods: list = field(default_factory=list)  # [(residues, mass, max_per_pep)]
    max_mods_per_pep: int = 0
    modmass_all: bool = False
    threads: int = 0


@dataclass
class SpecConfig:
    n_spectra: int = 1000
    seed: int = 0
    keep_frac: float = 0.6
    mz_sigma: float = 0.005
    peak_mode: int = F.GX_PEAKS_NOISE_UNIFORM
    noise_min: int = 20
    noise_max: int = 80
    peaks_min: int = 50
    peaks_max: int = 150
    power_alpha: float = 2.0
    noise_mz_lo: float = 150.0
    noise_mz_hi: float = 2000.0
    int_sigma: float = 1.0
    charge_min: int = 2
    charge_max: int = 2
    qalen: int = 100
    base_int: int = 100_000
    min_int: int = 1000
    dm_choices: tuple = (10.0,)


DB_PRESETS = {
    "tiny": DBConfig(n_peptides=2_000, seed=7),
    "small": DBConfig(n_peptides=50_000, seed=0),
    "open2m": DBConfig(n_peptides=2_000_000, seed=1, mods=[("M", 15.9949, 2)], max_mods_per_pep=2),
    # 50M peptides, Met oxidation (<= 2) and N-terminal acetylation (an extension), <= 2 mods per peptide:
    # ~244M entries, ~12.3G ions; build it with fragments=False and let the GPU build the fragments
    "c4": DBConfig(n_peptides=50_000_000, seed=4, mods=[("M", 15.9949, 2), ("^", 42.0106, 1)], max_mods_per_pep=2),
}

SPEC_PRESETS = {
    "tiny": SpecConfig(n_spectra=64, seed=7, dm_choices=(10.0,)),
    "small": SpecConfig(n_spectra=1000, seed=0, dm_choices=(10.0,)),
    "open2m": SpecConfig(n_spectra=100_000, seed=2, peak_mode=F.GX_PEAKS_TOTAL_UNIFORM, peaks_min=50,
                         peaks_max=150, charge_min=2, charge_max=3, dm_choices=(500.0,)),
    "uneven": SpecConfig(n_spectra=200_000, seed=3, peak_mode=F.GX_PEAKS_TOTAL_POWER, peaks_min=20,
                         peaks_max=2000, power_alpha=2.0, charge_min=2, charge_max=3, qalen=2048,
                         dm_choices=(10.0, 100.0, 500.0)),
    "c4": SpecConfig(n_spectra=1_000_000, seed=5, peak_mode=F.GX_PEAKS_TOTAL_POWER, peaks_min=20,
                     peaks_max=2000, power_alpha=2.0, charge_min=2, charge_max=3, qalen=2048,
                     dm_choices=(10.0, 100.0, 500.0)),
}


class Bucket:
    """numpy views of one length bucket (valid while the SlmIndex lives)."""

    def __init__(self, raw: F.gx_bucket, owner):
        self.raw = raw
        self._owner = owner
        self.peplen = raw.peplen
        self.n_peps = raw.n_peps
        self.n_entries = raw.n_entries
        self.speclen = raw.speclen
        self.n_ions = raw.n_ions
        self.nbins = raw.nbins

        def view(p, n, ct, dt):
            if n == 0 or not p:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).view(dt)

        self.mass = view(raw.mass, raw.n_entries, C.c_float, np.float32)
        self.seqid = view(raw.seqid, raw.n_entries, C.c_uint32, np.uint32)
        self.modnum = view(raw.modnum, raw.n_entries, C.c_uint32, np.uint32)
        self.sites = view(raw.sites, raw.n_entries, C.c_uint64, np.uint64)
        self.bA = view(raw.bA, raw.nbins + 1, C.c_uint32, np.uint32)
        self.iA = view(raw.iA, raw.n_ions, C.c_uint32, np.uint32)
        self.seqs = view(raw.seqs, raw.n_peps * raw.peplen, C.c_char, np.uint8)

    def sequence(self, psid: int) -> str:
        L = self.peplen
        s = int(self.seqid[psid]) * L
        return bytes(self.seqs[s:s + L]).decode()


class SlmIndex:
    """A finalized synthetic (or user-supplied) peptide index."""

    def __init__(self, cfg: DBConfig, peptides: list[str] | None = None, fragments: bool = True):
        """`fragments=False` stops after the mass-sorted entries (no bA/iA on
        the host): the fragment-ion arrays are then built on the GPU when a
        Searcher takes the index (gx_index_build)."""
        lib = F.host_lib()
        p = F.gx_db_params()
        p.min_len, p.max_len, p.maxz, p.scale = cfg.min_len, cfg.max_len, cfg.maxz, cfg.scale
        p.min_mass, p.max_mass = cfg.min_mass, cfg.max_mass
        p.max_mods_per_pep = cfg.max_mods_per_pep
        p.n_mod_types = len(cfg.mods)
        for i, (res, mass, mx) in enumerate(cfg.mods):
            p.mods[i].residues = res.encode()
            p.mods[i].mass = mass
            p.mods[i].max_per_pep = mx
        p.modmass_all = int(cfg.modmass_all)
        p.threads = cfg.threads
        self.cfg = cfg
        self._lib = lib
        h = C.c_void_p()
        F.check("gx_db_create", lib.gx_db_create(C.byref(p), C.byref(h)))
        self._h = h
        if peptides is not None:
            enc = "".join(peptides).encode()
            lens = np.array([len(s) for s in peptides], dtype=np.int32)
            kept = C.c_uint64()
            F.check("gx_db_add_peptides",
                    lib.gx_db_add_peptides(h, enc, F.ptr(lens), len(peptides), C.byref(kept)))
            self.n_kept = kept.value
        else:
            F.check("gx_db_generate_peptides", lib.gx_db_generate_peptides(h, cfg.seed, cfg.n_peptides))
            self.n_kept = cfg.n_peptides
        self.has_fragments = fragments
        if fragments:
            F.check("gx_db_finalize", lib.gx_db_finalize(h))
        else:
            F.check("gx_db_finalize_entries", lib.gx_db_finalize_entries(h))
        self.buckets: list[Bucket] = []
        for i in range(lib.gx_db_num_buckets(h)):
            b = F.gx_bucket()
            F.check("gx_db_get_bucket", lib.gx_db_get_bucket(h, i, C.byref(b)))
            self.buckets.append(Bucket(b, self))
        self.raw_array = (F.gx_bucket * len(self.buckets))(*[b.raw for b in self.buckets])

    def build_params(self) -> F.gx_index_build_params:
        """The parameters gx_index_build needs to build this index's fragment
        arrays (charges, scale, mass range, modification masses)."""
        bp = F.gx_index_build_params()
        F.check("gx_db_build_params", self._lib.gx_db_build_params(self._h, C.byref(bp)))
        return bp

    @property
    def n_entries(self) -> int:
        return sum(b.n_entries for b in self.buckets)

    @property
    def n_ions(self) -> int:
        return sum(b.n_ions for b in self.buckets)

    @property
    def maxz(self) -> int:
        return self.cfg.maxz

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.buckets = []
            self._lib.gx_db_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QueryBatch:
    """A CSR query batch (Queries<int>) held in numpy arrays."""

    def __init__(self, idx, moz, intensity, precurse, charges=None, rtimes=None, dm=None,
                 truth_bucket=None, truth_psid=None):
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32)
        self.moz = np.ascontiguousarray(moz, dtype=np.int32)
        self.intensity = np.ascontiguousarray(intensity, dtype=np.int32)
        self.precurse = np.ascontiguousarray(precurse, dtype=np.float32)
        n = len(self.precurse)
        self.charges = np.ascontiguousarray(charges if charges is not None else np.ones(n), dtype=np.int32)
        self.rtimes = np.ascontiguousarray(rtimes if rtimes is not None else np.arange(n), dtype=np.float32)
        self.dm = None if dm is None else np.ascontiguousarray(dm, dtype=np.float32)
        self.truth_bucket = truth_bucket
        self.truth_psid = truth_psid
        if len(self.idx) != n + 1 or int(self.idx[-1]) != len(self.moz) or len(self.moz) != len(self.intensity):
            raise ValueError("inconsistent CSR query batch")

    @property
    def n_spectra(self) -> int:
        return len(self.precurse)

    @property
    def n_peaks(self) -> int:
        return len(self.moz)

    def peaks(self, s: int):
        a, b = int(self.idx[s]), int(self.idx[s + 1])
        return self.moz[a:b], self.intensity[a:b]

    def subset(self, sel) -> "QueryBatch":
        sel = np.asarray(sel)
        lens = (self.idx[1:] - self.idx[:-1])[sel]
        idx = np.zeros(len(sel) + 1, dtype=np.uint32)
        idx[1:] = np.cumsum(lens)
        moz = np.concatenate([self.peaks(int(s))[0] for s in sel]) if len(sel) else np.zeros(0, np.int32)
        inten = np.concatenate([self.peaks(int(s))[1] for s in sel]) if len(sel) else np.zeros(0, np.int32)
        return QueryBatch(idx, moz, inten, self.precurse[sel], self.charges[sel], self.rtimes[sel],
                          None if self.dm is None else self.dm[sel],
                          None if self.truth_bucket is None else self.truth_bucket[sel],
                          None if self.truth_psid is None else self.truth_psid[sel])

    def as_gx_queries(self) -> F.gx_queries:
        """A borrowed gx_queries view of this batch (valid while it lives)."""
        q = F.gx_queries()
        q.n_spectra, q.n_peaks = self.n_spectra, self.n_peaks

        def pt(a, ct):
            return C.cast(a.ctypes.data, C.POINTER(ct)) if a is not None and a.size else None

        q.idx, q.moz, q.intensity = pt(self.idx, C.c_uint32), pt(self.moz, C.c_int32), pt(self.intensity, C.c_int32)
        q.precurse, q.charges, q.rtimes = (pt(self.precurse, C.c_float), pt(self.charges, C.c_int32),
                                           pt(self.rtimes, C.c_float))
        q.dm = pt(self.dm, C.c_float)
        return q

    def as_ffi(self, with_dm: bool = True) -> F.gx_query_batch:
        q = F.gx_query_batch()
        q.n_spectra = self.n_spectra
        q.n_peaks = self.n_peaks
        q.idx, q.moz, q.intensity, q.precurse = (F.ptr(self.idx), F.ptr(self.moz), F.ptr(self.intensity),
                                                 F.ptr(self.precurse))
        q.dm = F.ptr(self.dm) if (with_dm and self.dm is not None) else None
        return q


@dataclass
class PickParams:
    """Peak picking (gParams base_int, min_int, QALEN; argp.hpp:304-307)."""
    base_int: int = 100_000
    min_int: int = 1000
    qalen: int = 100

    def to_ffi(self) -> F.gx_pick_params:
        p = F.gx_pick_params()
        p.base_int, p.min_int, p.qalen = self.base_int, self.min_int, self.qalen
        return p


def _batch_from_gx(lib, out) -> QueryBatch:
    """Copy a library-owned gx_queries into a QueryBatch and destroy it."""
    try:
        q = out.contents
        n, npk = q.n_spectra, q.n_peaks

        def cp(pp, k, dt):
            if not pp:
                return None
            if k == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(pp, shape=(k,)).astype(dt, copy=True)

        return QueryBatch(cp(q.idx, n + 1, np.uint32), cp(q.moz, npk, np.int32), cp(q.intensity, npk, np.int32),
                          cp(q.precurse, n, np.float32), cp(q.charges, n, np.int32), cp(q.rtimes, n, np.float32),
                          cp(q.dm, n, np.float32), cp(q.truth_bucket, n, np.int32), cp(q.truth_psid, n, np.int32))
    finally:
        lib.gx_queries_destroy(out)


def _spec_params(cfg: SpecConfig) -> F.gx_spec_params:
    p = F.gx_spec_params()
    p.n_spectra = cfg.n_spectra
    p.keep_frac, p.mz_sigma = cfg.keep_frac, cfg.mz_sigma
    p.peak_mode, p.noise_min, p.noise_max = cfg.peak_mode, cfg.noise_min, cfg.noise_max
    p.peaks_min, p.peaks_max, p.power_alpha = cfg.peaks_min, cfg.peaks_max, cfg.power_alpha
    p.noise_mz_lo, p.noise_mz_hi, p.int_sigma = cfg.noise_mz_lo, cfg.noise_mz_hi, cfg.int_sigma
    p.charge_min, p.charge_max = cfg.charge_min, cfg.charge_max
    p.qalen, p.base_int, p.min_int = cfg.qalen, cfg.base_int, cfg.min_int
    p.n_dm = len(cfg.dm_choices)
    for i, d in enumerate(cfg.dm_choices):
        p.dm_choices[i] = d
    return p


def generate_raw_spectra(index: SlmIndex, cfg: SpecConfig) -> QueryBatch:
    """The spectra of generate_spectra(index, cfg) before peak picking, as an
    MS2 reader stores them (m/z*scale, intensity*100)."""
    lib = F.host_lib()
    p = _spec_params(cfg)
    out = C.POINTER(F.gx_queries)()
    F.check("gx_spectra_generate_raw", lib.gx_spectra_generate_raw(index._h, C.byref(p), cfg.seed, C.byref(out)))
    return _batch_from_gx(lib, out)


def read_ms2(path, scale: int = 100) -> QueryBatch:
    """Raw spectra of an MS2 text file (gx_ms2_read)."""
    lib = F.host_lib()
    out = C.POINTER(F.gx_queries)()
    F.check("gx_ms2_read", lib.gx_ms2_read(str(path).encode(), scale, C.byref(out)))
    return _batch_from_gx(lib, out)


def pick_peaks(batch: QueryBatch, pick: PickParams | None = None) -> QueryBatch:
    """Peak picking on the CPU (gx_queries_pick)."""
    lib = F.host_lib()
    pp = (pick or PickParams()).to_ffi()
    out = C.POINTER(F.gx_queries)()
    F.check("gx_queries_pick", lib.gx_queries_pick(C.byref(batch.as_gx_queries()), C.byref(pp), C.byref(out)))
    return _batch_from_gx(lib, out)


def write_pbin(path, batch: QueryBatch) -> None:
    lib = F.host_lib()
    F.check("gx_pbin_write", lib.gx_pbin_write(str(path).encode(), C.byref(batch.as_gx_queries())))


def read_pbin(path) -> QueryBatch:
    lib = F.host_lib()
    out = C.POINTER(F.gx_queries)()
    F.check("gx_pbin_read", lib.gx_pbin_read(str(path).encode(), C.byref(out)))
    return _batch_from_gx(lib, out)


def generate_spectra(index: SlmIndex, cfg: SpecConfig) -> QueryBatch:
    """Seeded synthetic spectra of index entries, peaks picked (gx_spectra_generate)."""
    lib = F.host_lib()
    p = _spec_params(cfg)
    out = C.POINTER(F.gx_queries)()
    F.check("gx_spectra_generate", lib.gx_spectra_generate(index._h, C.byref(p), cfg.seed, C.byref(out)))
    return _batch_from_gx(lib, out)
