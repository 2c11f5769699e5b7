"""GPU search: the host side of the reference's s3::search / s4::processResults.

`Searcher` owns a device context (s3::initialize, superstep3/kernel.cu:251-261)
and a device-resident index (s1::getdIndex, superstep1/kernel.cu:310-358).
`Searcher.search()` runs one query batch the way DSLIM_QuerySpectrum does
(dslim_query.cpp:749-1081) and returns one record per spectrum (the top hit's
hCell fields, cpsms and the tail-fit e-value) plus, on request, the 802-bin
survival histograms. Every call goes through libgicops_hip.so; there is no
CPU fallback: a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F
from .slm import PickParams, QueryBatch, SlmIndex


@dataclass
class SearchParams:
    """gParams subset (slm_dsts.h:349-431) with the argp derivations
    (argp.hpp:272-370): dF = (uint)(0.02*scale) bins, base intensities are
    the query generator's."""
    dF: int = 2
    scale: int = 100
    max_mass: int = 5000
    min_shp: int = 4
    min_cpsm: int = 4
    dM: float = 10.0
    expect_max: float = 20.0
    evalue_mode: int = F.GX_EVALUE_TAILFIT
    surrogate_conf: float = 0.5
    engine: int = F.GX_ENGINE_BATCHED
    config: int = -1
    use_graph: bool = False
    split: int = 0
    whole_bin: int = 0
    wide_stream: int = 0

    def to_ffi(self) -> F.gx_search_params:
        p = F.gx_search_params()
        p.dF, p.scale, p.max_mass = self.dF, self.scale, self.max_mass
        p.min_shp, p.min_cpsm, p.evalue_mode = self.min_shp, self.min_cpsm, self.evalue_mode
        p.dM, p.expect_max, p.surrogate_conf = self.dM, self.expect_max, self.surrogate_conf
        p.engine, p.config, p.use_graph = self.engine, self.config, int(self.use_graph)
        p.split, p.whole_bin, p.wide_stream = self.split, self.whole_bin, self.wide_stream
        return p


def stats_dict(st: F.gx_search_stats) -> dict:
    return {k: getattr(st, k) for k, _ in F.gx_search_stats._fields_}


class DeviceQueries:
    """A query batch resident in HBM (dQueries, superstep3/kernel.cu:192-206)."""

    def __init__(self, searcher: "Searcher", batch: QueryBatch, with_dm: bool = True, pick: PickParams | None = None):
        """Uploads `batch` as it is, or, with `pick`, treats it as raw spectra
        and picks their peaks on the GPU (gx_queries_preprocess)."""
        self._lib = searcher._lib
        self._s = searcher
        self.n_spectra = batch.n_spectra
        q = batch.as_ffi(with_dm)
        h = C.c_void_p()
        if pick is None:
            F.check("gx_queries_upload", self._lib.gx_queries_upload(searcher._ctx, C.byref(q), C.byref(h)))
        else:
            pp = pick.to_ffi()
            F.check("gx_queries_preprocess",
                    self._lib.gx_queries_preprocess(searcher._ctx, C.byref(q), C.byref(pp), C.byref(h)))
        self._h = h

    def shape(self) -> tuple[int, int, float]:
        """(n_spectra, n_peaks, device ms of the GPU peak picking or 0)."""
        ns, npk, ms = C.c_uint32(), C.c_uint64(), C.c_float()
        F.check("gx_queries_shape", self._lib.gx_queries_shape(self._h, C.byref(ns), C.byref(npk), C.byref(ms)))
        return ns.value, npk.value, ms.value

    def peaks(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(idx, moz, intensity) of the batch as the device holds it."""
        ns, npk, _ = self.shape()
        idx = np.zeros(ns + 1, np.uint32)
        moz = np.zeros(npk, np.int32)
        inten = np.zeros(npk, np.int32)
        F.check("gx_queries_download", self._lib.gx_queries_download(
            self._h, F.ptr(idx), F.ptr(moz) if npk else None, F.ptr(inten) if npk else None))
        return idx, moz, inten

    def free(self) -> None:
        if getattr(self, "_h", None):
            self._lib.gx_queries_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Searcher:
    def __init__(self, index: SlmIndex, device: int = 0, build_on_device: bool | None = None):
        """Uploads the host-built fragment-ion index, or builds it on the GPU
        from the index's mass-sorted entries (gx_index_build) when
        `build_on_device` is set; None builds on the GPU exactly when the host
        index carries no fragment arrays."""
        self._lib = F.hip_lib()
        ctx = C.c_void_p()
        F.check("gx_ctx_create", self._lib.gx_ctx_create(device, C.byref(ctx)))
        self._ctx = ctx
        self.index = index
        if build_on_device is None:
            build_on_device = not index.has_fragments
        if not build_on_device and not index.has_fragments:
            raise ValueError("the index has no host fragment arrays to upload; build it on the device")
        self.built_on_device = build_on_device
        d = C.c_void_p()
        if build_on_device:
            bp = index.build_params()
            st = self._lib.gx_index_build(ctx, index.raw_array, len(index.buckets), C.byref(bp), C.byref(d))
            name = "gx_index_build"
        else:
            st = self._lib.gx_index_upload(ctx, index.raw_array, len(index.buckets), index.maxz, C.byref(d))
            name = "gx_index_upload"
        if st:
            self._lib.gx_ctx_destroy(ctx)
            self._ctx = None
            F.check(name, st)
        self._didx = d
        self._children: list = []  # surrogates, closed before the context

    def device_info(self) -> tuple[str, int]:
        buf = C.create_string_buffer(256)
        ncu = C.c_int()
        F.check("gx_ctx_device_info", self._lib.gx_ctx_device_info(self._ctx, buf, 256, C.byref(ncu)))
        return buf.value.decode(), ncu.value

    def index_arrays(self, bucket: int) -> tuple[np.ndarray, np.ndarray]:
        """(bA, iA) of one bucket as the device index holds them."""
        b = self.index.buckets[bucket]
        bA = np.zeros(b.nbins + 1 if b.n_entries else 0, dtype=np.uint32)
        iA = np.zeros(b.n_ions, dtype=np.uint32)
        F.check("gx_index_download",
                self._lib.gx_index_download(self._didx, bucket, F.ptr(bA) if bA.size else None,
                                            F.ptr(iA) if iA.size else None))
        return bA, iA

    def read_bandwidth(self, segment_lines: int = 0) -> float:
        """Measured read bandwidth (GB/s): sequential (0) or random segments of
        1, 2, 4 or 8 consecutive 128-B lines (gx_measure_read_bandwidth)."""
        g = C.c_double()
        F.check("gx_measure_read_bandwidth", self._lib.gx_measure_read_bandwidth(self._ctx, segment_lines, C.byref(g)))
        return g.value

    def host_index(self) -> "HostIndex":
        """A host copy of the device index: the host index's entries with the
        device's bA/iA downloaded (for CPU-side checks of a GPU-built index)."""
        return HostIndex(self)

    def upload(self, batch: QueryBatch, with_dm: bool = True) -> DeviceQueries:
        return DeviceQueries(self, batch, with_dm)

    def preprocess(self, raw: QueryBatch, pick: PickParams | None = None, with_dm: bool = True) -> DeviceQueries:
        """Pick the peaks of raw spectra on the GPU into a device batch."""
        return DeviceQueries(self, raw, with_dm, pick or PickParams())

    def search_device(self, dq: DeviceQueries, params: SearchParams) -> dict:
        st = F.gx_search_stats()
        p = params.to_ffi()
        F.check("gx_search_device",
                self._lib.gx_search_device(self._ctx, self._didx, dq._h, C.byref(p), C.byref(st)))
        return stats_dict(st)

    def autotune(self, sample: DeviceQueries, params: SearchParams) -> dict:
        """Time every launch configuration x window split on `sample`, then the
        whole-bin threshold, and make the fastest this searcher's default
        (config=-1, split=0, whole_bin=0)."""
        r = F.gx_tune_result()
        p = params.to_ffi()
        F.check("gx_autotune", self._lib.gx_autotune(self._ctx, self._didx, sample._h, C.byref(p), C.byref(r)))
        return {"config": r.config, "split": r.split, "whole_bin": r.whole_bin,
                "tried": [(r.tried_config[i], r.tried_split[i], r.tried_whole_bin[i], r.tried_ms[i])
                          for i in range(r.n_tried)]}

    def download(self, dq: DeviceQueries, want_hist: bool = False):
        psm = np.zeros(dq.n_spectra, dtype=F.PSM_DTYPE)
        hist = np.zeros((dq.n_spectra, F.GX_HIST_BINS), dtype=np.uint32) if want_hist else None
        F.check("gx_results_download",
                self._lib.gx_results_download(self._ctx, dq._h, F.ptr(psm), F.ptr(hist)))
        return psm, hist

    def search(self, batch: QueryBatch, params: SearchParams, want_hist: bool = False, with_dm: bool = True):
        """Upload, search, model and download one batch -> (psms, hist | None, stats)."""
        psm = np.zeros(batch.n_spectra, dtype=F.PSM_DTYPE)
        hist = np.zeros((batch.n_spectra, F.GX_HIST_BINS), dtype=np.uint32) if want_hist else None
        st = F.gx_search_stats()
        q = batch.as_ffi(with_dm)
        p = params.to_ffi()
        F.check("gx_search", self._lib.gx_search(self._ctx, self._didx, C.byref(q), C.byref(p), F.ptr(psm),
                                                 F.ptr(hist), C.byref(st)))
        return psm, hist, stats_dict(st)

    def features(self, dq: DeviceQueries) -> dict:
        """Surrogate features and plain-fit targets of a searched batch
        (gx_features_download): feat (n,128), target (n,4) = {mark before the
        mark2 rule, mark2, n-1, servable}; loge (n,) = log10 of the plain
        e-value; valid, servable (n,) bool."""
        n = dq.n_spectra
        feat = np.zeros((n, F.MLP_FEAT), np.float32)
        target = np.zeros((n, F.MLP_TARGET), np.float32)
        loge = np.zeros(n, np.float32)
        valid = np.zeros(n, np.uint8)
        serv = np.zeros(n, np.uint8)
        F.check("gx_features_download", self._lib.gx_features_download(
            self._ctx, dq._h, F.fptr(feat), F.fptr(target), F.fptr(loge), F.u8ptr(valid), F.u8ptr(serv)))
        return {"feat": feat, "target": target, "loge": loge, "valid": valid.astype(bool),
                "servable": serv.astype(bool)}

    def surrogate_outputs(self, dq: DeviceQueries) -> np.ndarray:
        """After a learned-mode search: (n,4) raw outputs {mark, mark2,
        confidence logit} and 1.0 where the surrogate served the e-value."""
        out = np.zeros((dq.n_spectra, 4), np.float32)
        F.check("gx_surrogate_outputs_download",
                self._lib.gx_surrogate_outputs_download(self._ctx, dq._h, F.fptr(out)))
        return out

    def close(self) -> None:
        for ch in self._children:
            ch.close()
        self._children = []
        if getattr(self, "_didx", None):
            self._lib.gx_index_free(self._didx)
            self._didx = None
        if getattr(self, "_ctx", None):
            self._lib.gx_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostIndex:
    """Buckets with the host index's entry arrays and the device index's
    bA/iA copied to the host: `raw_array` (gx_bucket[]) and `buckets` have
    the shape SlmIndex has."""

    def __init__(self, searcher: "Searcher"):
        index = searcher.index
        self._index = index
        self.arrays = []
        self.buckets = index.buckets
        self.raw_array = (F.gx_bucket * len(index.buckets))()
        for i, b in enumerate(index.buckets):
            bA, iA = searcher.index_arrays(i)
            self.arrays.append((bA, iA))
            r = self.raw_array[i]
            C.memmove(C.byref(r), C.byref(b.raw), C.sizeof(F.gx_bucket))
            r.bA = bA.ctypes.data if bA.size else None
            r.iA = iA.ctypes.data if iA.size else None


def shard_bounds(n_entries: int, n_shards: int) -> list[tuple[int, int]]:
    """Contiguous psid ranges [a, e) of one bucket's mass-sorted entries, one per shard."""
    return [(n_entries * k // n_shards, n_entries * (k + 1) // n_shards) for k in range(n_shards)]


class ShardedSearcher:
    """The index-sharded search on one GPU: every bucket's mass-sorted entries
    split into `n_shards` contiguous psid ranges, each shard's fragment-ion
    index built on the GPU (gx_index_build), every batch searched against each
    shard in turn, and the shards' partial results (gx_partials_export)
    merged into PSMs with the tail fit on the summed histograms
    (gx_partials_merge). The merge of contiguous ranges equals the unsharded
    search exactly. This is what an index-sharded multi-GPU search does per
    rank (the reference's MPI model, lbe_internal.cpp:75-103 and
    dslim_score.cpp:199-387), with the exchange of partials through host
    memory."""

    def __init__(self, index: SlmIndex, n_shards: int, device: int = 0):
        if n_shards < 1:
            raise ValueError("n_shards >= 1")
        self._lib = F.hip_lib()
        ctx = C.c_void_p()
        F.check("gx_ctx_create", self._lib.gx_ctx_create(device, C.byref(ctx)))
        self._ctx = ctx
        self.index = index
        self.n_shards = n_shards
        self._children: list = []
        self._didx: list = []
        self.offsets: list[np.ndarray] = []
        bp = index.build_params()
        nb = len(index.buckets)
        for k in range(n_shards):
            views = (F.gx_bucket * nb)()
            off = np.zeros(nb, np.uint32)
            for i, b in enumerate(index.buckets):
                a, e = shard_bounds(b.n_entries, n_shards)[k]
                v = views[i]
                v.peplen, v.n_peps, v.speclen, v.nbins = b.peplen, b.n_peps, b.speclen, b.nbins
                v.n_entries = e - a
                v.n_ions = (e - a) * b.speclen
                v.seqs = b.raw.seqs
                if e > a:
                    v.mass = b.mass[a:].ctypes.data
                    v.seqid = b.seqid[a:].ctypes.data
                    v.modnum = b.modnum[a:].ctypes.data
                    v.sites = b.sites[a:].ctypes.data
                off[i] = a
            d = C.c_void_p()
            st = self._lib.gx_index_build(ctx, views, nb, C.byref(bp), C.byref(d))
            if st:
                self.close()
                F.check("gx_index_build", st)
            self._didx.append(d)
            self.offsets.append(off)

    def upload(self, batch: QueryBatch, with_dm: bool = True) -> DeviceQueries:
        return DeviceQueries(self, batch, with_dm)

    def search(self, batch: QueryBatch, params: SearchParams, want_hist: bool = False):
        """Search every shard, exchange the partials through host memory and
        merge -> (psms, merged hist | None, stats with per-shard and merge times)."""
        if params.engine != F.GX_ENGINE_BATCHED:
            raise ValueError("the sharded search exports the batched engine's partials")
        dq = self.upload(batch)
        n = batch.n_spectra
        recs = [np.zeros(n, F.PARTIAL_DTYPE) for _ in range(self.n_shards)]
        hists = [np.zeros((n, F.GX_HIST_BINS), np.uint32) for _ in range(self.n_shards)]
        p = params.to_ffi()
        shard_ms, ions = [], 0
        for k, d in enumerate(self._didx):
            st = F.gx_search_stats()
            F.check("gx_search_device", self._lib.gx_search_device(self._ctx, d, dq._h, C.byref(p), C.byref(st)))
            shard_ms.append(st.total_ms)
            ions += st.ions
            F.check("gx_partials_export", self._lib.gx_partials_export(
                self._ctx, d, dq._h, F.ptr(self.offsets[k]), F.ptr(recs[k]), F.ptr(hists[k]), 0))
        dq.free()
        psm, hist, ms = self.merge(recs, hists, params, want_hist)
        return psm, hist, {"shard_ms": shard_ms, "merge_ms": ms, "ions": ions}

    def merge(self, recs, hists, params: SearchParams, want_hist: bool = False):
        """gx_partials_merge over host partials -> (psms, hist | None, device ms)."""
        n = len(recs[0])
        k = len(recs)
        psm = np.zeros(n, dtype=F.PSM_DTYPE)
        hist = np.zeros((n, F.GX_HIST_BINS), np.uint32) if want_hist else None
        rp = (C.c_void_p * k)(*[r.ctypes.data for r in recs])
        hp = (C.c_void_p * k)(*[h.ctypes.data for h in hists])
        ms = C.c_float()
        p = params.to_ffi()
        F.check("gx_partials_merge", self._lib.gx_partials_merge(self._ctx, k, rp, hp, n, C.byref(p), 0, F.ptr(psm),
                                                                 F.ptr(hist), C.byref(ms)))
        return psm, hist, ms.value

    def close(self) -> None:
        for d in self._didx:
            self._lib.gx_index_free(d)
        self._didx = []
        if getattr(self, "_ctx", None):
            self._lib.gx_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

