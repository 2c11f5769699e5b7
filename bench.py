#!/usr/bin/env python3
"""Benchmark: spectra/sec of the fragment-ion-index open search on one MI355X.

Workload (default `open2m`): 2M synthetic tryptic peptides plus variable Met
oxidation (<=2 per peptide, ~5.6M index entries, ~290M fragment ions, charge-1
b/y, 0.01 Da bins), 100k query spectra (50-150 raw peaks, top 100 kept),
dF = 0.02 Da, open search dM = +-500 Da. A step is one search of the whole
batch (window, scan, scoring, histogram, top hit and e-value for every
spectrum) with the batch already resident in HBM.

The headline value is the learned+batched form: the batched engine
(autotuned launch configuration, one hipGraph per batch) with the learned
e-value surrogate, which is trained first on the GPU from 500k spectra of
held-out seeds searched with the plain tail fit. The plain tail fit judges
the learned run (top hits, e-value error) and the plain forms are timed the
same way beside it: batched with the plain fit, and the per-spectrum engine.

Prints one JSON line. `roofline.achieved` is the search kernel's algorithmic
bytes (4 B per window-restricted iA id, counted by the kernel itself) over
its device time measured with HIP events on the context stream; `traffic` is
the HBM bytes per launch from the committed rocprofv3 counter summary.
"""
from __future__ import annotations

import argparse
import json

import os
import statistics
import subprocess
import sys
import time
from pathlib import Path

# The CPU baseline's OpenMP threads are bound to consecutive cores. libgomp
# reads this when it loads (with the host library, before any search).
os.environ.setdefault("OMP_PROC_BIND", "close")

import numpy as np  # noqa: E402

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def ensure_built() -> None:
    from gicops_amd import _build

    if not (_build.LIBDIR / "libgicops_host.so").exists():
        _build.build_host()
    if not (_build.LIBDIR / "libgicops_hip.so").exists():
        _build.build_hip()
    if not (_build.ORACLE / "liboracle.so").exists():
        _build.build_oracle()


def algorithmic_bytes(st: dict, n_peaks: int) -> int:
    return 8 * n_peaks + 8 * st["bins"] + 4 * st["ions"] + 24 * st["spectra"]


def traffic_from_profiles(workload: str, config: int, whole_bin: int):
    """HBM bytes per launch of the search kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE, scaled by a copy calibration),
    when it was taken with the same launch configuration."""
    f = ROOT / "profiles" / f"search_spectra_{workload}_hbm_traffic.txt"
    if not f.exists():
        return None
    vals = {}
    for line in f.read_text().splitlines():
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    if int(vals.get("launch_config", -1)) != config:
        return None
    key = f"whole_bin_{whole_bin}_hbm_bytes_per_launch"
    return float(vals[key]) if key in vals else None


def cpu_topology() -> dict:
    """Sockets x cores x SMT of the host (lscpu), and the CPUs this process may use."""
    topo = {"cpus_usable": len(os.sched_getaffinity(0)), "nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "smt",
                "Model name": "model"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                topo[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    return topo


def cpu_baseline(index, batch, params_ffi, budget_s: float, repeats: int = 3):
    """The oracle (the CPU restatement of DSLIM_QuerySpectrum + ModelTailFit,
    dslim_query.cpp:749-1081 and expeRT.cpp:396-559) on a bounded sample of
    the same batch: OpenMP over spectra on every usable host CPU, threads
    bound close (OMP_PROC_BIND), median of `repeats` timed runs."""
    from oracle import pyoracle as O

    cores = len(os.sched_getaffinity(0))
    n = batch.n_spectra
    probe = batch.subset(np.arange(min(n, 64)))
    t = time.perf_counter()
    O.search(index, probe, params_ffi, threads=cores)
    dt = max(time.perf_counter() - t, 1e-3)
    rate = probe.n_spectra / dt
    m = int(min(n, max(probe.n_spectra, rate * budget_s)))
    sample = batch.subset(np.arange(m))
    rates = []
    for _ in range(repeats):
        t = time.perf_counter()
        O.search(index, sample, params_ffi, threads=cores)
        rates.append(m / (time.perf_counter() - t))
    return {"value": round(statistics.median(rates), 2), "unit": "spectra/s", "cores": cores, "kind": "port",
            "port_of": "dslim_query.cpp:749-1081 (DSLIM_QuerySpectrum) + expeRT.cpp:396-559 (ModelTailFit)",
            "runs_spectra_per_s": [round(r, 2) for r in rates],
            "topology": cpu_topology(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "sample": f"first {m} spectra of the same batch, OpenMP over spectra (schedule dynamic), "
                      f"median of {repeats} runs"}


def train_surrogate(s, index, sc, dM: float, args) -> dict:
    """Train the learned e-value model on the GPU from (histogram, plain-fit)
    pairs of spectra generated with held-out seeds (not the timed batch)."""
    import dataclasses

    from gicops_amd import slm
    from gicops_amd.search import SearchParams
    from gicops_amd.surrogate import Surrogate, TrainParams

    t = time.perf_counter()
    sur = Surrogate(s, seed=7)
    plain = SearchParams(dM=dM)
    left, seed, rows = args.train_spectra, 1000, 0
    while left > 0:
        n = min(left, 100_000)
        b = slm.generate_spectra(index, dataclasses.replace(sc, n_spectra=n, seed=seed))
        dq = s.upload(b)
        s.search_device(dq, plain)
        rows += sur.add_training(dq)
        dq.free()
        left -= n
        seed += 1
    t_data = time.perf_counter() - t
    t = time.perf_counter()
    rep = sur.train(TrainParams(batch=1024, epochs=args.train_epochs, lr=2e-3, lr_decay=0.7, conf_weight=0.5,
                                seed=1))
    t_train = time.perf_counter() - t
    sur.attach()
    log(f"surrogate: {rows} training pairs from {args.train_spectra} spectra ({t_data:.1f} s), trained "
        f"{rep['epochs']} epochs in {t_train:.1f} s; mse {rep['mse_mark'][0]:.4f}/{rep['mse_mark2'][0]:.4f} -> "
        f"{rep['mse_mark'][-1]:.4f}/{rep['mse_mark2'][-1]:.4f}, marks in their check windows "
        f"{rep['in_window'][-1]:.4f}")
    return {"train_pairs": rows, "train_spectra": args.train_spectra, "epochs": rep["epochs"],
            "train_s": round(t_train, 2), "data_s": round(t_data, 2),
            "mse_mark_first_last": [round(rep["mse_mark"][0], 5), round(rep["mse_mark"][-1], 5)],
            "mse_mark2_first_last": [round(rep["mse_mark2"][0], 5), round(rep["mse_mark2"][-1], 5)],
            "train_in_window_last": round(rep["in_window"][-1], 5), "conf_threshold": args.surrogate_conf}


def judge_learned(psm, ref, st, raw) -> dict:
    """The plain tail fit as the judge of the learned run: top hits must be
    identical; e-value error in relative log10 over the modeled spectra
    (a spectrum the plain fit leaves undefined counts as an error if the
    surrogate served it); how far the network's predicted marks were."""
    from gicops_amd import _ffi as F

    same = all(np.array_equal(psm[k], ref[k]) for k in ("psid", "idxoffset", "hyperscore", "cpsms", "hyp",
                                                        "sharedions"))
    modeled = (ref["flags"] & F.GX_PSM_MODELED) != 0
    sur = (psm["flags"] & F.GX_PSM_SURROGATE) != 0
    fit = (ref["flags"] & F.GX_PSM_FIT_OK) != 0
    ok = np.zeros(len(psm), bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        lr = np.log10(ref["evalue"][fit])
        ll = np.log10(psm["evalue"][fit])
        # equal e-values count as within (an e-value of 0 or inf has no finite log10 difference)
        ok[fit] = (psm["evalue"][fit] == ref["evalue"][fit]) | (np.abs(ll - lr) <= 1e-3 * np.maximum(np.abs(lr), 1e-3))
    ok[~fit] = ~sur[~fit]
    err = np.abs(ll - lr)[sur[fit]] if sur.any() else np.zeros(0)
    rel = (err / np.maximum(np.abs(lr[sur[fit]]), 1e-3)) if sur.any() else np.zeros(0)
    return {"top_hits_identical": bool(same), "modeled": int(modeled.sum()), "surrogate_served": int(sur.sum()),
            "coverage": round(float(sur.sum()) / max(1, int(modeled.sum())), 5),
            "fallbacks": int(st["fallbacks"]),
            "within_1e-3_rel_log10": round(float(ok[modeled].mean()), 5),
            "served_within_1e-3_rel_log10": round(float((rel <= 1e-3).mean()), 5) if len(rel) else None,
            "served_max_rel_log10_error": float(rel.max()) if len(rel) else None,
            "served_mu_beta_equal_plain": round(float(np.mean((psm["mu"][sur] == ref["mu"][sur]) &
                                                              (psm["beta"][sur] == ref["beta"][sur]))), 6)
            if sur.any() else None}


def calibrate_copy(mib: int) -> int:
    """A plain device-to-device copy of a known size, for scaling the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) of rocprofv3 PMC runs."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        hip.hipMemcpy(dst, src, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes_per_call": n, "calls": 3}), flush=True)
    return 0


def mlp_gemm(rows: int, calls: int = 5) -> int:
    """Only the surrogate's forward and backward passes (eight bf16 MFMA GEMMs
    per call: 128-256-256-16 forward; dW3, dZ2, dW2, dZ1, dW1 backward) on
    `rows` synthetic rows, for the GEMM kernel's throughput under a kernel
    trace. FLOPs per row: 2*(128*256 + 256*256 + 256*16) forward, and
    2*(256*16 + 16*256 + 2*256*256 + 128*256) backward."""
    ensure_built()
    from gicops_amd import slm
    from gicops_amd.search import Searcher
    from gicops_amd.surrogate import Surrogate

    index = slm.SlmIndex(slm.DB_PRESETS["tiny"])
    s = Searcher(index)
    sur = Surrogate(s, seed=3)
    rng = np.random.default_rng(0)
    feat = rng.standard_normal((rows, 128)).astype(np.float32)
    target = np.zeros((rows, 4), np.float32)
    target[:, 0] = rng.integers(0, 50, rows)
    target[:, 1] = target[:, 0] + rng.integers(1, 50, rows)
    target[:, 2] = 100.0
    target[:, 3] = 1.0
    norm = np.array([0.2, 0.1, 0.8, 0.1], np.float32)
    t = time.perf_counter()
    for _ in range(calls):
        sur.grad(feat, target, norm, 0.5)
    wall = time.perf_counter() - t
    flops = rows * (2 * (128 * 256 + 256 * 256 + 256 * 16) + 2 * (256 * 16 + 16 * 256 + 2 * 256 * 256 + 128 * 256))
    print(json.dumps({"mlp_gemm_rows": rows, "calls": calls, "gemm_flops_per_call": flops,
                      "wall_s_with_host_copies": round(wall, 3)}), flush=True)
    sur.close()
    s.close()
    index.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="open2m", choices=["open2m", "uneven", "small", "c4"],
                    help="c4 is 1M spectra of the uneven mix on the 50M-peptide index; run it with --steps 2 "
                         "--warmup 0 --no-compare-spectrum --config 0 (about 3 minutes)")
    ap.add_argument("--spectra", type=int, default=0, help="override the batch size")
    ap.add_argument("--cpu-budget", type=float, default=8.0, help="seconds of CPU baseline work per run (3 runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine", default="batched", choices=["batched", "spectrum"])
    ap.add_argument("--config", type=int, default=-1, help="batched launch configuration (-1: try all, keep the best)")
    ap.add_argument("--no-compare-spectrum", action="store_true",
                    help="skip the one run of the per-spectrum engine with the plain fit")
    ap.add_argument("--evalue", default="learned", choices=["learned", "plain"])
    ap.add_argument("--train-spectra", type=int, default=0,
                    help="spectra searched (held-out seeds) to make surrogate training pairs "
                         "(0: 500k, or 100k for the uneven and 50M-peptide workloads)")
    ap.add_argument("--train-epochs", type=int, default=6)
    ap.add_argument("--surrogate-conf", type=float, default=0.5,
                    help="confidence threshold: below it the plain fit runs without a window check")
    ap.add_argument("--whole-bin", type=int, default=0, help="whole-bin threshold (0: tuned)")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches instead of hipGraph replays")
    ap.add_argument("--no-bandwidth", action="store_true",
                    help="skip the read-bandwidth probe (for short profiler runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    ap.add_argument("--mlp-gemm", type=int, default=0, metavar="ROWS",
                    help="only run 5 forward+backward passes of the surrogate on ROWS synthetic rows and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.mlp_gemm:
        return mlp_gemm(args.mlp_gemm)
    if not args.train_spectra:
        args.train_spectra = 100_000 if args.workload in ("c4", "uneven") else 500_000
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); on more GPUs the search "
              "would shard by spectrum over a replicated index, with no communication until results are "
              "gathered", file=sys.stderr)
        return 2

    ensure_built()
    from gicops_amd import slm
    from gicops_amd.search import Searcher, SearchParams
    from oracle import pyoracle as O

    dbname = {"open2m": "open2m", "uneven": "open2m", "small": "small", "c4": "c4"}[args.workload]
    dM = 10.0 if args.workload == "small" else 500.0
    t = time.perf_counter()
    # the 50M-peptide database: entries on the host, fragment arrays built on the GPU
    index = slm.SlmIndex(slm.DB_PRESETS[dbname], fragments=(dbname != "c4"))
    log(f"index: {index.n_entries} entries, {index.n_ions} ions, host part built in {time.perf_counter() - t:.1f} s")
    sc = slm.SPEC_PRESETS[args.workload]
    if args.spectra:
        sc.n_spectra = args.spectra
    t = time.perf_counter()
    batch = slm.generate_spectra(index, sc)
    log(f"batch: {batch.n_spectra} spectra, {batch.n_peaks} peaks, generated in {time.perf_counter() - t:.1f} s")

    t = time.perf_counter()
    s = Searcher(index)
    log(f"device index ({'built on the GPU' if s.built_on_device else 'uploaded'}) in "
        f"{time.perf_counter() - t:.1f} s")
    name, ncu = s.device_info()
    log(f"device: {name}, {ncu} CUs")
    from gicops_amd import _ffi as F
    engine = F.GX_ENGINE_BATCHED if args.engine == "batched" else F.GX_ENGINE_SPECTRUM
    params = SearchParams(dM=dM, engine=engine, config=args.config, whole_bin=args.whole_bin)
    t = time.perf_counter()
    dq = s.upload(batch)
    log(f"query upload {time.perf_counter() - t:.2f} s")
    extra = {}
    if engine == F.GX_ENGINE_BATCHED and args.config < 0:
        # the project's autotuner on a sample of the batch (launch configuration x window split)
        ns = min(batch.n_spectra, 25000)
        sample = s.upload(batch.subset(np.arange(ns)))
        t = time.perf_counter()
        tune = s.autotune(sample, params)
        sample.free()
        log(f"autotune on {ns} spectra ({time.perf_counter() - t:.1f} s): config {tune['config']}, split "
            f"{tune['split']}, whole bins <= {tune['whole_bin']}; tried "
            + ", ".join(f"c{c}/s{sp}/w{w}: {ms:.1f} ms" for c, sp, w, ms in tune["tried"]))
        extra["autotune"] = {"config": tune["config"], "split": tune["split"], "whole_bin": tune["whole_bin"],
                             "tried_ms": {f"c{c}s{sp}w{w}": round(ms, 3) for c, sp, w, ms in tune["tried"]}}
    # one eager run of the plain form for the per-kernel device times (HIP
    # events); its results are also the judge of the learned e-values
    plain = SearchParams(dM=dM, engine=engine, config=args.config, whole_bin=args.whole_bin)
    st = s.search_device(dq, plain)
    log(f"eager plain: total {st['total_ms']:.1f} ms (plan {st['plan_ms']:.1f}, search {st['kernel_ms']:.1f}, "
        f"evalue {st['evalue_ms']:.1f}), {st['items']} items, {st['chunks']} chunks, config {st['config']}")
    eager = dict(st)
    judge, _ = s.download(dq)
    if engine == F.GX_ENGINE_BATCHED:
        # a new batch takes the eager path: plan on the device, host scheduling
        # of the work units, then the launches (buffers already sized)
        t = time.perf_counter()
        st2 = s.search_device(dq, plain)
        extra["new_batch_ms_per_step"] = round((time.perf_counter() - t) * 1000.0, 3)
        extra["new_batch_device_ms"] = round(st2["total_ms"], 3)
        log(f"eager plain again (a new batch's cost): {extra['new_batch_ms_per_step']:.1f} ms wall, "
            f"{st2['total_ms']:.1f} ms device")
    use_graph = engine == F.GX_ENGINE_BATCHED and not args.no_graph
    learned = args.evalue == "learned" and engine == F.GX_ENGINE_BATCHED
    if learned:
        extra["surrogate"] = train_surrogate(s, index, sc, dM, args)
        params = SearchParams(dM=dM, engine=engine, config=args.config, evalue_mode=F.GX_EVALUE_LEARNED,
                              surrogate_conf=args.surrogate_conf, whole_bin=args.whole_bin)
        # the e-value stage's device time (HIP events around it), eager runs of
        # the two forms alternated, median of 3 each
        ev_l, ev_p = [], []
        for _ in range(3 if batch.n_spectra <= 200_000 else 1):
            st = s.search_device(dq, params)
            ev_l.append(st["evalue_ms"])
            ev_p.append(s.search_device(dq, plain)["evalue_ms"])
        log(f"eager learned: total {st['total_ms']:.1f} ms (evalue stage {ev_l} ms against plain {ev_p} ms, "
            f"{st['fallbacks']} plain-fit fallbacks)")
        extra["learned_evalue_ms"] = round(statistics.median(ev_l), 4)
        extra["plain_evalue_ms"] = round(statistics.median(ev_p), 4)
    else:
        params = plain
        extra["plain_evalue_ms"] = round(eager["evalue_ms"], 4)
    params.use_graph = use_graph

    def timed(p, label):
        for i in range(args.warmup):
            st = s.search_device(dq, p)
            log(f"{label} warmup {i}: total {st['total_ms']:.1f} ms")
        t0 = time.perf_counter()
        for i in range(args.steps):
            st = s.search_device(dq, p)
            log(f"{label} step {i}: total {st['total_ms']:.1f} ms")
        return (time.perf_counter() - t0) * 1000.0 / args.steps, st

    ms_step, st_last = timed(params, "learned" if learned else "plain")
    psm, _ = s.download(dq)
    value = batch.n_spectra / (ms_step / 1000.0)
    if learned:
        raw = s.surrogate_outputs(dq)
        extra["judge"] = judge_learned(psm, judge, st_last, raw)
        log(f"judge: {extra['judge']}")
        # the plain tail-fit form, timed the same way
        pp = SearchParams(dM=dM, engine=engine, config=args.config, use_graph=use_graph, whole_bin=args.whole_bin)
        ms_plain, _ = timed(pp, "plain")
        extra["plain_batched_spectra_per_s"] = round(batch.n_spectra / (ms_plain / 1000.0), 1)
        extra["plain_batched_ms_per_step"] = round(ms_plain, 3)
        # learned and plain steps alternated, so that drift of the clock or
        # the device over the run falls on both alike
        al, ap = [], []
        for i in range(args.steps):
            for p_, acc in ((params, al), (pp, ap)):
                t0 = time.perf_counter()
                s.search_device(dq, p_)
                acc.append((time.perf_counter() - t0) * 1000.0)
        extra["alternated_ms_per_step"] = {"learned": round(statistics.mean(al), 3),
                                           "plain": round(statistics.mean(ap), 3),
                                           "learned_median": round(statistics.median(al), 3),
                                           "plain_median": round(statistics.median(ap), 3)}
        log(f"alternated steps: {extra['alternated_ms_per_step']}")
    if engine == F.GX_ENGINE_BATCHED and not args.no_compare_spectrum:
        pp = SearchParams(dM=dM, engine=F.GX_ENGINE_SPECTRUM)
        st = s.search_device(dq, pp)
        extra["plain_spectrum_engine_ms"] = round(st["total_ms"], 3)
        extra["plain_spectrum_engine_spectra_per_s"] = round(batch.n_spectra / (st["total_ms"] / 1000.0), 1)
        log(f"per-spectrum engine with the plain fit: {st['total_ms']:.1f} ms")
    st = eager
    avg_kms = st["kernel_ms"]
    abytes = algorithmic_bytes(st, batch.n_peaks)
    achieved = 4 * st["ions"] / (avg_kms / 1000.0) / 1e9  # = per-launch bytes / average launch time
    launches = max(1, int(st["chunks"])) if engine == F.GX_ENGINE_BATCHED else 1
    reported = int(((psm["flags"] & 8) > 0).sum())
    log(f"{value:.0f} spectra/s, kernel {avg_kms:.1f} ms, {achieved:.0f} GB/s algorithmic, "
        f"H={st['ions']} B={st['bins']} W={st['window']} cand={st['candidates']} reported={reported}")

    measured = {}
    try:
        if args.no_bandwidth:
            raise RuntimeError("skipped (--no-bandwidth)")
        measured = {"sequential": round(s.read_bandwidth(0), 1), "random_1x128B": round(s.read_bandwidth(1), 1),
                    "random_2x128B": round(s.read_bandwidth(2), 1)}
        log(f"measured read bandwidth: {measured}")
    except Exception as e:  # diagnostic only
        log(f"read bandwidth not measured: {e}")
    cpu = None
    if not args.no_cpu:
        # a GPU-built index is searched on the CPU through a host copy of its arrays
        cpu = cpu_baseline(index if index.has_fragments else s.host_index(), batch, O.default_params(dM=dM),
                           args.cpu_budget)
        log(f"cpu baseline: {cpu}")

    whole = args.whole_bin or extra.get("autotune", {}).get("whole_bin", 64)
    traffic = (traffic_from_profiles(args.workload, eager["config"], whole) if engine == F.GX_ENGINE_BATCHED
               else None)
    out = {
        "metric": ("spectra/sec, 2M-peptide open search: learned+batched vs plain tail-fit, 1 GPU"
                   if args.workload == "open2m" else f"spectra/sec, {args.workload} workload, 1 GPU"),
        "value": round(value, 1),
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "peptides": index.n_kept,
            "index_entries": index.n_entries,
            "index_ions": index.n_ions,
            "spectra": batch.n_spectra,
            "dM_da": dM,
            "dF_bins": params.dF,
            "evalue": "learned surrogate, plain tail-fit fallback" if learned else "plain tail fit",
            "engine": args.engine,
            "launch_config": eager["config"],
            "graph": bool(params.use_graph),
            "plan_ms": round(eager["plan_ms"], 3),
            **extra,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "search_spectra" if engine == F.GX_ENGINE_BATCHED else "search_kernel",
            "launches": launches,
            "avg_launch_ms": round(avg_kms / launches, 4),
            "algorithmic_bytes_per_launch": round(4 * st["ions"] / launches),
            "algorithmic_bytes_rule": "4 B per window-restricted iA id (H)",
            "job_algorithmic_bytes": abytes,
            "job_algorithmic_gbs": round(abytes / (ms_step / 1000.0) / 1e9, 1),
            # measured on this device (gx_measure_read_bandwidth), beside the published peak
            "measured_read_gbs": measured,
            "frac_of_measured_random_2x128B": (round(achieved / measured["random_2x128B"], 4)
                                               if measured.get("random_2x128B") else None),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
