"""Learned e-value surrogate on the GPU: the MFMA forward pass and the
backward pass against the CPU oracle, the features and targets of searched
batches, the fused learned stage against the oracle's restatement of it on
the kernel's own network outputs (forced acceptance), training on the GPU,
and the learned search mode against the plain fit (top hits identical, served
e-values within 1e-3 relative of log10 e, the rest bit-identical)."""
import dataclasses

import numpy as np
import pytest

from gicops_amd import _ffi as F
from gicops_amd import slm
from gicops_amd.search import SearchParams
from gicops_amd.surrogate import Surrogate, TrainParams
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _params(seed):
    rng = np.random.default_rng(seed)
    p = M.unpack(M.init_params(seed))
    p["b1"][:] = rng.normal(0, 0.1, M.HIDDEN)
    p["b2"][:] = rng.normal(0, 0.1, M.HIDDEN)
    p["b3"][:M.OUT] = rng.normal(0, 0.1, M.OUT)
    return M.pack(p).astype(np.float32)


@pytest.mark.parametrize("n", [1, 37, 300])
def test_mfma_forward_matches_oracle(searcher_open, n):
    say(f"case: MLP forward n={n}")
    w = _params(n)
    s = Surrogate(searcher_open, seed=None)
    s.set_params(w)
    x = np.random.default_rng(100 + n).normal(0, 0.5, (n, M.FEAT)).astype(np.float32)
    out = s.predict(x)
    ref = M.forward(w.astype(np.float64), x, quantize=True)["out"]
    assert np.all(out[:, M.OUT:] == 0.0)
    err = np.abs(out[:, :M.OUT] - ref[:, :M.OUT])
    assert err.max() <= 2e-3 * (1 + np.abs(ref[:, :M.OUT]).max()), err.max()
    s.close()


# 1500 rows: the weight-gradient GEMMs (K = rows) split K over blocks
@pytest.mark.parametrize("n", [50, 64, 1500])
def test_gradients_match_oracle(searcher_open, n):
    say(f"case: MLP gradient n={n}")
    w = _params(7 + n)
    rng = np.random.default_rng(n)
    x = rng.normal(0, 0.5, (n, M.FEAT)).astype(np.float32)
    m = rng.integers(20, 400, n)
    mark = (rng.random(n) * 0.5 * (m - 1)).astype(int)
    mark2 = np.maximum(mark + 1, ((0.7 + 0.3 * rng.random(n)) * (m - 1)).astype(int))
    target = np.stack([mark, mark2, m - 1, (rng.random(n) < 0.8)], 1).astype(np.float32)
    normv = np.array([0.25, 0.15, 0.95, 0.05], np.float32)
    s = Surrogate(searcher_open, seed=None)
    s.set_params(w)
    g, out, loss = s.grad(x, target, normv, 0.5)
    gr, outr, lossr = M.backward(w.astype(np.float64), x, target.astype(np.float64), normv.astype(np.float64), 0.5,
                                 quantize=True)
    assert np.allclose(out[:, :M.OUT], outr[:, :M.OUT], rtol=2e-3, atol=2e-3)
    # the confidence labels must agree for the gradients to be comparable
    assert abs(loss[3] - lossr[3]) <= 1, (loss, lossr)
    assert np.allclose(loss[:3], lossr[:3], rtol=1e-2, atol=1e-4)
    gp, rp = M.unpack(g.astype(np.float64)), M.unpack(gr)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        rel = np.linalg.norm(gp[k] - rp[k]) / max(np.linalg.norm(rp[k]), 1e-12)
        assert rel < 2e-2, (k, rel)
    assert np.all(gp["W3"][:, M.OUT:] == 0) and np.all(gp["b3"][M.OUT:] == 0)
    s.close()


def _batch(index, n, seed):
    return slm.generate_spectra(index, dataclasses.replace(slm.SPEC_PRESETS["open2m"], n_spectra=n, seed=seed))


def _plain(searcher, index, n, seed):
    dq = searcher.upload(_batch(index, n, seed))
    searcher.search_device(dq, SearchParams(dM=500.0))
    psm, hist = searcher.download(dq, want_hist=True)
    return dq, psm, hist


def test_features_match_oracle(open_index, searcher_open):
    say("case: surrogate features and targets of a searched batch")
    dq, psm, hist = _plain(searcher_open, open_index, 800, 301)
    f = searcher_open.features(dq)
    feat, target, loge, valid, serv = M.features(hist, psm)
    assert np.array_equal(f["valid"], valid) and np.array_equal(f["servable"], serv)
    assert serv.sum() > 400
    # the exact marks (double log10 on both sides) and the curve lengths
    assert np.array_equal(f["target"], target)
    # device and host log10f may differ in the last bit before the bf16 rounding
    ulp = np.abs(feat) * 2.0 ** -7 + 1e-30
    assert np.all(np.abs(f["feat"] - feat) <= ulp + 1e-7)
    m = valid
    assert np.allclose(f["loge"][m], loge[m], rtol=1e-6, atol=1e-6)
    dq.free()


def _check_learned(searcher, dq, ref_psm, ref_hist, w, normv, conf, label):
    """One learned search of a batch checked against the plain search of the
    same batch and the oracle's restatement of the stage on the kernel's own
    network outputs. Returns (served mask, raw outputs)."""
    learned = SearchParams(dM=500.0, evalue_mode=F.GX_EVALUE_LEARNED, surrogate_conf=conf)
    st = searcher.search_device(dq, learned)
    psm, hist = searcher.download(dq, want_hist=True)
    raw = searcher.surrogate_outputs(dq)
    # top hits, counts and histograms do not depend on the e-value mode
    assert np.array_equal(hist, ref_hist)
    for k in ("hyperscore", "psid", "idxoffset", "sharedions", "totalions", "cpsms", "hyp"):
        assert np.array_equal(psm[k], ref_psm[k]), k
    sur = (psm["flags"] & F.GX_PSM_SURROGATE) != 0
    assert np.array_equal(sur, raw[:, 3] > 0.5)
    modeled = (ref_psm["flags"] & F.GX_PSM_MODELED) != 0
    assert np.all(modeled[sur])
    assert st["fallbacks"] == int(modeled.sum() - sur.sum())
    # every spectrum not served keeps the plain fit bit for bit
    rest = ~sur
    for k in ("mu", "beta", "flags"):
        assert np.array_equal(psm[k][rest], ref_psm[k][rest]), k
    assert np.array_equal(psm["evalue"][rest].view(np.uint64), ref_psm["evalue"][rest].view(np.uint64))
    # the kernel's network outputs against the oracle's forward pass on the oracle's features
    feat, _, _, _, serv = M.features(ref_hist, ref_psm)
    ref_out = M.forward(w.astype(np.float64), feat, quantize=True)["out"]
    m = modeled & ((ref_psm["hyp"] > 0))
    err = np.abs(raw[m, :3] - ref_out[m, :3])
    assert err.max() <= 3e-3 * (1 + np.abs(ref_out[m, :3]).max()), err.max()
    # which spectra are served, and their results, from the same outputs
    o_served, o_mu, o_beta, o_ev, o_rep = M.serve(raw, ref_hist, ref_psm, normv, conf)
    assert np.array_equal(sur, o_served)
    assert np.all(np.abs(psm["mu"][sur].astype(np.int64) - o_mu[sur]) <= 1)
    assert np.all(np.abs(psm["beta"][sur].astype(np.int64) - o_beta[sur]) <= 1)
    # a served e-value is formed from its own truncated mu, beta
    for i in np.nonzero(sur)[0]:
        e = M.evalue_of(int(psm["mu"][i]), int(psm["beta"][i]), int(psm["hyp"][i]), int(psm["cpsms"][i]))
        assert psm["evalue"][i] == pytest.approx(e, rel=1e-12)
        assert bool(psm["flags"][i] & F.GX_PSM_REPORTED) == (psm["evalue"][i] < 20.0)
        assert psm["flags"][i] & F.GX_PSM_FIT_OK
    # the plain fit judges every served e-value
    fit = (ref_psm["flags"] & F.GX_PSM_FIT_OK) != 0
    assert np.all(fit[sur])
    le_ref = np.log10(ref_psm["evalue"][sur])
    le = np.log10(psm["evalue"][sur])
    rel = np.abs(le - le_ref) / np.maximum(np.abs(le_ref), 1e-3)
    assert np.all(rel <= 1e-3), rel.max()
    say(f"  {label}: served {int(sur.sum())} of {int(modeled.sum())} modeled ({int(serv.sum())} servable), "
        f"max rel. log10 e error {rel.max() if len(rel) else 0.0:.2e}, fallbacks {st['fallbacks']}")
    return sur, raw


def test_learned_stage_serves_and_matches_oracle(open_index, searcher_open):
    """The surrogate's accept branch, forced: (1) a network whose outputs are
    its biases (every spectrum predicts the same two fractions, confidence
    high), and (2) random weights with the confidence threshold at 0. Each
    served spectrum is checked against the oracle's restatement of the stage
    on the same outputs and against the plain fit."""
    say("case: learned stage, forced acceptance")
    dq, ref_psm, ref_hist = _plain(searcher_open, open_index, 3000, 321)
    s = Surrogate(searcher_open, seed=None)
    normv = np.array([0.0, 1.0, 0.0, 1.0], np.float32)
    p = M.unpack(np.zeros(M.N_PARAMS))
    p["b3"][:3] = [0.2, 0.98, 8.0]
    w = M.pack(p).astype(np.float32)
    s.set_params(w, normv)
    s.attach()
    sur, _ = _check_learned(searcher_open, dq, ref_psm, ref_hist, w, normv, 0.9, "constant outputs")
    assert sur.sum() > 50
    w = _params(5)
    s.set_params(w, np.array([0.25, 0.2, 0.95, 0.05], np.float32))
    sur, _ = _check_learned(searcher_open, dq, ref_psm, ref_hist, w, np.array([0.25, 0.2, 0.95, 0.05], np.float32),
                            0.0, "random weights, threshold 0")
    assert sur.sum() > 0
    dq.free()
    s.close()


def test_training_and_learned_search(open_index, searcher_open):
    """Train on batches searched with the plain fit, then search a held-out
    batch in learned mode: the trained network serves many more spectra than
    the untrained one at the same threshold, every served e-value is within
    1e-3 relative of the plain fit's log10 e, the rest are the plain fit's,
    and graph replays reproduce the eager result."""
    say("case: surrogate training + learned search")
    s = Surrogate(searcher_open, seed=3)
    for seed in (311, 312, 313):
        dq, _, _ = _plain(searcher_open, open_index, 6000, seed)
        assert s.add_training(dq) > 2000
        dq.free()
    dq, ref_psm, ref_hist = _plain(searcher_open, open_index, 4000, 399)
    # the untrained network, with the normalisation training will use
    s.train(TrainParams(batch=512, epochs=1, lr=0.0, seed=1))
    w0, normv = s.get_params()
    s.attach()
    sur0, _ = _check_learned(searcher_open, dq, ref_psm, ref_hist, w0, normv, 0.5, "untrained")
    rep = s.train(TrainParams(batch=512, epochs=12, lr=2e-3, lr_decay=0.8, conf_weight=0.5, seed=2))
    say(f"train: {rep['samples']} rows, mse {rep['mse_mark'][0]:.4f}/{rep['mse_mark2'][0]:.4f} -> "
        f"{rep['mse_mark'][-1]:.4f}/{rep['mse_mark2'][-1]:.4f}, in window {rep['in_window'][-1]:.3f}")
    assert rep["mse_mark"][-1] < rep["mse_mark"][0]
    w1, normv = s.get_params()
    sur1, raw = _check_learned(searcher_open, dq, ref_psm, ref_hist, w1, normv, 0.5, "trained")
    modeled = (ref_psm["flags"] & F.GX_PSM_MODELED) != 0
    assert sur1.sum() > 0.3 * modeled.sum()
    assert sur1.sum() > 2 * max(1, sur0.sum())
    # the same through a captured graph, replayed
    psm, _ = searcher_open.download(dq)
    gp = SearchParams(dM=500.0, evalue_mode=F.GX_EVALUE_LEARNED, surrogate_conf=0.5, use_graph=True)
    searcher_open.search_device(dq, gp)
    for _ in range(2):
        searcher_open.search_device(dq, gp)
        p2, _ = searcher_open.download(dq)
        assert p2.tobytes() == psm.tobytes()
    dq.free()
    s.close()


def test_learned_mode_requires_a_surrogate(open_index):
    from gicops_amd.search import Searcher

    sr = Searcher(open_index)
    batch = _batch(open_index, 50, 398)
    with pytest.raises(F.GxError) as e:
        sr.search(batch, SearchParams(dM=500.0, evalue_mode=F.GX_EVALUE_LEARNED))
    assert e.value.code == F.GX_ERR_NOT_READY
    with pytest.raises(F.GxError) as e:
        sr.search(batch, SearchParams(dM=500.0, evalue_mode=F.GX_EVALUE_LEARNED, engine=F.GX_ENGINE_SPECTRUM))
    assert e.value.code == F.GX_ERR_INVLD_PARAM
    sr.close()
