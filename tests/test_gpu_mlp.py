"""Learned e-value surrogate on the GPU: the MFMA forward pass and the
backward pass against the CPU oracle, the features of searched batches,
training on the GPU, and the learned search mode (top hits identical to the
plain mode, plain fit as fallback and judge)."""
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


@pytest.mark.parametrize("n", [50, 64])
def test_gradients_match_oracle(searcher_open, n):
    say(f"case: MLP gradient n={n}")
    w = _params(7 + n)
    rng = np.random.default_rng(n)
    x = rng.normal(0, 0.5, (n, M.FEAT)).astype(np.float32)
    target = np.stack([rng.normal(-0.03, 0.01, n), rng.normal(-2.0, 1.0, n)], 1).astype(np.float32)
    loge = rng.normal(-3.0, 2.0, n).astype(np.float32)
    normv = np.array([-0.03, 0.01, -2.0, 1.0], np.float32)
    s = Surrogate(searcher_open, seed=None)
    s.set_params(w)
    g, out, loss = s.grad(x, target, loge, normv, 0.5, 0.5)
    gr, outr, lossr = M.backward(w.astype(np.float64), x, target.astype(np.float64), loge.astype(np.float64),
                                 normv.astype(np.float64), 0.5, 0.5, quantize=True)
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


def test_features_match_oracle(open_index, searcher_open):
    say("case: surrogate features of a searched batch")
    batch = _batch(open_index, 800, 301)
    dq = searcher_open.upload(batch)
    searcher_open.search_device(dq, SearchParams(dM=500.0))
    psm, hist = searcher_open.download(dq, want_hist=True)
    f = searcher_open.features(dq)
    feat, anchor, target, loge, valid, sparse = M.features(hist, psm)
    assert np.array_equal(f["valid"], valid) and np.array_equal(f["sparse"], sparse)
    m = valid
    assert m.sum() > 500
    # device and host log10 may differ in the last bit before the bf16 rounding
    ulp = np.abs(feat) * 2.0 ** -7 + 1e-30
    assert np.all(np.abs(f["feat"] - feat) <= ulp + 1e-7)
    assert np.allclose(f["anchor"][m], anchor[m], rtol=1e-6, atol=1e-6)
    assert np.allclose(f["target"][m], target[m], rtol=1e-6, atol=1e-6)
    assert np.allclose(f["loge"][m], loge[m], rtol=1e-6, atol=1e-6)
    dq.free()


def _u_error(out, anchor, target, normv):
    u_pred = anchor + out[:, 1] * normv[3] + normv[2]
    return np.abs(u_pred - (anchor + target[:, 1]))


def test_training_and_learned_search(open_index, searcher_open):
    """Train on batches searched with the plain fit, then search a held-out
    batch in learned mode: top hits, counts and histograms are unchanged,
    every e-value the surrogate did not serve is the plain fit's, and at
    least 99% of the spectra stay within 1e-3 relative of log10 e."""
    say("case: surrogate training + learned search")
    s = Surrogate(searcher_open, seed=3)
    plain = SearchParams(dM=500.0)
    for seed in (311, 312, 313):
        dq = searcher_open.upload(_batch(open_index, 6000, seed))
        searcher_open.search_device(dq, plain)
        assert s.add_training(dq) > 4000
        dq.free()
    held = _batch(open_index, 4000, 399)
    dq = searcher_open.upload(held)
    searcher_open.search_device(dq, plain)
    ref_psm, ref_hist = searcher_open.download(dq, want_hist=True)
    f = searcher_open.features(dq)
    v = f["valid"]
    # the untrained network, with the normalisation training will use
    rep = s.train(TrainParams(batch=512, epochs=1, lr=0.0, seed=1))  # fixes the normalisation only
    w0, normv = s.get_params()
    untrained_err = _u_error(s.predict(f["feat"][v]), f["anchor"][v], f["target"][v], normv)
    rep = s.train(TrainParams(batch=512, epochs=12, lr=2e-3, lr_decay=0.8, conf_weight=0.5, tol=1e-3, seed=2))
    say(f"train: {rep['samples']} rows, mse_u {rep['mse_u'][0]:.4f} -> {rep['mse_u'][-1]:.4f}")
    assert rep["mse_u"][-1] < rep["mse_u"][0]
    trained_err = _u_error(s.predict(f["feat"][v]), f["anchor"][v], f["target"][v], normv)
    say(f"median |du|: untrained {np.median(untrained_err):.3f}, trained {np.median(trained_err):.3f}")
    assert np.median(trained_err) < 0.5 * np.median(untrained_err)

    s.attach()
    learned = SearchParams(dM=500.0, evalue_mode=F.GX_EVALUE_LEARNED, surrogate_conf=0.9)
    st = searcher_open.search_device(dq, learned)
    psm, hist = searcher_open.download(dq, want_hist=True)
    assert np.array_equal(hist, ref_hist)
    for k in ("hyperscore", "psid", "idxoffset", "sharedions", "totalions", "cpsms", "hyp"):
        assert np.array_equal(psm[k], ref_psm[k]), k
    sur = (psm["flags"] & F.GX_PSM_SURROGATE) != 0
    modeled = (ref_psm["flags"] & F.GX_PSM_MODELED) != 0
    assert np.all(modeled[sur])
    assert st["fallbacks"] == int(modeled.sum() - sur.sum())
    rest = ~sur
    for k in ("mu", "beta", "flags"):
        assert np.array_equal(psm[k][rest], ref_psm[k][rest]), k
    assert np.array_equal(psm["evalue"][rest].view(np.uint64), ref_psm["evalue"][rest].view(np.uint64))
    # the plain fit judges every modeled spectrum; where it is undefined, a
    # surrogate value counts as an error
    fit = (ref_psm["flags"] & F.GX_PSM_FIT_OK) != 0
    within = np.zeros(len(psm), bool)
    le_ref = np.log10(ref_psm["evalue"][fit])
    le = np.log10(psm["evalue"][fit])
    within[fit] = np.abs(le - le_ref) <= 1e-3 * np.maximum(np.abs(le_ref), 1e-3)
    within[~fit] = ~sur[~fit]
    frac = within[modeled].mean()
    say(f"learned: {sur.sum()} of {modeled.sum()} modeled spectra served by the surrogate, "
        f"{frac:.4f} within 1e-3 rel log10")
    assert frac >= 0.99

    # the same through a captured graph, replayed
    gp = dataclasses.replace(learned, use_graph=True)
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
