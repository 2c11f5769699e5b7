"""The surrogate's CPU oracle, pinned without the GPU: GELU against the normal
CDF, bf16 rounding against torch, the backward pass against central finite
differences of the float64 loss, and the survival curve against hand-computed
values and the tail-fit oracle (expeRT.cpp:414-559)."""
import math

import numpy as np
import pytest
import torch
from scipy.stats import norm

from gicops_amd import _ffi as F
from oracle import mlp_oracle as M
from oracle import pyoracle as O


def test_layout_matches_library_constants():
    assert M.N_PARAMS == 128 * 256 + 256 + 256 * 256 + 256 + 256 * 16 + 16
    assert (M.FEAT, M.HIDDEN, M.OUT_PAD) == (F.MLP_FEAT, F.MLP_HIDDEN, F.MLP_OUT_PAD)
    w = np.arange(M.N_PARAMS, dtype=np.float64)
    assert np.array_equal(M.pack(M.unpack(w)), w)


def test_gelu_is_z_times_normal_cdf():
    z = np.linspace(-8, 8, 1001)
    assert np.allclose(M.gelu(z), z * norm.cdf(z), rtol=1e-12, atol=1e-14)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(z), fd, rtol=1e-6, atol=1e-8)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 1000), rng.normal(0, 1e4, 1000),
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(x), ref)
    # ties to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert M.bf16(np.float32(1.0 + 2.0 ** -8))[()] == 1.0
    assert M.bf16(np.float32(1.0 + 3 * 2.0 ** -8))[()] == 1.0 + 2.0 ** -6


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    w = M.init_params(seed)
    p = M.unpack(w)
    p["b1"][:] = rng.normal(0, 0.1, M.HIDDEN)
    p["b2"][:] = rng.normal(0, 0.1, M.HIDDEN)
    p["b3"][:M.OUT] = rng.normal(0, 0.1, M.OUT)
    w = M.pack(p)
    x = rng.normal(0, 0.5, (n, M.FEAT))
    target = np.stack([rng.normal(-0.03, 0.01, n), rng.normal(-2.0, 1.0, n)], 1)
    loge = rng.normal(-1.0, 2.0, n)
    normv = np.array([-0.03, 0.01, -2.0, 1.0])
    return w, x, target, loge, normv


def test_backward_matches_finite_differences():
    n, cw, tol = 24, 0.7, 5e-2
    w, x, target, loge, normv = _problem(n, 3)
    g, out, _ = M.backward(w, x, target, loge, normv, cw, tol, quantize=False)
    # labels held fixed, as the gradient treats them
    _, _, y, _, _, _ = M.loss_terms(out, target, loge, normv, cw, tol)

    def loss(wv):
        o = M.forward(wv, x, quantize=False)["out"]
        return M.loss_terms(o, target, loge, normv, cw, tol, y=y)[-1]

    rng = np.random.default_rng(5)
    blocks = [(M.W1, M.B1), (M.B1, M.W2), (M.W2, M.B2), (M.B2, M.W3), (M.B3, M.B3 + M.OUT)]
    idx = np.concatenate([rng.integers(a, b, 12) for a, b in blocks])
    # W3's first three columns
    idx = np.concatenate([idx, M.W3 + rng.integers(0, M.HIDDEN, 12) * M.OUT_PAD + rng.integers(0, M.OUT, 12)])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (loss(wp) - loss(wm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (int(i), fd, g[i])
    # padding output columns get no gradient
    W3 = M.unpack(g)["W3"]
    assert np.all(W3[:, M.OUT:] == 0) and np.all(M.unpack(g)["b3"][M.OUT:] == 0)


def test_quantized_backward_is_close_to_exact():
    w, x, target, loge, normv = _problem(64, 9)
    g32, _, _ = M.backward(w, x, target, loge, normv, 0.5, 1e-2, quantize=True)
    g64, _, _ = M.backward(w, x, target, loge, normv, 0.5, 1e-2, quantize=False)
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 0.05


def _psm(hyp, cps, flags=2 | 4, mu=0, beta=0):
    r = np.zeros(1, dtype=F.PSM_DTYPE)
    r["hyp"], r["cpsms"], r["flags"], r["mu"], r["beta"] = hyp, cps, flags, mu, beta
    return r


def test_features_of_a_hand_histogram():
    h = np.zeros((1, M.HIST_BINS), np.uint32)
    h[0, [3, 4, 6, 9]] = [4, 3, 2, 1]  # cpsms 10, the top hit alone in bin 9
    feat, anchor, target, loge, valid, sparse = M.features(h, _psm(9, 10, mu=-250000, beta=-500000))
    # curve: bins 3..6, s = 1 - cumsum/10 = 0.6, 0.3, 0.3, 0.1
    ls = np.log10([0.6, 0.3, 0.3, 0.1])
    assert anchor[0] == np.float32(ls[-1])
    t = np.arange(M.CURVE_PTS) * 3 / (M.CURVE_PTS - 1)
    assert np.array_equal(feat[0, :M.CURVE_PTS], M.bf16((np.interp(t, np.arange(4), ls) / 4).astype(np.float32)))
    sc = [3 / 800, 6 / 800, 9 / 800, 3 / 400, 4 / 400, 1 / 4, ls[-1] / 4, ls[0]]
    assert np.array_equal(feat[0, M.CURVE_PTS:], M.bf16(np.array(sc, np.float32)))
    u = -0.25 * 9 - 0.5
    assert target[0, 0] == np.float32(-0.25) and target[0, 1] == np.float32(u - ls[-1])
    assert loge[0] == np.float32(1.0 + u)
    assert valid[0] and not sparse[0]
    # two non-empty bins below the top hit: sparse
    h2 = np.zeros((1, M.HIST_BINS), np.uint32)
    h2[0, [3, 6, 9]] = [5, 4, 1]
    assert M.features(h2, _psm(9, 10))[5][0]
    # not modeled: no features, sparse
    f3 = M.features(h, _psm(9, 10, flags=1))
    assert not f3[0].any() and f3[5][0] and not f3[4][0]


def test_curve_agrees_with_the_tail_fit_oracle():
    """With two curve bins the fit's line passes through both points
    (mark 0, mark2 1), so its slope is their difference and its value at end1
    is the anchor."""
    h = np.zeros(M.HIST_BINS, np.uint32)
    h[[10, 11, 30]] = [5, 3, 1]
    ok, slope, icept, mu, beta = O.tailfit(h, 30, 9)
    assert ok
    stt1, end1, ls, bad = M.curve(h.astype(np.int64), 30, 9)
    assert (stt1, end1, bad) == (10, 11, False)
    assert math.isclose(slope, ls[1] - ls[0], rel_tol=1e-12)
    assert math.isclose(slope * end1 + icept, ls[-1], rel_tol=1e-12, abs_tol=1e-12)
    # and on random tails, the fitted line at end1 stays within the curve's range
    rng = np.random.default_rng(1)
    for _ in range(200):
        h = np.zeros(M.HIST_BINS, np.uint32)
        hyp = int(rng.integers(60, 400))
        k = int(rng.integers(5, hyp - 1))
        h[:k] = rng.geometric(0.3, k) * (rng.random(k) < 0.7)
        h[hyp] = 1
        cps = int(h.sum())
        ok, slope, icept, _, _ = O.tailfit(h, hyp, cps)
        if not ok or np.count_nonzero(h[:hyp]) < 3:
            continue
        stt1, end1, ls, bad = M.curve(h.astype(np.int64), hyp, cps)
        assert not bad
        v = slope * end1 + icept
        assert ls.min() - 1.0 <= v <= ls.max() + 1.0


def test_predicted_evalue_closed_form():
    out = np.zeros((2, M.OUT_PAD))
    out[:, 1] = [0.5, -1.0]
    out[:, 2] = [0.0, 100.0]
    e, u, conf = M.predict_evalue(out, np.array([-2.0, -3.0]), _psm(9, 100).repeat(2), [0, 1, -1.0, 2.0])
    assert np.allclose(u, [-2.0 + 0.0, -3.0 - 3.0])
    assert np.allclose(e, 100 * 10.0 ** u)
    assert np.allclose(conf, [0.5, 1.0])
