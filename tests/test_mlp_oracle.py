"""The surrogate's CPU oracle, pinned without the GPU: GELU against the normal
CDF, bf16 rounding against torch, the backward pass against central finite
differences of the float64 loss, the curve and features against hand-computed
values, the window fit against the plain tail-fit oracle (expeRT.cpp:396-559),
and the closed-form window label against a brute-force scan of the windows."""
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
    assert (M.FEAT, M.HIDDEN, M.OUT_PAD, M.TARGET) == (F.MLP_FEAT, F.MLP_HIDDEN, F.MLP_OUT_PAD, F.MLP_TARGET)
    w = np.arange(M.N_PARAMS, dtype=np.float64)
    assert np.array_equal(M.pack(M.unpack(w)), w)


def test_gelu_is_torchs_tanh_gelu():
    z = np.linspace(-8, 8, 1001)
    ref = torch.nn.functional.gelu(torch.from_numpy(z), approximate="tanh").numpy()
    assert np.allclose(M.gelu(z), ref, rtol=1e-12, atol=1e-14)
    # within 1e-3 of the exact GELU z * Phi(z)
    assert np.max(np.abs(M.gelu(z) - z * norm.cdf(z))) < 1e-3
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
    m = rng.integers(20, 400, n)
    mark = (rng.random(n) * 0.5 * (m - 1)).astype(int)
    mark2 = np.maximum(mark + 1, ((0.7 + 0.3 * rng.random(n)) * (m - 1)).astype(int))
    serv = rng.random(n) < 0.8
    target = np.stack([mark, mark2, m - 1, serv.astype(float)], 1).astype(np.float64)
    normv = np.array([0.25, 0.15, 0.95, 0.05])
    return w, x, target, normv


def test_backward_matches_finite_differences():
    n, cw = 24, 0.7
    w, x, target, normv = _problem(n, 3)
    g, out, _ = M.backward(w, x, target, normv, cw, quantize=False)
    # labels held fixed, as the gradient treats them
    _, _, y, _, _, _ = M.loss_terms(out, target, normv, cw)

    def loss(wv):
        o = M.forward(wv, x, quantize=False)["out"]
        return M.loss_terms(o, target, normv, cw, y=y)[-1]

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
    w, x, target, normv = _problem(64, 9)
    g32, _, _ = M.backward(w, x, target, normv, 0.5, quantize=True)
    g64, _, _ = M.backward(w, x, target, normv, 0.5, quantize=False)
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 0.05


def _psm(hyp, cps, flags=2 | 4, mu=0, beta=0):
    r = np.zeros(1, dtype=F.PSM_DTYPE)
    r["hyp"], r["cpsms"], r["flags"], r["mu"], r["beta"] = hyp, cps, flags, mu, beta
    return r


def test_curve_and_features_of_a_hand_histogram():
    h = np.zeros((1, M.HIST_BINS), np.uint32)
    # bins 3..12 below the top hit (bin 20): counts give survival 1 - C/40
    h[0, 3:13] = [8, 6, 5, 0, 4, 3, 3, 2, 1, 1]
    h[0, 20] = 7  # the top hit's bin and 6 more candidates there: cpsms 40
    stt1, end1, c = M.curve_prefix(h[0].astype(np.int64), 20)
    assert (stt1, end1) == (3, 12)
    assert c.tolist() == [8, 14, 19, 19, 23, 26, 29, 31, 32, 33]
    assert M.servable_curve(c, 40)
    ls = M.log_survival(c, 40)
    assert np.allclose(ls, np.log10(1 - c / 40), rtol=0, atol=1e-15)
    # marks: t1 = ls0 + 0.22 (lsN - ls0), t2 = ls0 + 0.87 (lsN - ls0)
    t1 = ls[0] + (ls[-1] - ls[0]) * 0.22
    t2 = ls[0] + (ls[-1] - ls[0]) * 0.87
    f = int(np.nonzero(ls <= t1)[0][0])
    mark2 = int(np.nonzero(ls >= t2)[0][-1])
    assert M.exact_marks(c, 40) == (f, mark2, f - 1)
    feat, target, loge, valid, serv = M.features(h, _psm(20, 40, mu=-250000, beta=-500000))
    assert valid[0] and serv[0]
    assert target[0].tolist() == [f - 1, mark2, 9, 1.0]
    assert loge[0] == np.float32(math.log10(40) - 0.25 * 20 - 0.5)
    # curve points: linear interpolation of the prefix counts at t = j*9/119
    t = np.arange(M.CURVE_PTS) * 9 / (M.CURVE_PTS - 1)
    cnt = np.interp(t, np.arange(10), c)
    assert np.allclose(feat[0, :M.CURVE_PTS], np.log10(1 - cnt / 40) / 4, rtol=2 ** -7, atol=1e-6)
    sc = [3 / 800, 12 / 800, 20 / 800, 8 / 400, 10 / 400, math.log10(40) / 4, math.log10(1 - 33 / 40) / 4,
          math.log10(1 - 8 / 40)]
    assert np.allclose(feat[0, M.CURVE_PTS:], sc, rtol=2 ** -7, atol=1e-7)
    # too short a curve: not servable (the plain fit serves it)
    h2 = np.zeros((1, M.HIST_BINS), np.uint32)
    h2[0, [3, 5, 6, 9]] = [4, 3, 2, 1]
    assert not M.features(h2, _psm(9, 10))[4][0]
    # not modeled: no features
    f3 = M.features(h, _psm(20, 40, flags=1))
    assert not f3[0].any() and not f3[4][0] and not f3[3][0]


def _random_tail(rng):
    h = np.zeros(M.HIST_BINS, np.uint32)
    hyp = int(rng.integers(60, 700))
    k = int(rng.integers(10, hyp - 1))
    lo = int(rng.integers(0, k - 8))
    h[lo:k] = rng.geometric(rng.uniform(0.05, 0.6), k - lo) * (rng.random(k - lo) < rng.uniform(0.3, 1.0))
    h[hyp] = int(rng.integers(1, 4))
    return h, hyp, int(h.sum())


def test_window_fit_equals_the_tail_fit_oracle():
    """On servable curves the exact marks and the window's least-squares line
    give the plain tail fit's slope and intercept (orc_tailfit, a separate C
    restatement of expeRT::ModelTailFit): the marks, thresholds and window
    bounds agree with the reference's."""
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(600):
        h, hyp, cps = _random_tail(rng)
        stt1, end1, c = M.curve_prefix(h.astype(np.int64), hyp)
        if not M.servable_curve(c, cps):
            continue
        f, mark2, mark = M.exact_marks(c, cps)
        if f is None or f < 1 or mark < 0:
            continue
        ok, slope, icept, mu, beta = O.tailfit(h, hyp, cps)
        assert ok
        a, b = M.window_fit(M.log_survival(c, cps), stt1, mark, mark2)
        assert math.isclose(a, slope, rel_tol=1e-10, abs_tol=1e-13), (a, slope)
        assert math.isclose(b, icept, rel_tol=1e-10, abs_tol=1e-12), (b, icept)
        assert abs(int(a * 1e6) - mu) <= 1 and abs(int(b * 1e6) - beta) <= 1
        checked += 1
    assert checked > 300


def test_window_label_equals_the_window_scan():
    """snap_label (the training label, a closed form) agrees with the scan the
    kernel does on the curve (snap_windows) for every predicted index pair on
    random servable curves, and a successful scan returns the exact marks."""
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(200):
        h, hyp, cps = _random_tail(rng)
        stt1, end1, c = M.curve_prefix(h.astype(np.int64), hyp)
        if not M.servable_curve(c, cps):
            continue
        f, mark2, mark = M.exact_marks(c, cps)
        if f is None or f < 1 or mark < 0:
            continue
        n = len(c)
        ls = M.log_survival(c, cps)
        qs = sorted(set([0, n - 1, f - 1, f, mark2, max(0, mark2 - 20)] + list(rng.integers(0, n, 12))))
        for q1 in qs:
            for q2 in qs:
                got = M.snap_windows(ls, q1, q2)
                assert (got is not None) == M.snap_label(q1, q2, f - 1, mark2, n), (n, q1, q2, f, mark2)
                if got is not None:
                    assert got == (mark, mark2)
                checked += 1
    assert checked > 5000


def test_predicted_index_rounds_to_nearest_and_clamps():
    assert M.predicted_index(0.0, 1.0, 0.5, 101) == 50
    assert M.predicted_index(0.0, 1.0, 0.505, 101) == 50  # 50.5 rounds to even
    # float32(0.515) * 100 rounds to 51.5 in float32, and 51.5 to the even 52
    assert M.predicted_index(0.0, 1.0, 0.515, 101) == 52
    assert M.predicted_index(0.5, 0.02, 0.5, 101) == 51  # (0.5 * 0.02 + 0.5) * 100
    assert M.predicted_index(-9.0, 1.0, 0.0, 101) == 0
    assert M.predicted_index(9.0, 1.0, 0.0, 101) == 100
    assert M.predicted_index(float("nan"), 1.0, 0.0, 101) == 0


def test_serve_with_outputs_at_the_true_marks():
    """Outputs placed on a curve's true marks are served with the plain fit's
    window: mu and beta within one unit of orc_tailfit's, the e-value from the
    truncated integers. A low confidence logit or a short curve is not
    served."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        h, hyp, cps = _random_tail(rng)
        stt1, end1, c = M.curve_prefix(h.astype(np.int64), hyp)
        if not M.servable_curve(c, cps):
            continue
        f, mark2, mark = M.exact_marks(c, cps)
        if f is None or f < 1 or mark < 0:
            continue
        n = len(c)
        psm = _psm(hyp, cps, flags=2)
        out = np.array([[(f - 1) / (n - 1), mark2 / (n - 1), 5.0]])
        served, mu, beta, ev, rep = M.serve(out, h[None, :], psm, [0.0, 1.0, 0.0, 1.0], 0.9)
        assert served[0]
        ok, slope, icept, pmu, pbeta = O.tailfit(h, hyp, cps)
        assert abs(mu[0] - pmu) <= 1 and abs(beta[0] - pbeta) <= 1
        assert ev[0] == M.evalue_of(int(mu[0]), int(beta[0]), hyp, cps)
        out[0, 2] = -5.0  # confidence 0.0067 < 0.9
        assert not M.serve(out, h[None, :], psm, [0.0, 1.0, 0.0, 1.0], 0.9)[0][0]
