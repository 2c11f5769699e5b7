"""CPU restatement of the learned e-value surrogate - TEST INFRASTRUCTURE ONLY.

Imported by tests/ and bench.py's accuracy report, never by the gicops_amd
package. It restates, in numpy float64:

- the surrogate features of a searched spectrum: the log10 survival curve
  that expeRT::ModelTailFit regresses on (expeRT.cpp:414-475), resampled to
  120 points, and 8 scalars (curve bounds, top bin, length, candidate count,
  the curve's last value as the anchor, its first value);
- the targets: the plain fit's slope mu/1e6 and its fitted log10 survival at
  the top bin u = mu/1e6*hyp + beta/1e6 (expeRT.cpp:526-559; the e-value is
  cpsms * 10^u), the latter as a residual over the anchor;
- the MLP forward pass, loss and backward pass, with the same bf16 rounding
  points as the HIP kernels when `quantize=True` (inputs, weights, hidden
  activations, output gradient and hidden-layer gradients are rounded to
  bf16; accumulation, biases and master weights stay wide);
- the confidence routing.

It is pinned by central finite differences of the float64 loss (backward vs
forward), by GELU against the normal CDF, and by hand-computed features of
small histograms (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

FEAT, HIDDEN, OUT_PAD, OUT, CURVE_PTS, HIST_BINS = 128, 256, 16, 3, 120, 802
MIN_TAIL_BINS = 3
PSM_MODELED, PSM_FIT_OK = 2, 4

# flat parameter layout (row-major blocks, in this order)
W1 = 0
B1 = W1 + FEAT * HIDDEN
W2 = B1 + HIDDEN
B2 = W2 + HIDDEN * HIDDEN
W3 = B2 + HIDDEN
B3 = W3 + HIDDEN * OUT_PAD
N_PARAMS = B3 + OUT_PAD


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (round to nearest, ties to even) and
    return them as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(w: np.ndarray) -> dict:
    w = np.asarray(w, dtype=np.float64)
    return {"W1": w[W1:B1].reshape(FEAT, HIDDEN), "b1": w[B1:W2], "W2": w[W2:B2].reshape(HIDDEN, HIDDEN),
            "b2": w[B2:W3], "W3": w[W3:B3].reshape(HIDDEN, OUT_PAD), "b3": w[B3:N_PARAMS]}


def pack(d: dict) -> np.ndarray:
    return np.concatenate([d["W1"].ravel(), d["b1"], d["W2"].ravel(), d["b2"], d["W3"].ravel(), d["b3"]])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


# ---- features -------------------------------------------------------------
def curve(h: np.ndarray, hyp: int, cps: int):
    """The plain fit's log10 survival curve (expeRT.cpp:414-475): bins stt1
    (first non-empty) .. end1 (last non-empty below hyp; +1 when equal),
    s = 1 - cumsum/cpsms from stt1, s > 1 -> 0.999, s <= 0 -> the last value
    >= 1e-4 (or the last value), log10. Returns (stt1, end1, values, bad)."""
    nz = np.nonzero(h[:hyp])[0]
    end1 = int(nz[-1]) if len(nz) else hyp - 1
    stt1 = int(nz[0]) if len(nz) else 0
    if stt1 == end1:
        end1 += 1
    n = end1 - stt1 + 1
    seg = np.zeros(n, np.int64)
    top = min(end1 + 1, HIST_BINS)
    seg[:top - stt1] = h[stt1:top]
    sx = -(np.cumsum(seg) / cps - 1.0)
    sx[sx > 1.0] = 0.999
    good = np.nonzero(sx >= 1e-4)[0]
    rv = sx[good[-1]] if len(good) else sx[n - 1]
    sx = np.where(sx <= 0.0, rv, sx)
    with np.errstate(divide="ignore", invalid="ignore"):
        ls = np.log10(sx)
    bad = not np.all(np.isfinite(ls))
    ls = np.where(np.isfinite(ls), ls, -8.0)
    return stt1, end1, ls, bad


def features(hist: np.ndarray, psm: np.ndarray):
    """Features, anchors, targets, log10 e and the valid / sparse flags of
    searched spectra (hist: n x 802 counts, psm: gx_psm records)."""
    n = len(psm)
    feat = np.zeros((n, FEAT), np.float64)
    anchor = np.zeros(n, np.float64)
    target = np.zeros((n, 2), np.float64)
    loge = np.zeros(n, np.float64)
    valid = np.zeros(n, bool)
    sparse = np.ones(n, bool)
    for s in range(n):
        r = psm[s]
        hyp, cps, flags = int(r["hyp"]), max(int(r["cpsms"]), 0), int(r["flags"])
        modeled = bool(flags & PSM_MODELED) and cps > 0 and 0 < hyp < HIST_BINS
        fit = bool(flags & PSM_FIT_OK)
        u = int(r["mu"]) / 1e6 * hyp + int(r["beta"]) / 1e6
        if not modeled:
            target[s] = (np.float32(int(r["mu"]) / 1e6), np.float32(u))
            loge[s] = np.float32(math.log10(cps) + u) if fit and cps > 0 else 0.0
            continue
        h = hist[s].astype(np.int64)
        stt1, end1, ls, bad = curve(h, hyp, cps)
        m = len(ls)
        for j in range(CURVE_PTS):
            t = j * (m - 1) / (CURVE_PTS - 1) if m > 1 else 0.0
            i0 = min(int(t), m - 1)
            i1 = min(i0 + 1, m - 1)
            fr = t - i0
            feat[s, j] = np.float32((ls[i0] * (1.0 - fr) + ls[i1] * fr) / 4.0)
        lc = math.log10(cps)
        feat[s, CURVE_PTS:] = [np.float32(stt1 / 800.0), np.float32(end1 / 800.0), np.float32(hyp / 800.0),
                               np.float32((hyp - end1) / 400.0), np.float32(m / 400.0), np.float32(lc / 4.0),
                               np.float32(ls[-1] / 4.0), np.float32(ls[0])]
        anchor[s] = np.float32(ls[-1])
        target[s] = (np.float32(int(r["mu"]) / 1e6), np.float32(u - ls[-1]))
        loge[s] = np.float32(lc + u) if fit else 0.0
        valid[s] = fit
        sparse[s] = bad or np.count_nonzero(h[:hyp]) < MIN_TAIL_BINS
    return bf16(feat), anchor, target, loge, valid, sparse


# ---- network --------------------------------------------------------------
def forward(w: np.ndarray, x: np.ndarray, quantize: bool = True) -> dict:
    """Forward pass; returns the intermediates the backward pass needs."""
    p = unpack(w)
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    xb = q(x)
    W1b, W2b, W3b = q(p["W1"]), q(p["W2"]), q(p["W3"])
    z1 = xb @ W1b + p["b1"]
    h1 = q(gelu(z1))
    z2 = h1 @ W2b + p["b2"]
    h2 = q(gelu(z2))
    out = h2 @ W3b + p["b3"]
    return {"x": xb, "W1": W1b, "W2": W2b, "W3": W3b, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "out": out}


def loss_terms(out: np.ndarray, target: np.ndarray, loge: np.ndarray, norm, conf_weight: float, tol: float,
               y: np.ndarray | None = None):
    """Per-row errors, confidence labels and the mean loss
    mean(ea^2 + eu^2 + conf_weight * bce). `y` fixes the labels (for finite
    differences); by default they are |u_pred - u| <= tol * max(|loge|, 1e-3)."""
    a_shift, a_scale, u_shift, u_scale = (float(v) for v in norm)
    ta = (target[:, 0] - a_shift) / a_scale
    tu = (target[:, 1] - u_shift) / u_scale
    ea = out[:, 0] - ta
    eu = out[:, 1] - tu
    if y is None:
        y = (np.abs(eu) * u_scale <= tol * np.maximum(np.abs(loge), 1e-3)).astype(np.float64)
    p = 1.0 / (1.0 + np.exp(-out[:, 2]))
    bce = -(y * np.log(np.maximum(p, 1e-7)) + (1 - y) * np.log(np.maximum(1 - p, 1e-7)))
    n = len(out)
    mean_loss = float(np.sum(ea * ea + eu * eu + conf_weight * bce) / n)
    return ea, eu, y, p, bce, mean_loss


def backward(w: np.ndarray, x: np.ndarray, target: np.ndarray, loge: np.ndarray, norm, conf_weight: float,
             tol: float, quantize: bool = True):
    """Gradient of the mean loss w.r.t. the flat parameters, the raw outputs
    and the loss sums {sq. err slope, sq. err u, bce, rows within tol}."""
    f = forward(w, x, quantize)
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    out = f["out"]
    n = len(out)
    ea, eu, y, p, bce, _ = loss_terms(out, target, loge, norm, conf_weight, tol)
    dout = np.zeros((n, OUT_PAD))
    dout[:, 0] = 2.0 * ea / n
    dout[:, 1] = 2.0 * eu / n
    dout[:, 2] = conf_weight * (p - y) / n
    dout = q(dout)
    g = {"W3": f["h2"].T @ dout, "b3": dout.sum(0)}
    dz2 = q((dout @ f["W3"].T) * gelu_grad(f["z2"]))
    g["W2"] = f["h1"].T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = q((dz2 @ f["W2"].T) * gelu_grad(f["z1"]))
    g["W1"] = f["x"].T @ dz1
    g["b1"] = dz1.sum(0)
    sums = np.array([np.sum(ea * ea), np.sum(eu * eu), np.sum(bce), np.sum(y)])
    return pack(g), out, sums


def predict_evalue(out: np.ndarray, anchor: np.ndarray, psm: np.ndarray, norm):
    """Surrogate e-value cpsms * 10^u (u = anchor + predicted residual) and
    confidence of each row."""
    a_shift, a_scale, u_shift, u_scale = (float(v) for v in norm)
    u = np.asarray(anchor, np.float64) + out[:, 1] * u_scale + u_shift
    conf = 1.0 / (1.0 + np.exp(-out[:, 2]))
    e = psm["cpsms"].astype(np.float64) * np.power(10.0, u)
    return e, u, conf


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights and zero biases (the same shape rules as
    gx_mlp_init, a different random stream)."""
    rng = np.random.default_rng(seed)
    p = {"W1": rng.uniform(-1, 1, (FEAT, HIDDEN)) * math.sqrt(6.0 / (FEAT + HIDDEN)),
         "b1": np.zeros(HIDDEN),
         "W2": rng.uniform(-1, 1, (HIDDEN, HIDDEN)) * math.sqrt(6.0 / (2 * HIDDEN)),
         "b2": np.zeros(HIDDEN),
         "W3": np.zeros((HIDDEN, OUT_PAD)),
         "b3": np.zeros(OUT_PAD)}
    p["W3"][:, :OUT] = rng.uniform(-1, 1, (HIDDEN, OUT)) * math.sqrt(6.0 / (HIDDEN + OUT))
    return pack(p)
