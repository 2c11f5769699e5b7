"""CPU restatement of the learned e-value surrogate - TEST INFRASTRUCTURE ONLY.

Imported by tests/ and bench.py's accuracy report, never by the gicops_amd
package. It restates, in numpy:

- the tail fit's curve (expeRT.cpp:414-475): bounds stt1/end1 below the top
  hit's bin, exact prefix counts, log10 survival in the plain fit's double
  arithmetic;
- the surrogate's features in the kernels' single-precision arithmetic (the
  log10 survival of linearly interpolated prefix counts at 120 points, and 8
  scalars), rounded to bf16;
- the targets: the plain fit's marks (expeRT.cpp:477-515) - one before the
  first point at or below 22 % of the curve's drop, the last point at or
  above 87 % of it - the curve length and whether the surrogate may serve;
- the MLP forward pass, loss (squared errors of the two normalised mark
  positions, cross-entropy of the confidence against the window-check label)
  and backward pass, with the same bf16 rounding points as the HIP kernels
  when `quantize=True`;
- the learned e-value stage: predicted indexes, the 32-point check windows
  with the kernel's acceptance rules, the least-squares line on the window
  (Burkardt's llsq, expeRT.cpp:1241-1289) and the e-value from the truncated
  mu, beta (dslim_query.cpp:1023-1041).

It is pinned by central finite differences of the float64 loss, GELU against
the normal CDF, bf16 against torch, hand-computed curves and features, the
window fit against the plain tail-fit oracle (orc_tailfit) on random tails,
and the closed-form window label against a brute-force scan of the windows
(tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

FEAT, HIDDEN, OUT_PAD, OUT, CURVE_PTS, HIST_BINS = 128, 256, 16, 3, 120, 802
TARGET, MIN_CURVE, SNAP = 4, 8, 32
PSM_MODELED, PSM_FIT_OK, PSM_REPORTED, PSM_SURROGATE = 2, 4, 8, 16

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


GELU_K, GELU_C = 2.0 * math.sqrt(2.0 / math.pi), 0.044715


def gelu(z):
    """GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
    = z * sigmoid(2 sqrt(2/pi) (z + 0.044715 z^3))."""
    return 0.5 * z * (1.0 + np.tanh(0.5 * GELU_K * (z + GELU_C * z ** 3)))


def gelu_grad(z):
    s = 1.0 / (1.0 + np.exp(-GELU_K * (z + GELU_C * z ** 3)))
    return s + z * s * (1.0 - s) * GELU_K * (1.0 + 3.0 * GELU_C * z * z)


# ---- curve and features -------------------------------------------------------
def curve_prefix(h: np.ndarray, hyp: int):
    """Curve bounds of the tail fit (expeRT.cpp:414-452) and exact prefix
    counts: stt1 = first non-empty bin below hyp, end1 = last (hyp - 1 and 0
    if none), end1 + 1 if they are equal; C[i] = sum h[stt1..stt1+i]."""
    nz = np.nonzero(h[:hyp])[0]
    end1 = int(nz[-1]) if len(nz) else hyp - 1
    stt1 = int(nz[0]) if len(nz) else 0
    if stt1 == end1:
        end1 += 1
    n = end1 - stt1 + 1
    seg = np.zeros(n, np.int64)
    top = min(end1 + 1, HIST_BINS)
    seg[:top - stt1] = h[stt1:top]
    return stt1, end1, np.cumsum(seg)


def log_survival(c, cps: int) -> np.ndarray:
    """log10 of x = -(C/cpsms - 1) in the plain fit's order of operations
    (expeRT.cpp:453-475), for curves with no replaced values."""
    x = np.asarray(c, dtype=np.float64) / float(cps)
    x = x + (-1.0)
    x = -x
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.log10(x)


def _f32_log_survival(cnt, cps) -> np.float32:
    x = np.float32(1.0) - np.float32(cnt) / np.float32(cps)
    return np.log10(np.maximum(x, np.float32(1e-7))).astype(np.float32)


def feature_row(c: np.ndarray, stt1: int, end1: int, hyp: int, cps: int) -> np.ndarray:
    """The features of one modeled spectrum in the kernels' float32
    arithmetic (before the bf16 rounding)."""
    n = len(c)
    f = np.zeros(FEAT, np.float32)
    fc = np.float32(cps)
    for j in range(CURVE_PTS):
        t = np.float32(j * (n - 1)) / np.float32(CURVE_PTS - 1)
        i0 = min(int(t), n - 1)
        i1 = min(i0 + 1, n - 1)
        fr = np.float32(t - np.float32(i0))
        cnt = np.float32(np.float32(c[i0]) * (np.float32(1.0) - fr)) + np.float32(np.float32(c[i1]) * fr)
        f[j] = _f32_log_survival(cnt, fc) / np.float32(4.0)
    f[CURVE_PTS:] = [np.float32(stt1) / np.float32(800), np.float32(end1) / np.float32(800),
                     np.float32(hyp) / np.float32(800), np.float32(hyp - end1) / np.float32(400),
                     np.float32(n) / np.float32(400), np.log10(fc) / np.float32(4),
                     _f32_log_survival(c[-1], fc) / np.float32(4), _f32_log_survival(c[0], fc)]
    return f


def servable_curve(c: np.ndarray, cps: int) -> bool:
    """At least MIN_CURVE points and the last prefix count below cpsms (no
    survival value <= 0, so none is replaced and the curve is monotone)."""
    return len(c) >= MIN_CURVE and int(c[-1]) < cps


def exact_marks(c: np.ndarray, cps: int):
    """The plain fit's marks on a servable curve (expeRT.cpp:477-515):
    (f, mark2, mark) with f the first point at or below the 22 % threshold
    (None if none), mark2 the last point at or above the 87 % threshold and
    mark = f - 1, moved to mark2 - 1 when not below mark2."""
    ls = log_survival(c, cps)
    n = len(ls)
    s0 = ls[0]
    hgt = ls[n - 1] - s0
    t1, t2 = s0 + hgt * 0.22, s0 + hgt * 0.87
    below = np.nonzero(ls <= t1)[0]
    above = np.nonzero(ls >= t2)[0]
    f = int(below[0]) if len(below) else None
    mark2 = int(above[-1]) if len(above) else n - 1
    mark = (f if f is not None else 0) - 1
    if mark >= mark2:
        mark = mark2 - 1
    return f, mark2, mark


def features(hist: np.ndarray, psm: np.ndarray):
    """Features (bf16, as float64), targets (n x 4: mark before the mark2
    rule, mark2, n-1, servable), log10 of the plain e-value, and the valid /
    servable flags of searched spectra (hist: n x 802, psm: gx_psm records)."""
    n = len(psm)
    feat = np.zeros((n, FEAT), np.float32)
    target = np.zeros((n, TARGET), np.float32)
    loge = np.zeros(n, np.float32)
    valid = np.zeros(n, bool)
    serv = np.zeros(n, bool)
    for s in range(n):
        r = psm[s]
        hyp, cps, flags = int(r["hyp"]), max(int(r["cpsms"]), 0), int(r["flags"])
        modeled = bool(flags & PSM_MODELED) and cps > 0 and 0 < hyp < HIST_BINS
        fit = bool(flags & PSM_FIT_OK)
        u = int(r["mu"]) / 1e6 * hyp + int(r["beta"]) / 1e6
        if fit and cps > 0:
            loge[s] = np.float32(math.log10(cps) + u)
        if not modeled:
            continue
        valid[s] = fit
        stt1, end1, c = curve_prefix(hist[s].astype(np.int64), hyp)
        feat[s] = feature_row(c, stt1, end1, hyp, cps)
        mark = mark2 = 0
        ok = False
        if servable_curve(c, cps):
            f, mark2, ruled = exact_marks(c, cps)
            mark = (f if f is not None else 0) - 1
            ok = f is not None and f >= 1 and ruled >= 0
        target[s] = (mark, mark2, len(c) - 1, 1.0 if ok else 0.0)
        serv[s] = ok
    return bf16(feat), target, loge, valid, serv


# ---- window check -----------------------------------------------------------------
def predicted_index(o, scale, shift, n: int) -> int:
    """Index of a mark from a raw output (float32, as the kernels compute it)."""
    p = (np.float32(o) * np.float32(scale) + np.float32(shift)) * np.float32(n - 1)
    if not (p > 0):
        return 0
    if p >= np.float32(n - 1):
        return n - 1
    return int(np.rint(p))


def snap_base(q: int, n: int, back: int) -> int:
    return min(max(q - back, 0), max(n - SNAP, 0))


def snap_label(q1: int, q2: int, mark: int, mark2: int, n: int) -> bool:
    """Both marks lie in their check windows (closed form of the scan in
    `snap_windows`, for a monotone curve)."""
    f = mark + 1
    b1 = snap_base(q1, n, SNAP // 2 - 1)
    ok1 = f < b1 + SNAP and (f > b1 or b1 == 0)
    b2 = snap_base(q2, n, SNAP // 2)
    e2 = min(b2 + SNAP, n)
    ok2 = b2 <= mark2 < e2 and (mark2 < e2 - 1 or e2 == n)
    return ok1 and ok2


def snap_windows(ls: np.ndarray, q1: int, q2: int):
    """The kernel's check on the curve itself: the first point at or below t1
    found in the window around q1 (not on its first point unless that is the
    curve's first), the last point at or above t2 in the window around q2 (not
    on its last point unless that is the curve's last). Returns (mark, mark2)
    after the mark2 rule, or None."""
    n = len(ls)
    s0 = ls[0]
    hgt = ls[n - 1] - s0
    t1, t2 = s0 + hgt * 0.22, s0 + hgt * 0.87
    b1, b2 = snap_base(q1, n, SNAP // 2 - 1), snap_base(q2, n, SNAP // 2)
    w1 = [p for p in range(b1, min(b1 + SNAP, n)) if ls[p] <= t1]
    w2 = [p for p in range(b2, min(b2 + SNAP, n)) if ls[p] >= t2]
    if not w1 or not w2:
        return None
    f = w1[0]
    if (f == b1 and b1 > 0) or f == 0:
        return None
    mark2 = w2[-1]
    e2 = min(b2 + SNAP, n)
    if mark2 == e2 - 1 and e2 < n:
        return None
    mark = f - 1
    if mark >= mark2:
        mark = mark2 - 1
    if mark < 0:
        return None
    return mark, mark2


def window_fit(ls: np.ndarray, stt1: int, mark: int, mark2: int):
    """Least-squares line of log10 survival over x = stt1+mark .. stt1+mark2
    (expeRT.cpp:516-545), as (slope, intercept)."""
    y = ls[mark:mark2 + 1]
    x = np.arange(stt1 + mark, stt1 + mark2 + 1, dtype=np.float64)
    m = len(y)
    if m == 1:
        return 0.0, float(y[0])
    xbar, ybar = x.sum() / m, y.sum() / m
    top = np.sum((x - xbar) * (y - ybar))
    bot = np.sum((x - xbar) ** 2)
    a = top / bot
    return float(a), float(ybar - a * xbar)


def evalue_of(mu: int, beta: int, hyp: int, cps: int) -> float:
    """e = cpsms * 10^((mu/1e6) hyp + beta/1e6) (dslim_query.cpp:1023-1041)."""
    return float(10.0 ** ((mu / 1e6) * hyp + beta / 1e6) * cps)


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


def loss_terms(out: np.ndarray, target: np.ndarray, norm, conf_weight: float, y: np.ndarray | None = None):
    """Per-row errors, confidence labels and the mean loss
    mean(e0^2 + e1^2 + conf_weight * bce), e_k = output_k - (mark_k/(n-1) -
    shift_k)/scale_k over servable rows (0 elsewhere). `y` fixes the labels
    (for finite differences); by default a row is labelled 1 when it is
    servable and both marks lie in the check windows around the indexes its
    outputs predict."""
    a_shift, a_scale, u_shift, u_scale = (float(v) for v in norm)
    n = len(out)
    serv = target[:, 3] > 0.5
    nm1 = np.where(serv, target[:, 2], 1.0)
    e0 = np.where(serv, out[:, 0] - (target[:, 0] / nm1 - a_shift) / a_scale, 0.0)
    e1 = np.where(serv, out[:, 1] - (target[:, 1] / nm1 - u_shift) / u_scale, 0.0)
    if y is None:
        y = np.zeros(n)
        for i in np.nonzero(serv)[0]:
            m = int(target[i, 2]) + 1
            q1 = predicted_index(out[i, 0], a_scale, a_shift, m)
            q2 = predicted_index(out[i, 1], u_scale, u_shift, m)
            y[i] = 1.0 if snap_label(q1, q2, int(target[i, 0]), int(target[i, 1]), m) else 0.0
    p = 1.0 / (1.0 + np.exp(-out[:, 2]))
    bce = -(y * np.log(np.maximum(p, 1e-7)) + (1 - y) * np.log(np.maximum(1 - p, 1e-7)))
    mean_loss = float(np.sum(e0 * e0 + e1 * e1 + conf_weight * bce) / n)
    return e0, e1, y, p, bce, mean_loss


def backward(w: np.ndarray, x: np.ndarray, target: np.ndarray, norm, conf_weight: float, quantize: bool = True):
    """Gradient of the mean loss w.r.t. the flat parameters, the raw outputs
    and the loss sums {sq. err mark, sq. err mark2, bce, rows labelled 1}."""
    f = forward(w, x, quantize)
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    out = f["out"]
    n = len(out)
    e0, e1, y, p, bce, _ = loss_terms(out, target, norm, conf_weight)
    dout = np.zeros((n, OUT_PAD))
    dout[:, 0] = 2.0 * e0 / n
    dout[:, 1] = 2.0 * e1 / n
    dout[:, 2] = conf_weight * (p - y) / n
    dout = q(dout)
    g = {"W3": f["h2"].T @ dout, "b3": dout.sum(0)}
    dz2 = q((dout @ f["W3"].T) * gelu_grad(f["z2"]))
    g["W2"] = f["h1"].T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = q((dz2 @ f["W2"].T) * gelu_grad(f["z1"]))
    g["W1"] = f["x"].T @ dz1
    g["b1"] = dz1.sum(0)
    sums = np.array([np.sum(e0 * e0), np.sum(e1 * e1), np.sum(bce), np.sum(y)])
    return pack(g), out, sums


def serve(out: np.ndarray, hist: np.ndarray, psm: np.ndarray, norm, conf: float, expect_max: float = 20.0):
    """The learned stage's decisions and served results from the network's
    raw outputs (n x >=3) on spectra searched with the plain fit (`psm` gives
    the top hits): per spectrum whether the surrogate serves it, and its mu,
    beta, e-value and report flag where it does."""
    a_shift, a_scale, u_shift, u_scale = (float(v) for v in norm)
    logit = -np.inf if conf <= 0 else np.inf if conf >= 1 else math.log(conf / (1 - conf))
    n = len(psm)
    served = np.zeros(n, bool)
    mu = np.zeros(n, np.int64)
    beta = np.zeros(n, np.int64)
    ev = np.full(n, np.inf)
    rep = np.zeros(n, bool)
    for s in range(n):
        r = psm[s]
        hyp, cps, flags = int(r["hyp"]), int(r["cpsms"]), int(r["flags"])
        if not (flags & PSM_MODELED) or not 0 < hyp < HIST_BINS or not np.float32(out[s, 2]) >= np.float32(logit):
            continue
        stt1, end1, c = curve_prefix(hist[s].astype(np.int64), hyp)
        if not servable_curve(c, cps):
            continue
        m = len(c)
        q1 = predicted_index(out[s, 0], a_scale, a_shift, m)
        q2 = predicted_index(out[s, 1], u_scale, u_shift, m)
        ls = log_survival(c, cps)
        w = snap_windows(ls, q1, q2)
        if w is None:
            continue
        a, b = window_fit(ls, stt1, *w)
        am, bm = a * 1e6, b * 1e6
        if not (abs(am) < 2147483647.0 and abs(bm) < 2147483647.0):
            continue
        served[s] = True
        mu[s], beta[s] = int(am), int(bm)
        ev[s] = evalue_of(mu[s], beta[s], hyp, cps)
        rep[s] = ev[s] < expect_max
    return served, mu, beta, ev, rep


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
