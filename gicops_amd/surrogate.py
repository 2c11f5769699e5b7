"""Learned e-value surrogate: the host side of gx_mlp_* (libgicops_hip.so).

The plain e-value model is expeRT::ModelTailFit (expeRT.cpp:396-559): a
least-squares line through the log10 survival of the candidate-score
histogram between two marks (expeRT.cpp:477-515), evaluated at the top hit's
bin. The surrogate is a 128-256-256-3 GELU MLP that predicts where the two
marks lie on the curve and how sure it is. It is trained on the GPU from the
marks the plain fit itself finds. At search time (SearchParams.evalue_mode =
GX_EVALUE_LEARNED) one fused kernel runs the network, checks each predicted
mark against its defining inequality among the 32 points around it, and fits
the line on the window in parallel; the plain fit runs for every modeled
spectrum the surrogate does not serve. Every call goes through the HIP
library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F


@dataclass
class TrainParams:
    batch: int = 1024
    epochs: int = 8
    lr: float = 2e-3
    lr_decay: float = 0.7
    conf_weight: float = 0.5
    seed: int = 0

    def to_ffi(self) -> F.gx_train_params:
        t = F.gx_train_params()
        t.batch, t.epochs, t.lr, t.lr_decay = self.batch, self.epochs, self.lr, self.lr_decay
        t.conf_weight, t.seed = self.conf_weight, self.seed
        return t


def param_count() -> int:
    return int(F.hip_lib().gx_mlp_param_count())


class Surrogate:
    """A gx_mlp bound to a Searcher's device context."""

    def __init__(self, searcher, seed: int | None = 0):
        self._lib = searcher._lib
        self._s = searcher
        h = C.c_void_p()
        F.check("gx_mlp_create", self._lib.gx_mlp_create(searcher._ctx, C.byref(h)))
        self._h = h
        searcher._children.append(self)
        self.n = param_count()
        if seed is not None:
            F.check("gx_mlp_init", self._lib.gx_mlp_init(h, seed))

    # parameters -------------------------------------------------------------
    def get_params(self) -> tuple[np.ndarray, np.ndarray]:
        w = np.zeros(self.n, np.float32)
        nm = np.zeros(4, np.float32)
        F.check("gx_mlp_get_params", self._lib.gx_mlp_get_params(self._h, F.fptr(w), self.n, F.fptr(nm)))
        return w, nm

    def set_params(self, w: np.ndarray, norm: np.ndarray | None = None) -> None:
        w = np.ascontiguousarray(w, np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        F.check("gx_mlp_set_params", self._lib.gx_mlp_set_params(self._h, F.fptr(w), w.size, F.fptr(nm)))

    # training ---------------------------------------------------------------
    def add_training(self, dq) -> int:
        """Append the spectra of a batch searched with the plain tail fit."""
        n = C.c_uint64()
        F.check("gx_mlp_add_training", self._lib.gx_mlp_add_training(self._h, dq._h, C.byref(n)))
        return int(n.value)

    def train(self, tp: TrainParams) -> dict:
        rep = F.gx_train_report()
        t = tp.to_ffi()
        F.check("gx_mlp_train", self._lib.gx_mlp_train(self._h, C.byref(t), C.byref(rep)))
        e = rep.epochs
        return {"samples": rep.samples, "epochs": e,
                "mse_mark": list(rep.mse_mark[:e]), "mse_mark2": list(rep.mse_mark2[:e]),
                "bce": list(rep.bce[:e]), "in_window": list(rep.in_window[:e])}

    def grad(self, feat, target, norm, conf_weight: float):
        """Gradient of the mean loss on host rows (target n x 4) -> (grad, out (n,16), loss sums[4])."""
        feat = np.ascontiguousarray(feat, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        norm = np.ascontiguousarray(norm, np.float32)
        n = feat.shape[0]
        g = np.zeros(self.n, np.float32)
        out = np.zeros((n, F.MLP_OUT_PAD), np.float32)
        loss = np.zeros(4, np.float32)
        F.check("gx_mlp_grad", self._lib.gx_mlp_grad(self._h, F.fptr(feat), F.fptr(target), n, F.fptr(norm),
                                                     conf_weight, F.fptr(g), F.fptr(out), F.fptr(loss)))
        return g, out, loss

    def predict(self, feat) -> np.ndarray:
        feat = np.ascontiguousarray(feat, np.float32)
        out = np.zeros((feat.shape[0], F.MLP_OUT_PAD), np.float32)
        F.check("gx_mlp_predict", self._lib.gx_mlp_predict(self._h, F.fptr(feat), feat.shape[0], F.fptr(out)))
        return out

    def attach(self) -> None:
        """Make this the context's surrogate for GX_EVALUE_LEARNED searches."""
        F.check("gx_ctx_set_surrogate", self._lib.gx_ctx_set_surrogate(self._s._ctx, self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.gx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
