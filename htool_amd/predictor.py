"""Learned ACA rank predictor (no counterpart in the reference).

An 8 -> 64 -> 64 -> 1 GELU perceptron in FP32 maps the features of an
admissible block (ln m, ln n, ln d, ln diam_t, ln diam_s, eta, ln epsilon,
ln gap) to ln(1 + t), t the number of ACA cross terms the block needs. It is
trained on the GPU by libhtool_amd's own forward/backward kernels and Adam,
from the ACA records of earlier builds (`HMatrix.aca_records`), and used by
`HMatrix.build` only to presize each block's ACA workspace and to order the
blocks by expected work: ACA's stopping test still decides every rank, and a
block that outgrows its presized workspace runs again with the full one.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi

NFEAT = capi.MLP_NFEAT
NPARAMS = capi.MLP_NPARAMS


def targets_from_terms(terms) -> np.ndarray:
    """Training target of a record: ln(1 + terms needed)."""
    return np.log1p(np.asarray(terms, dtype=np.float64)).astype(np.float32)


class RankPredictor:
    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        capi.check(capi.lib().htg_predictor_create(C.c_ulonglong(seed), C.byref(h)))
        self._h = h

    def __del__(self):
        try:
            if self._h is not None and self._h.value:
                capi.lib().htg_predictor_destroy(self._h)
        except Exception:
            pass
        self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    # -- parameters -------------------------------------------------------------
    def get_params(self):
        p = np.zeros(NPARAMS, np.float32)
        m = np.zeros(NFEAT, np.float32)
        s = np.zeros(NFEAT, np.float32)
        capi.check(capi.lib().htg_predictor_get_params(self._h, capi.ptr(p), capi.ptr(m), capi.ptr(s)))
        return p, m, s

    def set_params(self, params=None, mean=None, std=None) -> None:
        f = lambda a, n: None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(n)  # noqa: E731
        p, m, s = f(params, NPARAMS), f(mean, NFEAT), f(std, NFEAT)
        capi.check(capi.lib().htg_predictor_set_params(self._h, capi.ptr(p), capi.ptr(m), capi.ptr(s)))

    def save(self, path: str) -> None:
        p, m, s = self.get_params()
        np.savez(path, params=p, mean=m, std=s)

    @classmethod
    def load(cls, path: str) -> "RankPredictor":
        d = np.load(path)
        r = cls()
        r.set_params(d["params"], d["mean"], d["std"])
        return r

    # -- training -----------------------------------------------------------------
    def loss_grad(self, feat, target):
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, NFEAT)
        target = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_double()
        grad = np.zeros(NPARAMS, np.float32)
        capi.check(capi.lib().htg_predictor_loss_grad(self._h, capi.ptr(feat), capi.ptr(target), feat.shape[0],
                                                      C.byref(loss), capi.ptr(grad)))
        return loss.value, grad

    def train(self, feat, target, epochs: int = 20, batch: int = 1024, lr: float = 2e-3, seed: int = 0,
              fit_norm: bool = True) -> np.ndarray:
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, NFEAT)
        target = np.ascontiguousarray(target, dtype=np.float32)
        hist = np.zeros(max(epochs, 1))
        capi.check(capi.lib().htg_predictor_train(self._h, capi.ptr(feat), capi.ptr(target), feat.shape[0], epochs,
                                                  batch, C.c_float(lr), C.c_ulonglong(seed), int(fit_norm),
                                                  capi.ptr(hist)))
        return hist[:epochs]

    # -- inference ----------------------------------------------------------------
    def predict_raw(self, feat) -> np.ndarray:
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, NFEAT)
        out = np.zeros(feat.shape[0], np.float32)
        capi.check(capi.lib().htg_predictor_predict(self._h, capi.ptr(feat), feat.shape[0], capi.ptr(out)))
        return out

    def predict_terms(self, feat) -> np.ndarray:
        return np.expm1(np.maximum(self.predict_raw(feat), 0.0))


def misprediction(pred_terms, terms, kmax=None, margin: float = 0.25) -> dict:
    """Rates of the library's presizing rule cap = min(ceil((1 + margin) t_hat) + 2,
    kmax) against the terms each block needed: `under_rate` is the fraction of
    blocks that would outgrow their workspace (a rerun), `ws_ratio` the reserved
    workspace terms over the needed ones."""
    t = np.asarray(terms, dtype=np.float64)
    cap = np.ceil((1.0 + margin) * np.asarray(pred_terms, dtype=np.float64)) + 2
    if kmax is not None:
        cap = np.minimum(cap, np.asarray(kmax, dtype=np.float64))
    if t.size == 0:
        return {"under_rate": 0.0, "ws_ratio": 0.0, "mean_abs_log_ratio": 0.0}
    return {"under_rate": float(np.mean(t > cap)),
            "ws_ratio": float(cap.sum() / max(t.sum(), 1.0)),
            "mean_abs_log_ratio": float(np.mean(np.abs(np.log1p(pred_terms) - np.log1p(t))))}
