"""Learned step / lambda controller (the EP_CTRL_* API of include/eplace_amd.h).

A 12 -> 64 -> 64 -> 2 MLP run on the device once per iteration of the
"learned" optimiser method: it proposes the first trial step (as a factor on
the plain Lipschitz proposal) and a factor on the reference's penalty
multiplier (zero outputs reproduce plain backtracking).  It is trained on
trajectories that the plain backtracking method itself produces, by the
library's own forward / backward / Adam kernels (ep_ctrl_train).

Training samples come from the optimiser log (EP_LOG_* columns): inputs are
the logged features of iteration k, targets are what plain backtracking
did from there:
    t0 = min(0, log(1/L' at the accepted trial / plain proposal 1/L)), so the
         controller only shrinks a step that plain backtracking would have
         had to retry;
    t1 = log(lambda after iteration k + c - 1 / lambda before iteration k)
         - log(plain multiplier of iteration k): the extra penalty growth
         that takes plain's growth over a lookahead of c iterations
         (fractional c interpolates the last one).  c = 1 gives 0 (the
         reference's updatePenaltyFactor); c > 1 trades HPWL for fewer
         iterations.
Iteration 0 is skipped (its plain proposal is the constant 1).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L

F, H, NO = L.EP_CTRL_FEATURES, L.EP_CTRL_HIDDEN, L.EP_CTRL_OUTPUTS
NW, NP = L.EP_CTRL_WEIGHTS, L.EP_CTRL_PARAMS
O_XMU, O_XSD, O_YMU, O_YSD = NW, NW + F, NW + 2 * F, NW + 2 * F + NO
PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "controller_params.npy")


def _ptr(a):
    return a.ctypes.data if a is not None else None


class Controller:
    """Device-resident controller parameters and their training state."""

    def __init__(self, params=None, seed: int = 1, device: int = 0):
        self.lib = L.load()
        p = None if params is None else np.ascontiguousarray(params, np.float32).reshape(-1)
        if p is not None and p.size != NP:
            raise ValueError(f"expected {NP} parameters, got {p.size}")
        h = C.c_void_p()
        rc = self.lib.ep_ctrl_create(int(device), _ptr(p), int(seed), C.byref(h))
        self.h = h.value
        if rc != L.EP_OK:
            msg = self.lib.ep_ctrl_last_error(self.h).decode() if self.h else "ep_ctrl_create failed"
            if self.h:
                self.lib.ep_ctrl_destroy(self.h)
            self.h = None
            raise L.EPError(rc, msg)

    @classmethod
    def load(cls, path: str = PARAMS_PATH, device: int = 0) -> "Controller":
        return cls(np.load(path, allow_pickle=False), device=device)

    def save(self, path: str) -> None:
        np.save(path, self.params)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ep_ctrl_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != L.EP_OK:
            raise L.EPError(rc, self.lib.ep_ctrl_last_error(self.h).decode())

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(NP, np.float32)
        self._check(self.lib.ep_ctrl_get_params(self.h, _ptr(out)))
        return out

    @params.setter
    def params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32).reshape(-1)
        if p.size != NP:
            raise ValueError(f"expected {NP} parameters, got {p.size}")
        self._check(self.lib.ep_ctrl_set_params(self.h, _ptr(p)))

    def set_normalization(self, xmu, xsd, ymu, ysd) -> None:
        p = self.params
        p[O_XMU:O_XMU + F] = xmu
        p[O_XSD:O_XSD + F] = xsd
        p[O_YMU:O_YMU + NO] = ymu
        p[O_YSD:O_YSD + NO] = ysd
        self.params = p

    def train(self, X, Y, steps: int, lr: float = 1e-3) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(self.lib.ep_ctrl_train(self.h, _ptr(X), _ptr(Y), len(X), int(steps), float(lr), _ptr(losses)))
        return losses[:steps]

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((len(X), NO), np.float32)
        self._check(self.lib.ep_ctrl_predict(self.h, _ptr(X), len(X), _ptr(Y)))
        return Y

    def loss_grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(NW, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.ep_ctrl_loss_grad(self.h, _ptr(X), _ptr(Y), len(X), _ptr(g), _ptr(loss)))
        return float(loss[0]), g


def samples_from_log(log: np.ndarray, lookahead: float = 1.0):
    """(X [n, 12], Y [n, 2]) training pairs from a plain backtracking log."""
    log = np.asarray(log, np.float64)
    whole = int(np.floor(lookahead))
    frac = lookahead - whole
    span = whole + (1 if frac > 0 else 0)
    if lookahead < 1.0 or len(log) < 1 + span:
        return np.zeros((0, F), np.float32), np.zeros((0, NO), np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        lm = np.log(log[:, L.EP_LOG_LAMBDA] / log[:, L.EP_LOG_LAMBDA_BEFORE])  # per-iteration log multiplier
        t0 = np.minimum(0.0, np.log(log[:, L.EP_LOG_STEP] / log[:, L.EP_LOG_STEP_PROP]))
    ks = np.arange(1, len(log) - span + 1)
    t1 = np.zeros(len(ks))
    for j in range(whole):
        t1 += lm[ks + j]
    if frac > 0:
        t1 += frac * lm[ks + whole]
    X = log[ks, L.EP_LOG_FEATURES:L.EP_LOG_FEATURES + F]
    Y = np.stack([t0[ks], t1 - lm[ks]], axis=1)
    ok = np.all(np.isfinite(X), axis=1) & np.all(np.isfinite(Y), axis=1)
    return X[ok].astype(np.float32), Y[ok].astype(np.float32)


def normalization(X, Y):
    xmu, xsd = X.mean(0), X.std(0)
    ymu, ysd = Y.mean(0), Y.std(0)
    return xmu, np.maximum(xsd, 1e-3), ymu, np.maximum(ysd, 1e-3)


def fit(X, Y, steps: int = 4000, lr: float = 2e-3, seed: int = 1, device: int = 0):
    """A controller trained on (X, Y) with the kernels' full-batch Adam."""
    ctrl = Controller(seed=seed, device=device)
    ctrl.set_normalization(*normalization(X, Y))
    losses = ctrl.train(X, Y, steps, lr)
    return ctrl, losses
