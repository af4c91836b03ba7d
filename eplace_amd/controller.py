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
    t0 = log(1/L' at the accepted trial / plain proposal 1/L) where plain
         backtracking retried the proposal (1/L' < 0.95 proposal), else 0:
         the controller only shrinks a step that plain backtracking would
         have had to retry;
    t1 = log(f_k / 0.95), f_k the floor under the reference's multiplier:
         with G_k plain's mean lambda growth per iteration over the next W
         iterations, f_k = `floor` where G_k <= g_lo, no floor (0.95) where
         G_k >= g_hi, linear in between, and likewise faded out unless HPWL
         is between scale_lo and scale_hi times 350000 or more.  The reference's multiplier
         1.05^(1 - dHPWL/350000) (eplace.cpp:846-883) uses an absolute HPWL
         reference, so on large designs HPWL noise keeps it near 1 for
         hundreds of iterations while lambda stalls; the floor lifts those
         stalls and leaves iterations where plain's lambda already grows
         alone.
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


def samples_from_log(log: np.ndarray, floor: float = 1.01, window: int = 20, g_lo: float = 1.0,
                     g_hi: float = 1.02, scale_lo: float = 300.0, scale_hi: float = 800.0):
    """(X [n, 12], Y [n, 2]) training pairs from a plain backtracking log."""
    log = np.asarray(log, np.float64)
    n = len(log)
    if n < 2:
        return np.zeros((0, F), np.float32), np.zeros((0, NO), np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = log[:, L.EP_LOG_STEP] / log[:, L.EP_LOG_STEP_PROP]
        # shorten only where plain backtracking retried (0.95 step > 1/L', nesterov.hpp:355)
        t0 = np.where(r < 0.95, np.log(r), 0.0)
        ks = np.arange(1, n)
        last = np.minimum(ks + window - 1, n - 1)
        growth = np.exp(np.log(log[last, L.EP_LOG_LAMBDA] / log[ks, L.EP_LOG_LAMBDA_BEFORE]) / (last - ks + 1))
    w = np.clip((g_hi - growth) / (g_hi - g_lo), 0.0, 1.0)
    # only where HPWL dwarfs the reference's absolute dHPWL constant
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = np.log(log[ks, L.EP_LOG_HPWL] / 350000.0)
    w = w * np.clip((scale - np.log(scale_lo)) / (np.log(scale_hi) - np.log(scale_lo)), 0.0, 1.0)
    f = 0.95 + w * (floor - 0.95)
    t1 = np.log(f / 0.95)
    X = log[ks, L.EP_LOG_FEATURES:L.EP_LOG_FEATURES + F]
    Y = np.stack([t0[ks], t1], axis=1)
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
