"""CPU restatement of the step / lambda controller network (FP64 numpy).

TEST INFRASTRUCTURE: imported only by tests/ as the checker of the
controller's device kernels.  The controller has no counterpart in the
reference; this file restates the library's own definition (parameter
layout of include/eplace_amd.h, EP_CTRL_*):
    x_n = (x - xmu) / xsd
    h1 = tanh(W1 x_n + b1),  h2 = tanh(W2 h1 + b2),  o = W3 h2 + b3
    y = o * ysd + ymu
    loss = mean over samples of sum over outputs (o - (y_target - ymu) / ysd)^2
and one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the weights.
Pinned by central finite differences of the loss (tests/test_controller_oracle.py).
"""
from __future__ import annotations

import numpy as np

F, H, NO = 12, 64, 2
oW1 = 0
ob1 = oW1 + H * F
oW2 = ob1 + H
ob2 = oW2 + H * H
oW3 = ob2 + H
ob3 = oW3 + NO * H
NW = ob3 + NO
oXmu, oXsd, oYmu, oYsd = NW, NW + F, NW + 2 * F, NW + 2 * F + NO
NP = oYsd + NO


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[oW1:ob1].reshape(H, F), b1=p[ob1:oW2], W2=p[oW2:ob2].reshape(H, H), b2=p[ob2:oW3],
                W3=p[oW3:ob3].reshape(NO, H), b3=p[ob3:NW], xmu=p[oXmu:oXsd], xsd=p[oXsd:oYmu],
                ymu=p[oYmu:oYsd], ysd=p[oYsd:NP])


def forward(p, X):
    """Normalised outputs o [n, NO] and the activations for backward."""
    q = unpack(p)
    xn = (np.asarray(X, np.float64) - q["xmu"]) / q["xsd"]
    h1 = np.tanh(xn @ q["W1"].T + q["b1"])
    h2 = np.tanh(h1 @ q["W2"].T + q["b2"])
    o = h2 @ q["W3"].T + q["b3"]
    return o, (xn, h1, h2)


def predict(p, X):
    q = unpack(p)
    o, _ = forward(p, X)
    return o * q["ysd"] + q["ymu"]


def loss_grad(p, X, Y):
    """Loss and its gradient with respect to the NW weights."""
    q = unpack(p)
    n = len(X)
    o, (xn, h1, h2) = forward(p, X)
    d = o - (np.asarray(Y, np.float64) - q["ymu"]) / q["ysd"]
    loss = float((d * d).sum() / n)
    do = 2.0 * d / n
    g = np.zeros(NW)
    g[oW3:ob3] = (do.T @ h2).reshape(-1)
    g[ob3:NW] = do.sum(0)
    dz2 = (do @ q["W3"]) * (1.0 - h2 * h2)
    g[oW2:ob2] = (dz2.T @ h1).reshape(-1)
    g[ob2:oW3] = dz2.sum(0)
    dz1 = (dz2 @ q["W2"]) * (1.0 - h1 * h1)
    g[oW1:ob1] = (dz1.T @ xn).reshape(-1)
    g[ob1:oW2] = dz1.sum(0)
    return loss, g


def adam_steps(p, X, Y, steps, lr):
    """`steps` full-batch Adam updates from zero moments; returns (params, losses)."""
    p = np.array(p, np.float64)
    m = np.zeros(NW)
    v = np.zeros(NW)
    losses = []
    for t in range(1, steps + 1):
        loss, g = loss_grad(p, X, Y)
        losses.append(loss)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        mh = m / (1 - 0.9 ** t)
        vh = v / (1 - 0.999 ** t)
        p[:NW] -= lr * mh / (np.sqrt(vh) + 1e-8)
    return p, np.array(losses)
