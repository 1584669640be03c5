"""CPU ORACLE -- test infrastructure only.

FP64 numpy restatement of the learned secular start's network and of its
training loss and gradient, the checker for the device training kernels
(hsspack_amd/csrc/hsdc_mlp.hip).  Nothing in the product imports it.

Network (weights flattened as W1(64x34, row = hidden unit), b1(64), W2(64x64),
b2(64), w3(64), b3):  h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = w3.h2 + b3.
Loss over a batch: L = (1/B) sum_s (y_s - t_s)^2, t_s = log(converged offset of the
root from its origin pole / classical starting offset).

Pinned by tests/test_mlp.py: the analytic gradient against central differences
of the loss, the forward pass against the C restatement (orc_mlp_forward).
There is no reference counterpart: HSSPACK's DLAED4 has no learned start.
"""
import numpy as np

NIN, NH = 34, 64
W1, B1 = 0, NH * NIN
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NH
NPARAM = B3 + 1


def unpack(w):
    w = np.asarray(w, np.float64)
    return (w[W1:B1].reshape(NH, NIN), w[B1:W2], w[W2:B2].reshape(NH, NH), w[B2:W3], w[W3:B3],
            w[B3])


def forward(w, x):
    """y (B,) for x (B, 34)."""
    a1, c1, a2, c2, a3, c3 = unpack(w)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ a1.T + c1, 0.0)
    h2 = np.maximum(h1 @ a2.T + c2, 0.0)
    return h2 @ a3 + c3


def loss_grad(w, x, t):
    """(loss, gradient (NPARAM,)) of the weighted squared error."""
    a1, c1, a2, c2, a3, c3 = unpack(w)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    B = len(t)
    z1 = x @ a1.T + c1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ a2.T + c2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ a3 + c3
    r = y - t
    loss = float(np.sum(r * r) / B)
    gy = 2.0 * r / B
    g = np.zeros(NPARAM)
    g[W3:B3] = h2.T @ gy
    g[B3] = gy.sum()
    d2 = np.outer(gy, a3) * (z2 > 0)
    g[W2:B2] = (d2.T @ h1).ravel()
    g[B2:W3] = d2.sum(0)
    d1 = (d2 @ a2) * (z1 > 0)
    g[W1:B1] = (d1.T @ x).ravel()
    g[B1:W2] = d1.sum(0)
    return loss, g


def loss(w, x, t):
    y = forward(w, x)
    t = np.asarray(t, np.float64)
    return float(np.mean((y - t) ** 2))
