"""TEST INFRASTRUCTURE ONLY: FP64 restatement of the warm-start network of
mlcp_amd/csrc/mlcp_mlp.hip (same parameter layout), used to check the GPU
kernels' loss and gradients.  Pinned by central finite differences
(tests/test_mlp_oracle.py) and by a closed-form special case (with zero
output weights the network is its input plus the output bias).

Parameters (one vector): W1 [nhid][nin+nemb], b1 [nhid], W2 [nhid][nhid],
b2 [nhid], W3 [nout][nhid], b3 [nout], E [nmodes][nemb], nout = nin.
Network: h1 = tanh(W1 [x; E[mode]] + b1), h2 = tanh(W2 h1 + b2),
y = x + W3 h2 + b3 (skip connection).  Loss 1/(2N) sum |y - t|^2.
"""
from __future__ import annotations

import numpy as np


def layout(nin, nemb, nmodes, nhid):
    nx = nin + nemb
    sizes = [("W1", (nhid, nx)), ("b1", (nhid,)), ("W2", (nhid, nhid)),
             ("b2", (nhid,)), ("W3", (nin, nhid)), ("b3", (nin,)),
             ("E", (nmodes, nemb))]
    out, off = {}, 0
    for name, shp in sizes:
        n = int(np.prod(shp))
        out[name] = (off, shp)
        off += n
    return out, off


def unpack(p, nin, nemb, nmodes, nhid):
    lay, n = layout(nin, nemb, nmodes, nhid)
    assert p.size == n
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in lay.items()}


def loss_grad(p, X, modes, T, nin, nemb, nmodes, nhid):
    """Loss and gradient (both FP64) of the batch (X, modes, T)."""
    p = np.asarray(p, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    P = unpack(p, nin, nemb, nmodes, nhid)
    N = X.shape[0]
    X0 = np.hstack([X, P["E"][modes]])
    H1 = np.tanh(X0 @ P["W1"].T + P["b1"])
    H2 = np.tanh(H1 @ P["W2"].T + P["b2"])
    Y = X + H2 @ P["W3"].T + P["b3"]
    R = Y - T
    loss = 0.5 * np.sum(R * R) / N
    dY = R / N
    g = {}
    g["W3"] = dY.T @ H2
    g["b3"] = dY.sum(0)
    dZ2 = (dY @ P["W3"]) * (1.0 - H2 * H2)
    g["W2"] = dZ2.T @ H1
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ P["W2"]) * (1.0 - H1 * H1)
    g["W1"] = dZ1.T @ X0
    g["b1"] = dZ1.sum(0)
    dX0 = dZ1 @ P["W1"]
    gE = np.zeros_like(P["E"])
    np.add.at(gE, modes, dX0[:, nin:])
    g["E"] = gE
    lay, n = layout(nin, nemb, nmodes, nhid)
    flat = np.zeros(n)
    for k, (o, s) in lay.items():
        flat[o:o + int(np.prod(s))] = g[k].ravel()
    return loss, flat


def forward(p, X, modes, nin, nemb, nmodes, nhid):
    P = unpack(np.asarray(p, dtype=np.float64), nin, nemb, nmodes, nhid)
    X0 = np.hstack([np.asarray(X, dtype=np.float64), P["E"][modes]])
    H1 = np.tanh(X0 @ P["W1"].T + P["b1"])
    H2 = np.tanh(H1 @ P["W2"].T + P["b2"])
    return np.asarray(X, dtype=np.float64) + H2 @ P["W3"].T + P["b3"]
