"""Learned warm start of the reduced H*Q fit: the MLP that the library trains
and runs on the GPU (mlcp_amd/csrc/mlcp_mlp.hip), its training pairs, and
its use by a session.

The reduced H*Q phase (PRODHV_ALS3, ALSPow.f90:784-865; the GPU twin is
PRODHV_ALS, intertw_mvp.cu:501-549) fits a rank-R CP vector F to H Psi by
ALS sweeps that start from F = Psi.  Term r of the fit therefore grows out
of term r of Psi, and the pair (column r of Psi on mode k, column r of the
fit on mode k) is a training example of "where the sweeps take a factor
column".  The network maps the first (normalised, with a learned embedding
of the mode) to the second (normalised, sign-aligned with the input); its
predictions replace Psi's columns as the fit's starting point.

The plain start stays the judge and the fallback: ``gain`` measures the
sweeps both starts need to the same fit-residual tolerance on held-out
blocks, and ``choose`` keeps the network only where it needs at least
half a sweep per state fewer.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from ._lib import check, lib
from .hoplist import CPBlock

NEMB = 16
NHID = 128


class MLP:
    """The network on the GPU: nin inputs (a factor column, zero-padded),
    a nemb-wide embedding per mode, two tanh layers of nhid, nin outputs."""

    def __init__(self, nin: int, nmodes: int, nemb: int = NEMB,
                 nhid: int = NHID, seed: int = 1):
        self.nin, self.nmodes, self.nemb, self.nhid = nin, nmodes, nemb, nhid
        h = C.c_void_p()
        check(lib().mlcp_mlp_create(C.byref(h), nin, nemb, nmodes, nhid,
                                    C.c_ulonglong(seed)), "mlcp_mlp_create")
        self.h = h
        self.np = int(lib().mlcp_mlp_nparams(h))

    def close(self):
        if self.h:
            lib().mlcp_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.np, dtype=np.float32)
        check(lib().mlcp_mlp_params(self.h, p, 0), "mlp_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.np
        check(lib().mlcp_mlp_params(self.h, p, 1), "mlp_params")

    @staticmethod
    def _set(X, modes, T=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        modes = np.ascontiguousarray(modes, dtype=np.int32)
        T = None if T is None else np.ascontiguousarray(T, dtype=np.float32)
        return X, modes, T

    def forward(self, X, modes) -> np.ndarray:
        X, modes, _ = self._set(X, modes)
        Y = np.zeros((X.shape[0], self.nin), dtype=np.float32)
        check(lib().mlcp_mlp_forward(self.h, X, modes, X.shape[0], Y),
              "mlp_forward")
        return Y

    def loss_grad(self, X, modes, T):
        X, modes, T = self._set(X, modes, T)
        g = np.zeros(self.np, dtype=np.float32)
        loss = C.c_double()
        check(lib().mlcp_mlp_loss_grad(self.h, X, modes, T, X.shape[0],
                                       C.byref(loss), g), "mlp_loss_grad")
        return loss.value, g

    def train(self, X, modes, T, epochs=20, batch=512, lr=1e-3, seed=7):
        X, modes, T = self._set(X, modes, T)
        losses = np.zeros(epochs)
        check(lib().mlcp_mlp_train(self.h, X, modes, T, X.shape[0], epochs,
                                   batch, C.c_float(lr), C.c_ulonglong(seed),
                                   losses), "mlp_train")
        return losses


def column_pairs(Q: CPBlock, HQ: CPBlock, nin: int):
    """Training rows from a block and its reduced H*Q (same ranks, term r of
    HQ grown from term r of Q): x = Q's column r on mode k normalised and
    zero-padded to nin, t = HQ's column normalised and sign-aligned with x,
    mode = k."""
    off = np.concatenate([[0], np.cumsum(Q.nbas)])
    D = Q.nbas.size
    N = Q.base.shape[0]
    X = np.zeros((N * D, nin), dtype=np.float32)
    T = np.zeros((N * D, nin), dtype=np.float32)
    modes = np.tile(np.arange(D, dtype=np.int32), N)
    for d in range(D):
        a = Q.base[:, off[d]:off[d + 1]]
        b = HQ.base[:, off[d]:off[d + 1]]
        a = a / np.maximum(np.linalg.norm(a, axis=1, keepdims=True), 1e-300)
        b = b / np.maximum(np.linalg.norm(b, axis=1, keepdims=True), 1e-300)
        sgn = np.where(np.sum(a * b, axis=1) < 0, -1.0, 1.0)[:, None]
        X[d::D, :a.shape[1]] = a
        T[d::D, :b.shape[1]] = b * sgn
    return X, modes, T


@dataclass
class Gain:
    plain_sweeps: np.ndarray      # per state, reduced H*Q sweeps to tol
    warm_sweeps: np.ndarray
    cap: int

    @property
    def mean_plain(self) -> float:
        return float(np.mean(np.where(self.plain_sweeps > 0,
                                      self.plain_sweeps, self.cap)))

    @property
    def mean_warm(self) -> float:
        return float(np.mean(np.where(self.warm_sweeps > 0,
                                      self.warm_sweeps, self.cap)))


def sweeps_to_tolerance(session, mlp: Optional[MLP], nals: int,
                        als_tol: float = 1e-8) -> np.ndarray:
    """Reduced H*Q sweeps per state of the session's current block until the
    squared fit residual changes by <= als_tol |F|^2 (cap nals; 0 in the
    result means the cap was reached), from the plain start (mlp None) or
    the network's."""
    form = session.FORM_TOL | (session.FORM_ML if mlp is not None else 0)
    session.set_warmstart(mlp)
    session.set_form(form, 1e-6, als_tol)
    try:
        session.prodhv(nals)
        return session.sweeps()[1].copy()
    finally:
        session.set_form(0)
        session.set_warmstart(None)


def gain(session, mlp: MLP, nals: int, als_tol: float = 1e-8) -> Gain:
    """The sweeps both starts need on the session's current block."""
    p = sweeps_to_tolerance(session, None, nals, als_tol)
    w = sweeps_to_tolerance(session, mlp, nals, als_tol)
    return Gain(p, w, nals)


# Reduced H*Q sweeps per state the network must save on the held-out block
# to be kept.  A smaller gain is within the run-to-run spread (measured:
# 6.53 against 6.55 and 6.32 against 6.31 sweeps on the 24-mode block), and
# a start that ALS leaves earlier ends on a different truncated fit: on the
# 24-mode block such a network moved the teacher-forced eigenvalues by up
# to 0.1 cm^-1, against 2.5e-3 from the plain start.
MIN_SWEEPS_SAVED = 0.5


def choose(g: Gain) -> bool:
    """Keep the network only if it saves at least MIN_SWEEPS_SAVED sweeps
    per state over the plain start on the held-out block (the plain start
    is the fallback)."""
    return g.mean_warm <= g.mean_plain - MIN_SWEEPS_SAVED


@dataclass
class Fit:
    """A network trained on a plain solve's pairs and what it gains."""
    net: Optional[MLP]
    rows: int
    losses: np.ndarray
    pred_err: float        # held-out mean |y/|y| - t|^2
    ident_err: float       # the same for the plain start (y = x)
    gain: Gain             # reduced H*Q sweeps to the tolerance, both starts
    used: bool             # the network is kept (else the plain start)


def train_on_solve(session, w, bounds, ncycle: int, epochs: int = 20,
                   batch: int = 512, lr: float = 1e-3, skip: int = 0,
                   als_tol: float = 1e-6, seed: int = 3) -> Fit:
    """Runs ``ncycle`` plain cycles on the session's block (GetBlockShift
    from ``bounds``), collects the (Psi, reduced H Psi) column pairs of
    every cycle from the ``skip``-th on, trains the network on all but the
    last cycle's pairs and judges it on the last: the held-out prediction
    error against the plain start's, and the reduced H*Q sweeps both starts
    need to ``als_tol`` on the final block.  The network is kept only if its
    held-out error is lower and it passes ``choose``."""
    from .intertwine import block_shift
    nin = int(np.max(session.Q.nbas))
    D = int(session.Q.nbas.size)
    Xs, Ms, Ts = [], [], []
    eig = np.asarray(session.Q.eigv, dtype=np.float64).copy()
    for c in range(ncycle):
        if c >= skip:
            Q = session.download()
            HQ = session.prodhv(w.nals)
            X, M_, T = column_pairs(Q, HQ, nin)
            Xs.append(X)
            Ms.append(M_)
            Ts.append(T)
        eig = session.cycle(eig, w.npow, w.nals, True, 1,
                            block_shift(eig, bounds))
    if len(Xs) < 2:
        raise ValueError("need at least two collected cycles")
    Xtr = np.concatenate(Xs[:-1])
    Mtr = np.concatenate(Ms[:-1])
    Ttr = np.concatenate(Ts[:-1])
    net = MLP(nin, D, seed=seed)
    losses = net.train(Xtr, Mtr, Ttr, epochs=epochs, batch=batch, lr=lr)
    Y = net.forward(Xs[-1], Ms[-1]).astype(np.float64)
    Y /= np.maximum(np.linalg.norm(Y, axis=1, keepdims=True), 1e-30)
    pred = float(np.mean(np.sum((Y - Ts[-1]) ** 2, axis=1)))
    ident = float(np.mean(np.sum((Xs[-1] - Ts[-1]) ** 2, axis=1)))
    g = gain(session, net, w.nals, als_tol)
    used = pred < ident and choose(g)
    return Fit(net, int(Xtr.shape[0]), losses, pred, ident, g, used)
