"""The learned warm start on the GPU (mlcp_amd/csrc/mlcp_mlp.hip through
the C-ABI): the network's loss, gradients and outputs against the FP64
oracle (oracle/mlp_oracle.py, pinned by finite differences on the CPU), and
its training by the library's own Adam kernels.

Tolerances (the GPU network is FP32): loss and outputs 1e-5 relative,
gradients 2e-4 of the largest gradient entry.
"""
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _case(seed, nin, nmodes, N, nemb=16, nhid=128, scale=0.3):
    from mlcp_amd.warmstart import MLP
    rng = np.random.default_rng(seed)
    net = MLP(nin, nmodes, nemb=nemb, nhid=nhid, seed=seed)
    p = (rng.normal(scale=scale, size=net.np)).astype(np.float32)
    net.set_params(p)
    X = rng.normal(size=(N, nin)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    T = rng.normal(size=(N, nin)).astype(np.float32)
    modes = rng.integers(0, nmodes, size=N).astype(np.int32)
    return net, p, X, modes, T, (nin, nemb, nmodes, nhid)


@pytest.mark.parametrize("nin,nmodes,N", [(12, 4, 64), (60, 8, 517),
                                          (40, 6, 1)])
def test_gradients_match_oracle(hip, nin, nmodes, N):
    say("mlp gradients", nin, nmodes, N)
    net, p, X, modes, T, dims = _case(11 + nin, nin, nmodes, N)
    loss, g = net.loss_grad(X, modes, T)
    lo, go = M.loss_grad(p.astype(np.float64), X, modes, T, *dims)
    say("loss", loss, "oracle", lo, "max |dg|",
        float(np.max(np.abs(g - go))), "max |g|", float(np.abs(go).max()))
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    assert np.max(np.abs(g - go)) <= 2e-4 * np.abs(go).max()
    assert np.array_equal(net.params(), p)   # gradients do not move params


def test_forward_matches_oracle(hip):
    net, p, X, modes, T, dims = _case(5, 60, 8, 300)
    Y = net.forward(X, modes)
    Yo = M.forward(p.astype(np.float64), X, modes, *dims)
    assert np.max(np.abs(Y - Yo)) <= 1e-5 * np.abs(Yo).max()


def test_training_fits_a_teacher_network(hip):
    """Targets from a fixed random network of the same shape: Adam on the
    GPU must cut the loss by an order of magnitude, and the loss it reports
    must be the oracle's at the trained parameters."""
    say("mlp training")
    teacher, pt, X, modes, _, dims = _case(21, 24, 4, 4096, scale=0.4)
    T = M.forward(pt.astype(np.float64), X, modes, *dims).astype(np.float32)
    from mlcp_amd.warmstart import MLP
    net = MLP(24, 4, seed=99)
    l0, _ = net.loss_grad(X, modes, T)
    losses = net.train(X, modes, T, epochs=30, batch=256, lr=2e-3, seed=3)
    l1, _ = net.loss_grad(X, modes, T)
    lo, _ = M.loss_grad(net.params().astype(np.float64), X, modes, T, *dims)
    say("loss", l0, "->", l1, "epochs", np.round(losses[[0, -1]], 6).tolist())
    assert l1 <= 0.1 * l0
    assert abs(l1 - lo) <= 1e-4 * abs(lo)
    assert np.all(np.isfinite(losses))
