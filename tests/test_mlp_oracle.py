"""The FP64 oracle of the warm-start network (oracle/mlp_oracle.py), pinned
on the CPU: its gradient against central finite differences of its loss,
and a closed form (zero hidden-to-output weights: the output is x + b3, so
the gradient of b3 is the mean residual)."""
import numpy as np

from oracle import mlp_oracle as M


def _case(seed, nin=7, nemb=3, nmodes=4, nhid=5, N=11):
    rng = np.random.default_rng(seed)
    _, n = M.layout(nin, nemb, nmodes, nhid)
    p = rng.normal(scale=0.5, size=n)
    X = rng.normal(size=(N, nin))
    T = rng.normal(size=(N, nin))
    modes = rng.integers(0, nmodes, size=N)
    return p, X, modes, T, (nin, nemb, nmodes, nhid)


def test_gradient_matches_finite_differences():
    p, X, modes, T, dims = _case(1)
    loss, g = M.loss_grad(p, X, modes, T, *dims)
    h = 1e-6
    fd = np.zeros_like(p)
    for i in range(p.size):
        q = p.copy()
        q[i] += h
        lp, _ = M.loss_grad(q, X, modes, T, *dims)
        q[i] -= 2 * h
        lm, _ = M.loss_grad(q, X, modes, T, *dims)
        fd[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.abs(g).max())
    # every parameter block is exercised (embedding rows of absent modes
    # have zero gradient, the others not)
    lay, _ = M.layout(*dims)
    for k, (o, s) in lay.items():
        assert np.any(g[o:o + int(np.prod(s))] != 0.0), k


def test_output_bias_closed_form():
    p, X, modes, T, dims = _case(2)
    lay, _ = M.layout(*dims)
    o, s = lay["W3"]
    p[o:o + int(np.prod(s))] = 0.0
    ob, sb = lay["b3"]
    b3 = p[ob:ob + sb[0]]
    loss, g = M.loss_grad(p, X, modes, T, *dims)
    R = X + b3[None, :] - T
    assert np.allclose(g[ob:ob + sb[0]], R.mean(0), rtol=1e-14, atol=1e-15)
    assert np.isclose(loss, 0.5 * np.sum(R * R) / X.shape[0], rtol=1e-14)
