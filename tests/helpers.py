"""Shared test inputs: seeded random term lists and CP blocks, and full-tensor
brute force used to pin the oracle on tiny problems."""
from __future__ import annotations

import numpy as np

from mlcp_amd.hoplist import CPBlock, HopList


def random_system(seed, nbas, ncoup=4, lam=0.05, ndiag_terms=True):
    """A separable diagonal part plus ``ncoup`` random two-mode couplings.

    Mirrors the operator classes of a top-layer term list: per-mode diagonal
    ("eigen") operators, identities, and dense symmetric couplings."""
    rng = np.random.default_rng(seed)
    nbas = [int(n) for n in nbas]
    ndof = len(nbas)
    terms, coefs = [], []
    if ndiag_terms:
        for d in range(ndof):
            ops = [None] * ndof
            ops[d] = np.diag(np.sort(rng.uniform(0.5, 1.5, nbas[d])) +
                             np.arange(nbas[d]))
            terms.append(ops)
            coefs.append(1.0)
    qs = []
    for d in range(ndof):
        a = rng.normal(size=(nbas[d], nbas[d]))
        qs.append((a + a.T) / 2 / np.sqrt(nbas[d]))
    for c in range(ncoup):
        a, b = rng.choice(ndof, size=2, replace=False)
        ops = [None] * ndof
        ops[a] = qs[a]
        ops[b] = qs[b] if c % 2 == 0 else qs[b] @ qs[b]
        terms.append(ops)
        coefs.append(lam * rng.normal())
    return HopList.from_terms(nbas, terms, coefs)


def random_block(seed, nbas, ranks, normalized=True):
    rng = np.random.default_rng(seed)
    nbas = np.asarray(nbas, dtype=np.int32)
    ranks = np.asarray(ranks, dtype=np.int32)
    nt, nD = int(ranks.sum()), int(nbas.sum())
    base = rng.normal(size=(nt, nD))
    coef = rng.uniform(0.5, 1.5, size=nt)
    if normalized:
        off = np.concatenate([[0], np.cumsum(nbas)])
        for d in range(nbas.size):
            nrm = np.linalg.norm(base[:, off[d]:off[d + 1]], axis=1)
            base[:, off[d]:off[d + 1]] /= nrm[:, None]
    return CPBlock(nbas, ranks, coef, base, np.zeros(ranks.size))


def full_tensor(nbas, coef, base):
    """Dense tensor sum_r c_r (x)_d f_r^d (tiny problems only)."""
    off = np.concatenate([[0], np.cumsum(nbas)])
    out = np.zeros([int(n) for n in nbas])
    for r in range(coef.size):
        t = np.array(coef[r])
        for d in range(len(nbas)):
            t = np.multiply.outer(t, base[r, off[d]:off[d + 1]])
        out += t
    return out


def full_hamiltonian(H: HopList):
    """Dense H = sum_t coef_t (x)_d h_{t,d} (tiny problems only)."""
    nb = [int(n) for n in H.nbas]
    ids = H.opid.reshape(H.T, H.ndof)
    N = int(np.prod(nb))
    out = np.zeros((N, N))
    for t in range(H.T):
        m = np.array([[H.coef[t]]])
        for d in range(H.ndof):
            m = np.kron(m, H.dense_op(ids[t, d] - 1, nb[d]))
        out += m
    return out
