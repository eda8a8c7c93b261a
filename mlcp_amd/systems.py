"""Synthetic systems and their top-layer term lists, in the reference's input
conventions (the callers' side of the hot path).

The reference builds the term list that crosses the boundary from a PES
(FFPES.f90), harmonic-oscillator operator matrices (OpFuncs.f90), a mode
tree (ModeComb.f90, layers.inp) and lower-layer solves (HamilSetup.f90,
Updater.f90, ModeH.f90).  This module produces the same wire format for:

* ``hh111``  6-D modified Henon-Heiles, H = sum 1/2 (p_i^2 + q_i^2)
             + lam sum_{i=1..5} (q_i^2 q_{i+1} - q_{i+1}^3 / 3), lam = 0.111803,
             dimensionless;
* ``qff``    seeded synthetic quartic force fields in cm^-1:
             H = sum w_i/2 (p_i^2 + q_i^2) + sum_{i<=j<=k} f_ijk q_i q_j q_k
             + sum_{i<j} f_iijj q_i^2 q_j^2 + sum_i f_iiii q_i^4
             (the f2 = w/2 convention of pes/f2*.dat).

Lower layers are solved exactly (dense diagonalisation of each group's
Hamiltonian in the product basis of its children, keeping the lowest
eigenvectors) instead of by the reference's CP-format solves; the top-layer
operators are the group eigenvalues (diagonal "eigen" operators) and the
group-projected products of primitive operators for every PES term that
spans more than one group.
"""
from __future__ import annotations

import heapq
from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

import numpy as np

from .hoplist import CPBlock, HopList

Monomial = Tuple[Tuple[int, int], ...]   # ((mode, power), ...)


def ho_q_powers(n: int, kmax: int = 4) -> List[np.ndarray]:
    """Exact truncated matrices of q^k (k = 0..kmax) in the first n
    harmonic-oscillator functions, dimensionless q = (a + a+)/sqrt(2)."""
    m = n + kmax
    q = np.zeros((m, m))
    for i in range(m - 1):
        q[i, i + 1] = q[i + 1, i] = np.sqrt((i + 1) / 2.0)
    out, p = [np.eye(n)], np.eye(m)
    for _ in range(kmax):
        p = p @ q
        out.append(p[:n, :n].copy())
    return out


def ho_p2(n: int) -> np.ndarray:
    """Truncated p^2 = -(a - a+)^2 / 2."""
    m = n + 2
    a = np.zeros((m, m))
    for i in range(m - 1):
        a[i, i + 1] = np.sqrt(i + 1)
    d = a - a.T
    return (-(d @ d) / 2.0)[:n, :n]


@dataclass
class PES:
    nmodes: int
    omega: np.ndarray                 # harmonic frequency per mode (h0 = w/2 (p2 + q2))
    terms: List[Tuple[Monomial, float]] = field(default_factory=list)
    units: str = "cm-1"


def hh111_pes(lam: float = 0.111803) -> PES:
    pes = PES(6, np.ones(6), units="dimensionless")
    for i in range(5):
        pes.terms.append((((i, 2), (i + 1, 1)), lam))
        pes.terms.append((((i + 1, 3),), -lam / 3.0))
    return pes


def qff_pes(nmodes: int, seed: int, wmin=300.0, wmax=3500.0, cubic_density=0.3,
            cubic_sigma=30.0, iijj_sigma=5.0, iiii_sigma=8.0) -> PES:
    """Seeded synthetic quartic force field (numpy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    omega = np.exp(rng.uniform(np.log(wmin), np.log(wmax), nmodes))
    pes = PES(nmodes, omega)
    for i in range(nmodes):
        for j in range(i, nmodes):
            for k in range(j, nmodes):
                if rng.uniform() < cubic_density:
                    pw: Dict[int, int] = {}
                    for m in (i, j, k):
                        pw[m] = pw.get(m, 0) + 1
                    pes.terms.append((tuple(sorted(pw.items())),
                                      rng.normal(0.0, cubic_sigma)))
    for i in range(nmodes):
        for j in range(i + 1, nmodes):
            pes.terms.append((((i, 2), (j, 2)), rng.normal(0.0, iijj_sigma)))
    for i in range(nmodes):
        pes.terms.append((((i, 4),), abs(rng.normal(0.0, iiii_sigma))))
    return pes


@dataclass
class Layer:
    """A node of the mode tree: a group of primitive modes with a basis of
    ``keep`` eigenvectors of its Hamiltonian."""
    modes: Tuple[int, ...]
    energies: np.ndarray
    vectors: np.ndarray            # columns: eigenvectors in the children's product basis
    children: List["Layer"] = field(default_factory=list)


def _kron_all(mats: Sequence[np.ndarray]) -> np.ndarray:
    out = np.array([[1.0]])
    for m in mats:
        out = np.kron(out, m)
    return out


class _Basis:
    """Operator matrices of monomials restricted to a tree node."""

    def __init__(self, pes: PES, nprim: int):
        self.pes = pes
        self.nprim = nprim
        self.qk = ho_q_powers(nprim, 4)
        self.p2 = ho_p2(nprim)
        # projected operators per (node, monomial restricted to the node):
        # the many PES terms that share a factor on a group share its matrix
        self._ops: Dict[Tuple[int, Monomial], np.ndarray] = {}

    def prim(self, mode: int, power: int) -> np.ndarray:
        return self.qk[power]

    def op(self, node: Layer, mono: Monomial) -> np.ndarray:
        """Matrix of prod_{(m,k) in mono, m in node} q_m^k in node's basis."""
        if not node.children:
            (m,) = node.modes
            mat = np.eye(self.nprim)
            for (mm, k) in mono:
                if mm == m:
                    mat = self.prim(mm, k)
            return mat
        inside = set(node.modes)
        key = (id(node), tuple(x for x in mono if x[0] in inside))
        got = self._ops.get(key)
        if got is None:
            mats = [self.op(c, mono) for c in node.children]
            full = _kron_all(mats)
            got = node.vectors.T @ full @ node.vectors
            self._ops[key] = got
        return got

    def leaf(self, mode: int) -> Layer:
        w = self.pes.omega[mode]
        h = 0.5 * w * (self.p2 + self.qk[2])
        for (mono, c) in self.pes.terms:
            if len(mono) == 1 and mono[0][0] == mode:
                h = h + c * self.qk[mono[0][1]]
        return Layer((mode,), np.diag(h).copy(), np.eye(self.nprim))

    def hamiltonian(self, node: Layer) -> np.ndarray:
        """Dense Hamiltonian of the terms entirely inside the node, in the
        product basis of its children."""
        kids = node.children
        dims = [c.vectors.shape[1] if c.children else self.nprim for c in kids]
        N = int(np.prod(dims))
        H = np.zeros((N, N))
        # each child's own Hamiltonian (diagonal in its eigenbasis for inner
        # nodes, the primitive one-mode Hamiltonian for leaves)
        for ci, c in enumerate(kids):
            if c.children:
                hc = np.diag(c.energies)
            else:
                m = c.modes[0]
                hc = 0.5 * self.pes.omega[m] * (self.p2 + self.qk[2])
                for (mono, co) in self.pes.terms:
                    if len(mono) == 1 and mono[0][0] == m:
                        hc = hc + co * self.qk[mono[0][1]]
            mats = [np.eye(dims[j]) for j in range(len(kids))]
            mats[ci] = hc
            H += _kron_all(mats)
        # terms inside the node that span more than one child
        inside = set(node.modes)
        for (mono, co) in self.pes.terms:
            ms = {m for (m, _) in mono}
            if not ms <= inside:
                continue
            owners = {ci for ci, c in enumerate(kids) if ms & set(c.modes)}
            if len(owners) < 2:
                continue
            mats = [self.op(c, mono) if ci in owners else np.eye(dims[ci])
                    for ci, c in enumerate(kids)]
            H += co * _kron_all(mats)
        return H


def build_tree(pes: PES, nprim: int, tree: Sequence[Sequence[int]],
               keeps: Sequence[int]) -> Tuple[_Basis, List[Layer]]:
    """``tree[l]`` = group sizes at layer l (in children of layer l-1);
    ``keeps[l]`` = eigenvectors kept per group at layer l.  Returns the top
    layer's groups (the modes of the term list)."""
    basis = _Basis(pes, nprim)
    nodes = [basis.leaf(m) for m in range(pes.nmodes)]
    for sizes, keep in zip(tree, keeps):
        if sum(sizes) != len(nodes):
            raise ValueError("group sizes must cover the layer below")
        new, pos = [], 0
        for sz in sizes:
            kids = nodes[pos:pos + sz]
            pos += sz
            node = Layer(tuple(m for k in kids for m in k.modes), None, None,
                         kids)
            H = basis.hamiltonian(node)
            w, v = np.linalg.eigh(H)
            v = v[:, :keep]
            # canonical signs: the largest-magnitude entry of every kept
            # eigenvector positive (LAPACK's signs depend on the BLAS
            # kernels and thread count, and they enter every operator)
            big = np.argmax(np.abs(v), axis=0)
            v = v * np.where(v[big, np.arange(v.shape[1])] < 0, -1.0, 1.0)
            node.energies = w[:keep].copy()
            node.vectors = v.copy()
            new.append(node)
        nodes = new
    return basis, nodes


def top_layer_hoplist(basis: _Basis, groups: List[Layer]) -> HopList:
    """Term list of the top layer: one eigen term per group, then every PES
    term that spans more than one group."""
    D = len(groups)
    nb = [g.energies.size for g in groups]
    terms, coefs = [], []
    for g in range(D):
        ops = [None] * D
        ops[g] = np.diag(groups[g].energies)
        terms.append(ops)
        coefs.append(1.0)
    for (mono, co) in basis.pes.terms:
        ms = {m for (m, _) in mono}
        owners = [g for g in range(D) if ms & set(groups[g].modes)]
        if len(owners) < 2:
            continue
        ops = [None] * D
        for g in owners:
            ops[g] = basis.op(groups[g], mono)
        terms.append(ops)
        coefs.append(co)
    return HopList.from_terms(nb, terms, coefs)


def lowest_products(energies: List[np.ndarray], nstates: int):
    """The nstates lowest sums sum_g E_g[i_g] and their index tuples
    (best-first enumeration; energies ascending per group)."""
    D = len(energies)
    start = (0,) * D
    heap = [(sum(e[0] for e in energies), start)]
    seen = {start}
    out = []
    while heap and len(out) < nstates:
        e, idx = heapq.heappop(heap)
        out.append((e, idx))
        for g in range(D):
            if idx[g] + 1 < energies[g].size:
                nxt = idx[:g] + (idx[g] + 1,) + idx[g + 1:]
                if nxt not in seen:
                    seen.add(nxt)
                    heapq.heappush(heap, (e - energies[g][idx[g]] +
                                          energies[g][idx[g] + 1], nxt))
    return out


def guess_block(energies: List[np.ndarray], ranks: Sequence[int],
                seed: int) -> CPBlock:
    """Initial block: product functions of the lowest group states
    (BuildProdFunctions, Guess.f90:421-448: base 5e-16 off the selected
    functions, coefficient 1) augmented to each state's rank with random
    terms (AugmentVWithRandom, MODVECVECML.f90:1378-1402: uniform(-0.5,0.5)
    entries, unit per-mode norms, coefficient 1e-12 * min coef)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = np.array([e.size for e in energies], dtype=np.int32)
    off = np.concatenate([[0], np.cumsum(nb)])
    B = len(ranks)
    prods = lowest_products(energies, B)
    rows, coefs = [], []
    for s in range(B):
        e, idx = prods[s]
        b = np.full(int(nb.sum()), 5e-16)
        for g, i in enumerate(idx):
            b[off[g] + i] = 1.0
        rows.append(b)
        coefs.append(1.0)
        for _ in range(int(ranks[s]) - 1):
            r = rng.uniform(size=int(nb.sum())) - 0.5
            for g in range(nb.size):
                r[off[g]:off[g + 1]] /= np.linalg.norm(r[off[g]:off[g + 1]])
            rows.append(r)
            coefs.append(1e-12)
    eig = np.array([p[0] for p in prods])
    return CPBlock(nb, np.asarray(ranks, dtype=np.int32), np.array(coefs),
                   np.vstack(rows), eig)


def spectral_upper_bound(H: HopList) -> float:
    """sum_t |c_t| prod_d ||h_{t,d}||_2 >= the largest eigenvalue of H."""
    ids = H.opid.reshape(H.T, H.ndof)
    norms = {}
    tot = 0.0
    for t in range(H.T):
        p = abs(H.coef[t])
        for d in range(H.ndof):
            op = ids[t, d] - 1
            if op not in norms:
                m = H.dense_op(op, int(H.nbas[d]))
                norms[op] = float(np.max(np.abs(np.linalg.eigvalsh(m))))
            p *= norms[op]
        tot += p
    return tot


@dataclass
class Workload:
    name: str
    H: HopList
    Q0: CPBlock
    bounds: Tuple[float, float]
    npow: int
    nals: int
    ncycle: int
    units: str
    groups: List[Layer] = None


def hh111(nstates=40, rank=20, seed=20261203) -> Workload:
    """6-D modified Henon-Heiles; 8 HO functions per mode; three 2-mode
    groups keeping 24; top block of 40 states at psirank 20 (npow 5,
    psinals 10, ncycle 10)."""
    pes = hh111_pes()
    basis, groups = build_tree(pes, 8, [[2, 2, 2]], [24])
    H = top_layer_hoplist(basis, groups)
    Q0 = guess_block([g.energies for g in groups], [rank] * nstates, seed)
    ub = spectral_upper_bound(H)
    return Workload("hh111-6d", H, Q0, (float(Q0.eigv[0]), ub), 5, 10, 10,
                    "dimensionless", groups)


def qff_workload(nmodes, nprim, group, keep, nstates, rank_lo, rank_hi, seed,
                 name, npow=5, nals=10, ncycle=10, tree=None,
                 keeps=None) -> Workload:
    """A synthetic force field solved up a mode tree: one level of groups of
    ``group`` modes keeping ``keep`` eigenvectors, or the levels ``tree``
    (group sizes per level, in children of the level below) keeping
    ``keeps``; the top block of ``nstates`` states with ranks uniform in
    rank_lo..rank_hi."""
    pes = qff_pes(nmodes, seed)
    if tree is None:
        if nmodes % group:
            raise ValueError("nmodes must be a multiple of the group size")
        tree, keeps = [[group] * (nmodes // group)], [keep]
    basis, groups = build_tree(pes, nprim, tree, keeps)
    H = top_layer_hoplist(basis, groups)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    ranks = rng.integers(rank_lo, rank_hi + 1, size=nstates)
    Q0 = guess_block([g.energies for g in groups], ranks, seed + 2)
    ub = spectral_upper_bound(H)
    return Workload(name, H, Q0, (float(Q0.eigv[0]), ub), npow, nals, ncycle,
                    "cm-1", groups)


def qff12(nstates=64, seed=20261204) -> Workload:
    """12-mode synthetic QFF, 10 HO functions per mode, 2-mode groups keep
    40, 64 states with ranks uniform in 8..48."""
    return qff_workload(12, 10, 2, 40, nstates, 8, 48, seed, "qff12-d6n40")


def qff24(nstates=120, rank=32, seed=20261205) -> Workload:
    """24-mode synthetic QFF, 12 HO functions per mode, 3-mode groups keep
    60, 120 states at psirank 32."""
    return qff_workload(24, 12, 3, 60, nstates, rank, rank, seed,
                        "qff24-d8n60")


def qff48(nstates=200, seed=20261206) -> Workload:
    """48-mode synthetic QFF, 10 HO functions per mode, a three-layer tree:
    pairs keep 40, pairs of pairs (quads) keep 120, top block of 200 states
    with ranks uniform in 16..64."""
    return qff_workload(48, 10, 0, 0, nstates, 16, 64, seed, "qff48-d12n120",
                        tree=[[2] * 24, [2] * 12], keeps=[40, 120])
