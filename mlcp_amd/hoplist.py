"""Term list (HopList) and CP-block host types, in the reference's wire format.

The reference passes the Hamiltonian to the GPU path as a flat term list
(``TYPE HopList``, HamilSetup.f90:37-40, built by ModeH.f90:909-1068):

* ``opid[t*ndof + d]``  1-based index into ``mptr`` of the operator that term
  ``t`` applies on mode ``d``;
* ``mptr[op]``          1-based start of that operator's matrix in ``mat``;
* ``mat``               packed upper triangles, column major: element (i, j),
  i <= j, 0-based, at ``i + j*(j+1)//2`` (Utils.f90:559-597);
* ``coef[t]``           term coefficient, applied on mode 0 only
  (MODHVECML.f90:196, intertw_mvp.cu:168-169).

A block of CP vectors uses the reference's Fortran ``qbase(nD, R*nbloc)``
layout (GPUIterate.f90:137-147): term ``r`` of the block occupies
``base[r, :]`` of a C-ordered ``(Rtot, nD)`` array, mode ``d`` at columns
``off[d]:off[d+1]``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

OP_IDENTITY = 0
OP_DIAGONAL = 1
OP_DENSE = 2


def pack_upper(m: np.ndarray) -> np.ndarray:
    """Packed upper triangle of a symmetric matrix (column major)."""
    n = m.shape[0]
    out = np.empty(n * (n + 1) // 2, dtype=np.float64)
    k = 0
    for j in range(n):
        out[k:k + j + 1] = m[:j + 1, j]
        k += j + 1
    return out


def unpack_upper(ap: np.ndarray, n: int) -> np.ndarray:
    """Dense symmetric matrix from a packed upper triangle (Utils.f90:599-647)."""
    m = np.zeros((n, n), dtype=np.float64)
    k = 0
    for j in range(n):
        m[:j + 1, j] = ap[k:k + j + 1]
        k += j + 1
    iu = np.triu_indices(n, 1)
    m[iu[1], iu[0]] = m[iu]
    return m


@dataclass
class HopList:
    nbas: np.ndarray          # int32 (ndof,)
    opid: np.ndarray          # int32 (T*ndof,) 1-based
    mptr: np.ndarray          # int32 (nops,)   1-based
    coef: np.ndarray          # float64 (T,)
    mat: np.ndarray           # float64 (msize,)

    def __post_init__(self):
        self.nbas = np.ascontiguousarray(self.nbas, dtype=np.int32)
        self.opid = np.ascontiguousarray(self.opid, dtype=np.int32)
        self.mptr = np.ascontiguousarray(self.mptr, dtype=np.int32)
        self.coef = np.ascontiguousarray(self.coef, dtype=np.float64)
        self.mat = np.ascontiguousarray(self.mat, dtype=np.float64)

    @property
    def ndof(self) -> int:
        return int(self.nbas.size)

    @property
    def T(self) -> int:
        return int(self.coef.size)

    @property
    def nops(self) -> int:
        return int(self.mptr.size)

    @property
    def msize(self) -> int:
        return int(self.mat.size)

    def op_mode(self) -> np.ndarray:
        """Mode each operator is used on (-1 if unused)."""
        mode = np.full(self.nops, -1, dtype=np.int64)
        ids = self.opid.reshape(self.T, self.ndof)
        for d in range(self.ndof):
            mode[ids[:, d] - 1] = d
        return mode

    def validate(self) -> None:
        """Checks that every index stays inside its array (raises ValueError)."""
        ndof, T = self.ndof, self.T
        if ndof < 1 or T < 1:
            raise ValueError("empty term list")
        if self.opid.size != T * ndof:
            raise ValueError("opid length must be T*ndof")
        if self.opid.min() < 1 or self.opid.max() > self.nops:
            raise ValueError("opid out of range")
        ids = self.opid.reshape(T, ndof)
        seen = np.full(self.nops, -1)
        for d in range(ndof):
            for op in np.unique(ids[:, d]):
                if seen[op - 1] not in (-1, d):
                    raise ValueError(f"operator {op} used on two modes")
                seen[op - 1] = d
        for op in range(self.nops):
            d = seen[op]
            if d < 0:
                continue
            n = int(self.nbas[d])
            lo = int(self.mptr[op]) - 1
            if lo < 0 or lo + n * (n + 1) // 2 > self.msize:
                raise ValueError(f"operator {op + 1} matrix out of range")

    def dense_op(self, op: int, n: int) -> np.ndarray:
        lo = int(self.mptr[op]) - 1
        return unpack_upper(self.mat[lo:lo + n * (n + 1) // 2], n)

    def classify(self) -> np.ndarray:
        """OP_IDENTITY / OP_DIAGONAL / OP_DENSE per operator, by exact test."""
        mode = self.op_mode()
        cls = np.full(self.nops, OP_DENSE, dtype=np.int32)
        for op in range(self.nops):
            if mode[op] < 0:
                continue
            m = self.dense_op(op, int(self.nbas[mode[op]]))
            off = m - np.diag(np.diag(m))
            if np.all(off == 0.0):
                cls[op] = OP_IDENTITY if np.all(np.diag(m) == 1.0) else OP_DIAGONAL
        return cls

    @staticmethod
    def from_terms(nbas: Sequence[int], terms: List[List[np.ndarray]],
                   coefs: Sequence[float]) -> "HopList":
        """Builds a term list from dense per-mode operators.

        ``terms[t][d]`` is an (n_d, n_d) symmetric matrix or ``None`` for the
        identity.  The operator order follows GetHopList (ModeH.f90:1016-1060):
        the ndof identities first, then every distinct matrix in order of first
        use, deduplicated per mode by exact equality."""
        nbas = [int(n) for n in nbas]
        ndof = len(nbas)
        mats: List[np.ndarray] = []
        mptr: List[int] = []
        ls = 1
        for d in range(ndof):
            ap = pack_upper(np.eye(nbas[d]))
            mats.append(ap)
            mptr.append(ls)
            ls += ap.size
        uniq: List[List[tuple]] = [[] for _ in range(ndof)]
        opid = []
        for t, ops in enumerate(terms):
            for d in range(ndof):
                m = ops[d]
                if m is None:
                    opid.append(d + 1)
                    continue
                ap = pack_upper(np.asarray(m, dtype=np.float64))
                found = None
                for (key, idx) in uniq[d]:
                    if key.shape == ap.shape and np.array_equal(key, ap):
                        found = idx
                        break
                if found is None:
                    mats.append(ap)
                    mptr.append(ls)
                    ls += ap.size
                    found = len(mptr)
                    uniq[d].append((ap, found))
                opid.append(found)
        return HopList(np.array(nbas), np.array(opid), np.array(mptr),
                       np.array(coefs, dtype=np.float64), np.concatenate(mats))


def hoplist_from_hamiltonian(nbas: Sequence[int], opcoef, opindx,
                             iops: List[np.ndarray], eops: List[np.ndarray],
                             cops: List[np.ndarray],
                             pops: List[np.ndarray]) -> HopList:
    """GetHopList (ModeH.f90:909-1068): the wire format of a mode
    Hamiltonian held as TYPE(Hamiltonian).

    ``opindx[k][j]`` (term k, mode j) is 0 for the identity, -e for eigen
    operator e of mode j, i > len(pops) for compound operator i - len(pops),
    otherwise primitive operator i (1-based into ``pops``); ``iops`` and
    ``eops`` hold one packed matrix per mode, ``cops``/``pops`` the compound
    and primitive ones.  Operators are numbered identities (1..D), eigen
    (D+1..2D), compound, then per mode the distinct primitives of that mode
    in order of first use (terms outer, modes inner)."""
    opindx = np.asarray(opindx, dtype=np.int64)
    nrk, ndof = opindx.shape
    npop, ncop = len(pops), len(cops)
    pptr = [[] for _ in range(ndof)]
    for k in range(nrk):
        for j in range(ndof):
            v = int(opindx[k, j])
            if 1 <= v <= npop and v not in pptr[j]:
                pptr[j].append(v)
    opid = np.zeros(nrk * ndof, dtype=np.int32)
    l = 0
    for k in range(nrk):
        psum = 0
        for j in range(ndof):
            v = int(opindx[k, j])
            if v == 0:
                opid[l] = j + 1
            elif v < 0:
                opid[l] = ndof + abs(v)
            elif v > npop:
                opid[l] = 2 * ndof + v - npop
            else:
                opid[l] = 2 * ndof + ncop + psum + pptr[j].index(v) + 1
            psum += len(pptr[j])
            l += 1
    mats = list(iops) + list(eops) + list(cops)
    for j in range(ndof):
        mats += [pops[v - 1] for v in pptr[j]]
    mptr = np.zeros(len(mats), dtype=np.int32)
    pos = 1
    for o, m in enumerate(mats):
        mptr[o] = pos
        pos += np.asarray(m).size
    mat = (np.concatenate([np.asarray(m, dtype=np.float64).ravel()
                           for m in mats]) if mats else np.zeros(0))
    return HopList(np.asarray(nbas, dtype=np.int32), opid, mptr,
                   np.asarray(opcoef, dtype=np.float64), mat)


@dataclass
class CPBlock:
    nbas: np.ndarray     # int32 (ndof,)
    ranks: np.ndarray    # int32 (nbloc,)
    coef: np.ndarray     # float64 (Rtot,)
    base: np.ndarray     # float64 (Rtot, nD)
    eigv: np.ndarray = field(default=None)  # float64 (nbloc,)

    def __post_init__(self):
        self.nbas = np.ascontiguousarray(self.nbas, dtype=np.int32)
        self.ranks = np.ascontiguousarray(self.ranks, dtype=np.int32)
        self.coef = np.ascontiguousarray(self.coef, dtype=np.float64)
        self.base = np.ascontiguousarray(self.base, dtype=np.float64)
        if self.eigv is None:
            self.eigv = np.zeros(self.ranks.size)
        self.eigv = np.ascontiguousarray(self.eigv, dtype=np.float64)
        if self.base.shape != (int(self.ranks.sum()), int(self.nbas.sum())):
            raise ValueError("base must be (sum(ranks), sum(nbas))")

    @property
    def nbloc(self) -> int:
        return int(self.ranks.size)

    @property
    def offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.ranks)]).astype(np.int64)

    def state(self, s: int):
        o = self.offsets
        return self.coef[o[s]:o[s + 1]], self.base[o[s]:o[s + 1]]

    def copy(self) -> "CPBlock":
        return CPBlock(self.nbas.copy(), self.ranks.copy(), self.coef.copy(),
                       self.base.copy(), self.eigv.copy())
