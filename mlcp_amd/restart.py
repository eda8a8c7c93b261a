"""Restart files of the eigensolver: the psi file of a block of states.

Mirrors the reference's ``SavePsi`` / ``SavePsiFile`` / ``ReadPsi`` /
``ReadPsiFile`` (Restart.f90:579-793) and their use in ``SolveHPsi``
(Solver.f90:340,370,432).  The file is Fortran unformatted sequential: every
record is framed by its byte length as a little-endian int32 before and
after it (the gfortran/flang default).  Records, in order:

    (isavi int32, nev int32)          solver cycle of the save, block size
    bounds(2) real64                  spectral range
    eigv(nev) real64                  eigenvalues
    delta(nev) real64                 relative change in the last cycle
    for each state i:
        (ndof int32, nrk int32)
        nbas(ndof) int32
        coef(nrk) real64
        base(nD, nrk) real64          column major, term after term

The block is written to ``<resfile>_1_psi.rst`` and again to
``<resfile>_2_psi.rst``; reading tries the first and falls back to the second
(a crash during one write leaves the other whole).  A read succeeds only if
every record is present and ``nev`` and each state's ``nbas`` match the
block being restarted; on failure nothing is replaced.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

from .hoplist import CPBlock


def _put(f, payload: bytes) -> None:
    n = struct.pack("<i", len(payload))
    f.write(n)
    f.write(payload)
    f.write(n)


def _records(raw: bytes):
    """Split a sequential unformatted file into its records; stops at the
    first damaged frame."""
    pos, out = 0, []
    while pos + 4 <= len(raw):
        (n,) = struct.unpack_from("<i", raw, pos)
        end = pos + 4 + n
        if n < 0 or end + 4 > len(raw):
            break
        (m,) = struct.unpack_from("<i", raw, end)
        if m != n:
            break
        out.append(raw[pos + 4:end])
        pos = end + 4
    return out


def save_psi_file(fnm: str, isavi: int, bounds, eigv, delta,
                  Q: CPBlock) -> None:
    """SavePsiFile (Restart.f90:759-793)."""
    eigv = np.asarray(eigv, dtype="<f8")
    delta = np.asarray(delta, dtype="<f8")
    nev = eigv.size
    if Q.nbloc != nev or delta.size != nev:
        raise ValueError("block, eigv and delta sizes differ")
    nb = np.asarray(Q.nbas, dtype="<i4")
    o = Q.offsets
    with open(fnm, "wb") as f:
        _put(f, struct.pack("<ii", int(isavi), nev))
        _put(f, np.asarray(bounds, dtype="<f8").reshape(2).tobytes())
        _put(f, eigv.tobytes())
        _put(f, delta.tobytes())
        for i in range(nev):
            r = int(Q.ranks[i])
            _put(f, struct.pack("<ii", nb.size, r))
            _put(f, nb.tobytes())
            _put(f, np.asarray(Q.coef[o[i]:o[i + 1]], dtype="<f8").tobytes())
            _put(f, np.asarray(Q.base[o[i]:o[i + 1]], dtype="<f8").tobytes())


@dataclass
class PsiRead:
    success: bool
    isavi: int = 0
    bounds: np.ndarray = None
    eigv: np.ndarray = None
    delta: np.ndarray = None
    block: CPBlock = None


def read_psi_file(fnm: str, nev: int, nbas) -> PsiRead:
    """ReadPsiFile (Restart.f90:614-729): every check of the reference, no
    partial result on failure."""
    if not os.path.exists(fnm):
        return PsiRead(False)
    nbas = np.asarray(nbas, dtype=np.int32)
    rec = _records(open(fnm, "rb").read())
    try:
        if len(rec) < 4 or len(rec[0]) < 8:
            return PsiRead(False)
        isavi, n = struct.unpack_from("<ii", rec[0])
        if n != nev or len(rec[1]) < 16:
            return PsiRead(False)
        bounds = np.frombuffer(rec[1][:16], dtype="<f8").copy()
        if len(rec[2]) < 8 * nev or len(rec[3]) < 8 * nev:
            return PsiRead(False)
        eigv = np.frombuffer(rec[2][:8 * nev], dtype="<f8").copy()
        delta = np.frombuffer(rec[3][:8 * nev], dtype="<f8").copy()
        nD = int(nbas.sum())
        ranks, coefs, bases = [], [], []
        k = 4
        for _ in range(nev):
            if k + 3 >= len(rec) or len(rec[k]) < 8:
                return PsiRead(False)
            ndof, nrk = struct.unpack_from("<ii", rec[k])
            if ndof != nbas.size or nrk < 1:
                return PsiRead(False)
            nb = np.frombuffer(rec[k + 1][:4 * ndof], dtype="<i4")
            if nb.size != ndof or not np.array_equal(nb, nbas):
                return PsiRead(False)
            if len(rec[k + 2]) < 8 * nrk or len(rec[k + 3]) < 8 * nrk * nD:
                return PsiRead(False)
            coefs.append(np.frombuffer(rec[k + 2][:8 * nrk], dtype="<f8"))
            bases.append(np.frombuffer(rec[k + 3][:8 * nrk * nD],
                                       dtype="<f8").reshape(nrk, nD))
            ranks.append(nrk)
            k += 4
    except struct.error:
        return PsiRead(False)
    Q = CPBlock(nbas, np.asarray(ranks, dtype=np.int32),
                np.concatenate(coefs).copy(), np.concatenate(bases).copy(),
                eigv.copy())
    return PsiRead(True, int(isavi), bounds, eigv, delta, Q)


def _is_none(resfile: str) -> bool:
    # the reference compares the first four characters case-insensitively
    return resfile[:4].lower() == "none"


def save_psi(resfile: str, isavi: int, bounds, eigv, delta, Q: CPBlock):
    """SavePsi (Restart.f90:733-755): both copies, unless resfile is
    'none'."""
    if _is_none(resfile):
        return
    for k in (1, 2):
        save_psi_file(f"{resfile}_{k}_psi.rst", isavi, bounds, eigv, delta, Q)


def read_psi(resfile: str, nev: int, nbas) -> PsiRead:
    """ReadPsi (Restart.f90:579-610): the first copy, else the second."""
    r = read_psi_file(f"{resfile}_1_psi.rst", nev, nbas)
    if not r.success:
        r = read_psi_file(f"{resfile}_2_psi.rst", nev, nbas)
    return r
