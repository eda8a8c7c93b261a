"""Force-field (QFF) files of the reference: ``pes/f<k><id>.dat``.

The reference reads a potential of up to ``ncp``-fold couplings from files
``pes/f1<id>.dat`` ... ``pes/f<ncp><id>.dat`` (``ReadFFHamiltonian``,
FFPES.f90:308-406); ``<id>`` is the 5-character system code.  Each line of
file k holds k mode numbers (1-based, a mode repeated for powers) and the
force constant, e.g. ``1 1 7 8  -3.2`` = -3.2 q1^2 q7 q8.  Reading stops at the
first line that list-directed input cannot read.  With ``divide`` every
constant is divided by the factorials of its mode powers
(``DistribModePower``, FFPES.f90:471-511).  The quadratic file holds
``i i w_i/2`` so that, with the kinetic term the reference adds, the harmonic
part is sum_i (w_i/2)(p_i^2 + q_i^2).

``write_pes_files`` writes a generated force field (mlcp_amd.systems.PES) in
this format and ``read_ff_hamiltonian`` mirrors the reference's reader, so a
system generated here is the system the reference reads.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass
from typing import List

import numpy as np


def ff_path(dirpath: str, k: int, sysid: str) -> str:
    if len(sysid) != 5:
        raise ValueError("the system code has exactly 5 characters")
    return os.path.join(dirpath, "pes", f"f{k}{sysid}.dat")


def pes_monomials(pes):
    """(modes 1-based with repetition, force constant) per term of a PES,
    the harmonic part as i i w_i/2."""
    out = [((i + 1, i + 1), float(pes.omega[i]) / 2.0)
           for i in range(pes.nmodes)]
    for mono, c in pes.terms:
        q = []
        for m, p in mono:
            q += [m + 1] * p
        out.append((tuple(sorted(q)), float(c)))
    return out


def write_pes_files(pes, sysid: str, dirpath: str) -> List[str]:
    """Writes the force field as pes/f<k><id>.dat (exact decimal form of
    every constant); returns the files written."""
    by_order = {}
    for q, c in pes_monomials(pes):
        by_order.setdefault(len(q), []).append((q, c))
    os.makedirs(os.path.join(dirpath, "pes"), exist_ok=True)
    files = []
    for k in sorted(by_order):
        p = ff_path(dirpath, k, sysid)
        with open(p, "w") as f:
            for q, c in by_order[k]:
                f.write(" ".join(f"{m:3d}" for m in q) + f"  {c!r}\n")
        files.append(p)
    return files


@dataclass
class Configs:
    """TYPE(Configs) of CPConfig.f90:13-16 for one coupling order."""
    qns: np.ndarray      # (n, k) int, 1-based modes
    coef: np.ndarray     # (n,)


_num = re.compile(r"[,\s]+")


def _fortran_real(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "E"))


def _read_items(lines, pos, k):
    """One list-directed read of k integers and a real starting at line
    ``pos``: values may continue onto following lines, the rest of the last
    line used is skipped.  Returns (qns, value, next line) or None when the
    read fails (end of file or an unreadable value)."""
    vals = []
    while len(vals) < k + 1:
        if pos >= len(lines):
            return None
        toks = [t for t in _num.split(lines[pos].strip()) if t]
        pos += 1
        for t in toks:
            if t.startswith("/"):
                return None
            vals.append(t)
            if len(vals) == k + 1:
                break
    try:
        qns = [int(v) for v in vals[:k]]
        x = _fortran_real(vals[k])
    except ValueError:
        return None
    return qns, x, pos


def distrib_mode_power(modes):
    """DistribModePower (FFPES.f90:471-511): [(mode, power)] in order of
    first appearance."""
    out = []
    for m in modes:
        for e in out:
            if e[0] == m:
                e[1] += 1
                break
        else:
            out.append([m, 1])
    return out


def read_ff_hamiltonian(dirpath: str, sysid: str, ncp: int,
                        divide: bool) -> List[Configs]:
    """ReadFFHamiltonian (FFPES.f90:308-406): the constants of orders
    1..ncp; an order without a file or without a readable line gets one zero
    configuration."""
    data = {}
    for k in range(1, ncp + 1):
        p = ff_path(dirpath, k, sysid)
        if not os.path.exists(p):
            continue
        lines = open(p).read().splitlines()
        rows, pos = [], 0
        while True:
            r = _read_items(lines, pos, k)
            if r is None:
                break
            rows.append(r[:2])
            pos = r[2]
        data[k] = rows
    out = []
    for k in range(1, ncp + 1):
        rows = data.get(k, [])
        n = max(1, len(rows))
        qns = np.zeros((n, k), dtype=np.int32)
        coef = np.zeros(n)
        for i, (q, c) in enumerate(rows):
            qns[i] = q
            if divide:
                for _, pw in distrib_mode_power(q):
                    c = c / float(math.factorial(pw))
            coef[i] = c
        out.append(Configs(qns, coef))
    return out
