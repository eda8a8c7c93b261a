"""Writes the teacher-forced fixtures of tests/golden/: a late block of a
benchmark workload's solve and the FP64 oracle's result of the next cycle
from exactly that block.

    python tests/golden/make_teacher.py <workload> <block.npz>

The block (ranks, coef, base, eigv after the cycle, bounds) comes from a
solve on the GPU (the trajectory itself is not compared: ALS sweeps amplify
rounding by x15-2000 per cycle on these blocks); the oracle then runs one
intertwine cycle (npow power sweeps, Gram-Schmidt, reduced H*Q, block
eigenproblem, updates) with the E-shift GetBlockShift gives for those
eigenvalues and bounds.  Saved as teacher_<workload>.npz: the input block,
the shift, the oracle's eigenvalues and its output coefficients.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(
    __file__))))
sys.path.insert(0, ROOT)

from mlcp_amd import systems  # noqa: E402
from mlcp_amd.hoplist import CPBlock  # noqa: E402
from mlcp_amd.intertwine import block_shift  # noqa: E402
from oracle.orc import C  # noqa: E402
from tests.golden.make_golden import fingerprint  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main(name, src):
    w = {"hh111": systems.hh111, "qff12": systems.qff12,
         "qff24": systems.qff24}[name]()
    z = np.load(src)
    Q = CPBlock(w.Q0.nbas, z["ranks"].astype(np.int32), z["coef"], z["base"],
                z["eigv"].copy())
    bounds = np.asarray(z["bounds"], dtype=np.float64)
    es = block_shift(Q.eigv, bounds)
    t0 = time.time()
    out = C.intertwine(w.H, Q, w.npow, w.nals, True, 1, es)
    print(f"{name}: oracle cycle {time.time() - t0:.1f} s, E0 "
          f"{out.eigv[0]:.6f}", flush=True)
    np.savez_compressed(
        os.path.join(HERE, f"teacher_{name}.npz"), ranks=Q.ranks,
        coef=Q.coef, base=Q.base, eigv_in=Q.eigv, bounds=bounds,
        eshift=np.float64(es), npow=w.npow, nals=w.nals,
        eigv_out=out.eigv, coef_out=out.coef,
        fp_mat_sum=fingerprint(w.H, w.Q0)["mat_sum"])


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
