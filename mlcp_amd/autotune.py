"""Autotuner of the library's launch parameters, with a persisted table.

The reference recomputes its launch shapes on every call from fixed formulas
(``PVVkernel::initnew``, modvecvec.cu:27-94).  Here the shapes that matter
are measured per problem on the device and kept in a text table that the
library reads when a session is created, so a later run (including the
Fortran drop-in, which never passes through Python) starts tuned.

Knobs (``Session.tune_set``; none of them changes a result bit):

* ``had_tile``  -- elements of the R x R Gram tile per cross-mode Hadamard
  workgroup (multiple of 64, at most what the operator Grams leave of LDS);
* ``had_split`` -- waves sharing a 64-element group of that tile, each over
  its own range of the mode's sub-slots (1, 2, 4, 8 or 16; had_tile/64 x
  had_split waves per workgroup, at most 16);
* ``gs_stride`` -- Gram-Schmidt workgroup stride (1: all CUs, 8: one XCD).

Table format, one line per problem::

    <key> <had_tile> <had_split> <gs_stride>

with the key of ``Session.tune_key()``
(``<arch>/D<ndof>/n<n_1>...../B<nbloc>/R<min>.<max>.<sum>/o<nops>/t<T>``).
The library reads ``$MLCP_HIP_TUNING`` if set (``none``: no table), else
``tuning_table.txt`` beside ``libmlcp_hip.so``.

    python -m mlcp_amd.autotune --workload qff24 [--table PATH] [--reps 2]
"""
from __future__ import annotations

import argparse
import os
import time
from typing import Dict, List, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning_table.txt")

SPLITS = (1, 2, 4, 8, 16)
STRIDES = (1, 8)


def read_table(path: str = TABLE) -> Dict[str, Tuple[int, int, int]]:
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        f = line.split()
        if len(f) != 4:
            raise ValueError(f"{path}: bad line {line!r}")
        out[f[0]] = (int(f[1]), int(f[2]), int(f[3]))
    return out


def write_entry(path: str, key: str, params: Tuple[int, int, int]) -> None:
    """Adds or replaces the entry of ``key``; other entries are kept."""
    if any(c.isspace() for c in key):
        raise ValueError("a key has no white space")
    t = read_table(path)
    t[key] = tuple(int(x) for x in params)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        f.write("# launch parameters per problem, written by mlcp_amd.autotune\n"
                "# <key> <had_tile> <had_split> <gs_stride>\n")
        for k in sorted(t):
            f.write(f"{k} {t[k][0]} {t[k][1]} {t[k][2]}\n")
    os.replace(tmp, path)


def candidates(had_tile_max: int) -> List[Tuple[int, int]]:
    """(had_tile, had_split) pairs of the Hadamard search."""
    tiles = [t for t in range(64, 257, 64) if t <= had_tile_max]
    return [(t, p) for t in tiles for p in SPLITS if (t // 64) * p <= 16]


def _measure(sess, eigv, npow, nals, eshift, reps):
    """Phase milliseconds of one cycle from the saved block (minimum over
    reps), see Session.phase_ms."""
    best = None
    for _ in range(reps):
        sess.restore()
        sess.cycle(eigv, npow, nals, True, 1, eshift)
        ms = sess.phase_ms()
        best = ms if best is None else np.minimum(best, ms)
    return best


def tune(sess, eigv, npow: int, nals: int, eshift: float, reps: int = 2,
         table: str = TABLE, persist: bool = True, log=None):
    """Coordinate search: the Hadamard knobs on the power and reduced H*Q
    phases, then the Gram-Schmidt stride on its phase.  The session ends on
    the best parameters with its block restored.  Returns (params, trials)."""
    sess.save()
    tmax = sess.tuning()["had_tile_max"]
    trials = []
    best = None
    for t, p in candidates(tmax):
        sess.tune_set(t, p, 0)
        ms = _measure(sess, eigv, npow, nals, eshift, reps)
        cost = float(ms[0] + ms[2])
        trials.append({"had_tile": t, "had_split": p, "ms": cost})
        if log:
            log(f"had_tile {t:3d} had_split {p}: {cost:9.3f} ms")
        if best is None or cost < best[0]:
            best = (cost, t, p)
    sess.tune_set(best[1], best[2], 0)
    bx = None
    for x in STRIDES:
        sess.tune_set(0, 0, x)
        ms = _measure(sess, eigv, npow, nals, eshift, reps)
        cost = float(ms[1])
        trials.append({"gs_stride": x, "ms": cost})
        if log:
            log(f"gs_stride {x}: {cost:9.3f} ms")
        if bx is None or cost < bx[0]:
            bx = (cost, x)
    sess.tune_set(0, 0, bx[1])
    sess.restore()
    params = (best[1], best[2], bx[1])
    if persist:
        write_entry(table, sess.tune_key(), params)
    return params, trials


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="qff12")
    ap.add_argument("--table", default=TABLE)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()
    from mlcp_amd import systems
    from mlcp_amd.intertwine import Session, block_shift
    w = getattr(systems, args.workload)()
    os.environ["MLCP_HIP_TUNING"] = "none"   # start from the defaults
    t0 = time.time()
    with Session(w.Q0, w.H) as s:
        print("key", s.tune_key(), "defaults", s.tuning(), flush=True)
        params, _ = tune(s, w.Q0.eigv, w.npow, w.nals,
                         block_shift(w.Q0.eigv, w.bounds), reps=args.reps,
                         table=args.table,
                         log=lambda m: print(m, flush=True))
        print("chosen", params, f"in {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
