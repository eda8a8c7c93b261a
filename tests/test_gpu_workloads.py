"""The two benchmark workloads at their full sizes, on the GPU, against the
FP64 oracle's eigenvalues after every cycle (tests/golden/workload_*.json,
written by make_golden.py with the oracle running the same cycles from the
same block).

  hh111  Henon-Heiles, 3 top modes of 24 functions, 40 states of rank 20
  qff12  quartic force field, 6 top modes of 40 functions, 64 states of
         ranks 8..48, 182 terms

ALS is a fixed-point iteration whose sweeps amplify a rounding-level
difference: measured on the MI355X, the 12-mode block's largest eigenvalue
difference is 1.8e-10 cm^-1 after the first cycle and grows by 15-2000x per
cycle (3.7e-7, 2.3e-5, 8.3e-4, 1.3e-2, ... 5.9 after ten), while the
Henon-Heiles block stays below 1.4e-5 for all ten.  So the 12-mode block is
held to per-cycle bounds for its first three cycles only, and the whole
trajectory of both blocks to what does not depend on the rounding:
eigenvalues finite and in ascending order.  The E-shift of a cycle is formed
from the previous cycle's eigenvalues (exact in cycle 0), so it is held to
twice their bound.

Tolerances (cm^-1, |w_gpu - w_oracle| over all states):
  hh111  every cycle                       1e-3
  qff12  cycle 0 / 1 / 2                   1e-7 / 1e-4 / 1e-2
  E-shift of cycle c > 0: 2x the bound of cycle c - 1; cycle 0: rel 1e-12
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

BOUNDS = {"hh111": [1e-3] * 10, "qff12": [1e-7, 1e-4, 1e-2]}


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _workload(name):
    from mlcp_amd import systems
    from tests.golden.make_golden import fingerprint
    w = {"hh111": systems.hh111, "qff12": systems.qff12}[name]()
    ref = json.load(open(os.path.join(ROOT, "tests", "golden",
                                      f"workload_{name}.json")))
    # the generator must still produce the block the oracle ran
    fp = fingerprint(w.H, w.Q0)
    for k, v in ref["fingerprint"].items():
        if isinstance(v, (list, float)):
            np.testing.assert_allclose(np.asarray(fp[k], float),
                                       np.asarray(v, float), rtol=1e-9, atol=0)
        else:
            assert fp[k] == v, k
    assert (w.ncycle, w.npow, w.nals) == (ref["ncycle"], ref["npow"], ref["nals"])
    return w, ref


@pytest.mark.parametrize("name", ["hh111", "qff12"])
def test_benchmark_workload_full_size(hip, name):
    from mlcp_amd.intertwine import Session, block_shift
    w, ref = _workload(name)
    say(name, "states", w.Q0.nbloc, "ranks", int(w.Q0.ranks.min()),
        int(w.Q0.ranks.max()), "cycles", w.ncycle)
    bounds = BOUNDS[name]
    with Session(w.Q0, w.H) as s:
        eig = w.Q0.eigv.copy()
        for c in range(w.ncycle):
            es = block_shift(eig, w.bounds)
            # the shift is formed from the previous cycle's eigenvalues
            tol_es = (1e-12 * abs(ref["eshift"][0]) if c == 0
                      else 2.0 * bounds[c - 1] if c <= len(bounds) else None)
            if tol_es is not None:
                assert abs(es - ref["eshift"][c]) <= tol_es, (c, es)
            eig = s.cycle(eig, w.npow, w.nals, True, 1, es)
            assert np.all(np.isfinite(eig))
            assert np.all(np.diff(eig) >= 0.0), f"cycle {c}: not ascending"
            if c < len(bounds):
                err = np.abs(eig - np.asarray(ref["eigv"][c])).max()
                say(name, "cycle", c, "max |dE|", f"{err:.3e}")
                assert err <= bounds[c], (c, err)
