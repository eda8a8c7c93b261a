"""The device-resident solve (mlcp_hip_session_solve): SolveHPsi's cycle
loop with GetBlockShift on the device and every cycle after the first
replayed as one captured HIP graph.  It must give, bit for bit, what the
host loop (mlcp_hip_session_cycle with the shift formed on the host) gives:
the same kernels in the same order on the same data, and a shift formula
with exact steps (2 a - b, a sum and a halving).  Also with the graph off
(MLCP_HIP_GRAPH=0) and in the tolerance form."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _host_loop(s, w, bounds, ncyc):
    from mlcp_amd.intertwine import block_shift
    eig = w.Q0.eigv.copy()
    hist, shifts = [], []
    for _ in range(ncyc):
        es = block_shift(eig, bounds)
        eig = s.cycle(eig, w.npow, w.nals, True, 1, es)
        hist.append(eig.copy())
        shifts.append(es)
    return eig, np.array(hist), np.array(shifts)


@pytest.mark.parametrize("name,graph,form", [("hh111", "1", 0),
                                             ("qff12", "1", 0),
                                             ("qff12", "0", 0),
                                             ("qff12", "1", 1)])
def test_device_solve_matches_host_loop(hip, monkeypatch, name, graph, form):
    from mlcp_amd.intertwine import Session
    from tests.test_gpu_workloads import _workload
    monkeypatch.setenv("MLCP_HIP_GRAPH", graph)
    w, ref = _workload(name)
    ncyc = 4
    say("device solve", name, "graph", graph, "form", form, "cycles", ncyc)
    with Session(w.Q0, w.H) as s:
        s.set_form(form, 1e-5, 1e-7)
        s.save()
        e_h, h_h, sh_h = _host_loop(s, w, ref["bounds"], ncyc)
        b_h = s.download()
        s.restore()
        e_d, h_d, sh_d = s.solve(w.Q0.eigv, ref["bounds"], ncyc, w.npow,
                                 w.nals)
        b_d = s.download()
        # a second solve replays the cached graph
        s.restore()
        e_2, h_2, _ = s.solve(w.Q0.eigv, ref["bounds"], ncyc, w.npow, w.nals)
    assert np.array_equal(sh_h, sh_d), (sh_h, sh_d)
    assert np.array_equal(h_h, h_d)
    assert np.array_equal(e_h, e_d) and np.array_equal(e_d, e_2)
    assert np.array_equal(h_d, h_2)
    assert np.array_equal(b_h.base, b_d.base)
    assert np.array_equal(b_h.coef, b_d.coef)
