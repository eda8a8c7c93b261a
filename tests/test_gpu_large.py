"""The 48-mode configuration at full size on the GPU: a three-layer tree
(pairs keep 40, quads keep 120), 12 top modes of 120 functions, 200 states
of ranks 16..64 and 6807 terms.  No oracle runs at this size (one
oracle sweep of one state costs ~1e12 flops in the reference's lowmem=3
form), so the checks are properties that hold whatever the rounding:

* after the power sweeps and Gram-Schmidt, the block's overlap matrix S is
  the identity to the ALS tolerance (|S - I| <= 0.05), and its diagonal is
  one to rounding (every state is renormalised);
* a full cycle (reduced H*Q, block eigenproblem, updates) returns finite,
  ascending eigenvalues inside the spectral bounds;
* the Rayleigh quotients of the new block lie inside the bounds too.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_qff48_full_size_properties(hip):
    from mlcp_amd import systems
    from mlcp_amd.intertwine import Session, block_shift, spectral_range
    w = systems.qff48()
    say("qff48 terms", w.H.T, "ops", w.H.nops, "states", w.Q0.nbloc,
        "ranks", int(w.Q0.ranks.min()), int(w.Q0.ranks.max()))
    assert (w.H.ndof, int(w.H.nbas[0]), w.Q0.nbloc) == (12, 120, 200)
    lo, hi = spectral_range(w.H, 50, 10, 5)
    say("spectral range", lo, hi)
    assert lo < w.Q0.eigv.min() < hi
    es = block_shift(w.Q0.eigv, (lo, hi))
    with Session(w.Q0, w.H) as s:
        s.save()
        s.power(w.npow, 1, es)
        s.ortho(w.nals)
        _, S, _, _ = s.gep(w.nals)
        S = S + np.triu(S, 1).T     # the upper triangle is returned
        off = np.abs(S - np.eye(w.Q0.nbloc))
        say("after Gram-Schmidt: max |S - I| off-diagonal",
            f"{(off - np.diag(np.diag(off))).max():.3e}",
            "diagonal", f"{np.abs(np.diag(S) - 1).max():.3e}")
        assert np.abs(np.diag(S) - 1.0).max() < 1e-10
        assert off.max() < 0.05
        s.restore()
        eig = s.cycle(w.Q0.eigv, w.npow, w.nals, True, 1, es)
        ms = s.phase_ms()
        rq = s.rayleigh()
    say("cycle phases ms", np.round(ms, 1).tolist(), "E0..2",
        np.round(eig[:3], 3).tolist())
    assert np.all(np.isfinite(eig)) and np.all(np.diff(eig) >= 0.0)
    assert lo <= eig.min() and eig.max() <= hi
    assert np.all(np.isfinite(rq)) and lo <= rq.min() and rq.max() <= hi
