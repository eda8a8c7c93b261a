"""The error channel on the GPU, through the C-ABI.

The reference never reads cuSOLVER's devInfo (linalg.cu:92-107) and has no
error channel to its Fortran host.  Here every kernel checks what it can
(a non-positive pivot of the regularised normal equations: bit 2; a NaN or
zero norm: bit 1; a non positive definite overlap matrix: bit 4; a
non-finite matrix or no convergence in the block eigensolver: bit 8), exits
normally, and the host call reports the bits through mlcp_hip_last_error()
and leaves the caller's arrays as they were.

Also: intwdevicemem_'s figure against the memory a session really holds.
"""
import re
import sys

import numpy as np
import pytest

from tests.helpers import random_block, random_system

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _call_intertwine(L, Q, H, npow=2, nals=3):
    base, coef, eig = Q.base.copy(), Q.coef.copy(), Q.eigv.copy()
    rc = L.mlcp_hip_intertwine(base, coef, Q.nbas, H.ndof, Q.nbloc, Q.ranks,
                               H.opid, H.mptr, H.coef, H.mat, H.T, H.nops,
                               H.msize, eig, npow, nals, 1, 1, 5.0)
    return rc, base, coef, eig


def test_nan_coefficient_is_reported_and_inputs_kept(hip):
    L = hip
    say("NaN coefficient")
    nb = [5, 6, 4]
    H = random_system(301, nb, ncoup=5)
    Q = random_block(302, nb, [4, 3, 5])
    Q.eigv[:] = [1.0, 2.0, 3.0]
    Q.coef[5] = np.nan
    rc, base, coef, eig = _call_intertwine(L, Q, H)
    msg = L.mlcp_hip_last_error().decode()
    assert rc == -1 and msg, (rc, msg)
    assert ("NaN" in msg) or ("pivot" in msg) or ("eigensolver" in msg), msg
    assert np.array_equal(base, Q.base)
    assert np.array_equal(coef[np.isfinite(Q.coef)], Q.coef[np.isfinite(Q.coef)])
    assert np.isnan(coef[5])
    assert np.array_equal(eig, Q.eigv)
    # the library is usable afterwards
    Q.coef[5] = 1.0
    rc, *_ = _call_intertwine(L, Q, H)
    assert rc == 0, L.mlcp_hip_last_error().decode()


def test_rank_deficient_block_is_regularised_like_the_oracle(hip):
    """A state whose terms are duplicated (rank-deficient CP
    representation): the normal equations prod_{d!=k} G_d are exactly
    singular and only the penalty 1e-10 max c (als.cu:248-306, ALSPow.f90)
    makes them solvable, on the GPU as in the oracle.  No error bit, no
    NaN, ascending eigenvalues.  The split of the solution among the
    duplicated terms is set by rounding at condition ~1e10 (LU in the
    oracle, Gauss-Jordan in the L D L^T association here), so the two sides
    agree only loosely after one cycle: 9.5e-2 measured, against
    eigenvalues of order 5 (held to 5% of the largest)."""
    from mlcp_amd.intertwine import intertwine_ragged
    from oracle.orc import C
    say("duplicated terms")
    nb = [5, 6, 4]
    H = random_system(303, nb, ncoup=5)
    Q = random_block(304, nb, [4, 4, 4])
    o = Q.offsets
    for s in range(Q.nbloc):
        Q.base[o[s] + 2] = Q.base[o[s]]
        Q.base[o[s] + 3] = Q.base[o[s] + 1]
        Q.coef[o[s] + 2] = Q.coef[o[s]]
        Q.coef[o[s] + 3] = Q.coef[o[s] + 1]
    Q.eigv[:] = [1.0, 2.0, 3.0]
    g = intertwine_ragged(Q, H, 2, 3, True, 1, 5.0)
    r = C.intertwine(H, Q, 2, 3, True, 1, 5.0)
    assert np.all(np.isfinite(g.eigv)) and np.all(np.isfinite(g.base))
    assert np.all(np.diff(g.eigv) >= 0.0)
    err = np.max(np.abs(g.eigv - r.eigv))
    say("duplicated terms: max |dE| vs oracle", f"{err:.2e}")
    assert err < 5e-2 * np.abs(r.eigv).max()


def test_eigensolver_nan_matrix(hip):
    """A NaN in QHQ with a positive definite S: the QL iteration cannot
    converge; the eigensolver reports bit 8 instead of returning NaNs."""
    L = hip
    say("gep NaN")
    n = 12
    rng = np.random.default_rng(5)
    a = rng.normal(size=(n, n))
    A = a + a.T
    A[2, 7] = A[7, 2] = np.nan
    B = np.eye(n)
    Ac = np.array(A.T, order="C")
    Bc = np.array(B.T, order="C")
    X = np.full((n, n), 3.0)
    w = np.full(n, -7.0)
    rc = L.mlcp_hip_gep(n, Ac, Bc, X, w)
    assert rc == -1 and "non-finite" in L.mlcp_hip_last_error().decode()
    assert np.all(X == 3.0) and np.all(w == -7.0)


def test_intwdevicemem_bounds_the_real_allocation(hip, capfd):
    """intwdevicemem_ (intertwining.cu:31-94) prints what a call needs;
    its total must cover what a session really holds after a cycle with
    update, and not overstate it by more than the slot bound allows."""
    import ctypes as C
    from mlcp_amd.intertwine import Session, gpu_memory
    from mlcp_amd import systems
    say("intwdevicemem_")
    w = systems.hh111()
    B0 = hip.mlcp_hip_device_bytes()
    with Session(w.Q0, w.H) as s:
        s.cycle(w.Q0.eigv, 1, 1, True, 1, 30.0)
        used = hip.mlcp_hip_device_bytes() - B0
    capfd.readouterr()
    gpu_memory(w.H, int(w.Q0.ranks[0]), w.Q0.nbloc, True)
    out = capfd.readouterr().out
    m = re.search(r"TOTAL\s*:\s*([0-9.]+) MB", out)
    assert m, out
    est = float(m.group(1)) * (1 << 20)
    say("intwdevicemem_ estimate", est, "held", used)
    assert used > 0
    assert used <= est * 1.001 and est <= 2.0 * used, (est, used)
    assert hip.mlcp_hip_device_bytes() == B0
