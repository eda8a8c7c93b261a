"""The C oracle reproduces the committed outputs of the compiled reference
routines (tests/golden/ref_*.npz, written by tests/golden/make_golden.py
from oracle/_ref) without the reference present."""
import glob
import os

import numpy as np
import pytest

from mlcp_amd.hoplist import CPBlock, HopList
from oracle.orc import C

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden",
                                    "ref_*.npz")))


def load(path):
    z = np.load(path)
    H = HopList(z["nbas"], z["opid"], z["mptr"], z["hcoef"], z["mat"])
    Q = CPBlock(z["nbas"], z["ranks"], z["coef"], z["base"])
    return z, H, Q


@pytest.fixture(params=FIX, ids=[os.path.basename(f) for f in FIX])
def fx(request):
    return load(request.param)


def test_fixtures_present():
    assert len(FIX) >= 2


def test_pow3(fx):
    z, H, Q = fx
    o = Q.offsets
    for s in range(Q.nbloc):
        c, b = C.als_pow3(H, *Q.state(s), 3, 1, 3.5)
        assert np.array_equal(c, z["pow_coef"][o[s]:o[s + 1]])
        assert np.array_equal(b, z["pow_base"][o[s]:o[s + 1]])


def test_prodhv(fx):
    z, H, Q = fx
    o = Q.offsets
    for s in range(Q.nbloc):
        c, b = C.prodhv_als3(H, *Q.state(s), 3)
        assert np.array_equal(c, z["hq_coef"][o[s]:o[s + 1]])
        assert np.array_equal(b, z["hq_base"][o[s]:o[s + 1]])


def test_ortho(fx):
    z, H, Q = fx
    g = C.als_ortho3(Q, 4)
    assert np.array_equal(g.coef, z["gs_coef"])
    assert np.array_equal(g.base, z["gs_base"])


def test_qhq_overlaps(fx):
    z, H, Q = fx
    assert np.array_equal(C.getqhq3(H, Q, 3), z["qhq"])
    assert np.array_equal(C.getoverlaps(Q), z["s"])


def test_updates(fx):
    z, H, Q = fx
    o = Q.offsets
    for s in range(Q.nbloc):
        c, b = C.sumlcvec3(*Q.state(s), Q, z["U"][:, s], 3)
        assert np.array_equal(c, z["up_coef"][o[s]:o[s + 1]])
        assert np.array_equal(b, z["up_base"][o[s]:o[s + 1]])


def test_rayleigh(fx):
    z, H, Q = fx
    assert np.array_equal(C.rayleigh(H, Q), z["rq"])


def test_reduc_als(fx):
    z, H, Q = fx
    c, b = C.reduc_als(H.nbas, z["G_coef"], z["G_base"], z["F_coef"],
                       z["F_base"], 4)
    assert np.array_equal(c, z["red_coef"])
    assert np.array_equal(b, z["red_base"])


def test_cycle_eigenvalues(fx):
    """DSYGV is replaced by Cholesky + Jacobi in the oracle: eigenvalues to
    rounding, vectors up to sign."""
    z, H, Q = fx
    o = C.intertwine(H, Q, 3, 4, True, 1, 5.0)
    assert np.max(np.abs(o.eigv - z["cyc_eigv"])) < 1e-12 * np.abs(z["cyc_eigv"]).max()


def test_update_opermat(fx):
    z, H, Q = fx
    modes = H.op_mode()
    for o in range(H.nops):
        d = int(modes[o])
        nq = int(H.nbas[d]) * (int(H.nbas[d]) + 1) // 2
        x = H.mat[H.mptr[o] - 1:H.mptr[o] - 1 + nq]
        assert np.array_equal(C.update_opermat(Q, d, x), z["qxq"][o]), o
