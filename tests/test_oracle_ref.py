"""The C oracle against the compiled reference routines (oracle/_ref, built
from the reference's Fortran sources by oracle/ref/Makefile).

Where the oracle keeps the reference's loop order the agreement is bitwise;
the block eigensolver (DSYGV there, Cholesky + Jacobi here) agrees to
rounding and up to the sign of each eigenvector."""
import numpy as np
import pytest

from oracle.orc import C, Ref
from tests.helpers import full_tensor, random_block, random_system

pytestmark = pytest.mark.skipif(not Ref.available(),
                                reason="compiled reference not built")

SYSTEMS = [
    ("3mode", [5, 6, 4], [4, 4, 4], 5),
    ("ragged", [5, 6, 4], [3, 5, 2, 4], 6),
    ("5mode", [3, 4, 3, 5, 4], [3, 4, 3], 8),
]


def _pow_block(H, Q, npow, eshift):
    o = Q.copy()
    off = Q.offsets
    for s in range(Q.nbloc):
        c, b = C.als_pow3(H, *Q.state(s), npow, 1, eshift)
        o.coef[off[s]:off[s + 1]] = c
        o.base[off[s]:off[s + 1]] = b
    return o


@pytest.mark.parametrize("name,nb,ranks,nc", SYSTEMS)
def test_als_pow3_bitwise(name, nb, ranks, nc):
    H = random_system(101, nb, ncoup=nc)
    Q = random_block(102, nb, ranks)
    o = _pow_block(H, Q, 3, 3.5)
    r = Ref.als_pow3(H, Q, 3, 1, 3.5)
    assert np.array_equal(o.coef, r.coef)
    assert np.array_equal(o.base, r.base)


@pytest.mark.parametrize("name,nb,ranks,nc", SYSTEMS)
def test_prodhv_als3_bitwise(name, nb, ranks, nc):
    H = random_system(111, nb, ncoup=nc)
    Q = random_block(112, nb, ranks)
    r = Ref.prodhv_als3(H, Q, 4)
    off = Q.offsets
    for s in range(Q.nbloc):
        c, b = C.prodhv_als3(H, *Q.state(s), 4)
        assert np.array_equal(c, r.coef[off[s]:off[s + 1]])
        assert np.array_equal(b, r.base[off[s]:off[s + 1]])


@pytest.mark.parametrize("name,nb,ranks,nc", SYSTEMS)
def test_als_ortho3_bitwise(name, nb, ranks, nc):
    H = random_system(121, nb, ncoup=nc)
    Q = random_block(122, nb, ranks)
    o = C.als_ortho3(Q, 5)
    r = Ref.als_ortho3(H, Q, 5)
    assert np.array_equal(o.coef, r.coef)
    assert np.array_equal(o.base, r.base)


@pytest.mark.parametrize("name,nb,ranks,nc", SYSTEMS)
def test_qhq_overlaps_bitwise(name, nb, ranks, nc):
    H = random_system(131, nb, ncoup=nc)
    Q = random_block(132, nb, ranks)
    M, S = Ref.getqhq_overlaps(H, Q, 3)
    assert np.array_equal(C.getqhq3(H, Q, 3), M)
    assert np.array_equal(C.getoverlaps(Q), S)


@pytest.mark.parametrize("name,nb,ranks,nc", SYSTEMS)
def test_updatevecs2_bitwise(name, nb, ranks, nc):
    rng = np.random.default_rng(141)
    H = random_system(141, nb, ncoup=nc)
    Q = random_block(142, nb, ranks)
    U = rng.normal(size=(Q.nbloc, Q.nbloc))
    r = Ref.updatevecs2(H, Q, U, 3)
    off = Q.offsets
    for s in range(Q.nbloc):
        c, b = C.sumlcvec3(*Q.state(s), Q, U[:, s], 3)
        assert np.array_equal(c, r.coef[off[s]:off[s + 1]])
        assert np.array_equal(b, r.base[off[s]:off[s + 1]])


@pytest.mark.parametrize("nb", [[5, 6, 4], [3, 4, 3, 5]])
def test_reduc_als_bitwise(nb):
    """ndim <= 3 rebuilds the normal equations, ndim > 3 up/downdates them
    by division (REDUCTIONML.f90:618-667); both branches are restated."""
    H = random_system(151, nb, ncoup=3)
    F = random_block(152, nb, [3])
    G = random_block(153, nb, [7])
    Q = random_block(154, nb, [3, 7])
    Q.coef[:3], Q.base[:3] = F.coef, F.base
    Q.coef[3:], Q.base[3:] = G.coef, G.base
    rc, rb = Ref.reduc_als(H, Q, 4)
    c, b = C.reduc_als(np.asarray(nb, dtype=np.int32), G.coef, G.base,
                       F.coef, F.base, 4)
    assert np.array_equal(c, rc)
    assert np.array_equal(b, rb)


@pytest.mark.parametrize("name,nb,ranks,nc", SYSTEMS)
def test_rayleigh_bitwise(name, nb, ranks, nc):
    H = random_system(161, nb, ncoup=nc)
    Q = random_block(162, nb, ranks)
    assert np.array_equal(C.rayleigh(H, Q), Ref.rayleigh(H, Q))


@pytest.mark.parametrize("n", [2, 5, 17])
def test_dsygv_vs_reference(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n))
    A = a + a.T
    b = rng.normal(size=(n, n))
    B = np.eye(n) * n + b @ b.T
    H = random_system(1, [2, 2], ncoup=1)
    Q = random_block(2, [2, 2], [1] * n)
    wr, Xr = Ref.dsygv(H, Q, A, B)
    w, X = C.dsygv(A, B)
    assert np.max(np.abs(w - wr)) < 1e-12 * np.abs(wr).max()
    sgn = np.sign(np.sum(X * Xr, axis=0))
    assert np.max(np.abs(X - Xr * sgn)) < 1e-10


@pytest.mark.parametrize("name,nb,ranks,nc", SYSTEMS[:2])
def test_full_cycle_vs_reference(name, nb, ranks, nc):
    """One Iterate cycle (styp=-1, lowmem=3, update): eigenvalues to rounding,
    each new vector equal up to the eigenvector sign."""
    H = random_system(171, nb, ncoup=nc)
    Q = random_block(172, nb, ranks)
    r = Ref.cycle(H, Q, 3, 4, 1, 5.0)
    o = C.intertwine(H, Q, 3, 4, True, 1, 5.0)
    assert np.max(np.abs(o.eigv - r.eigv)) < 1e-12 * np.abs(r.eigv).max()
    for s in range(Q.nbloc):
        a = full_tensor(Q.nbas, *o.state(s))
        b = full_tensor(Q.nbas, *r.state(s))
        sg = np.sign(np.sum(a * b))
        assert np.linalg.norm(a - sg * b) < 1e-9 * np.linalg.norm(b)


def test_cycle_without_update_bitwise():
    nb = [5, 6, 4]
    H = random_system(181, nb, ncoup=5)
    Q = random_block(182, nb, [4, 4, 4])
    r = Ref.cycle(H, Q, 2, 3, 0, 4.0)
    o = C.intertwine(H, Q, 2, 3, False, 1, 4.0)
    assert np.array_equal(o.coef, r.coef)
    assert np.array_equal(o.base, r.base)


def _op_packed(H, o):
    d = int(H.op_mode()[o])
    nq = int(H.nbas[d]) * (int(H.nbas[d]) + 1) // 2
    return d, H.mat[H.mptr[o] - 1:H.mptr[o] - 1 + nq]


@pytest.mark.parametrize("name,nb,ranks,nc", SYSTEMS)
def test_update_opermat_bitwise(name, nb, ranks, nc):
    """UpdateOperMat (Updater.f90:109-153) of every operator on its mode."""
    H = random_system(111, nb, ncoup=nc)
    Q = random_block(112, nb, ranks)
    ref = Ref.update_opermat(H, Q)
    for o in range(H.nops):
        d, x = _op_packed(H, o)
        assert np.array_equal(C.update_opermat(Q, d, x), ref[o]), o
