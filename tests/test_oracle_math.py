"""The FP64 oracle against what the mathematics fixes: dense brute force on
tiny problems, fixed points, exact recoveries, scipy's eigensolver.  A
dropped term, a wrong index or a transposed operand in the oracle fails one
of these (they do not restate the oracle's formulas)."""
import numpy as np
import pytest
import scipy.linalg as sla

from mlcp_amd.hoplist import CPBlock, HopList, pack_upper
from oracle.orc import C
from tests.helpers import (full_hamiltonian, full_tensor, random_block,
                           random_system)

NB = [4, 3, 5]


def test_dspmv_matches_dense():
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 24):
        a = rng.normal(size=(n, n))
        m = a + a.T
        x = rng.normal(size=n)
        y = C.dspmv(pack_upper(m), x, 1.7)
        assert np.allclose(y, 1.7 * m @ x, rtol=1e-13, atol=1e-13)


def test_prodvv_is_full_tensor_inner_product():
    A = random_block(2, NB, [3])
    B = random_block(3, NB, [4])
    ta = full_tensor(A.nbas, A.coef, A.base)
    tb = full_tensor(B.nbas, B.coef, B.base)
    got = C.prodvv(A.nbas, A.coef, A.base, B.coef, B.base)
    assert abs(got - np.sum(ta * tb)) < 1e-12 * np.sum(np.abs(ta * tb))
    assert abs(C.prodvv(A.nbas, A.coef, A.base) - np.sum(ta * ta)) < 1e-12


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("recalc", [False, True])
def test_buildbjk_brute_force(k, recalc):
    """bjk(l, m) = < F_l^(not k) x e_m^(k), (H - E) Psi_old >  and
    BB(i, l) = prod_{d != k} <F_l^d, Fold_i^d>  (or <F_l, F_i> if recalc)."""
    H = random_system(4, NB, ncoup=4)
    Fo = random_block(5, NB, [3])
    F = random_block(6, NB, [3]) if recalc else Fo
    E = 2.5
    BB, bjk = C.buildbjk(H, Fo.coef, Fo.base, F.coef, F.base, k, 1, E, recalc)
    Hd = full_hamiltonian(H)
    psi = full_tensor(Fo.nbas, Fo.coef, Fo.base).ravel()
    phi = ((Hd - E * np.eye(Hd.shape[0])) @ psi).reshape(NB)
    off = np.concatenate([[0], np.cumsum(NB)])
    R = F.coef.size
    for l in range(R):
        t = phi
        for d in reversed(range(3)):
            if d == k:
                continue
            t = np.tensordot(t, F.base[l, off[d]:off[d + 1]], axes=([d], [0]))
        assert np.allclose(bjk[l], t.ravel(), rtol=1e-11, atol=1e-12)
    other = F if recalc else Fo
    for i in range(R):
        for l in range(R):
            p = 1.0
            for d in range(3):
                if d != k:
                    p *= F.base[l, off[d]:off[d + 1]] @ other.base[i, off[d]:off[d + 1]]
            assert abs(BB[i, l] - p) < 1e-13


def separable_system():
    """Only per-mode diagonal terms: product states are exact eigenstates."""
    rng = np.random.default_rng(7)
    terms, coefs = [], []
    for d, n in enumerate(NB):
        ops = [None] * 3
        ops[d] = np.diag(np.sort(rng.uniform(0, 3, n)) + 3 * d)
        terms.append(ops)
        coefs.append(1.0)
    return HopList.from_terms(NB, terms, coefs)


def test_als_pow3_fixed_point_of_separable_h():
    H = separable_system()
    nD = sum(NB)
    base = np.zeros((1, nD))
    off = np.concatenate([[0], np.cumsum(NB)])
    for d in range(3):
        base[0, off[d]] = 1.0       # lowest product state
    c, b = C.als_pow3(H, np.array([1.0]), base, 4, 1, 20.0)
    t = full_tensor(NB, c, b)
    t0 = full_tensor(NB, np.array([1.0]), base)
    assert np.linalg.norm(np.abs(t) - t0) < 1e-12


def test_prodhv_als3_exact_for_one_term():
    """H = h (x) I (x) I keeps the rank: the reduced H Psi is exact up to the
    1e-10 relative penalty."""
    rng = np.random.default_rng(8)
    a = rng.normal(size=(NB[0], NB[0]))
    H = HopList.from_terms(NB, [[a + a.T, None, None]], [1.3])
    F = random_block(9, NB, [3])
    c, b = C.prodhv_als3(H, F.coef, F.base, 2)
    want = (full_hamiltonian(H) @ full_tensor(NB, F.coef, F.base).ravel())
    got = full_tensor(NB, c, b).ravel()
    assert np.linalg.norm(got - want) < 1e-8 * np.linalg.norm(want)


def test_ortho3_orthonormalises():
    """With enough rank the ALS Gram-Schmidt block is orthonormal."""
    Q = random_block(10, [3, 3, 3], [9, 9, 9])
    o = C.als_ortho3(Q, 20)
    S = np.array([[C.prodvv(o.nbas, *o.state(i), *o.state(j))
                   for j in range(3)] for i in range(3)])
    assert np.max(np.abs(S - np.eye(3))) < 1e-6


def test_gep_matches_scipy():
    H = random_system(11, NB, ncoup=4)
    Q = C.als_ortho3(random_block(12, NB, [6, 6, 6, 6]), 8)
    M = C.getqhq3(H, Q, 4)
    S = C.getoverlaps(Q)
    Ms = np.triu(M) + np.triu(M, 1).T
    Ss = np.triu(S) + np.triu(S, 1).T
    w, X = C.dsygv(M, S)
    ws = sla.eigh(Ms, Ss, eigvals_only=True)
    assert np.allclose(w, ws, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(Ms @ X - Ss @ X * w)) < 1e-11
    assert np.max(np.abs(X.T @ Ss @ X - np.eye(4))) < 1e-11


def test_getoverlaps_upper_triangle():
    Q = random_block(13, NB, [2, 3, 4])
    S = C.getoverlaps(Q)
    T = [full_tensor(NB, *Q.state(i)) for i in range(3)]
    for i in range(3):
        for j in range(3):
            want = np.sum(T[i] * T[j]) if j >= i else 0.0
            assert abs(S[i, j] - want) < 1e-12


def test_sumlcvec3_recovers_single_vector():
    Q = random_block(14, NB, [3, 3, 3])
    lc = np.array([0.0, 1.0, 0.0])
    c, b = C.sumlcvec3(*Q.state(1), Q, lc, 3)
    t = full_tensor(NB, c, b)
    want = full_tensor(NB, *Q.state(1))
    want /= np.linalg.norm(want)
    assert np.linalg.norm(t - want) < 1e-8


def test_reduc_als_fixed_point():
    G = random_block(15, NB, [3])
    c, b = C.reduc_als(G.nbas, G.coef, G.base, G.coef, G.base, 3)
    assert np.linalg.norm(full_tensor(NB, c, b) - full_tensor(NB, G.coef, G.base)) \
        < 1e-10


def test_rayleigh_full_tensor():
    H = random_system(16, NB, ncoup=4)
    Q = random_block(17, NB, [2, 3])
    rq = C.rayleigh(H, Q)
    Hd = full_hamiltonian(H)
    for s in range(2):
        v = full_tensor(NB, *Q.state(s)).ravel()
        assert abs(rq[s] - v @ Hd @ v / (v @ v)) < 1e-11 * abs(rq[s])


def test_harmonic_limit_keeps_exact_eigenstates():
    """lambda = 0 Henon-Heiles: the product guesses are exact eigenstates,
    so a full cycle returns E = sum (n_i + 1/2) exactly (3, 4, 4, ...)."""
    from mlcp_amd import systems as S
    pes = S.hh111_pes(0.0)
    basis, groups = S.build_tree(pes, 6, [[2, 2, 2]], [10])
    H = S.top_layer_hoplist(basis, groups)
    Q = S.guess_block([g.energies for g in groups], [4] * 8, 5)
    out = C.intertwine(H, Q, 3, 4, True, 1, 0.5 * (2 * Q.eigv[-1] - Q.eigv[-2] + 12))
    assert np.allclose(np.sort(out.eigv), [3, 4, 4, 4, 4, 4, 4, 5], atol=1e-8)


@pytest.mark.parametrize("imode", [0, 1, 2])
def test_update_opermat_full_tensor(imode):
    """QXQ(i,j) = psi_j . (I x .. x X x .. x I) psi_i for j >= i, packed
    upper by columns (i + j(j+1)/2); for a symmetric X the packed matrix is
    the symmetric Q^T X Q."""
    rng = np.random.default_rng(20 + imode)
    n = NB[imode]
    a = rng.normal(size=(n, n))
    X = a + a.T
    Q = random_block(21, NB, [2, 3, 2, 4])
    got = C.update_opermat(Q, imode, pack_upper(X))
    ops = [np.eye(m) for m in NB]
    ops[imode] = X
    big = ops[0]
    for m in ops[1:]:
        big = np.kron(big, m)
    psi = [full_tensor(NB, *Q.state(s)).ravel() for s in range(Q.nbloc)]
    for i in range(Q.nbloc):
        for j in range(i, Q.nbloc):
            want = psi[j] @ big @ psi[i]
            assert abs(got[i + j * (j + 1) // 2] - want) < 1e-12 * max(1, abs(want))
