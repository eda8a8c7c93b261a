"""Term-list (HopList) bookkeeping: bit-exact round trips and validation."""
import numpy as np
import pytest

from mlcp_amd import systems
from mlcp_amd.hoplist import (OP_DENSE, OP_DIAGONAL, OP_IDENTITY, HopList,
                              pack_upper, unpack_upper)


def test_pack_roundtrip_bitexact():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 40):
        a = rng.normal(size=(n, n))
        m = a + a.T
        ap = pack_upper(m)
        assert ap.size == n * (n + 1) // 2
        assert np.array_equal(pack_upper(unpack_upper(ap, n)), ap)
        assert np.array_equal(unpack_upper(ap, n), m)


def test_packed_index_convention():
    """element (i,j), i <= j, at i + j(j+1)/2 (Utils.f90:559-597)."""
    m = np.arange(16.0).reshape(4, 4)
    m = m + m.T
    ap = pack_upper(m)
    for j in range(4):
        for i in range(j + 1):
            assert ap[i + j * (j + 1) // 2] == m[i, j]


def test_from_terms_layout_and_dedup():
    n = [3, 4]
    q0 = np.diag([1.0, 2.0, 3.0])
    a = np.ones((4, 4))
    H = HopList.from_terms(n, [[q0, None], [None, a], [q0, a]], [1.0, 2.0, 3.0])
    # identities first (opid 1..ndof), then distinct matrices in first use
    assert list(H.opid) == [3, 2, 1, 4, 3, 4]
    assert list(H.mptr[:2]) == [1, 1 + 6]
    assert H.nops == 4 and H.msize == 6 + 10 + 6 + 10
    assert list(H.classify()) == [OP_IDENTITY, OP_IDENTITY, OP_DIAGONAL,
                                  OP_DENSE]
    H.validate()
    for t in range(3):
        for d in range(2):
            op = H.opid[t * 2 + d] - 1
            m = H.dense_op(op, n[d])
            want = [[q0, None], [None, a], [q0, a]][t][d]
            want = np.eye(n[d]) if want is None else want
            assert np.array_equal(m, want)


@pytest.mark.parametrize("bad", ["range", "twomodes", "mat"])
def test_validate_rejects(bad):
    H = HopList.from_terms([3, 3], [[np.eye(3) * 2, None]], [1.0])
    opid, mptr, mat = H.opid.copy(), H.mptr.copy(), H.mat.copy()
    if bad == "range":
        opid[0] = 99
    elif bad == "twomodes":
        opid[1] = opid[0]
    else:
        mptr[-1] = mat.size
    with pytest.raises(ValueError):
        HopList(H.nbas, opid, mptr, H.coef, mat).validate()


def test_hh111_term_list_shape():
    w = systems.hh111()
    assert w.H.T == 5 and w.H.ndof == 3
    assert list(w.H.nbas) == [24, 24, 24]
    assert w.H.nops == 10 and w.H.msize == 10 * 300
    w.H.validate()
    assert w.Q0.nbloc == 40 and set(w.Q0.ranks) == {20}


def test_qff_generator_is_seeded():
    a = systems.qff_pes(6, 3)
    b = systems.qff_pes(6, 3)
    assert np.array_equal(a.omega, b.omega)
    assert a.terms == b.terms
    assert all(300 <= w <= 3500 for w in a.omega)
