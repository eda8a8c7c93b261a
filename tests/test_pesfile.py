"""Force-field files and the term-list wire format against the reference's
own readers/builders: ReadFFHamiltonian (FFPES.f90:308-406) and GetHopList
(ModeH.f90:909-1068), through committed outputs of the compiled reference
(tests/golden/ffhl_ref.npz) and, where it is built, the reference itself."""
import os

import numpy as np
import pytest

from mlcp_amd import systems
from mlcp_amd.hoplist import hoplist_from_hamiltonian
from mlcp_amd.pesfile import (distrib_mode_power, pes_monomials,
                              read_ff_hamiltonian, write_pes_files)
from oracle.orc import Ref
from tests.golden.make_golden import FF_ODD, hoplist_case

Z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ffhl_ref.npz"))


def _tree(tmp_path, files):
    for rel, text in files.items():
        p = tmp_path / rel
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(text)


def test_generated_files_parse_like_the_reference(tmp_path):
    files = {f"pes/f{k}qff05.dat": str(Z[f"ff_text_{k}"]) for k in range(1, 5)
             if str(Z[f"ff_text_{k}"])}
    _tree(tmp_path, files)
    for divide in (0, 1):
        got = read_ff_hamiltonian(str(tmp_path), "qff05", 4, bool(divide))
        for k, c in enumerate(got, 1):
            assert np.array_equal(c.qns, Z[f"ff_qns_{divide}_{k}"])
            assert np.array_equal(c.coef, Z[f"ff_coef_{divide}_{k}"])


def test_list_directed_edge_cases_like_the_reference(tmp_path):
    """A value list continued on the next line, comma separators, a D
    exponent, and reading stopping at an unreadable line."""
    _tree(tmp_path, {"pes/f3odd05.dat": FF_ODD})
    got = read_ff_hamiltonian(str(tmp_path), "odd05", 3, True)
    for k, c in enumerate(got, 1):
        assert np.array_equal(c.qns, Z[f"odd_qns_{k}"])
        assert np.array_equal(c.coef, Z[f"odd_coef_{k}"])
    assert got[2].coef.size == 4 and got[0].coef.size == 1   # zero config


def test_writer_round_trip_exact(tmp_path):
    pes = systems.qff_pes(6, 7)
    write_pes_files(pes, "abcde", str(tmp_path))
    got = read_ff_hamiltonian(str(tmp_path), "abcde", 4, False)
    want = {}
    for q, c in pes_monomials(pes):
        want.setdefault(len(q), []).append((list(q), c))
    for k, rows in want.items():
        assert got[k - 1].qns.tolist() == [q for q, _ in rows]
        assert np.array_equal(got[k - 1].coef, [c for _, c in rows])


def test_distrib_mode_power():
    assert distrib_mode_power([1, 1, 7, 8]) == [[1, 2], [7, 1], [8, 1]]
    assert distrib_mode_power([3, 2, 3, 3]) == [[3, 3], [2, 1]]


def test_hoplist_from_hamiltonian_like_the_reference():
    nbas, opcoef, opindx, iops, eops, cops, pops = hoplist_case()
    hl = hoplist_from_hamiltonian(nbas, opcoef, opindx, iops, eops,
                                  [x for _, x in cops], [x for _, x in pops])
    assert np.array_equal(hl.opid, Z["hl_opid"])
    assert np.array_equal(hl.mptr, Z["hl_mptr"])
    assert np.array_equal(hl.mat, Z["hl_mat"])
    assert np.array_equal(hl.coef, Z["hl_coef"])
    hl.validate()


def test_hoplist_numbering_by_hand():
    """D = 2: identities 1,2; eigen 3,4; no compound; primitives of mode 1
    in first-use order (pops 2 then 1) as 5,6, of mode 2 (pops 3) as 7."""
    one = np.array([1.0])
    tri = np.array([1.0, 0.0, 1.0])
    hl = hoplist_from_hamiltonian(
        [1, 2], [1.0, 2.0, 3.0], [[2, 0], [-1, 3], [1, -2]],
        [one, tri], [one, tri], [], [one * 5, one * 6, tri * 7])
    assert hl.opid.tolist() == [5, 2, 3, 7, 6, 4]
    assert hl.mptr.tolist() == [1, 2, 5, 6, 9, 10, 11]
    assert hl.mat[8] == 6.0 and hl.mat[9] == 5.0


@pytest.mark.skipif(not Ref.available(), reason="compiled reference not built")
def test_reference_reads_our_force_field(tmp_path):
    pes = systems.qff_pes(4, 11)
    files = write_pes_files(pes, "zz004", str(tmp_path))
    tree = {os.path.relpath(f, tmp_path): open(f).read() for f in files}
    ref = Ref.read_ff(tree, "zz004", 4, False)
    got = read_ff_hamiltonian(str(tmp_path), "zz004", 4, False)
    for (q, c), g in zip(ref, got):
        assert np.array_equal(q, g.qns) and np.array_equal(c, g.coef)
