"""Restart (psi) files: the reference's record layout (SavePsiFile /
ReadPsiFile, Restart.f90:614-793) byte for byte, its validation, and the
first-copy/second-copy fallback of ReadPsi."""
import glob
import os

import numpy as np
import pytest

from mlcp_amd.hoplist import CPBlock
from mlcp_amd.restart import read_psi, read_psi_file, save_psi, save_psi_file
from oracle.orc import Ref
from tests.helpers import random_block, random_system

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = sorted(glob.glob(os.path.join(GOLD, "ref_*_psi.rst")))


def golden_case(path):
    name = os.path.basename(path)[len("ref_"):-len("_psi.rst")]
    z = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    Q = CPBlock(z["nbas"], z["ranks"], z["coef"], z["base"])
    Q.eigv = np.arange(1, Q.nbloc + 1) * 1.25
    delta = np.linspace(-1e-3, 1e-3, Q.nbloc)
    return Q, delta


def test_golden_present():
    assert len(CASES) >= 2


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_writer_matches_reference_bytes(path, tmp_path):
    """The file the reference's SavePsiFile wrote, reproduced byte for
    byte."""
    Q, delta = golden_case(path)
    f = tmp_path / "x.rst"
    save_psi_file(str(f), 7, [0.5, 99.0], Q.eigv, delta, Q)
    assert f.read_bytes() == open(path, "rb").read()


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_reader_reads_reference_file(path):
    Q, delta = golden_case(path)
    r = read_psi_file(path, Q.nbloc, Q.nbas)
    assert r.success and r.isavi == 7
    assert np.array_equal(r.bounds, [0.5, 99.0])
    assert np.array_equal(r.eigv, Q.eigv) and np.array_equal(r.delta, delta)
    assert np.array_equal(r.block.ranks, Q.ranks)
    assert np.array_equal(r.block.coef, Q.coef)
    assert np.array_equal(r.block.base, Q.base)


def test_round_trip_ragged(tmp_path):
    Q = random_block(3, [4, 6, 3, 5], [2, 7, 1, 4, 3])
    Q.eigv = np.linspace(1, 2, Q.nbloc)
    f = str(tmp_path / "r.rst")
    save_psi_file(f, 12, [-1.0, 8.0], Q.eigv, np.full(Q.nbloc, 1e-9), Q)
    r = read_psi_file(f, Q.nbloc, Q.nbas)
    assert r.success and r.isavi == 12
    assert np.array_equal(r.block.ranks, Q.ranks)
    assert np.array_equal(r.block.base, Q.base)


def test_rejects_mismatch_and_damage(tmp_path):
    Q = random_block(4, [4, 3], [2, 3])
    f = str(tmp_path / "m.rst")
    save_psi_file(f, 1, [0, 1], [1.0, 2.0], [0.0, 0.0], Q)
    assert not read_psi_file(f, 3, Q.nbas).success        # nev differs
    assert not read_psi_file(f, 2, [4, 4]).success        # nbas differs
    assert not read_psi_file(f, 2, [4, 3, 2]).success     # ndof differs
    raw = open(f, "rb").read()
    open(f, "wb").write(raw[:-9])                          # cut short
    assert not read_psi_file(f, 2, Q.nbas).success
    assert not read_psi_file(str(tmp_path / "absent"), 2, Q.nbas).success


def test_fallback_to_second_copy_and_none(tmp_path, monkeypatch):
    Q = random_block(5, [3, 3], [2, 2])
    pre = str(tmp_path / "job")
    save_psi(pre, 4, [0, 1], [1.0, 2.0], [0.0, 0.0], Q)
    one, two = pre + "_1_psi.rst", pre + "_2_psi.rst"
    assert open(one, "rb").read() == open(two, "rb").read()
    open(one, "wb").write(b"\x05\x00")                     # damaged copy
    r = read_psi(pre, 2, Q.nbas)
    assert r.success and r.isavi == 4
    # the reference tests the first four characters of the name
    monkeypatch.chdir(tmp_path)
    save_psi("NONE", 1, [0, 1], [1.0, 2.0], [0.0, 0.0], Q)
    assert not glob.glob(str(tmp_path / "NONE*"))


@pytest.mark.skipif(not Ref.available(), reason="compiled reference not built")
def test_reference_reads_and_rejects_like_us(tmp_path):
    H = random_system(7, [4, 3, 5], ncoup=4)
    Q = random_block(8, [4, 3, 5], [3, 2, 4])
    Q.eigv = np.array([1.0, 2.0, 3.0])
    f = str(tmp_path / "ours.rst")
    save_psi_file(f, 3, [0.25, 9.0], Q.eigv, [1e-4, 2e-4, 3e-4], Q)
    data = open(f, "rb").read()
    isavi, ok, b, e, d, blk = Ref.read_psi_file(H, Q, data, Q.ranks)
    assert ok and isavi == 3 and np.array_equal(blk.base, Q.base)
    assert np.array_equal(d, [1e-4, 2e-4, 3e-4]) and np.array_equal(b, [0.25, 9.0])
    # damaged: both refuse
    _, ok2, *_ = Ref.read_psi_file(H, Q, data[:-9], Q.ranks)
    open(f, "wb").write(data[:-9])
    assert not ok2 and not read_psi_file(f, 3, Q.nbas).success
