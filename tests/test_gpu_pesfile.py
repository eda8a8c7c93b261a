"""Term lists from the reference's own inputs through the GPU path.

* A force field in the reference's file format (written by
  ``write_pes_files`` and read back by the restatement of
  ReadFFHamiltonian, FFPES.f90:308-406) becomes a primitive-basis term list
  (kinetic part plus one term per force constant, q^k in the
  harmonic-oscillator basis) equal, bit for bit, to the one built from the field in memory; one
  drop-in cycle (``intertwinecpp_``) on it matches the FP64 oracle.
* The term list of GetHopList's numbering (ModeH.f90:909-1068; the CPU test
  ``test_hoplist_from_hamiltonian_like_the_reference`` pins it to the
  compiled reference's output) runs the same cycle.

Tolerances as in test_gpu_parity: eigenvalues 1e-9 relative, represented
tensors 1e-9 relative.
"""
import sys

import numpy as np
import pytest

from mlcp_amd.hoplist import HopList
from tests.helpers import full_tensor, random_block

pytestmark = pytest.mark.gpu

TOL = 1e-9


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _primitive_hoplist(nmodes, omega, monomials, nprim):
    """h = sum_d w_d/2 p_d^2 + sum c prod_d q_d^k: the kinetic part and one
    term per force constant (the files hold the harmonic potential as
    q_i q_i w_i/2), q^k exact in the first nprim oscillator functions."""
    from mlcp_amd.systems import ho_p2, ho_q_powers
    qp, p2 = ho_q_powers(nprim, 4), ho_p2(nprim)
    terms, coefs = [], []
    for d in range(nmodes):
        ops = [None] * nmodes
        ops[d] = 0.5 * omega[d] * p2
        terms.append(ops)
        coefs.append(1.0)
    for pw, c in monomials:
        ops = [None] * nmodes
        for m, k in pw:
            ops[m] = qp[k]
        terms.append(ops)
        coefs.append(float(c))
    return HopList.from_terms([nprim] * nmodes, terms, coefs)


def _monomials_from_files(configs):
    """(mode, power) tuples in order of first appearance (DistribModePower)
    with their constants, from ReadFFHamiltonian's configurations (modes
    1-based); all-zero padding configurations are dropped."""
    from mlcp_amd.pesfile import distrib_mode_power
    out = []
    for cfg in configs:
        for q, c in zip(cfg.qns, cfg.coef):
            if not np.any(q):
                continue
            out.append((tuple((m - 1, k) for m, k in
                              distrib_mode_power(q.tolist())), c))
    return out


def _cycle_against_oracle(H, Q, eshift):
    from mlcp_amd.intertwine import gpu_intertwine
    from oracle.orc import C
    g = gpu_intertwine(Q, H, 2, 3, True, 1, eshift, noise=0.0)
    o = C.intertwine(H, Q, 2, 3, True, 1, eshift)
    dw = np.max(np.abs(g.eigv - o.eigv)) / np.abs(o.eigv).max()
    dt = max(_rel(full_tensor(g.nbas, *g.state(s)),
                  full_tensor(o.nbas, *o.state(s))) for s in range(Q.nbloc))
    return dw, dt


def test_force_field_files_through_the_dropin(hip, tmp_path):
    from mlcp_amd import systems
    from mlcp_amd.pesfile import (distrib_mode_power, pes_monomials,
                                  read_ff_hamiltonian, write_pes_files)
    nmodes, nprim = 4, 6
    pes = systems.qff_pes(nmodes, 31)
    write_pes_files(pes, "gp004", str(tmp_path))
    configs = read_ff_hamiltonian(str(tmp_path), "gp004", 4, False)
    from_files = _monomials_from_files(configs)
    in_memory = [(tuple((m - 1, k) for m, k in distrib_mode_power(list(q))),
                  c) for q, c in pes_monomials(pes)]
    H = _primitive_hoplist(nmodes, pes.omega, from_files, nprim)
    H2 = _primitive_hoplist(nmodes, pes.omega, in_memory, nprim)
    assert np.array_equal(H.opid, H2.opid) and np.array_equal(H.mptr, H2.mptr)
    assert np.array_equal(H.mat, H2.mat) and np.array_equal(H.coef, H2.coef)
    Q = random_block(33, [nprim] * nmodes, [4, 4, 4, 4])
    Q.eigv[:] = np.sort(np.random.default_rng(34).uniform(2000, 6000, 4))
    eshift = systems.spectral_upper_bound(H)
    dw, dt = _cycle_against_oracle(H, Q, eshift)
    say("force field", H.T, "terms, shift", f"{eshift:.1f}",
        "eigenvalues", f"{dw:.2e}", "tensors", f"{dt:.2e}")
    assert dw < TOL and dt < TOL


def test_reference_numbered_hoplist_through_the_dropin(hip):
    from mlcp_amd.hoplist import hoplist_from_hamiltonian
    from tests.golden.make_golden import hoplist_case
    nbas, opcoef, opindx, iops, eops, cops, pops = hoplist_case()
    H = hoplist_from_hamiltonian(nbas, opcoef, opindx, iops, eops,
                                 [x for _, x in cops], [x for _, x in pops])
    H.validate()
    Q = random_block(35, nbas, [3, 3, 3])
    Q.eigv[:] = [1.0, 2.0, 3.0]
    dw, dt = _cycle_against_oracle(H, Q, 30.0)
    say("GetHopList numbering", H.T, "terms, eigenvalues", f"{dw:.2e}",
        "tensors", f"{dt:.2e}")
    assert dw < TOL and dt < TOL
