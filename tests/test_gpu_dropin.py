"""The drop-in boundary end to end: a Fortran host (oracle/_ref/host_driver,
compiled with amdflang against the reference's modules) packs the block as
GPU_Intertwine does (GPUIterate.f90:130-180) and calls IntertwineCPP as an
implicit external with the reference's argument list (:162-167), which links
to this library's intertwinecpp_.  The result is checked against the FP64
oracle run on the same packed inputs (1e-14 added to every coefficient)."""
import sys

import numpy as np
import pytest

from oracle.orc import C, Ref
from tests.helpers import full_tensor, random_block, random_system

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Ref.host_available(),
                                 reason="reference host wrapper not built")]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("update", [1, 0])
def test_fortran_intertwinecpp_call_through_the_library(hip, update):
    say("Fortran IntertwineCPP call -> intertwinecpp_, update", update)
    nb = [5, 6, 4]
    H = random_system(131, nb, ncoup=5)
    Q = random_block(132, nb, [4, 4, 4, 4])
    Q.eigv[:] = [1.0, 2.0, 3.0, 4.0]
    got = Ref.host_intertwine(H, Q, 2, 3, update, 1, 6.0)
    Qn = Q.copy()
    Qn.coef = Qn.coef + 1e-14
    want = C.intertwine(H, Qn, 2, 3, bool(update), 1, 6.0)
    if update:
        assert np.max(np.abs(got.eigv - want.eigv)) < 1e-9 * np.abs(want.eigv).max()
    else:
        assert np.array_equal(got.eigv, Q.eigv)
    for s in range(Q.nbloc):
        a = full_tensor(Q.nbas, *got.state(s))
        b = full_tensor(Q.nbas, *want.state(s))
        assert np.linalg.norm(a - b) <= 1e-8 * np.linalg.norm(b), s
