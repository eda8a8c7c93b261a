"""GPU parity: every phase of the HIP path against the FP64 oracle, through
the C-ABI, on seeded inputs small enough for the oracle and for full-tensor
comparison.

Tolerances (FP64 everywhere; the HIP path sums in a different order and
solves the regularised normal equations by Gauss-Jordan elimination in the
L D L^T association where the reference uses LU, so agreement is to
rounding amplified by a few ALS sweeps):
  represented tensors   ||T_gpu - T_orc|| / ||T_orc|| <= 1e-9
  eigenvalues           |w_gpu - w_orc| <= 1e-9 * max(1, |w|)
"""
import sys

import numpy as np
import pytest

from tests.helpers import full_tensor, random_block, random_system

pytestmark = pytest.mark.gpu

TOL_T = 1e-9
TOL_W = 1e-9


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def tensors(Q):
    return [full_tensor(Q.nbas, *Q.state(s)) for s in range(Q.nbloc)]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = [
    ("uniform", [5, 6, 4], [4, 4, 4], 5),
    ("ragged", [5, 6, 4], [3, 6, 4, 2], 6),
    ("fourmode", [4, 3, 5, 4], [5, 3, 4], 7),
    # ranks above 32 (the four-row-tile factorisation and solve), with
    # bases large enough that the normal equations stay definite
    ("highrank", [16, 14, 18], [40, 20, 48, 33], 6),
    # the library's limits: rank 64 (four row tiles) and 128 basis functions
    # per mode (the widest register tiling)
    ("maxsize", [128, 100, 72], [64, 64, 37, 50], 6),
    # degenerate shapes: a mode of one basis function and rank-1 states, a
    # block of one state, two modes (prod_{d != k} G_d is a single Gram)
    ("edge", [1, 3, 2], [1, 2, 1], 3),
    ("single", [4, 5, 3], [5], 4),
    ("twomode", [6, 5], [3, 4, 2], 3),
]


@pytest.mark.parametrize("name,nb,ranks,ncoup", CASES)
def test_power_phase(hip, name, nb, ranks, ncoup):
    from mlcp_amd.intertwine import Session
    from oracle.orc import C
    say("power", name, nb, ranks)
    H = random_system(11, nb, ncoup=ncoup)
    Q = random_block(12, nb, ranks)
    with Session(Q, H) as s:
        s.power(3, 1, 4.0)
        g = s.download()
    for i in range(Q.nbloc):
        c, b = C.als_pow3(H, *Q.state(i), 3, 1, 4.0)
        e = rel(full_tensor(Q.nbas, *g.state(i)), full_tensor(Q.nbas, c, b))
        assert e < TOL_T, (i, e)


@pytest.mark.parametrize("name,nb,ranks,ncoup", CASES)
def test_ortho_phase(hip, name, nb, ranks, ncoup):
    from mlcp_amd.intertwine import Session
    from oracle.orc import C
    say("ortho", name, nb, ranks)
    H = random_system(21, nb, ncoup=ncoup)
    Q = random_block(22, nb, ranks)
    with Session(Q, H) as s:
        s.ortho(4)
        g = s.download()
    o = C.als_ortho3(Q, 4)
    for a, b in zip(tensors(g), tensors(o)):
        assert rel(a, b) < TOL_T


@pytest.mark.parametrize("name,nb,ranks,ncoup", CASES)
def test_prodhv_phase(hip, name, nb, ranks, ncoup):
    from mlcp_amd.intertwine import Session
    from oracle.orc import C
    say("prodhv", name, nb, ranks)
    H = random_system(31, nb, ncoup=ncoup)
    Q = random_block(32, nb, ranks)
    with Session(Q, H) as s:
        hq = s.prodhv(3)
    for i in range(Q.nbloc):
        c, b = C.prodhv_als3(H, *Q.state(i), 3)
        e = rel(full_tensor(Q.nbas, *hq.state(i)), full_tensor(Q.nbas, c, b))
        assert e < TOL_T, (i, e)


@pytest.mark.parametrize("name,nb,ranks,ncoup", CASES)
def test_gep_phase(hip, name, nb, ranks, ncoup):
    from mlcp_amd.intertwine import Session
    from oracle.orc import C
    say("gep", name, nb, ranks)
    H = random_system(41, nb, ncoup=ncoup)
    Q = C.als_ortho3(random_block(42, nb, ranks), 6)
    with Session(Q, H) as s:
        M, S, U, w = s.gep(3)
    Mo = C.getqhq3(H, Q, 3)
    So = C.getoverlaps(Q)
    iu = np.triu_indices(Q.nbloc)
    assert np.max(np.abs(M[iu] - Mo[iu])) < 1e-9 * np.max(np.abs(Mo))
    assert np.max(np.abs(S[iu] - So[iu])) < 1e-9
    wo, Uo = C.dsygv(Mo, So)
    assert np.max(np.abs(w - wo)) < TOL_W * max(1, np.max(np.abs(wo)))
    assert np.max(np.abs(U - Uo)) < 1e-7


@pytest.mark.parametrize("name,nb,ranks,ncoup", CASES)
def test_update_phase(hip, name, nb, ranks, ncoup):
    from mlcp_amd.intertwine import Session
    from oracle.orc import C
    say("update", name, nb, ranks)
    rng = np.random.default_rng(51)
    H = random_system(51, nb, ncoup=ncoup)
    Q = random_block(52, nb, ranks)
    U = rng.normal(size=(Q.nbloc, Q.nbloc))
    with Session(Q, H) as s:
        s.update(U, 3)
        g = s.download()
    for i in range(Q.nbloc):
        c, b = C.sumlcvec3(*Q.state(i), Q, U[:, i], 3)
        e = rel(full_tensor(Q.nbas, *g.state(i)), full_tensor(Q.nbas, c, b))
        assert e < TOL_T, (i, e)


def test_power_and_ortho_ill_conditioned(hip):
    """Power and Gram-Schmidt phases on the 3-mode Henon-Heiles top block,
    whose ALS normal equations reach condition numbers near 1e10.  With the
    Schur complements formed in the L D L^T association both phases agreed
    with the oracle to a few 1e-11 on the first generated block (Gauss-Jordan
    through explicitly inverted 4 x 4 pivot blocks gave 2.3e-10 and 6.0e-10
    there); the block generated with canonical eigenvector signs measures
    1.05e-10 (power, state 14).  The bound is 2e-10."""
    from mlcp_amd import systems
    from mlcp_amd.intertwine import Session, block_shift
    from oracle.orc import C
    w = systems.hh111()
    H, Q = w.H, w.Q0.copy()
    say("power+ortho ill-conditioned", list(Q.nbas), int(Q.ranks.max()), Q.nbloc)
    es = block_shift(w.Q0.eigv, w.bounds)
    with Session(Q, H) as s:
        s.power(w.npow, 1, es)
        g = s.download()
    P = Q.copy()
    ep = []
    for i in range(Q.nbloc):
        c, b = C.als_pow3(H, *Q.state(i), w.npow, 1, es)
        ep.append(rel(full_tensor(Q.nbas, *g.state(i)),
                      full_tensor(Q.nbas, c, b)))
        pc, pb = P.state(i)
        pc[:], pb[:] = c, b
    with Session(P, H) as s:
        s.ortho(w.nals)
        g = s.download()
    o = C.als_ortho3(P, w.nals)
    eo = [rel(a, b) for a, b in zip(tensors(g), tensors(o))]
    say("ill-conditioned: max rel power", f"{max(ep):.3e}", "ortho",
        f"{max(eo):.3e}")
    assert max(ep) < 2e-10, ("power", int(np.argmax(ep)), max(ep))
    assert max(eo) < 2e-10, ("ortho", int(np.argmax(eo)), max(eo))


def test_update_phase_ill_conditioned(hip):
    """The update of the 3-mode Henon-Heiles top block after one power +
    Gram-Schmidt pass: during its ALS sweeps the normal equations reach
    condition numbers near 2e10 (smallest eigenvalue ~7e-10 against a penalty
    of 1e-10).  The solve must keep them positive definite and agree with the
    oracle's LU solve; the tolerance is the represented-tensor bound widened
    by that conditioning."""
    from mlcp_amd import systems
    from mlcp_amd.intertwine import Session, block_shift
    from oracle.orc import C
    w = systems.hh111()
    H, Q = w.H, w.Q0.copy()
    say("update ill-conditioned", list(Q.nbas), int(Q.ranks.max()), Q.nbloc)
    es = block_shift(w.Q0.eigv, w.bounds)
    for i in range(Q.nbloc):
        c, b = C.als_pow3(H, *Q.state(i), w.npow, 1, es)
        qc, qb = Q.state(i)
        qc[:], qb[:] = c, b
    Q = C.als_ortho3(Q, w.nals)
    _, U = C.dsygv(C.getqhq3(H, Q, w.nals), C.getoverlaps(Q))
    with Session(Q, H) as s:
        s.update(U, w.nals)
        g = s.download()
    for i in range(Q.nbloc):
        c, b = C.sumlcvec3(*Q.state(i), Q, U[:, i], w.nals)
        e = rel(full_tensor(Q.nbas, *g.state(i)), full_tensor(Q.nbas, c, b))
        assert e < 1e-7, (i, e)


# 3: the smallest order with a Householder step; 131 / 132: the largest
# order whose working matrix is held in LDS and the smallest in global
# memory; 256: the largest block
@pytest.mark.parametrize("n", [1, 2, 3, 7, 33, 64, 130, 131, 132, 200, 256])
def test_gep_solver(hip, n):
    from mlcp_amd.intertwine import gpu_gep
    from oracle.orc import C
    say("gep solver", n)
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n))
    A = a + a.T + np.diag(np.arange(n) * 3.0)
    b = rng.normal(size=(n, n)) * 0.1
    B = np.eye(n) + b @ b.T
    w, X = gpu_gep(A, B)
    wo, Xo = C.dsygv(A, B)
    assert np.max(np.abs(w - wo)) < 1e-10 * max(1, np.max(np.abs(wo)))
    # residual and B-orthonormality
    assert np.max(np.abs(A @ X - B @ X * w)) < 1e-9 * max(1, np.max(np.abs(w)))
    assert np.max(np.abs(X.T @ B @ X - np.eye(n))) < 1e-10
    assert np.max(np.abs(X - Xo)) < 1e-7


@pytest.mark.parametrize("kind", ["triple", "diagonal", "identity"])
def test_gep_solver_degenerate(hip, kind):
    """Repeated eigenvalues (each three times, in a random basis), an
    already diagonal matrix (every Householder reflection skipped) and A = B
    (all eigenvalues 1): the eigenvectors are not unique, so the check is
    the eigenvalues against the oracle, the residual and B-orthonormality."""
    from mlcp_amd.intertwine import gpu_gep
    from oracle.orc import C
    say("gep degenerate", kind)
    n = 48
    rng = np.random.default_rng(11)
    b = rng.normal(size=(n, n)) * 0.1
    B = np.eye(n) + b @ b.T
    if kind == "triple":
        q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        A = q @ np.diag(np.repeat(np.arange(n // 3) * 7.0 - 20.0, 3)) @ q.T
        A = 0.5 * (A + A.T)
    elif kind == "diagonal":
        A = np.diag(rng.normal(size=n) * 10.0)
        B = np.eye(n)
    else:
        A = B.copy()
    w, X = gpu_gep(A, B)
    wo, _ = C.dsygv(A, B)
    scale = max(1, np.max(np.abs(wo)))
    assert np.max(np.abs(w - wo)) < 1e-10 * scale
    assert np.max(np.abs(A @ X - B @ X * w)) < 1e-9 * scale
    assert np.max(np.abs(X.T @ B @ X - np.eye(n))) < 1e-10


@pytest.mark.parametrize("name,nb,ranks,ncoup", CASES[:2])
def test_full_cycle(hip, name, nb, ranks, ncoup):
    from mlcp_amd.intertwine import intertwine_ragged
    from oracle.orc import C
    say("cycle", name, nb, ranks)
    H = random_system(61, nb, ncoup=ncoup)
    Q = random_block(62, nb, ranks)
    g = intertwine_ragged(Q, H, 3, 4, True, 1, 5.0)
    o = C.intertwine(H, Q, 3, 4, True, 1, 5.0)
    assert np.max(np.abs(g.eigv - o.eigv)) < TOL_W * max(1, np.abs(o.eigv).max())
    for a, b in zip(tensors(g), tensors(o)):
        assert rel(a, b) < 1e-8


def test_intertwinecpp_dropin(hip):
    """The reference entry point, uniform rank, through GPU_Intertwine's
    packing (noise 0 so that the oracle sees the same coefficients)."""
    from mlcp_amd.intertwine import gpu_intertwine
    from oracle.orc import C
    say("intertwinecpp_")
    nb = [5, 6, 4]
    H = random_system(71, nb, ncoup=5)
    Q = random_block(72, nb, [4, 4, 4, 4])
    Q.eigv[:] = [1.0, 2.0, 3.0, 4.0]
    g = gpu_intertwine(Q, H, 2, 3, True, 1, 6.0, noise=0.0)
    o = C.intertwine(H, Q, 2, 3, True, 1, 6.0)
    assert np.max(np.abs(g.eigv - o.eigv)) < TOL_W * np.abs(o.eigv).max()
    # update = False keeps eigv and only orthogonalises
    g2 = gpu_intertwine(Q, H, 2, 3, False, 1, 6.0, noise=0.0)
    o2 = C.intertwine(H, Q, 2, 3, False, 1, 6.0)
    assert np.array_equal(g2.eigv, Q.eigv)
    for a, b in zip(tensors(g2), tensors(o2)):
        assert rel(a, b) < 1e-8


def test_alscpp_dropin(hip):
    from mlcp_amd.intertwine import gpu_als
    from oracle.orc import C
    say("alscpp_")
    nb = np.array([5, 6, 4], dtype=np.int32)
    G = random_block(81, nb, [9])
    F = random_block(82, nb, [3])
    c, b = gpu_als(F.coef, F.base, G.coef, G.base, nb, 5)
    co, bo = C.reduc_als(nb, G.coef, G.base, F.coef, F.base, 5)
    assert rel(full_tensor(nb, c, b), full_tensor(nb, co, bo)) < TOL_T


@pytest.mark.parametrize("name,nb,ranks,ncoup", CASES)
def test_rayleigh_quotients(hip, name, nb, ranks, ncoup):
    from mlcp_amd.intertwine import Session
    from oracle.orc import C
    say("rayleigh", name, nb, ranks)
    H = random_system(91, nb, ncoup=ncoup)
    Q = random_block(92, nb, ranks)
    with Session(Q, H) as s:
        rq = s.rayleigh()
    ro = C.rayleigh(H, Q)
    assert np.max(np.abs(rq - ro)) < 1e-11 * np.abs(ro).max()


def test_spectral_range_matches_oracle(hip):
    from mlcp_amd.intertwine import spectral_probes, spectral_range
    from oracle.orc import C
    say("spectral range")
    H = random_system(93, [5, 6, 4], ncoup=5)
    lo, hi = spectral_range(H, 3, npow=4, ncyc=2, seed=5)
    a, b = spectral_probes(H.nbas, 3, 5)
    l, h = C.rayleigh(H, a)[0], C.rayleigh(H, b)[0]
    for _ in range(2):
        ca, ba = C.als_pow3(H, a.coef, a.base, 4, 1, 0.5 * (l + h))
        cb, bb = C.als_pow3(H, b.coef, b.base, 4, 1, 0.0)
        a.coef, a.base, b.coef, b.base = ca, ba, cb, bb
        l, h = C.rayleigh(H, a)[0], C.rayleigh(H, b)[0]
    pad = 0.001 * (h - l)
    assert abs(lo - (l - pad)) < 1e-8 * abs(h) and abs(hi - (h + pad)) < 1e-8 * abs(h)


@pytest.mark.parametrize("name,nb,ranks,ncoup", CASES)
def test_update_operators(hip, name, nb, ranks, ncoup):
    """UpdateOperMat of every operator of a term list, on the GPU, against
    the oracle (which matches the compiled reference bitwise)."""
    from mlcp_amd.intertwine import update_operators
    from oracle.orc import C
    say("update operators", name, nb, ranks)
    H = random_system(95, nb, ncoup=ncoup)
    Q = random_block(96, nb, ranks)
    modes = H.op_mode()
    ops = []
    for o in range(H.nops):
        n = int(H.nbas[modes[o]])
        ops.append(H.mat[H.mptr[o] - 1:H.mptr[o] - 1 + n * (n + 1) // 2])
    got = update_operators(Q, modes, ops)
    for o in range(H.nops):
        want = C.update_opermat(Q, int(modes[o]), ops[o])
        assert np.max(np.abs(got[o] - want)) < 1e-12 * max(1.0, np.max(np.abs(want))), o


def test_session_rejects_out_of_range_shapes(hip):
    """Ranks above 64, bases above 128 functions and an empty block are
    refused with a message before anything is launched."""
    from mlcp_amd.hoplist import CPBlock
    from mlcp_amd.intertwine import Session
    say("session rejects shapes")
    for nb, ranks, what in (([4, 5], [65], "rank"), ([4, 129], [3], "nbas")):
        H = random_system(61, nb, ncoup=2)
        Q = random_block(62, nb, ranks)
        with pytest.raises(RuntimeError, match=what):
            Session(Q, H)
    H = random_system(61, [4, 5], ncoup=2)
    empty = CPBlock(np.array([4, 5]), np.zeros(0, dtype=np.int32),
                    np.zeros(0), np.zeros((0, 9)), np.zeros(0))
    with pytest.raises(RuntimeError, match="nbloc"):
        Session(empty, H)


def test_update_operators_rejects_bad_mode(hip):
    from mlcp_amd.intertwine import update_operators
    say("update operators: bad mode")
    Q = random_block(97, [4, 3], [2, 2])
    with pytest.raises(RuntimeError, match="opmode"):
        update_operators(Q, [5], [np.zeros(10)])


def test_solve_restart_resumes(hip, tmp_path):
    """SolveHPsi's restart files: two cycles saved, then a restart resumes
    at cycle 3 and lands where an uninterrupted three-cycle solve lands."""
    from mlcp_amd.intertwine import Session, solve
    say("solve restart")
    nb = [5, 6, 4]
    H = random_system(98, nb, ncoup=5)
    Q = random_block(99, nb, [4, 4, 4, 4])
    Q.eigv[:] = [1.0, 2.0, 3.0, 4.0]
    bounds = (0.0, 40.0)
    with Session(Q, H) as s:
        full = solve(s, Q.eigv, bounds, 3, 2, 3, 0.0)
    pre = str(tmp_path / "job")
    with Session(Q, H) as s:
        first = solve(s, Q.eigv, bounds, 2, 2, 3, 0.0, resfile=pre)
    assert first.cycles == 2
    with Session(Q, H) as s:
        rest = solve(s, Q.eigv, bounds, 3, 2, 3, 0.0, resfile=pre,
                     restart=True)
    assert rest.cycles == 3 and len(rest.history) == 1
    assert np.max(np.abs(rest.eigv - full.eigv)) < 1e-10 * np.abs(full.eigv).max()


@pytest.mark.parametrize("ncoup", [6, 400])
def test_tuning_is_bitwise_neutral(hip, ncoup):
    """The launch parameters (Hadamard tile and slot split, Gram-Schmidt
    stride) change no result bit: a cycle from the same block gives the same
    eigenvalues and factors under defaults and under other parameters.  With
    400 couplings a part's term records span several 64-record windows,
    which start at different records under different slot splits."""
    from mlcp_amd.intertwine import Session
    say("tuning neutral", ncoup)
    nb = [5, 6, 4]
    H = random_system(71, nb, ncoup=ncoup)
    Q = random_block(72, nb, [6, 3, 5, 4])
    Q.eigv[:] = [1.0, 2.0, 3.0, 4.0]
    with Session(Q, H) as s:
        s.save()
        d = s.tuning()
        e0 = s.cycle(Q.eigv, 3, 3, True, 1, 5.0)
        b0 = s.download()
        for t, p, x in [(64, 1, 8), (128, 2, 1), (64, 8, 1), (64, 16, 1),
                        (d["had_tile_max"], 1, 8)]:
            s.restore()
            s.tune_set(t, p, x)
            assert s.tuning()["had_tile"] == t
            e1 = s.cycle(Q.eigv, 3, 3, True, 1, 5.0)
            b1 = s.download()
            assert np.array_equal(e0, e1), (t, p, x)
            assert np.array_equal(b0.base, b1.base) and np.array_equal(b0.coef, b1.coef)


def test_tuning_rejects_bad_values(hip):
    from mlcp_amd.intertwine import Session
    say("tuning rejects")
    nb = [5, 6, 4]
    with Session(random_block(74, nb, [4, 4]), random_system(73, nb, ncoup=4)) as s:
        for bad in [(100, 0, 0), (0, 3, 0), (0, 32, 0), (0, 0, 4), (1024, 0, 0),
                    (256, 8, 0)]:
            with pytest.raises(RuntimeError):
                s.tune_set(*bad)


def test_autotune_persists_and_applies(hip, tmp_path, monkeypatch):
    """The tuner writes the chosen parameters under the session's key, and a
    new session on the same problem starts from them (the library reads the
    table named by MLCP_HIP_TUNING)."""
    from mlcp_amd import autotune
    from mlcp_amd.intertwine import Session
    say("autotune")
    nb = [5, 6, 4]
    H = random_system(75, nb, ncoup=6)
    Q = random_block(76, nb, [6, 3, 5, 4])
    Q.eigv[:] = [1.0, 2.0, 3.0, 4.0]
    table = str(tmp_path / "tune.txt")
    monkeypatch.setenv("MLCP_HIP_TUNING", table)
    with Session(Q, H) as s:
        key = s.tune_key()
        params, trials = autotune.tune(s, Q.eigv, 2, 2, 5.0, reps=1, table=table)
    assert autotune.read_table(table)[key] == params
    assert len(trials) > 2
    with Session(Q, H) as s2:
        t = s2.tuning()
        assert (t["had_tile"], t["had_split"], t["gs_stride"]) == params
