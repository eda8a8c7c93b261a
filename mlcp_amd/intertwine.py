"""Host-side mirror of the reference's GPU wrappers (GPUIterate.f90) and of
the solver cycle that calls them (Solver.f90), over libmlcp_hip.so.

* ``gpu_intertwine``  GPU_Intertwine, GPUIterate.f90:89-187 -> intertwinecpp_
* ``gpu_als``         GPU_ALS_AB,     GPUIterate.f90:212-285 -> alscpp_
* ``gpu_memory``      Get_GPU_Memory, GPUIterate.f90:63-85   -> intwdevicemem_
* ``Session``         device-resident block for repeated cycles
* ``block_shift``     GetBlockShift, BlockPower.f90:20-45
* ``solve``           the cycle loop of SolveHPsi, Solver.f90:389-442
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List

import numpy as np

from ._lib import check, lib
from .hoplist import CPBlock, HopList


def _ci(v):
    return C.byref(C.c_int(int(v)))


def gpu_memory(H: HopList, rk: int, nbloc: int, update: bool = True,
               ncpu: int = 1) -> None:
    """Get_GPU_Memory: prints the device memory estimate."""
    lib().intwdevicemem_(_ci(rk), _ci(H.ndof), _ci(nbloc),
                         _ci(H.T * H.ndof), _ci(H.nops), _ci(H.msize),
                         _ci(ncpu), _ci(1 if update else 0), H.nbas)


def gpu_intertwine(Q: CPBlock, H: HopList, npow: int, nals: int,
                   update: bool = True, ishift: int = 1, eshift: float = 0.0,
                   noise: float = 1e-14, ncpu: int = 1) -> CPBlock:
    """GPU_Intertwine: one intertwinecpp_ call on a uniform-rank block.

    As in the reference, ``noise`` is added to every coefficient before the
    call (GPUIterate.f90:107,143) and the rank of every state must equal the
    first state's (GPUIterate.f90:152-158).  Returns the new block; its eigv
    holds the GEP eigenvalues if ``update``."""
    rk = int(Q.ranks[0])
    if np.any(Q.ranks != rk):
        raise ValueError("GPU_Intertwine(): bad rank (all states must share "
                         "the rank of the first)")
    out = Q.copy()
    out.coef = out.coef + noise
    L = lib()
    L.intertwinecpp_(out.base, out.coef, out.nbas, _ci(rk), _ci(H.ndof),
                     _ci(Q.nbloc), H.opid, H.mptr, H.coef, H.mat,
                     _ci(H.T * H.ndof), _ci(H.nops), _ci(H.msize), out.eigv,
                     _ci(npow), _ci(ncpu), _ci(nals), _ci(1 if update else 0),
                     _ci(ishift), C.byref(C.c_double(eshift)))
    err = L.mlcp_hip_last_error().decode()
    if err:
        raise RuntimeError(f"intertwinecpp_ failed: {err}")
    return out


def intertwine_ragged(Q: CPBlock, H: HopList, npow: int, nals: int,
                      update: bool = True, ishift: int = 1,
                      eshift: float = 0.0) -> CPBlock:
    """One cycle on a block with per-state ranks (mlcp_hip_intertwine)."""
    out = Q.copy()
    check(lib().mlcp_hip_intertwine(out.base, out.coef, out.nbas, H.ndof,
                                    out.nbloc, out.ranks, H.opid, H.mptr,
                                    H.coef, H.mat, H.T, H.nops, H.msize,
                                    out.eigv, npow, nals, int(update), ishift,
                                    eshift), "mlcp_hip_intertwine")
    return out


def gpu_als(F_coef, F_base, G_coef, G_base, nbas, nals: int):
    """GPU_ALS_AB: F (rank rF) <- ALS fit of G after nals sweeps."""
    rF, rG = F_coef.size, G_coef.size
    coef = np.ascontiguousarray(np.concatenate([F_coef, G_coef]))
    base = np.ascontiguousarray(np.vstack([F_base, G_base]))
    nbas = np.ascontiguousarray(nbas, dtype=np.int32)
    L = lib()
    L.alscpp_(coef, base, nbas, _ci(nbas.size), _ci(rF), _ci(rG), _ci(nals))
    err = L.mlcp_hip_last_error().decode()
    if err:
        raise RuntimeError(f"alscpp_ failed: {err}")
    return coef[:rF].copy(), base[:rF].copy()


def gpu_gep(A: np.ndarray, B: np.ndarray):
    """Generalized eigenproblem on the GPU (upper triangles used)."""
    n = A.shape[0]
    a = np.array(A.T, order="C", copy=True)
    b = np.array(B.T, order="C", copy=True)
    X = np.zeros((n, n))
    w = np.zeros(n)
    check(lib().mlcp_hip_gep(n, a, b, X, w), "mlcp_hip_gep")
    return w, X.T.copy()


def update_operators(Q: CPBlock, opmode, ops) -> np.ndarray:
    """Operator update for the layer above the solved one (UpdateH's loop
    over the primitive operators of the solved mode, Updater.f90:53-107,
    each through UpdateOperMat, :109-153), on the GPU.

    ``opmode[o]`` is the 0-based mode of Q that operator o acts on and
    ``ops[o]`` its packed upper triangle (n x n, column-wise
    ``i + j(j+1)/2``).  Returns one row per operator: the packed
    ``nbloc x nbloc`` matrix ``<Q_j, X_o Q_i>`` (i <= j), i.e. the operator
    in the basis of the block's states."""
    opmode = np.asarray(opmode, dtype=np.int32) + 1
    mats = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in ops]
    mptr = np.zeros(len(mats), dtype=np.int32)
    pos = 1
    for o, x in enumerate(mats):
        mptr[o] = pos
        pos += x.size
    mat = np.concatenate(mats) if mats else np.zeros(0)
    B = Q.nbloc
    out = np.zeros((len(mats), B * (B + 1) // 2))
    check(lib().mlcp_hip_update_opers(
        Q.base, Q.coef, Q.nbas, len(Q.nbas), B, Q.ranks, len(mats), opmode,
        mptr, mat, mat.size, out), "mlcp_hip_update_opers")
    return out


class Session:
    """A block and term list resident on the GPU across cycles."""

    def __init__(self, Q: CPBlock, H: HopList):
        self.H = H
        self.Q = Q.copy()
        h = C.c_void_p()
        check(lib().mlcp_hip_session_create(
            C.byref(h), self.Q.base, self.Q.coef, self.Q.nbas, H.ndof,
            self.Q.nbloc, self.Q.ranks, H.opid, H.mptr, H.coef, H.mat, H.T,
            H.nops, H.msize), "mlcp_hip_session_create")
        self.h = h

    def close(self):
        if self.h:
            lib().mlcp_hip_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def cycle(self, eigv, npow, nals, update=True, ishift=1, eshift=0.0):
        ev = np.ascontiguousarray(eigv, dtype=np.float64).copy()
        check(lib().mlcp_hip_session_cycle(self.h, ev, npow, nals,
                                           int(update), ishift, eshift),
              "mlcp_hip_session_cycle")
        return ev

    def solve(self, eigv, bounds, ncycle, npow, nals):
        """ncycle cycles with the E-shift formed on the device
        (mlcp_hip_session_solve): returns (final eigenvalues, per-cycle
        eigenvalues (ncycle x nbloc), per-cycle shifts)."""
        B = self.Q.nbloc
        ev = np.ascontiguousarray(eigv, dtype=np.float64).copy()
        hist = np.zeros(ncycle * (B + 1))
        check(lib().mlcp_hip_session_solve(
            self.h, ev, np.ascontiguousarray(bounds, dtype=np.float64),
            ncycle, npow, nals, hist), "mlcp_hip_session_solve")
        hist = hist.reshape(ncycle, B + 1)
        return ev, hist[:, :B].copy(), hist[:, B].copy()

    def download(self) -> CPBlock:
        out = self.Q.copy()
        check(lib().mlcp_hip_session_download(self.h, out.base, out.coef),
              "mlcp_hip_session_download")
        return out

    def upload(self, Q: CPBlock):
        q = Q.copy()
        check(lib().mlcp_hip_session_upload(self.h, q.base, q.coef),
              "mlcp_hip_session_upload")

    def save(self):
        check(lib().mlcp_hip_session_save(self.h), "session_save")

    def restore(self):
        check(lib().mlcp_hip_session_restore(self.h), "session_restore")

    KERNELS = {"op_hadamard": 1, "op_solve": 2, "op_refresh": 3,
               "gs_terms": 4, "gs_reduce": 5, "lc_rhs": 6, "lc_solve": 7,
               "lc_refresh": 8, "pair_inner": 9, "gep": 10}

    def probe(self, kernel: str, max_launches: int = 16384):
        kid = self.KERNELS[kernel] if kernel else 0
        check(lib().mlcp_hip_session_probe(self.h, kid, max_launches),
              "session_probe")

    def probe_read(self):
        ms, n = C.c_double(), C.c_long()
        check(lib().mlcp_hip_session_probe_read(self.h, C.byref(ms),
                                                C.byref(n)), "probe_read")
        return ms.value, n.value

    def phase_ms(self):
        ms = np.zeros(5)
        lib().mlcp_hip_session_phase_ms(self.h, ms)
        return ms

    def power(self, npow, ishift=1, eshift=0.0):
        check(lib().mlcp_hip_session_power(self.h, npow, ishift, eshift),
              "power")

    def ortho(self, nals):
        check(lib().mlcp_hip_session_ortho(self.h, nals), "ortho")

    def prodhv(self, nals) -> CPBlock:
        out = self.Q.copy()
        check(lib().mlcp_hip_session_prodhv(self.h, nals, out.base, out.coef),
              "prodhv")
        return out

    def gep(self, nals):
        B = self.Q.nbloc
        M, S, U, w = (np.zeros((B, B)), np.zeros((B, B)), np.zeros((B, B)),
                      np.zeros(B))
        check(lib().mlcp_hip_session_gep(self.h, nals, M, S, U, w), "gep")
        # column-major device matrices -> numpy (i, j)
        return M.T.copy(), S.T.copy(), U.T.copy(), w

    def rayleigh(self):
        rq = np.zeros(self.Q.nbloc)
        check(lib().mlcp_hip_session_rayleigh(self.h, rq), "rayleigh")
        return rq

    def update(self, U, nals):
        u = np.ascontiguousarray(np.asarray(U, dtype=np.float64).T)
        check(lib().mlcp_hip_session_update(self.h, u, nals), "update")

    FORM_TOL = 1
    FORM_MIXED = 2
    FORM_ML = 4

    def set_form(self, form=0, gs_tol=1e-5, als_tol=1e-7):
        """Form of the cycle (bits): 0 = the reference's fixed sweep counts
        in FP64; FORM_TOL = Gram-Schmidt, the reduced H*Q and the updates
        stop at a residual tolerance, with the fixed counts as the cap;
        FORM_MIXED = the power sweeps apply the terms on the FP32 matrix
        cores; FORM_ML = the reduced H*Q fit starts from the warm-start
        network's factors (see set_warmstart)."""
        check(lib().mlcp_hip_session_set_form(self.h, int(form),
                                              float(gs_tol), float(als_tol)),
              "set_form")

    def set_warmstart(self, mlp):
        """The network (mlcp_amd.warmstart.MLP) the FORM_ML form uses, or
        None.  The session keeps a reference; the network must outlive its
        use."""
        self._ws = mlp
        check(lib().mlcp_hip_session_set_warmstart(
            self.h, mlp.h if mlp is not None else None), "set_warmstart")

    def sweeps(self) -> np.ndarray:
        """(3, nbloc): per state, the sweeps Gram-Schmidt, the reduced H*Q
        and the updates needed in their last phase when they stopped at the
        tolerance (0: the fixed count ran)."""
        out = np.zeros(3 * self.Q.nbloc, dtype=np.int32)
        check(lib().mlcp_hip_session_sweeps(self.h, out), "sweeps")
        return out.reshape(3, self.Q.nbloc)

    def tune_set(self, had_tile=0, had_split=0, gs_stride=0):
        """Launch parameters (0 keeps a value); results do not depend on
        them bitwise."""
        check(lib().mlcp_hip_session_tune_set(self.h, int(had_tile),
                                              int(had_split), int(gs_stride)),
              "tune_set")

    def tuning(self) -> dict:
        out = np.zeros(4, dtype=np.int32)
        check(lib().mlcp_hip_session_tune_get(self.h, out), "tune_get")
        return {"had_tile": int(out[0]), "had_split": int(out[1]),
                "gs_stride": int(out[2]), "had_tile_max": int(out[3])}

    def tune_key(self) -> str:
        buf = C.create_string_buffer(512)
        check(lib().mlcp_hip_session_tune_key(self.h, buf, 512), "tune_key")
        return buf.value.decode()


def spectral_probes(nbas, rk: int, seed: int):
    """The two start vectors of GetSpectralRange (Solver.f90:582-606): the
    product of each mode's first (lowest) and last (highest) basis function,
    base 1e-15 elsewhere, padded to rank rk with random unit-norm terms of
    coefficient 1e-7 (the same random terms in both)."""
    nbas = np.asarray(nbas, dtype=np.int32)
    off = np.concatenate([[0], np.cumsum(nbas)])
    nD = int(nbas.sum())
    rng = np.random.Generator(np.random.PCG64(seed))
    pad = []
    for _ in range(rk - 1):
        x = rng.uniform(size=nD) - 0.5
        for d in range(nbas.size):
            x[off[d]:off[d + 1]] /= np.linalg.norm(x[off[d]:off[d + 1]])
        pad.append(x)
    out = []
    for first in (True, False):
        b = np.full((1, nD), 1e-15)
        for d in range(nbas.size):
            b[0, off[d] if first else off[d + 1] - 1] = 1.0
        base = np.vstack([b] + pad) if pad else b
        coef = np.array([1.0] + [1e-7] * (rk - 1))
        out.append(CPBlock(nbas, [rk], coef, base))
    return out


def spectral_range(H: HopList, rk: int, npow: int = 10, ncyc: int = 5,
                   seed: int = 1):
    """GetSpectralRange (Solver.f90:563-654) on the GPU: ALS-guided power
    iterations on the two probe vectors (shift 0 for the top of the spectrum,
    the mid-range shift for the bottom), Rayleigh quotients as the bounds,
    padded by 0.1% of the range."""
    lo_v, hi_v = spectral_probes(H.nbas, rk, seed)
    with Session(lo_v, H) as sl, Session(hi_v, H) as sh:
        lo, hi = sl.rayleigh()[0], sh.rayleigh()[0]
        for _ in range(ncyc):
            sl.power(npow, 1, 0.5 * (lo + hi))
            sh.power(npow, 1, 0.0)
            lo, hi = sl.rayleigh()[0], sh.rayleigh()[0]
    pad = 0.001 * (hi - lo)
    return lo - pad, hi + pad


def block_shift(eigv: np.ndarray, bounds) -> float:
    """GetBlockShift (BlockPower.f90:20-45)."""
    nev = eigv.size
    a = 2 * eigv[-1] - eigv[-2] if nev > 1 else 2 * eigv[0] - bounds[0]
    return 0.5 * (a + bounds[1])


@dataclass
class SolveResult:
    eigv: np.ndarray
    history: List[np.ndarray] = field(default_factory=list)
    cycles: int = 0
    converged: bool = False
    block: CPBlock = None


def solve(session: Session, eigv0, bounds, ncycle, npow, nals, solvtol,
          update=True, resfile="none", restart=False) -> SolveResult:
    """The cycle loop of SolveHPsi (Solver.f90:307-442) on a resident block:
    E-shift from GetBlockShift, one intertwine cycle, relative rms delta
    convergence test.

    Restart files as in the reference (Solver.f90:340,370,432): with
    ``restart`` the block, eigenvalues, deltas, spectral bounds and the
    cycle count are read from ``<resfile>_{1,2}_psi.rst`` (cycles up to the
    saved one are skipped; the saved ranks must be the session's); without a
    successful read the initial block is saved as cycle 0.  After every
    cycle the block is saved, unless ``resfile`` is 'none'."""
    from .restart import read_psi, save_psi
    eigv = np.asarray(eigv0, dtype=np.float64).copy()
    nev = eigv.size
    bounds = np.asarray(bounds, dtype=np.float64).copy()
    delta = np.zeros(nev)
    ist = 0
    got = None
    if restart:
        got = read_psi(resfile, nev, session.Q.nbas)
    if got is not None and got.success:
        if not np.array_equal(got.block.ranks, session.Q.ranks):
            raise ValueError("restart file ranks differ from the session's")
        session.upload(got.block)
        ist, bounds, eigv, delta = (got.isavi, got.bounds.copy(),
                                    got.eigv.copy(), got.delta.copy())
    else:
        save_psi(resfile, 0, bounds, eigv, delta, _with_eigv(
            session.download() if resfile[:4].lower() != "none" else None,
            eigv))
    res = SolveResult(eigv=eigv)
    res.cycles = ist
    for i in range(1, ncycle + 1):
        if i <= ist:
            continue
        eshift = block_shift(eigv, bounds)
        old = eigv.copy()
        eigv = session.cycle(eigv, npow, nals, update, 1, eshift)
        res.history.append(eigv.copy())
        res.cycles = i
        delta = (old - eigv) / np.abs(eigv)
        rms = np.sqrt(np.sum(delta ** 2)) / nev
        if resfile[:4].lower() != "none":
            save_psi(resfile, i, bounds, eigv, delta,
                     _with_eigv(session.download(), eigv))
        if rms <= solvtol:
            res.converged = True
            break
    res.eigv = eigv
    return res


def _with_eigv(Q, eigv):
    if Q is not None:
        Q.eigv = np.asarray(eigv, dtype=np.float64).copy()
    return Q
