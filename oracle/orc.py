"""TEST INFRASTRUCTURE ONLY: Python access to the FP64 oracle.

Two checkers live here:

* ``C``   -- the C restatement ``oracle/mlcp_oracle.c`` (built by
  ``oracle/Makefile`` into ``oracle/libmlcp_oracle.so``);
* ``Ref`` -- the compiled reference Fortran routines driven by
  ``oracle/ref/ref_driver.f90`` (built into ``oracle/_ref/ref_driver`` from the
  reference sources; absent where the sources are absent).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; the HIP library never does.
"""
from __future__ import annotations

import ctypes as C_
import os
import subprocess
import tempfile

import numpy as np

from mlcp_amd.hoplist import CPBlock, HopList

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmlcp_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
# the same routines built with -fopenmp (states of a block in parallel, as
# the reference's own build runs them)
REF_DRIVER_OMP = os.path.join(HERE, "_ref", "ref_driver_omp")
HOST_DRIVER = os.path.join(HERE, "_ref", "host_driver")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i, _d = C_.c_int, C_.c_double
_H = [_i, _ip, _i, _ip, _ip, _dp, _dp]        # ndof,nbas,T,opid,mptr,hcoef,mat

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH}")
        L = C_.CDLL(LIB_PATH)
        L.orc_dspmv_u.argtypes = [_i, _d, _dp, _dp, _dp]
        L.orc_dspmv_u.restype = None
        L.orc_prodvv.argtypes = [_i, _ip, _i, _dp, _dp, _i, _dp, _dp]
        L.orc_prodvv.restype = _d
        L.orc_buildbjk.argtypes = _H + [_i, _dp, _dp, _dp, _dp, _i, _i, _d,
                                        _i, _dp, _dp]
        L.orc_als_pow3.argtypes = _H + [_i, _dp, _dp, _i, _i, _d]
        L.orc_prodhv_als3.argtypes = _H + [_i, _dp, _dp, _dp, _dp, _i, _i, _d]
        L.orc_normbase.argtypes = [_i, _ip, _i, _dp, _dp]
        L.orc_normbase.restype = None
        L.orc_normcoef.argtypes = [_i, _ip, _i, _dp, _dp]
        L.orc_normcoef.restype = None
        L.orc_als_ortho3.argtypes = [_i, _ip, _i, _ip, _dp, _dp, _i]
        L.orc_getqhq3.argtypes = _H + [_i, _ip, _dp, _dp, _i, _dp]
        L.orc_getoverlaps.argtypes = [_i, _ip, _i, _ip, _dp, _dp, _dp]
        L.orc_getoverlaps.restype = None
        L.orc_dsygv.argtypes = [_i, _dp, _dp, _dp]
        L.orc_als_sumlcvec3.argtypes = [_i, _ip, _i, _dp, _dp, _i, _ip, _dp,
                                        _dp, _dp, _i]
        L.orc_intertwine.argtypes = [_i, _ip, _i, _ip, _dp, _dp, _i, _ip, _ip,
                                     _dp, _dp, _dp, _i, _i, _i, _i, _d]
        L.orc_reduc_als.argtypes = [_i, _ip, _i, _dp, _dp, _i, _dp, _dp, _i]
        L.orc_rayleigh.argtypes = _H + [_i, _ip, _dp, _dp, _dp]
        L.orc_rayleigh.restype = None
        L.orc_update_opermat.argtypes = [_i, _ip, _i, _ip, _dp, _dp, _i, _dp,
                                         _dp]
        _lib = L
    return _lib


def _h(H: HopList):
    return (H.ndof, H.nbas, H.T, H.opid, H.mptr, H.coef, H.mat)


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


class C:
    """The C restatement.  Block arguments are copied; results returned."""

    @staticmethod
    def dspmv(ap, x, fac=1.0):
        y = np.zeros_like(x)
        lib().orc_dspmv_u(x.size, fac, np.ascontiguousarray(ap),
                          np.ascontiguousarray(x), y)
        return y

    @staticmethod
    def prodvv(nbas, fcoef, fbase, gcoef=None, gbase=None):
        nbas = np.ascontiguousarray(nbas, dtype=np.int32)
        if gcoef is None:
            return lib().orc_prodvv(nbas.size, nbas, fcoef.size, fcoef, fbase,
                                    fcoef.size, fcoef, fbase)
        return lib().orc_prodvv(nbas.size, nbas, fcoef.size, fcoef, fbase,
                                gcoef.size, gcoef, gbase)

    @staticmethod
    def buildbjk(H, ocoef, obase, coef, base, k, ishift, eshift, recalc):
        R = coef.size
        BB = np.zeros((R, R))            # column major via transpose at end
        bjk = np.zeros((int(H.nbas[k]), R))
        _chk(lib().orc_buildbjk(*_h(H), R, ocoef, obase, coef, base, k,
                                ishift, eshift, int(recalc), BB, bjk),
             "buildbjk")
        return BB.T.copy(), bjk.T.copy()   # Fortran BB(i,l), bjk(l,j)

    @staticmethod
    def als_pow3(H, coef, base, nitn, ishift=1, eshift=0.0):
        c, b = coef.copy(), base.copy()
        _chk(lib().orc_als_pow3(*_h(H), c.size, c, b, nitn, ishift, eshift),
             "als_pow3")
        return c, b

    @staticmethod
    def prodhv_als3(H, coef, base, nitn, ishift=0, eshift=0.0):
        c, b = np.empty_like(coef), np.empty_like(base)
        _chk(lib().orc_prodhv_als3(*_h(H), coef.size, coef, base, c, b, nitn,
                                   ishift, eshift), "prodhv_als3")
        return c, b

    @staticmethod
    def normcoef(nbas, coef, base):
        c, b = coef.copy(), base.copy()
        lib().orc_normcoef(nbas.size, nbas, c.size, c, b)
        return c, b

    @staticmethod
    def als_ortho3(Q: CPBlock, nitn) -> CPBlock:
        out = Q.copy()
        _chk(lib().orc_als_ortho3(out.nbas.size, out.nbas, out.nbloc,
                                  out.ranks, out.coef, out.base, nitn),
             "als_ortho3")
        return out

    @staticmethod
    def getqhq3(H, Q: CPBlock, nitn):
        n = Q.nbloc
        M = np.zeros((n, n))
        _chk(lib().orc_getqhq3(*_h(H), n, Q.ranks, Q.coef, Q.base, nitn, M),
             "getqhq3")
        return M.T.copy()

    @staticmethod
    def getoverlaps(Q: CPBlock):
        n = Q.nbloc
        S = np.zeros((n, n))
        lib().orc_getoverlaps(Q.nbas.size, Q.nbas, n, Q.ranks, Q.coef, Q.base,
                              S)
        return S.T.copy()

    @staticmethod
    def dsygv(A, B):
        """A, B: (n, n) with upper triangles set.  Returns (w, X)."""
        n = A.shape[0]
        a = np.array(A.T, order="C", copy=True)
        b = np.array(B.T, order="C", copy=True)
        w = np.zeros(n)
        _chk(lib().orc_dsygv(n, a, b, w), "dsygv")
        return w, a.T.copy()

    @staticmethod
    def sumlcvec3(F_coef, F_base, Q: CPBlock, lc, nitn):
        c, b = F_coef.copy(), F_base.copy()
        _chk(lib().orc_als_sumlcvec3(Q.nbas.size, Q.nbas, c.size, c, b,
                                     Q.nbloc, Q.ranks, Q.coef, Q.base,
                                     np.ascontiguousarray(lc, dtype=np.float64),
                                     nitn), "sumlcvec3")
        return c, b

    @staticmethod
    def intertwine(H, Q: CPBlock, npow, nals, update=True, ishift=1,
                   eshift=0.0) -> CPBlock:
        out = Q.copy()
        _chk(lib().orc_intertwine(out.nbas.size, out.nbas, out.nbloc,
                                  out.ranks, out.coef, out.base, H.T, H.opid,
                                  H.mptr, H.coef, H.mat, out.eigv, npow, nals,
                                  int(update), ishift, eshift), "intertwine")
        return out

    @staticmethod
    def reduc_als(nbas, gcoef, gbase, fcoef, fbase, nitn):
        c, b = fcoef.copy(), fbase.copy()
        _chk(lib().orc_reduc_als(nbas.size, nbas, gcoef.size, gcoef, gbase,
                                 c.size, c, b, nitn), "reduc_als")
        return c, b

    @staticmethod
    def rayleigh(H, Q: CPBlock):
        rq = np.zeros(Q.nbloc)
        lib().orc_rayleigh(*_h(H), Q.nbloc, Q.ranks, Q.coef, Q.base, rq)
        return rq


    @staticmethod
    def update_opermat(Q: CPBlock, imode, xpacked):
        """UpdateOperMat: packed upper <Q_j, X Q_i> (j >= i) for the packed
        upper operator X on mode imode (0-based)."""
        B = Q.nbloc
        out = np.zeros(B * (B + 1) // 2)
        _chk(lib().orc_update_opermat(len(Q.nbas), Q.nbas, B, Q.ranks, Q.coef,
                                      Q.base, int(imode),
                                      np.ascontiguousarray(xpacked,
                                                           dtype=np.float64),
                                      out), "update_opermat")
        return out


class Ref:
    """The compiled reference routines, through oracle/_ref/ref_driver."""

    MAGIC = 1296843600
    # driver binary used by run() when none is named (ref_driver, or
    # REF_DRIVER_OMP for runs with OMP_NUM_THREADS > 1)
    DRIVER = REF_DRIVER

    @staticmethod
    def available(omp: bool = False) -> bool:
        return os.path.exists(REF_DRIVER_OMP if omp else REF_DRIVER)

    @staticmethod
    def run(mode, H: HopList, Q: CPBlock, npow=0, nals=0, update=0, ishift=1,
            eshift=0.0, extra=None, files=None, collect=None, tree=None,
            driver=None):
        """Runs the driver; ``files`` maps suffixes to bytes written next to
        its input first, ``collect`` names suffixes of its output to return
        (then the result is (values, {suffix: bytes})), ``tree`` maps paths
        relative to the driver's working directory to bytes."""
        extra = np.zeros(0) if extra is None else np.ascontiguousarray(
            extra, dtype=np.float64).ravel(order="F")
        hdr = np.array([Ref.MAGIC, mode, H.ndof, Q.nbloc, H.T, H.nops, H.msize,
                        npow, nals, update, ishift, extra.size], dtype=np.int32)
        got = {}
        with tempfile.TemporaryDirectory() as td:
            fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            for suf, data in (files or {}).items():
                with open(fin + suf, "wb") as f:
                    f.write(data)
            for rel, data in (tree or {}).items():
                os.makedirs(os.path.dirname(os.path.join(td, rel)),
                            exist_ok=True)
                with open(os.path.join(td, rel), "wb") as f:
                    f.write(data)
            with open(fin, "wb") as f:
                f.write(hdr.tobytes())
                f.write(np.float64(eshift).tobytes())
                for a in (H.nbas, Q.ranks, H.opid, H.mptr):
                    f.write(np.ascontiguousarray(a, dtype=np.int32).tobytes())
                for a in (H.coef, H.mat, Q.coef, Q.base, Q.eigv, extra):
                    f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            p = subprocess.run([driver or Ref.DRIVER, fin, fout],
                               capture_output=True, text=True, timeout=600,
                               cwd=td)
            if p.returncode != 0 or not os.path.exists(fout):
                raise RuntimeError(f"ref_driver failed: {p.stdout}{p.stderr}")
            raw = open(fout, "rb").read()
            for suf in (collect or []):
                got[suf] = open(fout + suf, "rb").read()
        status = np.frombuffer(raw[:4], dtype=np.int32)[0]
        if status != 0:
            raise RuntimeError(f"ref_driver status {status}")
        vals = np.frombuffer(raw[4:], dtype=np.float64).copy()
        return (vals, got) if collect else vals

    @staticmethod
    def _block(vals, Q: CPBlock):
        nt, nD = Q.base.shape
        out = Q.copy()
        out.coef = vals[:nt].copy()
        out.base = vals[nt:nt + nt * nD].reshape(nt, nD).copy()
        return out, vals[nt + nt * nD:]

    @staticmethod
    def als_pow3(H, Q, npow, ishift=1, eshift=0.0):
        return Ref._block(Ref.run(1, H, Q, npow=npow, ishift=ishift,
                                  eshift=eshift), Q)[0]

    @staticmethod
    def prodhv_als3(H, Q, nals, ishift=0, eshift=0.0):
        return Ref._block(Ref.run(2, H, Q, nals=nals, ishift=ishift,
                                  eshift=eshift), Q)[0]

    @staticmethod
    def als_ortho3(H, Q, nals):
        return Ref._block(Ref.run(3, H, Q, nals=nals), Q)[0]

    @staticmethod
    def getqhq_overlaps(H, Q, nals):
        v = Ref.run(4, H, Q, nals=nals)
        n = Q.nbloc
        return (v[:n * n].reshape(n, n, order="F"),
                v[n * n:2 * n * n].reshape(n, n, order="F"))

    @staticmethod
    def cycle(H, Q, npow, nals, update=1, eshift=0.0):
        out, rest = Ref._block(Ref.run(5, H, Q, npow=npow, nals=nals,
                                       update=update, eshift=eshift), Q)
        out.eigv = rest[:Q.nbloc].copy()
        return out

    @staticmethod
    def updatevecs2(H, Q, U, nals):
        return Ref._block(Ref.run(6, H, Q, nals=nals, extra=U), Q)[0]

    @staticmethod
    def reduc_als(H, Q, nals):
        """Q holds F (state 0) then G (state 1); returns reduced F."""
        v = Ref.run(7, H, Q, nals=nals)
        rF, nD = int(Q.ranks[0]), Q.base.shape[1]
        return v[:rF].copy(), v[rF:rF + rF * nD].reshape(rF, nD).copy()

    @staticmethod
    def rayleigh(H, Q):
        return Ref.run(8, H, Q)[:Q.nbloc].copy()

    @staticmethod
    def update_opermat(H, Q):
        """UpdateOperMat of every operator of the term list on its mode, in
        operator order: one packed nbloc x nbloc matrix per operator."""
        v = Ref.run(10, H, Q)
        m = Q.nbloc * (Q.nbloc + 1) // 2
        return [v[o * m:(o + 1) * m].copy() for o in range(H.nops)]

    @staticmethod
    def save_psi_file(H, Q, isavi, bounds, delta) -> bytes:
        """SavePsiFile (Restart.f90:759-793) of Q: the file's bytes."""
        extra = np.concatenate([np.asarray(bounds, float).ravel(),
                                np.asarray(delta, float).ravel()])
        _, got = Ref.run(11, H, Q, npow=int(isavi), extra=extra,
                         collect=[".psi"])
        return got[".psi"]

    @staticmethod
    def read_psi_file(H, Q, data: bytes, ranks):
        """ReadPsiFile (Restart.f90:614-729) of a file's bytes into a block
        shaped like Q; ``ranks`` are the ranks the file holds."""
        v = Ref.run(12, H, Q, files={".psi": data})
        B = Q.nbloc
        isavi, ok = int(v[0]), bool(v[1])
        bounds, eigv, delta = v[2:4], v[4:4 + B], v[4 + B:4 + 2 * B]
        rest = v[4 + 2 * B:]
        ranks = np.asarray(ranks if ok else Q.ranks, dtype=np.int32)
        N, nD = int(ranks.sum()), int(np.sum(Q.nbas))
        out = CPBlock(Q.nbas, ranks, rest[:N].copy(),
                      rest[N:N + N * nD].reshape(N, nD).copy(), eigv.copy())
        return isavi, ok, bounds.copy(), eigv.copy(), delta.copy(), out

    @staticmethod
    def host_available() -> bool:
        return os.path.exists(HOST_DRIVER)

    @staticmethod
    def host_intertwine(H, Q, npow, nals, update, ishift, eshift):
        """A Fortran host calling IntertwineCPP as the reference's
        GPU_Intertwine does (GPUIterate.f90:130-180), linked against the HIP
        library (oracle/_ref/host_driver): one intertwine cycle of a
        uniform-rank block on the GPU."""
        v = Ref.run(0, H, Q, npow=npow, nals=nals, update=int(update),
                    ishift=ishift, eshift=eshift, driver=HOST_DRIVER)
        out, rest = Ref._block(v, Q)
        out.eigv = rest[:Q.nbloc].copy()
        return out

    @staticmethod
    def _dummy():
        nb = np.array([2], dtype=np.int32)
        H = HopList(nb, np.array([1], np.int32), np.array([1], np.int32),
                    np.array([1.0]), np.array([1.0, 0.0, 1.0]))
        Q = CPBlock(nb, np.array([1], np.int32), np.array([1.0]),
                    np.array([[1.0, 0.0]]), np.array([0.0]))
        return H, Q

    @staticmethod
    def read_ff(files: dict, sysid: str, ncp: int, divide: bool):
        """ReadFFHamiltonian (FFPES.f90:308-406) on the given pes/ files
        ({relative path: text}): [(qns (n,k), coef (n))] for k = 1..ncp."""
        H, Q = Ref._dummy()
        v = Ref.run(13, H, Q, npow=ncp, update=1 if divide else 0,
                    extra=np.array([float(ord(c)) for c in sysid]),
                    tree={k: t.encode() for k, t in files.items()})
        out, pos = [], 0
        for k in range(1, ncp + 1):
            n = int(v[pos])
            qns = v[pos + 1:pos + 1 + n * k].reshape(n, k, order="F")
            coef = v[pos + 1 + n * k:pos + 1 + n * k + n]
            out.append((qns.astype(np.int32), coef.copy()))
            pos += 1 + n * k + n
        return out

    @staticmethod
    def get_hoplist(nbas, opcoef, opindx, iops, eops, cops, pops):
        """GetHopList (ModeH.f90:909-1068) of a Hamiltonian in the
        reference's own form; operator matrices are (mode 1-based, packed)
        pairs except iops/eops, which are per mode."""
        nbas = np.asarray(nbas, dtype=np.int32)
        D = nbas.size
        opindx = np.asarray(opindx, dtype=np.int32)
        modes = list(range(1, D + 1)) * 2 + [m for m, _ in cops] + \
            [m for m, _ in pops]
        mats = list(iops) + list(eops) + [x for _, x in cops] + \
            [x for _, x in pops]
        mat = np.concatenate([np.asarray(x, float).ravel() for x in mats])
        H = HopList(nbas, opindx.ravel(), np.array(modes, np.int32),
                    np.asarray(opcoef, float), mat)
        Q = CPBlock(nbas, np.array([1], np.int32), np.array([1.0]),
                    np.zeros((1, int(nbas.sum()))), np.array([0.0]))
        v = Ref.run(14, H, Q, npow=len(cops), nals=len(pops))
        n1, n2, n3 = int(v[0]), int(v[1]), int(v[2])
        pos = 3
        opid = v[pos:pos + n1].astype(np.int32)
        mptr = v[pos + n1:pos + n1 + n2].astype(np.int32)
        m = v[pos + n1 + n2:pos + n1 + n2 + n3].copy()
        coef = v[pos + n1 + n2 + n3:].copy()
        return HopList(nbas, opid, mptr, coef, m)

    @staticmethod
    def dsygv(H, Q, A, B):
        n = Q.nbloc
        v = Ref.run(9, H, Q, extra=np.concatenate([A.ravel(order="F"),
                                                   B.ravel(order="F")]))
        return v[:n].copy(), v[n:n + n * n].reshape(n, n, order="F")
