"""Writes the committed fixtures of tests/golden/.

    python tests/golden/make_golden.py fixtures    # needs oracle/_ref (built
                                                   # from the reference sources)
    python tests/golden/make_golden.py workloads   # FP64 oracle solves

fixtures:  small seeded inputs and the outputs of the compiled reference
           routines (ALS_POW3, PRODHV_ALS3, ALS_ORTHO_3, GetQHQ3/GetOverlaps,
           UpdateVecs2, reduc_ALS, RayleighQuotient2, one Iterate cycle,
           UpdateOperMat of every operator), saved as ref_<case>.npz, and
           the reference's SavePsiFile of the block as ref_<case>_psi.rst;
           ReadFFHamiltonian of force-field files and GetHopList of a mode
           Hamiltonian as ffhl_ref.npz.  The C oracle must reproduce them without
           the reference present.
workloads: the spectral bounds (GetSpectralRange) and the eigenvalues of
           the FP64 oracle after each cycle of the top-layer solve of the
           benchmark workloads (GetBlockShift + one intertwine cycle per
           cycle), with a fingerprint of the generated inputs; saved as
           workload_<name>.json.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(
    __file__))))
sys.path.insert(0, ROOT)

from mlcp_amd import systems  # noqa: E402
from mlcp_amd.intertwine import block_shift  # noqa: E402
from oracle.orc import C, Ref  # noqa: E402
from tests.helpers import random_block, random_system  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def fingerprint(H, Q):
    """Rounding-tolerant identity of generated inputs."""
    return {
        "T": H.T, "nops": H.nops, "msize": H.msize,
        "nbas": [int(x) for x in H.nbas],
        "ranks_sum": int(Q.ranks.sum()),
        "mat_sum": float(np.sum(H.mat)), "mat_abs": float(np.sum(np.abs(H.mat))),
        "coef_sum": float(np.sum(H.coef)),
        "q_base_abs": float(np.sum(np.abs(Q.base))),
        "eig0": [float(x) for x in Q.eigv[:4]],
    }


def fixtures():
    if not Ref.available():
        raise SystemExit("oracle/_ref/ref_driver is not built")
    cases = [("d3", [5, 6, 4], [4, 4, 4], 5), ("d4r", [4, 3, 5, 4], [5, 3, 4], 7)]
    for name, nb, ranks, nc in cases:
        H = random_system(501, nb, ncoup=nc)
        Q = random_block(502, nb, ranks)
        rng = np.random.default_rng(503)
        U = rng.normal(size=(Q.nbloc, Q.nbloc))
        pw = Ref.als_pow3(H, Q, 3, 1, 3.5)
        hq = Ref.prodhv_als3(H, Q, 3)
        gs = Ref.als_ortho3(H, Q, 4)
        M, S = Ref.getqhq_overlaps(H, Q, 3)
        up = Ref.updatevecs2(H, Q, U, 3)
        rq = Ref.rayleigh(H, Q)
        cyc = Ref.cycle(H, Q, 3, 4, 1, 5.0)
        F = random_block(504, nb, [3])
        G = random_block(505, nb, [7])
        QF = random_block(506, nb, [3, 7])
        QF.coef[:3], QF.base[:3] = F.coef, F.base
        QF.coef[3:], QF.base[3:] = G.coef, G.base
        rc, rb = Ref.reduc_als(H, QF, 4)
        qxq = np.array(Ref.update_opermat(H, Q))
        Qp = Q.copy()
        Qp.eigv = np.arange(1, Q.nbloc + 1) * 1.25
        psi = Ref.save_psi_file(H, Qp, 7, [0.5, 99.0],
                                np.linspace(-1e-3, 1e-3, Q.nbloc))
        with open(os.path.join(HERE, f"ref_{name}_psi.rst"), "wb") as f:
            f.write(psi)
        np.savez_compressed(
            os.path.join(HERE, f"ref_{name}.npz"),
            nbas=H.nbas, opid=H.opid, mptr=H.mptr, hcoef=H.coef, mat=H.mat,
            ranks=Q.ranks, coef=Q.coef, base=Q.base, U=U,
            pow_coef=pw.coef, pow_base=pw.base, hq_coef=hq.coef,
            hq_base=hq.base, gs_coef=gs.coef, gs_base=gs.base, qhq=M, s=S,
            up_coef=up.coef, up_base=up.base, rq=rq, cyc_coef=cyc.coef,
            cyc_base=cyc.base, cyc_eigv=cyc.eigv, F_coef=F.coef,
            F_base=F.base, G_coef=G.coef, G_base=G.base, red_coef=rc,
            red_base=rb, qxq=qxq)
        print("wrote", name)
    ff_and_hoplist()


def hoplist_case():
    """A mode Hamiltonian in the reference's own form (identity, eigen,
    compound and primitive operators; ModeH.f90:909-1068 input)."""
    from mlcp_amd.hoplist import pack_upper
    rng = np.random.default_rng(3)
    nbas = [3, 4, 2]
    D = len(nbas)

    def rs(n):
        a = rng.normal(size=(n, n))
        return pack_upper(a + a.T)
    iops = [pack_upper(np.eye(n)) for n in nbas]
    eops = [pack_upper(np.diag(rng.uniform(1, 5, n))) for n in nbas]
    cops = [(2, rs(4)), (1, rs(3))]
    pops = [(1, rs(3)), (3, rs(2)), (2, rs(4)), (1, rs(3)), (2, rs(4))]
    npop = len(pops)
    prim = {m: [i + 1 for i, (mm, _) in enumerate(pops) if mm == m]
            for m in range(1, D + 1)}
    comp = {m: [npop + i + 1 for i, (mm, _) in enumerate(cops) if mm == m]
            for m in range(1, D + 1)}
    T = 9
    opindx = np.zeros((T, D), dtype=np.int32)
    for k in range(T):
        for j in range(D):
            r = rng.integers(0, 4)
            if r == 1:
                opindx[k, j] = -(j + 1)
            elif r == 2 and prim[j + 1]:
                opindx[k, j] = rng.choice(prim[j + 1])
            elif r == 3 and comp[j + 1]:
                opindx[k, j] = comp[j + 1][0]
    return nbas, rng.normal(size=T), opindx, iops, eops, cops, pops


FF_ODD = """  1   1   2  0.5
  2
  2 3   -1.25D+01
3,3,3,4.0
  1 1 2 7.0
  1 x 2 9.0
  4 4 4 1.0
"""


def ff_and_hoplist():
    """The reference's ReadFFHamiltonian on generated and oddly formatted
    force-field files, and its GetHopList: ffhl_ref.npz."""
    import tempfile
    from mlcp_amd.pesfile import write_pes_files
    pes = systems.qff_pes(5, 123)
    out = {}
    with tempfile.TemporaryDirectory() as td:
        files = write_pes_files(pes, "qff05", td)
        tree = {os.path.relpath(f, td): open(f).read() for f in files}
    for k in range(1, 5):
        key = f"pes/f{k}qff05.dat"
        out[f"ff_text_{k}"] = np.array(tree.get(key, ""))
    for divide in (0, 1):
        for k, (q, c) in enumerate(Ref.read_ff(tree, "qff05", 4, divide), 1):
            out[f"ff_qns_{divide}_{k}"] = q
            out[f"ff_coef_{divide}_{k}"] = c
    odd = {"pes/f3odd05.dat": FF_ODD}
    for k, (q, c) in enumerate(Ref.read_ff(odd, "odd05", 3, True), 1):
        out[f"odd_qns_{k}"] = q
        out[f"odd_coef_{k}"] = c
    nbas, opcoef, opindx, iops, eops, cops, pops = hoplist_case()
    hl = Ref.get_hoplist(nbas, opcoef, opindx, iops, eops, cops, pops)
    out.update(hl_opid=hl.opid, hl_mptr=hl.mptr, hl_mat=hl.mat,
               hl_coef=hl.coef)
    np.savez_compressed(os.path.join(HERE, "ffhl_ref.npz"), **out)
    print("wrote ffhl")


def spectral_range_oracle(H, rk, npow=10, ncyc=5, seed=1):
    """GetSpectralRange (Solver.f90:563-654) on the FP64 oracle: the probe
    vectors of ``spectral_probes`` (the product code's), ALS_POW3 with shift
    0 for the top and the mid-range shift for the bottom, Rayleigh quotients
    as the bounds, padded by 0.1% of the range.  The two probes run in two
    threads (the oracle's calls release the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    from mlcp_amd.intertwine import spectral_probes
    a, b = spectral_probes(H.nbas, rk, seed)
    lo, hi = C.rayleigh(H, a)[0], C.rayleigh(H, b)[0]

    def step(v, shift):
        c, bb = C.als_pow3(H, v.coef, v.base, npow, 1, shift)
        v.coef, v.base = c, bb
        return C.rayleigh(H, v)[0]

    with ThreadPoolExecutor(2) as ex:
        for i in range(ncyc):
            f1 = ex.submit(step, a, 0.5 * (lo + hi))
            f2 = ex.submit(step, b, 0.0)
            lo, hi = f1.result(), f2.result()
            print(f"  spectral range {i + 1}: [{lo:.6f}, {hi:.6f}]", flush=True)
    pad = 0.001 * (hi - lo)
    return float(lo - pad), float(hi + pad)


def solve_oracle(w, bounds, ncycle, on_cycle=None):
    Q = w.Q0.copy()
    eig = Q.eigv.copy()
    hist, shifts = [], []
    for c in range(ncycle):
        es = block_shift(eig, bounds)
        t0 = time.time()
        Q = C.intertwine(w.H, Q, w.npow, w.nals, True, 1, es)
        eig = Q.eigv.copy()
        hist.append([float(x) for x in eig])
        shifts.append(float(es))
        print(f"  {w.name} cycle {c + 1}: {time.time() - t0:.1f} s, "
              f"E0 {eig[0]:.6f}", flush=True)
        if on_cycle:
            on_cycle(hist, shifts)
    return hist, shifts


# cycles of the oracle's trajectory kept per workload (the 24-mode block
# takes about 45 minutes per oracle cycle with 4 threads)
GOLDEN_CYCLES = {"hh111": 10, "qff12": 10, "qff24": 2}


def workloads(names):
    """Per workload: the spectral bounds of GetSpectralRange on the oracle
    (rank min(50, psirank), 10 power sweeps, 5 cycles; Solver.f90:343-348)
    and the oracle's eigenvalues and E-shifts after each of the first
    GOLDEN_CYCLES cycles, written after every cycle."""
    makers = {"hh111": systems.hh111, "qff12": systems.qff12,
              "qff24": systems.qff24}
    for n in names:
        w = makers[n]()
        path = os.path.join(HERE, f"workload_{n}.json")
        t0 = time.time()
        rk = min(50, int(w.Q0.ranks.max()))
        bounds = spectral_range_oracle(w.H, rk)
        print(f"  {w.name} spectral range {bounds} "
              f"({time.time() - t0:.1f} s)", flush=True)
        out = {"workload": w.name, "units": w.units, "ncycle": w.ncycle,
               "npow": w.npow, "nals": w.nals, "bounds": list(bounds),
               "spectral_rank": rk, "fingerprint": fingerprint(w.H, w.Q0),
               "eshift": [], "eigv": []}

        def write(hist, shifts):
            out["eshift"], out["eigv"] = shifts, hist
            with open(path + ".tmp", "w") as f:
                json.dump(out, f, indent=1)
            os.replace(path + ".tmp", path)

        solve_oracle(w, bounds, GOLDEN_CYCLES[n], write)
        print("wrote", n)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "fixtures"
    if what == "fixtures":
        fixtures()
    elif what == "workloads":
        workloads(sys.argv[2:] or ["hh111", "qff12", "qff24"])
    else:
        raise SystemExit(__doc__)
