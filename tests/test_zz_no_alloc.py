"""mlcp_hip_gep_noalloc (the block eigensolver with its work matrices in the
library's fixed device array) against mlcp_hip_gep (the same eigensolver on
memory the library allocates), through the C-ABI.

Bar: bitwise equal eigenvalues and eigenvectors (one workgroup, the same
kernel and the same inputs), on two calls in a row, the second smaller, so
the second reuses the array the first left behind.  Also: sizes beyond the
array are refused before any launch, and a failing case reports its error
bit in *dinfo without writing the outputs.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _problem(seed, n):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n))
    A = a + a.T
    m = rng.normal(size=(n, n))
    B = m @ m.T / n + np.eye(n)
    return A, B


class _Hip:
    """The caller's device memory, from the HIP runtime the library itself
    uses (libamdhip64 of /opt/rocm, already loaded by the library)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(arr.nbytes)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p),
                                 C.c_size_t(arr.nbytes), 1) == 0   # H2D
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p,
                                 C.c_size_t(out.nbytes), 2) == 0   # D2H
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(L, A, B, fill_x=None):
    n = A.shape[0]
    h = _Hip()
    try:
        dA = h.put(A.T)
        dB = h.put(B.T)
        x0 = np.full(n * n, np.nan if fill_x is None else fill_x)
        w0 = np.full(n, -7.0)
        i0 = np.full(1, -1, dtype=np.int32)
        dX, dw, di = h.put(x0), h.put(w0), h.put(i0)
        rc = L.mlcp_hip_gep_noalloc(n, dA, dB, dX, dw, di, None)
        assert rc == 0, L.mlcp_hip_last_error().decode()
        X = h.get(dX, x0).reshape(n, n).T.copy()
        return int(h.get(di, i0)[0]), h.get(dw, w0), X
    finally:
        h.free()


def test_gep_noalloc_matches_gep(hip):
    from mlcp_amd.intertwine import gpu_gep
    L = hip
    for n, seed in ((150, 11), (37, 12)):
        say("gep_noalloc", n)
        A, B = _problem(seed, n)
        w_ref, X_ref = gpu_gep(A, B)
        info, w, X = _noalloc(L, A, B)
        assert info == 0
        assert np.array_equal(w, w_ref), np.max(np.abs(w - w_ref))
        assert np.array_equal(X, X_ref), np.max(np.abs(X - X_ref))
        # and it is the generalized eigenproblem's solution
        r = A @ X - B @ X * w
        assert np.max(np.abs(r)) < 1e-9 * np.abs(w).max()


def test_gep_noalloc_refuses_and_reports(hip):
    L = hip
    say("gep_noalloc refusal", 257)
    h = _Hip()
    try:
        z = h.put(np.zeros(4))
        rc = L.mlcp_hip_gep_noalloc(257, z, z, z, z, z, None)
    finally:
        h.free()
    assert rc == -1 and "256" in L.mlcp_hip_last_error().decode()
    say("gep_noalloc, B not positive definite")
    A, B = _problem(13, 20)
    B[3, 3] = -5.0
    info, w, X = _noalloc(L, A, B, fill_x=3.0)
    assert info == 4
    assert np.all(X == 3.0) and np.all(w == -7.0)
