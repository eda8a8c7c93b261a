"""The C-ABI library: it loads, exports every symbol include/mlcp_hip.h
declares, and the product package never reaches the oracle."""
import ast
import os

from mlcp_amd._lib import declared_symbols, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    L = lib()
    names = declared_symbols()
    assert "intertwinecpp_" in names and "alscpp_" in names
    assert "intwdevicemem_" in names
    for n in names:
        getattr(L, n)


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "mlcp_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            path = os.path.join(dirpath, f)
            if f.endswith(".py"):
                tree = ast.parse(open(path).read())
                for node in ast.walk(tree):
                    if isinstance(node, ast.Import):
                        assert not any(a.name.startswith("oracle")
                                       for a in node.names), path
                    if isinstance(node, ast.ImportFrom):
                        assert not (node.module or "").startswith("oracle"), path
            if f.endswith((".hip", ".h", ".cpp")):
                assert "mlcp_oracle" not in open(path).read(), path


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import mlcp_amd._lib as L
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(L, "_lib", None)
    try:
        L.lib()
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("no error without the library")


# A drop-in call with a rank above the library's limit fails in argument
# validation, before any device call, so these run without a GPU.
_CHILD = r'''
import ctypes as C, sys
import numpy as np
L = C.CDLL(sys.argv[1])
mode = sys.argv[2]
if mode == "return":
    L.mlcp_hip_set_error_mode(0)
L.mlcp_hip_last_error.restype = C.c_char_p
ci = lambda v: C.byref(C.c_int(v))
nbas = np.array([4, 4, 4], dtype=np.int32)
rk, nb = 65, 2
base = np.ones(rk * nb * 12)
coef = np.ones(rk * nb)
eig = np.array([1.0, 2.0])
opid = np.ones(3, dtype=np.int32)
mptr = np.ones(1, dtype=np.int32)
hc = np.ones(1)
mat = np.ones(10)
which = sys.argv[3]
if which == "intertwinecpp_":
    L.intertwinecpp_(base.ctypes, coef.ctypes, nbas.ctypes, ci(rk), ci(3),
                     ci(nb), opid.ctypes, mptr.ctypes, hc.ctypes, mat.ctypes,
                     ci(3), ci(1), ci(10), eig.ctypes, ci(1), ci(1), ci(1),
                     ci(1), ci(1), C.byref(C.c_double(5.0)))
else:
    L.alscpp_(coef.ctypes, base.ctypes, nbas.ctypes, ci(3), ci(rk), ci(rk),
              ci(2))
assert eig.tolist() == [1.0, 2.0] and np.all(base == 1.0) and np.all(coef == 1.0)
print("RETURNED", L.mlcp_hip_last_error().decode())
'''


def _run_child(mode, which):
    import subprocess
    import sys
    from mlcp_amd._lib import LIB_PATH
    env = dict(os.environ)
    env.pop("MLCP_HIP_ON_ERROR", None)
    return subprocess.run([sys.executable, "-c", _CHILD, LIB_PATH, mode,
                           which], capture_output=True, text=True, env=env,
                          timeout=120)


def test_dropin_failure_ends_the_process_like_the_reference():
    """intertwinecpp_/alscpp_ have no return code; the reference exits(1)
    on any error (utils.cu:111-119).  So does this library by default."""
    for which in ("intertwinecpp_", "alscpp_"):
        p = _run_child("default", which)
        assert p.returncode == 1, (which, p.returncode, p.stdout, p.stderr)
        assert "rank" in p.stderr and "RETURNED" not in p.stdout, p.stderr


def test_dropin_failure_returns_when_asked():
    """mlcp_hip_set_error_mode(0): the call returns, host arrays untouched,
    the message in mlcp_hip_last_error()."""
    for which in ("intertwinecpp_", "alscpp_"):
        p = _run_child("return", which)
        assert p.returncode == 0, (which, p.stdout, p.stderr)
        assert "RETURNED" in p.stdout and "rank" in p.stdout, p.stdout
