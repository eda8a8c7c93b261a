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
