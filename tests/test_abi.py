"""The C ABI library loads and exports every entry point include/faust_hip.h declares (no GPU calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="faust_hip.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    # functions only: type names (fh_..._t) in function-pointer members are not symbols
    return sorted(n for n in set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", txt)) if not n.endswith("_t"))


def test_header_symbols_exported():
    from faust_amd import _lib
    lib = C.CDLL(_lib.LIB_PATH)
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["faust_gm.h", "faust_hip.h"]
    names = [n for h in headers for n in _declared(h)]
    assert len(names) >= 38
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from faust_amd import _lib
    assert set(_declared()) == set(_lib.exported_names())


def test_no_cpu_fallback_without_library(tmp_path, monkeypatch):
    from faust_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    import pytest
    with pytest.raises(ImportError):
        _lib.lib()


def test_version_and_error_string():
    from faust_amd import _lib
    L = _lib.lib()
    assert b"gfx950" in L.fh_version()
    assert L.fh_last_error() is not None


def test_argument_errors_without_gpu():
    # argument validation happens before any device call
    from faust_amd import _lib
    L = _lib.lib()
    assert L.fh_faust_size(None, None) == _lib.FH_ERR_ARG
    assert b"null" in L.fh_last_error()


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus 2 exits with a message, before any GPU work."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
