"""The C ABI: the library loads, exports every function include/octhip.h declares, the
ctypes signatures agree with the header, and calls fail cleanly before initialisation
(no device computation happens in these tests)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from octhip import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "octhip.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"\bint\s+(octhip_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.S):
        args = [a.strip() for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        decls[m.group(1)] = args
    return decls


def test_header_and_binding_agree():
    d = declared()
    assert set(d) == set(_lib.SIGNATURES), set(d) ^ set(_lib.SIGNATURES)
    for name, args in d.items():
        assert len(args) == len(_lib.SIGNATURES[name]), name


def test_library_exports_every_declared_symbol():
    assert _lib.LIB_PATH.exists(), "build the library first"
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = set(declared()) - exported
    assert not missing, missing
    lib = _lib.lib()
    for name in declared():
        assert getattr(lib, name) is not None


def test_calls_before_init_fail_cleanly():
    lib = _lib.lib()
    if _lib._initialised:
        pytest.skip("library already initialised in this process")
    p = C.c_void_p()
    assert lib.octhip_malloc(C.byref(p), 16) == _lib.ERR_STATE
    assert "octhip_init" in _lib.last_error()
    h = C.c_void_p()
    assert lib.octhip_nl_operator_create(25, None, None, 1, None, None, 1, 1, None, C.byref(h)) == _lib.ERR_STATE
    maj, mnr = C.c_int(), C.c_int()
    assert lib.octhip_version(C.byref(maj), C.byref(mnr)) == 0
    assert lib.octhip_finalize() == 0


def test_errors_raise():
    with pytest.raises(_lib.OcthipError):
        _lib.check(_lib.ERR_INVALID)


def test_product_does_not_import_oracle():
    pkg = ROOT / "octopus-debian-package_amd"
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.h")):
        if f.name == "build.py":
            continue
        t = f.read_text()
        assert "import oracle" not in t and "liboracle" not in t, f
