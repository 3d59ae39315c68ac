"""The C-ABI library loads and exports every symbol include/hpup.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import hpup

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(REPO, "include", "hpup.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hpup_[a-zA-Z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hpup.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(hpup.EXPORTS) == names


def test_version_and_method_parsing():
    lib = hpup.lib()
    assert lib.hpup_version() >= 100
    for name, v in [(b"euler", 0), (b"Midpoint", 1), (b"rk2", 2), (b"RK4", 3), (b"heun", -1)]:
        assert lib.hpup_parse_method(name) == v
