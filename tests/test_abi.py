"""The C-ABI library loads and exports every entry point include/cpvs.h declares (CPU only:
no compute call is made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "cpvs.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cpvs_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    assert "cpvs_build" in names and "cpvs_container_evaluate" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from cpvs_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from cpvs_amd import _lib
    assert sorted(_lib.exported_names()) == _declared()


def test_library_is_gfx950_hip():
    data = open(os.path.join(ROOT, "cpvs_amd", "libcpvs_amd.so"), "rb").read()
    assert b"gfx950" in data
    assert b"__hip_fatbin" in data or b".hip_fatbin" in data


def test_oracle_header_and_library():
    import oracle
    lib = oracle.lib()
    for name in ("or_create", "or_traverse_cpu", "or_evaluate", "or_pyramid_build"):
        assert hasattr(lib, name)


def test_status_codes_match_header():
    from cpvs_amd import _lib
    src = open(os.path.join(ROOT, "include", "cpvs.h")).read()
    for name, value in [("CPVS_OK", 0), ("CPVS_EINVAL", 1), ("CPVS_ENOMEM", 2), ("CPVS_EHIP", 3),
                        ("CPVS_EINTERNAL", 4), ("CPVS_ESTATE", 5), ("CPVS_EBUDGET", 6)]:
        assert re.search(rf"#define {name} {value}\b", src)
        assert getattr(_lib, name) == value
