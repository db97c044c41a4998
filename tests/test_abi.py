"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes as C
import glob
import os
import re

import pytest

from katric_amd import _ffi
from tests.conftest import ROOT


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(kt_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("kt_create", "kt_destroy", "kt_count", "kt_orient", "kt_strerror", "kt_gen_rmat"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    lib = C.CDLL(_ffi.LIB_PATH)
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    assert name in _ffi.SIGNATURES, f"{name} has no ctypes signature"


def test_status_strings_and_no_device_behaviour():
    lib = _ffi.lib()
    assert lib.kt_strerror(0) == b"ok"
    assert b"invalid" in lib.kt_strerror(-1)
    assert lib.kt_strerror(-99) == b"unknown status"
    # null arguments are rejected before any device work
    assert lib.kt_count(None, 0, None, None, 0, None, None) == -1
    assert lib.kt_create(None, 0, 0) == -1


def test_device_code_targets_gfx950():
    from katric_amd import build as kb

    targets = kb.bundle_targets(_ffi.LIB_PATH)
    assert targets and all(t.endswith("--gfx950") for t in targets), targets


def test_no_oracle_in_product():
    """The product library and package never reference the oracle."""
    blob = open(_ffi.LIB_PATH, "rb").read()
    assert b"or_run_shmem" not in blob and b"liboracle" not in blob
    for f in glob.glob(os.path.join(ROOT, "katric_amd", "**", "*.py"), recursive=True):
        src = open(f).read()
        assert "import oracle" not in src and "from oracle" not in src, f
