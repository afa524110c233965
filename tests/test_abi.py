"""The C-ABI library: builds, loads, exports every declared symbol, and its pure entry points
agree with the host and the oracle.  No device call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from geophyinv import _lib
from geophyinv.fields import FIELDS, field_shape, has_field

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpi.h")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "geophyinv.jl_amd", "csrc")], check=True)
    return _lib.load()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpi_\w+)\s*\(", text)))


def test_header_declares_the_bound_api():
    assert set(declared_functions()) == set(_lib.SIGNATURES), declared_functions()


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (gpi_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    for f in declared_functions():
        assert getattr(lib, f) is not None


def test_abi_version(lib):
    import re
    hdr = open(os.path.join(ROOT, "include", "gpi.h")).read()
    assert lib.gpi_abi_version() == int(re.search(r"#define GPI_ABI_VERSION (\d+)", hdr).group(1))


@pytest.mark.parametrize("physics", [0, 1])
@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_field_shapes_agree(lib, physics, ndim, order):
    n = (37, 29 if ndim == 3 else 1, 41)
    for name in FIELDS:
        fid = FIELDS.index(name)
        out = (C.c_int64 * 3)()
        rc = lib.gpi_field_shape(ndim, order, physics, (C.c_int64 * 3)(*n), fid, out)
        if not has_field(physics, ndim, name):
            assert rc == -1, name
            assert O.field_shape(physics, ndim, order, n, name) is None
            continue
        assert rc == 0, name
        host = field_shape(physics, ndim, order, (n[0], n[2]) if ndim == 2 else n, name)
        host3 = (host[0], 1, host[1]) if ndim == 2 else host
        assert tuple(out) == host3, name
        assert O.field_shape(physics, ndim, order, n, name) == host3, name


def test_field_shape_rejects_bad_arguments(lib):
    out = (C.c_int64 * 3)()
    n = (C.c_int64 * 3)(10, 1, 10)
    assert lib.gpi_field_shape(4, 2, 0, n, 0, out) == -1
    assert lib.gpi_field_shape(2, 3, 0, n, 0, out) == -1
    assert lib.gpi_field_shape(2, 2, 2, n, 0, out) == -1
    assert lib.gpi_field_shape(2, 2, 0, n, 99, out) == -1
