"""The drop-in C-ABI library loads and exports every entry point its headers declare (no GPU needed)."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "honee-4aiad_amd", "lib")


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"//.*", "", txt)
    names = set(re.findall(r"^\s*(?:int|Ceed|const char \*|CeedOperator)\s+(Ceed\w+|HoneeMi\w+)\s*\(", txt, re.M))
    return sorted(names)


@pytest.mark.parametrize("header", ["ceed.h", "honee_mi355x.h"])
def test_libceed_exports_every_declared_symbol(header):
    lib = ct.CDLL(os.path.join(LIB, "libceed.so"), mode=ct.RTLD_GLOBAL)
    names = _declared(header)
    assert len(names) > (90 if header == "ceed.h" else 5)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_sentinels_exported():
    lib = ct.CDLL(os.path.join(LIB, "libceed.so"), mode=ct.RTLD_GLOBAL)
    for n in ("CEED_VECTOR_ACTIVE", "CEED_VECTOR_NONE", "CEED_BASIS_NONE", "CEED_ELEMRESTRICTION_NONE", "CEED_QFUNCTION_NONE",
              "CEED_STRIDES_BACKEND", "CEED_REQUEST_IMMEDIATE"):
        assert ct.c_void_p.in_dll(lib, n) is not None


def test_setup_driver_exports():
    ct.CDLL(os.path.join(LIB, "libceed.so"), mode=ct.RTLD_GLOBAL)
    lib = ct.CDLL(os.path.join(LIB, "libhonee_setup.so"))
    txt = open(os.path.join(ROOT, "honee-4aiad_amd", "csrc", "honee_setup.h")).read()
    names = re.findall(r"\b(HoneeMi\w+)\s*\(", txt)
    assert names
    for n in names:
        assert hasattr(lib, n), n


def test_unsupported_resource_is_an_error_not_a_crash():
    lib = ct.CDLL(os.path.join(LIB, "libceed.so"), mode=ct.RTLD_GLOBAL)
    ceed = ct.c_void_p()
    assert lib.CeedInit(b"/cpu/self", ct.byref(ceed)) != 0
