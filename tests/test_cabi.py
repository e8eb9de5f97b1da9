"""The drop-in C-ABI library loads and exports every entry point its headers declare (no GPU needed).

The symbol diff: tests/golden/honee_ceed_calls.txt lists every libCEED function and CEED_* constant HONEE's sources
name (written from the reference's src/, problems/, examples/ and include/ by tests/golden/make_ceed_call_list.py);
each function must be a dynamic symbol of libceed.so (nm -D) and each constant must be declared by include/ceed.h,
include/ceed/types.h or include/ceed/backend.h.
"""
import ctypes as ct
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "honee-4aiad_amd", "lib")


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"//.*", "", txt)
    names = set(re.findall(r"^\s*(?:int|Ceed|const char \*|CeedOperator)\s+(Ceed\w+|HoneeMi\w+)\s*\(", txt, re.M))
    return sorted(names)


@pytest.mark.parametrize("header", ["ceed.h", "ceed/backend.h", "honee_mi355x.h"])
def test_libceed_exports_every_declared_symbol(header):
    lib = ct.CDLL(os.path.join(LIB, "libceed.so"), mode=ct.RTLD_GLOBAL)
    names = _declared(header)
    assert len(names) > {"ceed.h": 90, "ceed/backend.h": 2, "honee_mi355x.h": 5}[header]
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


def _honee_names():
    rows = [l.split() for l in open(os.path.join(ROOT, "tests", "golden", "honee_ceed_calls.txt")) if l.strip() and not l.startswith("#")]
    return sorted(n for k, n in rows if k == "function"), sorted(n for k, n in rows if k == "constant")


def test_every_libceed_function_honee_calls_is_exported():
    funcs, _ = _honee_names()
    assert len(funcs) > 90
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIB, "libceed.so")], check=True, capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [n for n in funcs if n not in exported]
    assert not missing, missing


def test_every_constant_honee_names_is_declared():
    _, consts = _honee_names()
    txt = "".join(open(os.path.join(ROOT, "include", h)).read() for h in ("ceed.h", "ceed/types.h", "ceed/backend.h"))
    missing = [n for n in consts if not re.search(r"\b" + n + r"\b", txt)]
    assert not missing, missing


def test_init_failure_message_without_a_ceed():
    lib = ct.CDLL(os.path.join(LIB, "libceed.so"), mode=ct.RTLD_GLOBAL)
    ceed = ct.c_void_p()
    assert lib.CeedInit(b"/gpu/hip/mi355x:scatter=colored", ct.byref(ceed)) != 0
    msg = ct.c_char_p()
    assert lib.CeedGetErrorMessage(None, ct.byref(msg)) == 0
    assert msg.value and b"scatter" in msg.value
