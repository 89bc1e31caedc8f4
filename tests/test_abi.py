"""The C-ABI library without a GPU: it loads, exports every function that
include/gigpu.h declares, answers the host-only calls, and the Fortran
ISO_C_BINDING interface (gammcor-integrals_amd/fortran) compiles, links and
runs against it."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gigpu.h")
LIBDIR = os.path.join(ROOT, "gammcor-integrals_amd", "gigpu")
LIB = os.path.join(LIBDIR, "libgigpu.so")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gi_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libgigpu.so is not built (run __graft_entry__.build())")
    return C.CDLL(LIB)


def test_exports_every_declared_function(lib):
    names = header_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_only_calls(lib):
    lib.gi_chol_accuracy_tau.argtypes = [C.c_int64, C.POINTER(C.c_double)]
    lib.gi_last_error.restype = C.c_char_p
    tau = C.c_double()
    for acc, ref in ((1, 1e-5), (2, 1e-6), (3, 1e-7), (4, 1e-10)):
        assert lib.gi_chol_accuracy_tau(acc, C.byref(tau)) == 0
        assert tau.value == ref
    assert lib.gi_chol_accuracy_tau(0, C.byref(tau)) != 0
    assert lib.gi_last_error() == b"Invalid accuracy setting"


def test_fortran_interface(tmp_path):
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        pytest.skip("no Fortran compiler")
    src = os.path.join(ROOT, "gammcor-integrals_amd", "fortran", "gigpu_interface.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "abi_check.f90")
    r = subprocess.run([fc, "-c", src], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exe = str(tmp_path / "abi_check")
    r = subprocess.run([fc, prog, "gigpu_interface.o", "-L" + LIBDIR, "-lgigpu", "-Wl,-rpath," + LIBDIR,
                        "-Wl,-rpath,/opt/rocm/lib", "-o", exe], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fortran abi ok" in r.stdout


def test_fortran_interface_declares_every_function():
    src = open(os.path.join(ROOT, "gammcor-integrals_amd", "fortran", "gigpu_interface.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(gi_[a-z0-9_]+)"\)', src))
    missing = [n for n in header_functions() if n not in bound]
    assert not missing, missing
