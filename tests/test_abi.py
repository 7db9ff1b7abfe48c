"""The C ABI library loads and exports every entry point include/ddla_native.h declares.
No compute call is made (this runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

from ddla_amd import native


def test_library_present():
    assert os.path.exists(native.LIB_PATH), "run python -m ddla_amd.build"


def test_exports_every_header_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    declared = native.header_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_table_matches_header():
    assert sorted(native.SIGNATURES) == sorted(native.header_symbols())


def test_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for s in native.header_symbols():
        assert s in names  # unmangled


def test_no_device_reports_error_not_crash():
    # without a GPU the init call must fail with a status, never abort the process
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = native.lib()
    st = lib.ddla_init(0)
    assert st in (native.ERR_NO_DEVICE, native.ERR_HIP)
    assert native.last_error()


def test_header_compiles_as_c():
    src = '#include "ddla_native.h"\nint main(void){ddla_mpir_opts o; (void)o; return 0;}\n'
    inc = os.path.join(os.path.dirname(native.HEADER))
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-x", "c", "-", "-o",
                        "/dev/null"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
