"""CPU checks of the C ABI: the HIP library loads and exports every function
include/kapsel_gpu.h declares, and fails cleanly (no exit, no crash) when no
gfx950 device is present.  No compute call is made here."""
import ctypes
import os
import subprocess

import pytest

from icep_kapsel_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    declared = abi.declared_symbols()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [d for d in declared if d not in exported]
    assert not missing, missing


def test_library_is_gfx950_only():
    out = subprocess.run(["strings", abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_ctypes_struct_layout_matches_header():
    # kapsel_params / kapsel_time sizes as compiled by the C compiler
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "kapsel_gpu.h"
int main(void){printf("%zu %zu %zu %zu\n", sizeof(kapsel_params), offsetof(kapsel_params, precision),
  offsetof(kapsel_params, dielectric_cst), sizeof(kapsel_time)); return 0;}
'''
    tmp = "/tmp/kapsel_layout_check"
    with open(tmp + ".c", "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), tmp + ".c", "-o", tmp], check=True)
    vals = [int(v) for v in subprocess.run([tmp], capture_output=True, text=True).stdout.split()]
    assert vals == [ctypes.sizeof(abi.Params), abi.Params.precision.offset, abi.Params.dielectric_cst.offset,
                    ctypes.sizeof(abi.Time)]


def test_create_without_device_or_with_bad_params_returns_error():
    lib = abi.load_library()
    ctx = ctypes.c_void_p()
    bad = abi.make_params(24)  # not a power of two
    rc = lib.kapsel_create(ctypes.byref(bad), ctypes.byref(ctx))
    assert rc == -1 and b"power of two" in lib.kapsel_last_error()
    p = abi.make_params(32)
    rc = lib.kapsel_create(ctypes.byref(p), ctypes.byref(ctx))
    if rc == 0:
        pytest.skip("a GPU is present; the no-device path is not reachable")
    assert rc in (-5, -2)
    assert lib.kapsel_last_error()


def test_null_context_is_rejected():
    lib = abi.load_library()
    assert lib.kapsel_synchronize(None) == -1
    assert lib.kapsel_fft_r2c(None, 0) == -1
