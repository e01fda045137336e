"""The C ABI: libcfdx.so loads and exports every function include/cfdx.h declares, the
layouts match the reference's packed structs, and argument errors come back as codes
(CPU only: nothing here launches a kernel)."""
import ctypes as C
import os
import re

import numpy as np

import cfdx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "cfdx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cfdx_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) > 30
    lib = C.CDLL(cfdx.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(cfdx.EXPORTS), set(names) ^ set(cfdx.EXPORTS)


def test_layouts():
    assert cfdx.FLOW_DT.itemsize == 72 and cfdx.GRID_DT.itemsize == 56
    # 7 doubles, precision / viscous / sgs / pad (int32), c_s
    assert C.sizeof(cfdx.Params) == 7 * 8 + 4 * 4 + 8 and cfdx.Params.c_s.offset == 72
    assert C.sizeof(cfdx.BlockDesc) == 64 and cfdx.BlockDesc.aoa.offset == 56   # 13 int32, pad, double
    assert cfdx.GRID_DT.fields["N0"][1] == 24 and cfdx.GRID_DT.fields["loc"][1] == 52


def test_default_params_are_reference_constants():
    p = cfdx.default_params()
    assert (p.mach, p.reyl, p.pr, p.free_t, p.c_mts, p.c_t) == (1.5, 500.0, 0.72, 273.0, 0.03, 10.0)
    assert p.eps == 10e-8 and p.precision == cfdx.FP64 and p.viscous == 1


def test_error_codes_without_device():
    L = cfdx._lib
    assert L.cfdx_geom_periodic_table(0, 4, 4, None, None) == cfdx.E_INVALID
    assert b"periodic_table" in L.cfdx_last_error()
    assert L.cfdx_uq(None, None, None, None, None, None, None, 10) == cfdx.E_INVALID
    assert L.cfdx_block_step(None, 0.5, 0.01, 1, None) == cfdx.E_INVALID
    assert L.cfdx_math_eval(None, 0, None, None, 4) == cfdx.E_INVALID
    d = cfdx.BlockDesc(0, 4, 4, 7, 0)
    assert L.cfdx_block_bytes(C.byref(d)) == -1
    d = cfdx.BlockDesc(64, 64, 64, 7, 0)
    assert L.cfdx_block_bytes(C.byref(d)) > 64 ** 3 * 8 * 50


def test_device_count_is_safe_without_gpu():
    assert cfdx.device_count() >= 0
