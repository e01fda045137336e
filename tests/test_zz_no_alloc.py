"""The right-hand side entry point that allocates nothing (GPU).

cfdx_block_rhs_noalloc, called through the C ABI, writes into a caller's device buffer the same
L as cfdx_block_rhs (same kernels: bitwise), on two calls in a row on one device, the second with
a smaller block (a bounded C-grid block, so the gather of owned nodes past the ghost layers is
covered), and a mixed-precision block; a block past the fixed scratch array's size is refused
with CFDX_E_CAPACITY before any work.
"""
import ctypes as C

import numpy as np
import pytest

import cfdx
from cfdx import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


def periodic_block(ctx, n, precision):
    b = cfdx.Block(n, n, n, precision=precision, ctx=ctx)
    b.init_case(**workloads.tgv_case_kw(n))
    prm = cfdx.default_params(mach=0.1, reyl=1600.0)
    b.configure(0.1, 0.0, prm)
    return b, prm


def cgrid_block(ctx):
    d, g = cfdx.cgrid_desc(96, 32, 8, naca=(0.02, 0.4, 0.12), aoa_deg=2.0)
    _, geo = cfdx.cgrid_geometry(d, g)
    b = cfdx.Block(96, 32, 8, ctx=ctx, desc=d)
    b.set_geometry_soa(geo)
    prm = cfdx.default_params(mach=0.5, reyl=1e5)
    b.configure(0.5, 0.0, prm)
    b.set_flow(cfdx.freestream(96 * 32 * 8, 0.5, 2.0))
    b.step(3, 0.5, b.cfl_dt(0.5), prm)          # a state that is not uniform
    return b, prm


def noalloc_rhs(ctx, b, prm):
    """L through the C ABI into a device buffer pre-filled with NaN (every value must be written)."""
    out = ctx.to_device(np.full((5, b.n), np.nan))
    rc = cfdx._lib.cfdx_block_rhs_noalloc(b.h, C.byref(prm), C.c_void_p(out.ptr))
    assert rc == 0, cfdx._lib.cfdx_last_error().decode()
    ctx.sync()
    L = out.numpy()
    out.free()
    return L


def test_noalloc_rhs_equals_rhs_two_calls(ctx):
    print("rhs_noalloc: 64^3 TGV FP64, then a 96x32x8 C-grid block", flush=True)
    for make in (lambda: periodic_block(ctx, 64, cfdx.FP64), lambda: cgrid_block(ctx)):
        b, prm = make()
        want = b.rhs(prm)
        got = noalloc_rhs(ctx, b, prm)
        b.close()
        assert np.isfinite(want).all()
        assert np.array_equal(got, want)


def test_noalloc_rhs_mixed(ctx):
    print("rhs_noalloc: 48^3 TGV mixed", flush=True)
    b, prm = periodic_block(ctx, 48, cfdx.MIXED)
    want = b.rhs(prm)
    got = noalloc_rhs(ctx, b, prm)
    b.close()
    assert np.array_equal(got, want)


def test_noalloc_rhs_refuses_blocks_past_the_scratch_array(ctx):
    b, prm = periodic_block(ctx, 120, cfdx.FP64)      # 1,728,000 stored nodes > 1,677,721
    out = ctx.to_device(np.zeros(8))
    rc = cfdx._lib.cfdx_block_rhs_noalloc(b.h, C.byref(prm), C.c_void_p(out.ptr))
    out.free()
    b.close()
    assert rc == cfdx.E_CAPACITY and "64 MiB" in cfdx._lib.cfdx_last_error().decode()
