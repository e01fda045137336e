"""Residual monitor and the iteration driver (GPU)."""
import math

import numpy as np
import pytest

import cfdx
from cfdx import driver

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p")


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


def tgv(ctx, precision=cfdx.FP64, n=(20, 16, 12)):
    b = cfdx.Block(*n, precision=precision, ctx=ctx)
    b.init_case(grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05 * 2 * math.pi / n[0], phases=(0.2, 0.9, 1.7),
                init="tgv", mach=0.3)
    return b


@pytest.mark.parametrize("precision", [cfdx.FP64, cfdx.MIXED], ids=["fp64", "mixed"])
def test_monitor_matches_host_difference(ctx, precision):
    prm = cfdx.default_params(mach=0.3, reyl=800.0)
    b = tgv(ctx, precision)
    dt = b.cfl_dt(0.5)
    b.set_monitor(True)
    b.set_tiles((0, 1, 2, 3))   # partial tiles along i with the (256 x 1) node tile
    for _ in range(3):
        before = b.get_flow()
        b.step(1, 0.3, dt, prm)
        after = b.get_flow()
        du, dv, de = b.residual()
        e0 = before["p"] / (0.4 * before["rho"])
        e1 = after["p"] / (0.4 * after["rho"])
        want = (np.abs(after["u"] - before["u"]).max(), np.abs(after["v"] - before["v"]).max(), np.abs(e1 - e0).max())
        tol = 1e-9 if precision == cfdx.FP64 else 1e-4
        for got, w in zip((du, dv, de), want):
            assert abs(got - w) <= tol * w, (got, w)


def test_monitor_bounded_block(ctx):
    """A walled box: the shell kernels' monitor slots take part. The residual covers every owned node
    (side nodes with their value before the side conditions), so it is at least the host difference
    over the nodes the side conditions leave alone; it does not depend on the node tile shape."""
    prm = cfdx.default_params(mach=0.4, reyl=800.0)
    d = cfdx.BlockDesc(14, 12, 10, 0, cfdx.FP64)
    for q, bc in enumerate((1, 3, 1, 2, 4, 4)):
        d.bc[q] = bc
    xyz, geo = cfdx.box_geometry(d, (1.3, 1.1, 0.9), 0.02, (0.3, 1.2, 2.1))
    own = cfdx.owned_index(d)
    x, y = xyz[own, 0], xyz[own, 1]
    flow = cfdx.freestream(own.size, 0.4, 0.0)
    flow["u"] *= 1 + 0.2 * np.exp(-((y - 0.15) ** 2) / 0.01) * np.sin(3 * x)   # a bump near the j- wall
    res = []
    for tiles in ((0, 0, 0, 0), (0, 0, 0, 3)):
        b = cfdx.Block(14, 12, 10, ctx=ctx, desc=cfdx.BlockDesc.from_buffer_copy(d))
        b.set_geometry_soa(geo)
        b.configure(0.4, 0.0, prm)
        b.set_flow(flow)
        dt = 0.5 * b.cfl_dt(0.5)
        b.set_tiles(tiles)
        b.set_monitor(True)
        before = b.get_flow()
        b.step(1, 0.4, dt, prm)
        after = b.get_flow()
        res.append(b.residual())
        b.close()
    assert res[0] == res[1]
    dims, off = cfdx.padded(d)
    i, j, k = (own % dims[0]) - off[0], ((own // dims[0]) % dims[1]) - off[1], (own // (dims[0] * dims[1])) - off[2]
    inner = (i > 0) & (i < 13) & (j > 0) & (j < 11) & (k > 0) & (k < 9)
    e0 = before["p"] / (0.4 * before["rho"])
    e1 = after["p"] / (0.4 * after["rho"])
    host = (np.abs(after["u"] - before["u"])[inner].max(), np.abs(after["v"] - before["v"])[inner].max(),
            np.abs(e1 - e0)[inner].max())
    for got, w in zip(res[0], host):
        assert np.isfinite(got) and got >= w * (1 - 1e-9), (got, w)


def test_driver_restart_resume(ctx, tmp_path):
    prm = cfdx.default_params(mach=0.3, reyl=800.0)
    ref = tgv(ctx)
    dt = ref.cfl_dt(0.5)
    lines = []
    r = driver.Run(ref, 0.3, dt, prm, outdir=str(tmp_path / "a"), restart_every=2, tecplot_every=3,
                   xyz=np.zeros((ref.n, 3)), log=lines.append)
    r.run(5)
    assert len(r.history) == 5 and lines[0].startswith("iteration = 1 max_div_u = ")
    want = ref.get_flow()
    # restart files in turn: slot 1 at step 2, slot 2 at step 4
    assert cfdx.restart_latest(str(tmp_path / "a"), ref.n) == (2, 4)
    # resume from step 4 in a new block and finish step 5: the same state bit for bit
    b = tgv(ctx)
    r2 = driver.Run(b, 0.3, dt, prm, outdir=str(tmp_path / "a"), restart_every=2, log=lines.append)
    assert r2.resume() == 4
    r2.run(5)
    got = b.get_flow()
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f
    assert (tmp_path / "a" / "nodefile_3.dat").exists()


def test_nonfinite_check(ctx):
    """With the check on, a step whose state turns NaN (a zero density in the initial state) returns
    CFDX_E_NONFINITE; a healthy block steps on."""
    prm = cfdx.default_params(mach=0.3, reyl=800.0)
    b = tgv(ctx)
    dt = b.cfl_dt(0.5)
    b.set_check(True)
    b.step(2, 0.3, dt, prm)
    flow = b.get_flow()
    flow["rho"][100] = 0.0
    b.set_flow(flow)
    with pytest.raises(cfdx.CfdxError) as e:
        b.step(1, 0.3, dt, prm)
    assert e.value.code == cfdx.E_NONFINITE and "after step 3" in str(e.value)
    b.close()


def test_params_refusals(ctx):
    """Steps refuse an unknown SGS switch, a negative Smagorinsky constant and the Euler switch
    while a model is on (CFDX_E_INVALID, message, nothing launched)."""
    b = tgv(ctx)
    dt = b.cfl_dt(0.5)
    bad = [cfdx.default_params(sgs=7), cfdx.default_params(sgs="smagorinsky", c_s=-0.1),
           cfdx.default_params(reyl=math.inf, c_mts=0.0, viscous=0, sgs="smagorinsky", c_s=0.17)]
    for prm in bad:
        with pytest.raises(cfdx.CfdxError) as e:
            b.step(1, 0.3, dt, prm)
        assert e.value.code == cfdx.E_INVALID
    # the Euler switch with the Smagorinsky model off is accepted
    b.step(1, 0.3, dt, cfdx.default_params(reyl=math.inf, viscous=0, sgs="smagorinsky", c_s=0.0))
    b.close()
    with pytest.raises(TypeError):
        cfdx.default_params(cs=0.2)
