"""Device-side case set-up and the serving loop (GPU).

- a block initialised on the device (geometry of the map, half-point metrics, initial state)
  gives the same right-hand side as the same case built on the host and uploaded;
- the device CFL time step equals the host formula;
- every request served in a mix of uneven, staggered arrivals with a memory budget that forces
  queueing gives bitwise the result it gets alone (plain FP64 and mixed).
"""
import math

import numpy as np
import pytest

import cfdx
from cfdx import serve

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p")


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


CASES = [
    dict(ni=16, nj=12, nk=10, grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05 * 2 * math.pi / 16,
         phases=(0.3, 1.1, 2.0), init="tgv", mach=0.1),
    dict(ni=12, nj=14, nk=8, grid="box", length=(10.0,) * 3, init="vortex", mach=0.5, beta=5.0, center=(5.0, 5.0)),
    dict(ni=10, nj=10, nk=12, grid="sinusoid", length=(10.0, 10.0, 10.0), amp=0.2, phases=(2.0, 0.5, 1.0),
         init="vortex", mach=0.3, beta=5.0, center=(4.0, 6.0)),
]


@pytest.mark.parametrize("kw", CASES, ids=["tgv-sinusoid", "vortex-box", "vortex-sinusoid"])
def test_device_init_matches_host(ctx, kw):
    ni, nj, nk = kw["ni"], kw["nj"], kw["nk"]
    hk = {k: v for k, v in kw.items() if k not in ("ni", "nj", "nk")}
    case = cfdx.periodic_case(ni, nj, nk, **hk)
    prm = cfdx.default_params(mach=kw["mach"], reyl=1600.0)
    host = cfdx.Block(ni, nj, nk, ctx=ctx)
    host.set_case(case)
    dev = cfdx.Block(ni, nj, nk, ctx=ctx)
    dev.init_case(**hk)
    fh, fd = host.get_flow(), dev.get_flow()
    for f in FIELDS:
        scale = max(np.abs(fh[f]).max(), 1e-300)
        assert np.abs(fd[f] - fh[f]).max() <= 1e-14 * scale, f
    # the inputs differ in the last bits of sin / cos (device vs host libm). The RHS is a difference
    # of face fluxes up to ~(E + p)|u| ~ 1e3 |L| at Mach 0.1, so an input difference of a few ulp
    # shows up at ~1e-12 of max |L|; the bound is 1e-11 of the largest RHS component.
    Lh, Ld = host.rhs(prm), dev.rhs(prm)
    scale = np.abs(Lh).max()
    for c in range(5):
        err = np.abs(Ld[c] - Lh[c]).max() / scale
        assert err <= 1e-11, (c, err)
    # the same stable step (convective + viscous radius) on both sides, with the block's parameters
    dev.configure(kw["mach"], 0.0, prm)
    dt_host = cfdx.cfl_dt(case, 0.5, prm)
    dt_dev = dev.cfl_dt(0.5)
    assert abs(dt_dev - dt_host) <= 1e-13 * dt_host, (dt_dev, dt_host)


def small_stream():
    sizes = [(16, 12, 10), (12, 12, 12), (20, 10, 8), (10, 14, 12), (16, 16, 8), (12, 10, 14)]
    reqs = []
    for rid, (ni, nj, nk) in enumerate(sizes):
        tgv = rid % 2 == 0
        reqs.append(serve.Request(rid, ni, nj, nk, nsteps=6 + 2 * rid, arrival=3 * rid,
                                  grid="sinusoid", length=(2 * math.pi,) * 3 if tgv else (10.0,) * 3,
                                  phases=(0.1 * rid, 0.7, 1.3), init="tgv" if tgv else "vortex",
                                  mach=0.1 + 0.08 * rid, reyl=500.0 + 700.0 * rid, viscous=rid != 3))
    return reqs


@pytest.mark.parametrize("precision", [cfdx.FP64, cfdx.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_serving_alone_vs_mix(ctx, precision, graph):
    reqs = small_stream()
    # a budget of three of the largest blocks forces queueing and admission as slots free up
    big = max(cfdx.Block.bytes_for(r.ni, r.nj, r.nk, precision) for r in reqs)
    srv = serve.Server(precision=precision, ctx=ctx, reserve_bytes=0, budget_bytes=3 * big, chunk=4, graph=graph)
    res = srv.run(reqs, clock="steps")
    srv.close()
    assert sorted(res) == [r.rid for r in reqs]
    assert 2 <= srv.peak_active <= 3
    for r in reqs:
        alone = serve.run_alone(r, precision, ctx)
        assert res[r.rid].dt == alone.dt
        for f in FIELDS:
            assert np.array_equal(res[r.rid].flow[f], alone.flow[f]), (r.rid, f)
        assert all(np.isfinite(res[r.rid].flow[f]).all() for f in FIELDS)
    rep = serve.report(reqs, res)
    assert rep["p99_s"] >= rep["p50_s"] > 0


def test_serving_sjf_admits_least_work_first(ctx):
    """policy "sjf": of requests queued together, the one with the least work (nodes x steps) is
    admitted first; results stay bitwise the request alone."""
    reqs = small_stream()
    for r in reqs:
        r.arrival = 0
    # the largest block's bytes: about one block at a time, so admission order shows
    big = max(cfdx.Block.bytes_for(r.ni, r.nj, r.nk, cfdx.FP64) for r in reqs)
    srv = serve.Server(precision=cfdx.FP64, ctx=ctx, reserve_bytes=0, budget_bytes=big, chunk=4, policy="sjf")
    res = srv.run(reqs, clock="steps")
    srv.close()
    work = {r.rid: r.nodes * r.nsteps for r in reqs}
    for a in reqs:
        for b in reqs:
            if work[a.rid] < work[b.rid]:
                assert res[a.rid].admitted <= res[b.rid].admitted, (a.rid, b.rid)
    assert len({res[r.rid].admitted for r in reqs}) >= 3
    for r in reqs:
        alone = serve.run_alone(r, cfdx.FP64, ctx)
        for f in FIELDS:
            assert np.array_equal(res[r.rid].flow[f], alone.flow[f]), (r.rid, f)
    with pytest.raises(ValueError):
        serve.Server(precision=cfdx.FP64, ctx=ctx, policy="lifo")


@pytest.mark.parametrize("clock", ["steps", "wall"])
def test_serving_failure_releases_blocks(ctx, clock):
    """A request that fails in its set-up while another is on the GPU: run() raises, every block
    of the run is released and the batch is empty, and the same server then serves a valid
    stream with the results of each request alone (wall clock: the set-up fails on the set-up
    thread)."""
    reqs = small_stream()[:2]
    reqs[1].arrival = 2 if clock == "steps" else 0.02
    reqs[1].init = "no-such-state"
    srv = serve.Server(precision=cfdx.FP64, ctx=ctx, reserve_bytes=0, chunk=2)
    with pytest.raises(Exception):
        srv.run(reqs, clock=clock)
    assert len(srv.batch) == 0
    good = small_stream()[2:4]
    if clock == "wall":
        for r in good:
            r.arrival = 0.0
    res = srv.run(good, clock=clock)
    srv.close()
    for r in good:
        alone = serve.run_alone(r, cfdx.FP64, ctx)
        for f in FIELDS:
            assert np.array_equal(res[r.rid].flow[f], alone.flow[f]), (r.rid, f)


def test_serving_wall_clock(ctx):
    reqs = small_stream()
    for r in reqs:
        r.arrival = 0.01 * r.rid
    srv = serve.Server(precision=cfdx.MIXED, ctx=ctx, reserve_bytes=0, chunk=3)
    res = srv.run(reqs, clock="wall")
    srv.close()
    rep = serve.report(reqs, res)
    print("serving (wall clock):", rep)
    assert rep["requests"] == len(reqs) and rep["cell_updates_per_s"] > 0
    assert all(res[r.rid].finished >= res[r.rid].admitted >= r.arrival - 1e-9 for r in reqs)
    # wall clock: blocks set up, read back and released on the set-up thread beside the chunks
    for r in reqs:
        alone = serve.run_alone(r, cfdx.MIXED, ctx)
        for f in FIELDS:
            assert np.array_equal(res[r.rid].flow[f], alone.flow[f]), (r.rid, f)


def test_serving_wall_clock_set_up_ahead(ctx):
    """Wall clock with a node cap of about two requests: blocks are set up ahead on the set-up
    thread (up to twice the cap) and join the batch when the cap allows; every result is bitwise
    the request alone and the batch never holds more than the cap (or one request)."""
    reqs = small_stream()
    for r in reqs:
        r.arrival = 0.0
    cap = 2 * max(r.nodes for r in reqs)
    srv = serve.Server(precision=cfdx.FP64, ctx=ctx, reserve_bytes=0, chunk=2, max_nodes=cap)
    res = srv.run(reqs, clock="wall")
    srv.close()
    assert sorted(res) == [r.rid for r in reqs]
    assert 2 <= srv.peak_active < len(reqs)
    for r in reqs:
        assert res[r.rid].finished >= res[r.rid].admitted >= 0.0
        alone = serve.run_alone(r, cfdx.FP64, ctx)
        assert res[r.rid].dt == alone.dt
        for f in FIELDS:
            assert np.array_equal(res[r.rid].flow[f], alone.flow[f]), (r.rid, f)


def test_serving_airfoils_alone_vs_mix(ctx):
    """NACA C-grid requests (geometry built on the host by the worker threads as they arrive) and
    periodic requests in one batch: each result is bitwise the request alone."""
    reqs = [
        serve.Request(0, 48, 16, 4, nsteps=8, arrival=0, grid="cgrid", naca=(0.02, 0.4, 0.12), aoa_deg=2.0,
                      mach=0.5, reyl=2e5),
        serve.Request(1, 16, 12, 10, nsteps=6, arrival=2, grid="sinusoid", length=(2 * math.pi,) * 3,
                      phases=(0.2, 0.7, 1.3), init="tgv", mach=0.2, reyl=800.0),
        serve.Request(2, 56, 20, 6, nsteps=10, arrival=3, grid="cgrid", naca=(0.0, 0.4, 0.15), aoa_deg=-3.0,
                      mach=0.4, reyl=5e5),
        serve.Request(3, 48, 16, 8, nsteps=5, arrival=7, grid="cgrid", naca=(0.04, 0.4, 0.09), aoa_deg=8.0,
                      mach=0.65, reyl=1e6),
    ]
    srv = serve.Server(precision=cfdx.MIXED, ctx=ctx, reserve_bytes=0, chunk=3, prep_workers=2)
    res = srv.run(reqs, clock="steps")
    srv.close()
    assert sorted(res) == [r.rid for r in reqs] and srv.peak_active >= 2
    for r in reqs:
        alone = serve.run_alone(r, cfdx.MIXED, ctx)
        assert res[r.rid].dt == alone.dt
        for f in FIELDS:
            assert np.array_equal(res[r.rid].flow[f], alone.flow[f]), (r.rid, f)
            assert np.isfinite(res[r.rid].flow[f]).all()


def test_block_memory_pool(ctx):
    """Destroyed blocks return their memory to a per-device pool (no hipFree, which waits for the
    device); mem_info counts it as free, a new block of the same size reuses it, zeroed and
    giving the same results, and pool_trim frees it."""
    ctx.pool_trim()
    f0, _ = ctx.mem_info()
    need = cfdx.Block.bytes_for(64, 64, 64, cfdx.FP64)

    def run():
        b = cfdx.Block(64, 64, 64, ctx=ctx)
        b.init_case(grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.01, phases=(0.1, 0.2, 0.3), init="tgv", mach=0.2)
        prm = cfdx.default_params(mach=0.2, reyl=800.0)
        b.step(2, 0.2, b.cfl_dt(0.5), prm)
        return b, b.get_flow()

    b1, r1 = run()
    f1, _ = ctx.mem_info()
    assert f0 - f1 >= need
    b1.close()
    f2, _ = ctx.mem_info()
    assert abs(f2 - f0) <= 64 << 20
    b2, r2 = run()
    for f in FIELDS:
        assert np.array_equal(r1[f], r2[f])
    b2.close()
    ctx.pool_trim()
    f3, _ = ctx.mem_info()
    assert abs(f3 - f0) <= 64 << 20
