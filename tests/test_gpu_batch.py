"""Mixed precision and batched execution of the structured fast path (GPU).

- mixed form (float fluxes / reconstruction / viscous terms, double RK state and metrics) vs the
  plain FP64 form from the same state: per-field relative L2 <= 1e-5 after N steps;
- a batch of uneven blocks with different run parameters gives every block exactly the result
  it gets alone (same arithmetic per node), with and without hipGraph replay, over repeated
  replays with host work and synchronisations between them.
"""
import math

import numpy as np
import pytest

import cases
import cfdx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


def field_errors(a, b):
    """Per-field relative L2 and Linf of a against b for rho, u, v, w, p. The velocity
    components are scaled by the norm of the whole velocity field (a component can vanish)."""
    out = {}
    vel = np.sqrt(sum(np.asarray(b[c]) ** 2 for c in "uvw"))
    for f in ("rho", "u", "v", "w", "p"):
        d = a[f] - b[f]
        ref = vel if f in "uvw" else np.abs(b[f])
        out[f] = (float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300)), float(np.abs(d).max() / max(ref.max(), 1e-300)))
    return out


def tgv_case(n, mach=0.1):
    h = 2 * math.pi / n
    return cfdx.periodic_case(n, n, n, grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05 * h,
                              phases=(0.3, 1.1, 2.0), init="tgv", mach=mach)


@pytest.mark.parametrize("visc", [True, False])
def test_mixed_vs_fp64(ctx, visc):
    case = tgv_case(32)
    prm = cfdx.default_params(mach=0.1, reyl=1600.0) if visc else cfdx.euler_params()
    dt = cfdx.cfl_dt(case, 0.5)
    res = {}
    for prec in (cfdx.FP64, cfdx.MIXED):
        b = cfdx.Block(32, 32, 32, precision=prec, ctx=ctx)
        b.set_case(case)
        b.step(20, 0.1, dt, prm)
        res[prec] = b.get_flow()
    err = field_errors(res[cfdx.MIXED], res[cfdx.FP64])
    print("mixed vs fp64 after 20 steps (rel L2, rel Linf):", err)
    assert all(e[0] <= 1e-5 for e in err.values()), err
    assert all(np.isfinite(res[cfdx.MIXED][f]).all() for f in ("rho", "u", "v", "w", "p"))


@pytest.mark.parametrize("precision", [cfdx.FP64, cfdx.MIXED])
def test_batch_equals_alone(ctx, precision):
    specs = [((16, 12, 10), "tgv", 0.2, 800.0), ((12, 16, 8), "vortex", 0.5, None), ((20, 10, 12), "tgv", 0.3, 3000.0)]
    alone, blocks = [], []
    for (ni, nj, nk), init, mach, re in specs:
        if init == "tgv":
            case = cases.periodic(ni, nj, nk, init="tgv", mach=mach, noise=0.05)
            prm = cfdx.default_params(mach=mach, reyl=re)
        else:
            case = cases.periodic(ni, nj, nk, kind="box", init="vortex", mach=mach, noise=0.0)
            prm = cfdx.euler_params()
        dt = 0.5 * cfdx.cfl_dt(case, 1.0)
        a = cfdx.Block(ni, nj, nk, precision=precision, ctx=ctx)
        a.set_case(case)
        a.step(4, mach, dt, prm)
        alone.append(a.get_flow())
        b = cfdx.Block(ni, nj, nk, precision=precision, ctx=ctx)
        b.set_case(case)
        b.configure(mach, dt, prm)
        blocks.append(b)
    for graph in (False, True):
        q = cfdx.Batch(precision, ctx=ctx)
        for b, ((ni, nj, nk), init, mach, re) in zip(blocks, specs):
            q.add(b)
        # restart every block from its initial state
        for b, ((ni, nj, nk), init, mach, re) in zip(blocks, specs):
            case = cases.periodic(ni, nj, nk, init="tgv", mach=mach, noise=0.05) if init == "tgv" else \
                cases.periodic(ni, nj, nk, kind="box", init="vortex", mach=mach, noise=0.0)
            b.set_case(case)
        for rep in range(4):       # four replays of one step, host work in between
            q.step(1, graph=graph)
            ctx.sync()
            _ = np.zeros(1000).sum()
        for b, want in zip(blocks, alone):
            got = b.get_flow()
            for f in ("rho", "u", "v", "w", "p"):
                assert np.array_equal(got[f], want[f]), (graph, f)
        q.close()


def test_batch_persistent_grid(ctx):
    """A persistent launch (grid capped below the tile count) gives the same result."""
    case = cases.periodic(24, 20, 16, init="tgv", mach=0.2, noise=0.05)
    prm = cfdx.default_params(mach=0.2, reyl=1000.0)
    dt = 0.5 * cfdx.cfl_dt(case, 1.0)
    a = cfdx.Block(24, 20, 16, ctx=ctx)
    a.set_case(case)
    a.step(3, 0.2, dt, prm)
    want = a.get_flow()
    b = cfdx.Block(24, 20, 16, ctx=ctx)
    b.set_case(case)
    b.configure(0.2, dt, prm)
    q = cfdx.Batch(cfdx.FP64, ctx=ctx)
    q.add(b)
    q.set_grid(37)
    q.step(3, graph=True)
    got = b.get_flow()
    for f in ("rho", "u", "v", "w", "p"):
        assert np.array_equal(got[f], want[f]), f
