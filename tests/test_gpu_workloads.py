"""The batched, served and mixed-precision forms at the benchmark workloads, against the oracle (GPU).

- C4: the 64 seeded cubes (edges 32..128, vortex / Taylor-Green, Mach and Re draws) in one mixed
  batch replaying its captured RK3-step graph give every cube bitwise its result alone; the FP64 form
  of a seeded subset of the cubes, batched, matches the oracle's steps at 1e-12;
- C5: a seeded segment of the NACA 4-digit request stream on the ladder grid sizes, served by
  continuous batching, gives every request bitwise its result alone (mixed), and the FP64 server's
  results match the oracle through the side conditions of tests/bounded.py at 1e-12;
- mixed precision against the oracle's FP64 state (not the FP64 GPU path): the 32^3 curvilinear
  Taylor-Green vortex and a NACA C-grid, per-field relative L2 <= 1e-5;
- NS: the C2 flow at 256^3, one RK3 step against the oracle (1e-12), discrete conservation of the
  right-hand side and a finite state after further steps.
"""
import math

import numpy as np
import pytest

import bounded
import cfdx
from cfdx import serve, workloads

pytestmark = pytest.mark.gpu
FIELDS = ("rho", "u", "v", "w", "p")
TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


def max_rel(got, want):
    """max over fields of max |got - want| / max |want| (velocity components against the largest
    velocity magnitude, since a component can vanish)."""
    vel = np.sqrt(sum(want[c] ** 2 for c in "uvw")).max()
    return max(float(np.abs(got[f] - want[f]).max() / (vel if f in "uvw" else np.abs(want[f]).max())) for f in FIELDS)


def l2_errors(a, b):
    vel = np.sqrt(sum(b[c] ** 2 for c in "uvw"))
    return {f: float(np.linalg.norm(a[f] - b[f]) / np.linalg.norm(vel if f in "uvw" else b[f])) for f in FIELDS}


def cube_block(ctx, spec, precision, host_case=None):
    b = cfdx.Block(spec.edge, spec.edge, spec.edge, precision=precision, ctx=ctx)
    if host_case is not None:
        b.set_case(host_case)
    else:
        b.init_case(**spec.case_kw())
    prm = cfdx.default_params(mach=spec.mach, reyl=spec.reyl)
    b.configure(spec.mach, 0.0, prm)
    b.configure(spec.mach, b.cfl_dt(0.5), prm)
    return b, prm


def test_c4_batch_equals_alone(ctx):
    specs = workloads.c4_specs()
    nodes = sum(s.edge ** 3 for s in specs)
    print(f"C4: {len(specs)} cubes, {nodes} nodes, mixed, captured graph, 2 steps", flush=True)
    q = cfdx.Batch(cfdx.MIXED, ctx)
    blocks = []
    for s in specs:
        b, _ = cube_block(ctx, s, cfdx.MIXED)
        q.add(b)
        blocks.append(b)
    q.step(2, True)
    ctx.sync()
    got = [b.get_flow() for b in blocks]
    q.close()
    for b in blocks:
        b.close()
    for n, (s, g) in enumerate(zip(specs, got)):
        b, prm = cube_block(ctx, s, cfdx.MIXED)
        b.step(2, s.mach, b.dt, prm)
        want = b.get_flow()
        b.close()
        for f in FIELDS:
            assert np.array_equal(g[f], want[f]), (n, s, f)
        assert all(np.isfinite(g[f]).all() for f in FIELDS)


def test_c4_fp64_subset_vs_oracle(ctx, orc):
    import pyoracle
    specs = workloads.c4_specs()
    # a seeded subset: the first vortex and the first two Taylor-Green cubes of edge <= 64
    small = [s for s in specs if s.edge <= 64]
    subset = [next(s for s in small if not s.tgv)] + [s for s in small if s.tgv][:2]
    print("C4 FP64 subset vs oracle:", [(s.edge, "tgv" if s.tgv else "vortex", round(s.mach, 3), round(s.reyl)) for s in subset],
          flush=True)
    q = cfdx.Batch(cfdx.FP64, ctx)
    runs = []
    for s in subset:
        case = cfdx.periodic_case(s.edge, s.edge, s.edge, **s.case_kw())
        b, prm = cube_block(ctx, s, cfdx.FP64, host_case=case)
        q.add(b)
        runs.append((s, case, b))
    q.step(2, True)
    ctx.sync()
    for s, case, b in runs:
        ref = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in case.items()}
        orc.steps(ref, 2, s.mach, b.dt, pyoracle.reference_params(mach=s.mach, reyl=s.reyl))
        e = max_rel(b.get_flow(), ref["FLOW"][1:])
        print(f"  edge {s.edge}: {e:.2e}", flush=True)
        assert e <= TOL, (s, e)
    q.close()
    for _, _, b in runs:
        b.close()


def c5_segment(nsteps):
    """The first six requests of the seeded NACA stream on the ladder sizes, arriving one step apart
    on the step clock."""
    reqs = serve.airfoil_stream(6, 20.0, nsteps=nsteps)
    for n, r in enumerate(reqs):
        r.arrival = float(n)
    return reqs


def test_c5_ladder_served_equals_alone(ctx):
    reqs = c5_segment(4)
    print("C5 segment:", [(r.ni, r.nj, r.nk) for r in reqs], flush=True)
    srv = serve.Server(precision=cfdx.MIXED, ctx=ctx, chunk=2, graph=True)
    res = srv.run(reqs, clock="steps")
    srv.close()
    for r in reqs:
        alone = serve.run_alone(r, cfdx.MIXED, ctx)
        assert res[r.rid].finite and alone.finite
        for f in FIELDS:
            assert np.array_equal(res[r.rid].flow[f], alone.flow[f]), (r.rid, f)


def test_c5_ladder_fp64_served_vs_oracle(ctx, orc):
    import pyoracle
    reqs = c5_segment(3)
    srv = serve.Server(precision=cfdx.FP64, ctx=ctx, chunk=3, graph=True)
    res = srv.run(reqs, clock="steps")
    srv.close()
    for r in reqs:
        prep = serve.prepare(r)
        rc = bounded.reference_case(prep["desc"], prep["geo"], prep["flow"], r.mach)
        want = bounded.oracle_steps(orc, rc, r.nsteps, r.mach, res[r.rid].dt,
                                    pyoracle.reference_params(mach=r.mach, reyl=r.reyl))
        e = max_rel(res[r.rid].flow, want)
        print(f"  request {r.rid} {r.ni}x{r.nj}x{r.nk}: {e:.2e}", flush=True)
        assert e <= TOL, (r.rid, e)


def test_mixed_tgv32_vs_oracle(ctx, orc):
    import pyoracle
    kw = workloads.tgv_case_kw(32)
    case = cfdx.periodic_case(32, 32, 32, **kw)
    prm = cfdx.default_params(mach=0.1, reyl=1600.0)
    dt = cfdx.cfl_dt(case, 0.5, prm)
    ref = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in case.items()}
    orc.steps(ref, 20, 0.1, dt, pyoracle.reference_params(mach=0.1, reyl=1600.0))
    b = cfdx.Block(32, 32, 32, precision=cfdx.MIXED, ctx=ctx)
    b.set_case(case)
    b.step(20, 0.1, dt, prm)
    err = l2_errors(b.get_flow(), ref["FLOW"][1:])
    b.close()
    print("mixed 32^3 TGV, 20 steps, rel L2 vs oracle FP64:", {f: f"{e:.2e}" for f, e in err.items()}, flush=True)
    assert all(e <= 1e-5 for e in err.values()), err


def test_mixed_cgrid_vs_oracle(ctx, orc):
    import pyoracle
    d, g = cfdx.cgrid_desc(128, 64, 16, naca=(0.02, 0.4, 0.12), aoa_deg=2.0)
    _, geo = cfdx.cgrid_geometry(d, g)
    mach, reyl = 0.5, 1e5
    prm = cfdx.default_params(mach=mach, reyl=reyl)
    flow = cfdx.freestream(128 * 64 * 16, mach, 2.0)
    b = cfdx.Block(128, 64, 16, precision=cfdx.MIXED, ctx=ctx, desc=d)
    b.set_geometry_soa(geo)
    b.configure(mach, 0.0, prm)
    b.set_flow(flow)
    dt = b.cfl_dt(0.5)
    b.step(10, mach, dt, prm)
    got = b.get_flow()
    b.close()
    rc = bounded.reference_case(d, geo, flow, mach)
    want = bounded.oracle_steps(orc, rc, 10, mach, dt, pyoracle.reference_params(mach=mach, reyl=reyl))
    err = l2_errors(got, want)
    print("mixed 128x64x16 C-grid, 10 steps, rel L2 vs oracle FP64:", {f: f"{e:.2e}" for f, e in err.items()}, flush=True)
    assert all(e <= 1e-5 for e in err.values()), err


def test_ns256_step_conservation_finite(ctx, orc):
    import pyoracle
    n = 256
    print("NS 256^3 TGV: host case", flush=True)
    case = cfdx.periodic_case(n, n, n, **workloads.tgv_case_kw(n))
    prm = cfdx.default_params(mach=0.1, reyl=1600.0)
    dt = cfdx.cfl_dt(case, 0.5, prm)
    b = cfdx.Block(n, n, n, ctx=ctx)
    b.set_case(case)
    L = b.rhs(prm)
    total, scale = np.abs(L.sum(axis=1)), np.abs(L).sum(axis=1)
    print("NS conservation |sum L| / sum |L|:", total / scale, flush=True)
    assert (total <= 1e-11 * scale).all(), (total, scale)
    del L
    b.step(1, 0.1, dt, prm)
    got = b.get_flow()
    print("NS oracle step", flush=True)
    orc.steps(case, 1, 0.1, dt, pyoracle.reference_params(mach=0.1, reyl=1600.0))
    e = max_rel(got, case["FLOW"][1:])
    print(f"NS 256^3 one RK3 step vs oracle: {e:.2e}", flush=True)
    assert e <= TOL, e
    del case
    b.step(9, 0.1, dt, prm)
    assert b.check_finite()
    b.close()
