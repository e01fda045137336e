"""The benchmark configurations at their full sizes (GPU).

- C1 (64^3 isentropic vortex, Euler, Cartesian through the metric path): its 100 RK3 steps
  against the oracle, 1e-12 relative;
- C2 (128^3 curvilinear Taylor-Green, MTS LES): one RK3 step against the oracle, 1e-12 relative
  (velocity components against the largest velocity);
  the right-hand side against the oracle to 1e-12 of the face-flux scale (at Mach 0.1 on this
  grid L is a difference of fluxes about 1e3 times larger, so the order of the additions shows
  at ~1e-10 of max |L|); discrete conservation on the periodic grid (the face fluxes and the
  viscous node fluxes telescope: sum over nodes of L = 0 up to the rounding of the sums) in FP64
  and mixed;
- C3 (384x128x32 NACA 0012 C-grid): 20 steps from the free stream in mixed precision against
  FP64 (per-field relative L2 <= 1e-5) with a finite state;
- C2's full 200 steps against the oracle (about two minutes of CPU oracle work on 16 cores),
  1e-12 relative.
"""
import math

import numpy as np
import pytest

import cases
import cfdx

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


def c2_case():
    n = 128
    h = 2 * math.pi / n
    phases = tuple(np.random.default_rng(20261202).uniform(0, 2 * math.pi, 3))
    return cfdx.periodic_case(n, n, n, grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05 * h, phases=phases,
                              init="tgv", mach=0.1)


def test_c1_full_size_steps(ctx, orc):
    import pyoracle
    print("C1 64^3 vortex, 100 steps vs oracle", flush=True)
    case = cfdx.periodic_case(64, 64, 64, grid="box", length=(10.0,) * 3, init="vortex", mach=0.5)
    dt = cfdx.cfl_dt(case, 0.5)
    ref = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in case.items()}
    orc.steps(ref, 100, 0.5, dt, pyoracle.reference_params(reyl=math.inf, c_mts=0.0))
    b = cfdx.Block(64, 64, 64, ctx=ctx)
    b.set_case(case)
    b.step(100, 0.5, dt, cfdx.euler_params())
    err = cases.rel_err(cases.flow_matrix(b.get_flow()), cases.flow_matrix(ref["FLOW"][1:]))
    b.close()
    assert err.max() <= TOL, err


def flux_scale(case) -> float:
    """max over nodes of det |grad xi_d| (p + rho |u|^2): the size of the face fluxes."""
    f = case["FLOW"][1:]
    m = case["metric"][1:, :9].reshape(-1, 3, 3)
    g = np.sqrt((m ** 2).sum(axis=2)).max(axis=1)
    return float((case["det"][1:] * g * (f["p"] + f["rho"] * (f["u"] ** 2 + f["v"] ** 2 + f["w"] ** 2))).max())


def test_c2_full_size_step_rhs_and_conservation(ctx, orc):
    import pyoracle
    print("C2 128^3 TGV, one step and the RHS vs oracle, conservation", flush=True)
    case = c2_case()
    prm_o = pyoracle.reference_params(mach=0.1, reyl=1600.0)
    prm = cfdx.default_params(mach=0.1, reyl=1600.0)
    dt = cfdx.cfl_dt(case, 0.5)
    ref = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in case.items()}
    orc.steps(ref, 1, 0.1, dt, prm_o)
    b = cfdx.Block(128, 128, 128, ctx=ctx)
    b.set_case(case)
    b.step(1, 0.1, dt, prm)
    got, want = b.get_flow(), ref["FLOW"][1:]
    b.close()
    # w starts at zero in the Taylor-Green state: the velocity components are measured against the
    # largest velocity magnitude
    vel = np.sqrt(sum(want[c] ** 2 for c in "uvw")).max()
    for f in ("rho", "u", "v", "w", "p"):
        e = np.abs(got[f] - want[f]).max() / (vel if f in "uvw" else np.abs(want[f]).max())
        assert e <= TOL, (f, e)
    L_ref = orc.rhs(case, prm_o)[1:]
    fscale = flux_scale(case)
    for precision in (cfdx.FP64, cfdx.MIXED):
        b = cfdx.Block(128, 128, 128, precision=precision, ctx=ctx)
        b.set_case(case)
        L = b.rhs(prm)
        b.close()
        if precision == cfdx.FP64:
            e = np.abs(L.T - L_ref).max(axis=0) / fscale
            print("RHS error / flux scale", e, "max |L| / flux scale", np.abs(L_ref).max() / fscale, flush=True)
            assert (e <= TOL).all(), e
        total, scale = np.abs(L.sum(axis=1)), np.abs(L).sum(axis=1)
        print("conservation", precision, total / scale, flush=True)
        assert (total <= 1e-11 * scale).all(), (precision, total, scale)


def test_c2_200_steps(ctx, orc):
    import pyoracle
    print("C2 128^3 TGV, 200 steps vs oracle", flush=True)
    case = c2_case()
    dt = cfdx.cfl_dt(case, 0.5)
    ref = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in case.items()}
    for chunk in range(10):   # 20 steps at a time, with a progress line
        orc.steps(ref, 20, 0.1, dt, pyoracle.reference_params(mach=0.1, reyl=1600.0))
        print(f"oracle: {20 * (chunk + 1)} steps", flush=True)
    b = cfdx.Block(128, 128, 128, ctx=ctx)
    b.set_case(case)
    b.step(200, 0.1, dt, cfdx.default_params(mach=0.1, reyl=1600.0))
    got, want = b.get_flow(), ref["FLOW"][1:]
    b.close()
    vel = np.sqrt(sum(want[c] ** 2 for c in "uvw")).max()
    for f in ("rho", "u", "v", "w", "p"):
        e = np.abs(got[f] - want[f]).max() / (vel if f in "uvw" else np.abs(want[f]).max())
        print(f"C2 200 steps {f}: {e:.2e}", flush=True)
        assert e <= TOL, (f, e)


def test_c3_full_size_mixed_vs_fp64(ctx):
    print("C3 384x128x32 C-grid, 20 steps mixed vs FP64", flush=True)
    d, g = cfdx.cgrid_desc(384, 128, 32, naca=(0.0, 0.4, 0.12), aoa_deg=2.0)
    _, geo = cfdx.cgrid_geometry(d, g)
    prm = cfdx.default_params(mach=0.5, reyl=1e5)
    f0 = cfdx.freestream(384 * 128 * 32, 0.5, 2.0)
    out, dt = {}, None
    for precision in (cfdx.FP64, cfdx.MIXED):
        b = cfdx.Block(384, 128, 32, precision=precision, ctx=ctx, desc=d)
        b.set_geometry_soa(geo)
        b.configure(0.5, 0.0, prm)
        b.set_flow(f0)
        if dt is None:
            dt = b.cfl_dt(0.5)
        b.configure(0.5, dt, prm)
        b.step(20, 0.5, dt, prm)
        out[precision] = b.get_flow()
        b.close()
    a, r = out[cfdx.MIXED], out[cfdx.FP64]
    vel = np.sqrt(sum(r[c] ** 2 for c in "uvw"))
    for f in ("rho", "u", "v", "w", "p"):
        assert np.isfinite(a[f]).all() and np.isfinite(r[f]).all()
        ref = vel if f in "uvw" else np.abs(r[f])
        e = np.linalg.norm(a[f] - r[f]) / np.linalg.norm(ref)
        assert e <= 1e-5, (f, e)
