"""Stable time step of viscous blocks (CPU, the oracle kernels).

cfdx_block_cfl_dt and its host mirror cfdx.cfl_dt add a viscous radius to the convective one. The
regression: a low-Re walled C-grid (48 x 16 x 4, wall spacing 1e-4 chord, Re 1200) run through the
reference scheme on the oracle diverges within a few steps at the convective CFL-0.5 step, and stays
finite at the viscous-bounded one. At Re 1e5 on the same grid the viscous radius changes the step by
a modest factor only.
"""
import numpy as np

import bounded
import cfdx
import pyoracle

FIELDS = ("rho", "u", "v", "w", "p")


def cgrid_rc(mach):
    d, g = cfdx.cgrid_desc(48, 16, 4, naca=(0.02, 0.4, 0.12), n_wake=12, aoa_deg=3.0)
    xyz, geo = cfdx.cgrid_geometry(d, g)
    own = cfdx.owned_index(d)
    return bounded.reference_case(d, geo, cfdx.freestream(own.size, mach, 3.0), mach)


def finite(flow):
    return all(np.isfinite(flow[f]).all() for f in FIELDS)


def test_low_re_cgrid_needs_the_viscous_bound(orc):
    mach, re = 0.3, 1200.0
    rc = cgrid_rc(mach)
    prm = cfdx.default_params(mach=mach, reyl=re)
    dt_conv = cfdx.cfl_dt(rc, 0.5)
    dt_visc = cfdx.cfl_dt(rc, 0.5, prm)
    assert dt_visc < 0.1 * dt_conv, (dt_visc, dt_conv)
    oprm = pyoracle.reference_params(mach=mach, reyl=re)
    assert not finite(bounded.oracle_steps(orc, rc, 6, mach, dt_conv, oprm))
    assert finite(bounded.oracle_steps(orc, rc, 6, mach, dt_visc, oprm))


def test_high_re_step_changes_little():
    mach = 0.3
    rc = cgrid_rc(mach)
    dt_conv = cfdx.cfl_dt(rc, 0.5)
    dt_visc = cfdx.cfl_dt(rc, 0.5, cfdx.default_params(mach=mach, reyl=1e5))
    assert 0.5 * dt_conv < dt_visc < dt_conv, (dt_visc, dt_conv)


def test_euler_params_leave_the_convective_step():
    rc = cgrid_rc(0.5)
    assert cfdx.cfl_dt(rc, 0.5, cfdx.euler_params()) == cfdx.cfl_dt(rc, 0.5)
