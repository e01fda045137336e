"""Bounded blocks on the GPU: side conditions, loc-code branches, the C-grid wake cut.

- parity: the block path against the oracle kernels driven through the same side conditions
  (tests/bounded.py) on a NACA C-grid and on walled boxes, 2 RK3 steps, FP64, 1e-12 relative;
- mixed precision against FP64 on the C-grid (per-field relative L2 <= 1e-5);
- a batch holding bounded and periodic blocks gives each block bitwise its result alone.
"""
import math

import numpy as np
import pytest

import bounded
import cfdx

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p")


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


def perturbed(xyz_owned, mach, aoa):
    f = cfdx.freestream(xyz_owned.shape[0], mach, aoa)
    x, y, z = xyz_owned[:, 0], xyz_owned[:, 1], xyz_owned[:, 2]
    f["u"] *= 1 + 0.05 * np.sin(x) * np.cos(y)
    f["v"] += 0.02 * np.cos(0.8 * x + 0.1 * z)
    f["w"] += 0.01 * np.sin(0.6 * y)
    f["p"] *= 1 + 0.02 * np.cos(0.7 * x + 0.3 * y)
    f["rho"] *= 1 + 0.01 * np.sin(0.5 * y)
    return f


def cgrid_case():
    d, g = cfdx.cgrid_desc(48, 16, 4, naca=(0.02, 0.4, 0.12), n_wake=12, aoa_deg=2.0)
    xyz, geo = cfdx.cgrid_geometry(d, g)
    return d, xyz, geo, 0.5, 2.0, dict(mach=0.5, reyl=1e5)


def box_case(periodic, bc, dims, mach=0.4, reyl=800.0):
    d = cfdx.BlockDesc(*dims, periodic, cfdx.FP64)
    for q in range(6):
        d.bc[q] = bc[q]
    xyz, geo = cfdx.box_geometry(d, (1.3, 1.1, 0.9), 0.02, (0.3, 1.2, 2.1))
    return d, xyz, geo, mach, 0.0, dict(mach=mach, reyl=reyl)


CASES = {
    "cgrid": cgrid_case,
    "box-walls-j": lambda: box_case(4, (3, 2, 1, 1, 0, 0), (14, 12, 6)),
    "box-all-bounded": lambda: box_case(0, (4, 4, 1, 2, 1, 3), (10, 9, 8)),
    # the smallest bounded directions (7 nodes: the two-node shells leave 3 interior nodes) and
    # a periodic k of 3
    "box-minimal": lambda: box_case(4, (1, 3, 2, 1, 0, 0), (7, 7, 3)),
}


def run_block(ctx, d, geo, flow, mach, dt, prm, steps, precision=cfdx.FP64):
    """A block of descriptor d: geometry, flow (side conditions applied), `steps` RK3 steps; dt
    defaults to half the CFL-0.5 step of the initial state."""
    dd = cfdx.BlockDesc.from_buffer_copy(d)
    b = cfdx.Block(d.ni, d.nj, d.nk, precision=precision, ctx=ctx, desc=dd)
    b.set_geometry_soa(geo)
    b.configure(mach, 0.0, prm)       # the side conditions of set_flow use the block's Mach
    b.set_flow(flow)
    if dt is None:
        dt = 0.5 * b.cfl_dt(0.5)
    b.configure(mach, dt, prm)
    if steps:
        b.step(steps, mach, dt, prm)
    return b, dt


@pytest.mark.parametrize("name", list(CASES))
def test_bounded_parity(ctx, orc, name):
    d, xyz, geo, mach, aoa, pk = CASES[name]()
    own = cfdx.owned_index(d)
    flow = perturbed(xyz[own], mach, aoa)
    prm = cfdx.default_params(**pk)
    b, dt = run_block(ctx, d, geo, flow, mach, None, prm, 2)
    got = b.get_flow()
    import pyoracle
    rc = bounded.reference_case(d, geo, flow, mach)
    want = bounded.oracle_steps(orc, rc, 2, mach, dt, pyoracle.reference_params(**pk))
    vel = np.sqrt(sum(want[c] ** 2 for c in "uvw")).max()
    for f in FIELDS:
        scale = vel if f in "uvw" else np.abs(want[f]).max()
        err = np.abs(got[f] - want[f]).max() / scale
        print(f"{name} {f}: {err:.2e}")
        assert err <= 1e-12, (name, f, err)
    assert all(np.isfinite(got[f]).all() for f in FIELDS)


def test_cgrid_mixed_vs_fp64(ctx):
    d, xyz, geo, mach, aoa, pk = cgrid_case()
    own = cfdx.owned_index(d)
    flow = perturbed(xyz[own], mach, aoa)
    prm = cfdx.default_params(**pk)
    b64, dt = run_block(ctx, d, geo, flow, mach, None, prm, 10)
    b32, _ = run_block(ctx, d, geo, flow, mach, dt, prm, 10, precision=cfdx.MIXED)
    a, r = b32.get_flow(), b64.get_flow()
    vel = np.sqrt(sum(r[c] ** 2 for c in "uvw"))
    errs = {}
    for f in FIELDS:
        ref = vel if f in "uvw" else np.abs(r[f])
        errs[f] = np.linalg.norm(a[f] - r[f]) / np.linalg.norm(ref)
    print("cgrid mixed vs fp64 rel L2:", {f: f"{e:.2e}" for f, e in errs.items()})
    assert all(e <= 1e-5 for e in errs.values()), errs


@pytest.mark.parametrize("precision", [cfdx.FP64, cfdx.MIXED])
def test_batch_with_bounded_blocks(ctx, precision):
    """Bounded and periodic blocks in one batch give the bits of each block run alone. The
    periodic block alone runs the node kernels with non-temporal accesses, in the batch (a list
    with bounded blocks) with plain ones: a cache hint, so the bits must not change."""
    prm = cfdx.default_params(mach=0.5, reyl=1e5)
    alone, blocks = [], []
    for d, xyz, geo, mach, aoa, pk in (cgrid_case(), box_case(0, (4, 4, 1, 2, 1, 3), (10, 9, 8))):
        own = cfdx.owned_index(d)
        flow = perturbed(xyz[own], mach, aoa)
        p = cfdx.default_params(**pk)
        b, dt = run_block(ctx, d, geo, flow, mach, None, p, 3, precision)
        alone.append(b.get_flow())
        b2, _ = run_block(ctx, d, geo, flow, mach, dt, p, 0, precision)
        blocks.append(b2)

    def tgv():
        b = cfdx.Block(12, 10, 8, precision=precision, ctx=ctx)
        b.init_case(grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05, phases=(0.1, 0.2, 0.3), init="tgv", mach=0.3)
        return b
    ref_per = tgv()
    dtp = ref_per.cfl_dt(0.5)
    ref_per.step(3, 0.3, dtp, prm)
    alone.append(ref_per.get_flow())
    per = tgv()
    per.configure(0.3, dtp, prm)
    blocks.append(per)
    q = cfdx.Batch(precision, ctx)
    for b in blocks:
        q.add(b)
    q.step(3, True)
    for b, want in zip(blocks, alone):
        got = b.get_flow()
        for f in FIELDS:
            assert np.array_equal(got[f], want[f]), f
    q.close()


def test_geometry_not_finite_refused(ctx):
    d, xyz, geo, mach, aoa, pk = box_case(4, (3, 2, 1, 1, 0, 0), (14, 12, 6))
    bad = geo.copy()
    bad[7 * (geo.size // 22) + 123] = np.nan      # one metric value of one node
    b = cfdx.Block(d.ni, d.nj, d.nk, ctx=ctx, desc=cfdx.BlockDesc.from_buffer_copy(d))
    with pytest.raises(cfdx.CfdxError) as e:
        b.set_geometry_soa(bad)
    assert e.value.code == cfdx.E_NONFINITE
    b.set_geometry_soa(geo)                       # a finite geometry is taken afterwards
    b.close()


def test_cgrid_trailing_edge_follows_the_tip_rule(ctx):
    """The reference holds its tip nodes (Solver.cpp:660-680) and its wall nodes at singular
    points of unstructured meshes (Flow_Initialisation.cpp:1194-1273) at rest with p and T of
    the neighbour off the wall and rho = 1.4 M^2 p / T. On the C-grid the trailing-edge nodes
    (j = 0, i = cut_i0 and cut_i1, where the wall range meets the wake cut) are ordinary wall
    nodes of the block's j- side: after RK3 steps from a perturbed state they obey the same rule,
    and the wake-cut row j = 0 next to them holds one state on both sides of the cut."""
    d, xyz, geo, mach, aoa, kw = cgrid_case()
    prm = cfdx.default_params(**kw)
    b, _ = run_block(ctx, d, geo, perturbed(xyz[cfdx.owned_index(d)], mach, aoa), mach, None, prm, 4)
    f = b.get_flow()
    b.close()
    ni, nj, nk = d.ni, d.nj, d.nk
    at = lambda i, j, k: i + ni * (j + nj * k)
    M2 = 1.4 * mach * mach
    for k in range(nk):
        for i in (d.cut_i0, d.cut_i1):
            te, up = f[at(i, 0, k)], f[at(i, 1, k)]
            assert te["u"] == 0.0 and te["v"] == 0.0 and te["w"] == 0.0, (i, k)
            assert abs(te["p"] - up["p"]) <= 1e-12 * up["p"], (i, k)
            T_up = M2 * up["p"] / up["rho"]
            assert abs(M2 * te["p"] / te["rho"] - T_up) <= 1e-12 * T_up, (i, k)
        # the wake cut beside the trailing edge: node (i, 0) of the lower half is node (ni-1-i, 0)
        for i in (d.cut_i0 - 1, d.cut_i0 - 2):
            lo, hi = f[at(i, 0, k)], f[at(ni - 1 - i, 0, k)]
            for c in FIELDS:
                assert lo[c] == hi[c], (i, k, c)


@pytest.mark.parametrize("tiles", [(0, 0, 0, 1, 3), (2, 3, 1, 0, 2), (1, 2, 2, 3, 64)])
def test_cgrid_tiles_and_strips_same_bits(ctx, tiles):
    """The C-grid block (bounded, wake cut, shell tiles) gives bitwise the state of the default tiles
    with other tile variants and node-tile strips, including the plane-staged viscous tiles."""
    d, xyz, geo, mach, aoa, kw = cgrid_case()
    prm = cfdx.default_params(**kw)
    flow = perturbed(xyz[cfdx.owned_index(d)], mach, aoa)
    ref, dt = run_block(ctx, d, geo, flow, mach, None, prm, 3)
    want = ref.get_flow()
    ref.close()
    dd = cfdx.BlockDesc.from_buffer_copy(d)
    b = cfdx.Block(d.ni, d.nj, d.nk, ctx=ctx, desc=dd)
    b.set_geometry_soa(geo)
    b.configure(mach, 0.0, prm)
    b.set_flow(flow)
    b.set_tiles(tiles)
    b.step(3, mach, dt, prm)
    got = b.get_flow()
    b.close()
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), (tiles, f)
