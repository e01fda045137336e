"""Geometry of bounded blocks (no GPU): compact metrics with the reference closures, the NACA
C-grid and its wake cut."""
import math

import numpy as np
import pytest

import bounded
import cfdx


def box_desc(ni, nj, nk, periodic, bc=(0, 0, 0, 0, 0, 0)):
    d = cfdx.BlockDesc(ni, nj, nk, periodic, cfdx.FP64)
    for q in range(6):
        d.bc[q] = bc[q]
    return d


def test_padded_layout():
    d = box_desc(10, 9, 8, 4, (3, 2, 1, 1, 0, 0))
    assert cfdx.padded(d) == ((16, 15, 8), (3, 3, 0))
    own = cfdx.owned_index(d)
    assert own.size == 10 * 9 * 8 and own[0] == 3 + 16 * 3 and len(set(own)) == own.size


def test_cartesian_metrics_exact():
    # a box with every direction bounded: linear coordinates -> x_zeta = h exactly, closures included
    d = box_desc(9, 8, 7, 0, (1, 1, 2, 3, 4, 4))
    L = (2.0, 3.0, 5.0)
    xyz, geo = cfdx.box_geometry(d, L, 0.0)
    N = xyz.shape[0]
    h = np.array([L[0] / 9, L[1] / 8, L[2] / 7])
    g = geo.reshape(22, N)
    assert np.allclose(g[0], h.prod(), rtol=1e-13)
    for q in range(3):
        assert np.allclose(g[1 + 4 * q], 1.0 / h[q], rtol=1e-13)     # zeta_x, eta_y, xi_z
    own = cfdx.owned_index(d)
    for q in range(3):   # face metric rows and dets on the computed faces
        assert np.allclose(g[10 + 4 * q][own], 1.0 / h[q], rtol=1e-12)
        assert np.allclose(g[19 + q][own], h.prod(), rtol=1e-12)


def test_compact_closure_orders():
    # quadratic coordinate along a bounded direction: the closures (3rd / 4th order) and the
    # interior scheme are exact for it, so x_zeta = d/ds of the mapping at every stored node
    d = box_desc(12, 3, 3, 6, (1, 1, 0, 0, 0, 0))
    xyz, geo = cfdx.box_geometry(d, (1.0, 1.0, 1.0), 0.0)
    dims, off = cfdx.padded(d)
    N = xyz.shape[0]
    I = (np.arange(N) % dims[0]) - off[0]
    s = I.astype(float)
    xyz2 = xyz.copy()
    xyz2[:, 0] = 0.3 * s + 0.01 * s * s
    geo2 = np.zeros(22 * N)
    import ctypes as C
    L = np.array([1.0, 1.0, 1.0])
    cfdx._check(cfdx._lib.cfdx_geom_block_metrics(C.byref(d), cfdx._ptr(np.ascontiguousarray(xyz2)), cfdx._ptr(L),
                                                  cfdx._ptr(geo2)))
    g = geo2.reshape(22, N)
    xz = 0.3 + 0.02 * s
    assert np.allclose(1.0 / g[1], xz, rtol=1e-11)


def test_cgrid_symmetric_and_cut():
    d, g = cfdx.cgrid_desc(64, 24, 4, naca=(0.0, 0.4, 0.12), n_wake=14)
    xyz, geo = cfdx.cgrid_geometry(d, g)
    L = bounded.Layout(d)
    ni, nj, nk = L.n
    gm = geo.reshape(22, L.N)
    own = cfdx.owned_index(d)
    assert np.isfinite(geo).all() and (gm[0][own] > 0).all()
    i, j = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
    a, b = L.g(i, j, 1), L.g(ni - 1 - i, j, 1)
    # NACA 0012: the grid mirrors through y = 0 under i -> ni-1-i
    assert np.allclose(xyz[a, 0], xyz[b, 0], atol=1e-12) and np.allclose(xyz[a, 1], -xyz[b, 1], atol=1e-12)
    # wake cut: ghost rows are the nodes across, row 0 is shared, eta flips sign across the cut
    wi = np.arange(d.cut_i0)
    for m in (1, 2, 3):
        assert np.array_equal(xyz[L.g(wi, -m, 0)], xyz[L.g(ni - 1 - wi, m, 0)])
    assert np.array_equal(xyz[L.g(wi, 0, 0)], xyz[L.g(ni - 1 - wi, 0, 0)])
    r0, r1 = L.g(wi, 0, 0), L.g(ni - 1 - wi, 0, 0)
    assert np.allclose(gm[5][r0], -gm[5][r1], rtol=1e-9)          # eta_y
    assert np.allclose(gm[0][r0], gm[0][r1], rtol=1e-9)           # det
    # loc codes: the section's wall row and the outer sides are side nodes; the cut row is interior
    c = bounded.loc_codes(L)
    assert (c[L.g(np.arange(d.cut_i0, d.cut_i1 + 1), 0, 0)] == 1).all()
    assert (c[L.g(np.arange(1, d.cut_i0), 0, 0)] == 0).all()
    assert (c[L.g(5, nj - 1, 0)] == 1) and (c[L.g(0, 5, 0)] == 1) and (c[L.g(5, -1, 0)] == 0)
    assert c[L.g(ni // 2, -1, 0)] == 100


@pytest.mark.parametrize("dims", [(128, 64, 16), (192, 96, 24), (256, 128, 32), (384, 128, 32), (48, 16, 8)])
def test_cgrid_ladder_valid(dims):
    """Every stored node (ghosts included) has a positive Jacobian over the range of sections the
    airfoil request stream draws (thickness 8-18 %, camber 0-4 %) and the serving grid sizes."""
    ni, nj, nk = dims
    for naca in ((0.0, 0.4, 0.08), (0.04, 0.4, 0.18), (0.02, 0.4, 0.12)):
        d, g = cfdx.cgrid_desc(ni, nj, nk, naca=naca, n_wake=12 if ni == 48 else None)
        xyz, geo = cfdx.cgrid_geometry(d, g)
        N = xyz.shape[0]
        assert np.isfinite(geo).all() and (geo[:N] > 0).all(), (dims, naca)


def test_cgrid_plane_replication_matches_full():
    d, g = cfdx.cgrid_desc(48, 16, 8, naca=(0.03, 0.4, 0.15), n_wake=12, aoa_deg=3.0)
    x1, g1 = cfdx.cgrid_geometry(d, g)
    x2, g2 = cfdx.cgrid_geometry(d, g, full=True)
    assert np.allclose(x1, x2, rtol=0, atol=1e-15)
    N = x1.shape[0]
    a, b = g1.reshape(22, N), g2.reshape(22, N)
    # det, the node metrics, the face rows and the face dets, each against its group's scale (the
    # xi_x, xi_y components vanish up to round-off of the spanwise closures)
    for lo, hi in ((0, 1), (1, 10), (10, 19), (19, 22)):
        scale = np.abs(b[lo:hi]).max()
        assert np.abs(a[lo:hi] - b[lo:hi]).max() <= 1e-12 * scale, (lo, hi)
