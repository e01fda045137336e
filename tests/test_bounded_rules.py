"""The oracle's side conditions (tests/bounded.py apply_bcs) against the face rules of the
reference, written out here as closed forms (no GPU):

- wall (Flow_Initialisation.cpp:952-988): boundary node u = v = w = 0, p and T of the first node
  in (so rho = gamma M^2 p / T equals that node's rho), ghost layer m = node m in with the velocity
  reversed;
- outlet (:91-136): boundary node takes u, v, w, p, T of the first node in; the 3 ghost layers copy
  the boundary node;
- inlet / far field (:384-418): the free stream (u = cos a, v = sin a, p = 1 / (gamma M^2), T = 1)
  on the boundary node and the ghost layers (the reference's u = 1, v = 0 at a = 0);
- extrapolation (:640-700): boundary node = first node in, ghost layer m = node m in;
- C-grid wake cut: ghost row -m of a wake column i is row m of column ni-1-i; row 0 of the lower
  half is row 0 of the upper half; the stage vectors carry the same Q with the ghost's det.
Sides are applied i-, i+, j-, j+, k-, k+; the checks use face nodes away from the edges.
"""
import math

import numpy as np
import pytest

import bounded
import cfdx

WALL, FARFIELD, OUTFLOW, EXTRAP, WALL_CUT = 1, 2, 3, 4, 5
PRIM = ("rho", "u", "v", "w", "p")


def random_flow(n, mach, seed):
    rng = np.random.default_rng(seed)
    f = cfdx.freestream(n, mach, 0.0)
    f["rho"] *= 1 + 0.1 * rng.uniform(-1, 1, n)
    f["u"] = 0.6 + 0.2 * rng.uniform(-1, 1, n)
    f["v"] = 0.2 * rng.uniform(-1, 1, n)
    f["w"] = 0.2 * rng.uniform(-1, 1, n)
    f["p"] *= 1 + 0.1 * rng.uniform(-1, 1, n)
    return f


def box(bc, dims=(8, 9, 7), aoa=0.0):
    d = cfdx.BlockDesc(*dims, 0, cfdx.FP64)
    for q in range(6):
        d.bc[q] = bc[q]
    d.aoa = aoa
    xyz, geo = cfdx.box_geometry(d, (1.3, 1.1, 0.9), 0.02, (0.3, 1.2, 2.1))
    return d, geo


def side_nodes(L, side):
    """(boundary node ids, [first 3 ids in], [3 ghost ids]) of the face interior of a side."""
    q, hi = side // 2, side % 2
    n = L.n
    o1, o2 = [x for x in range(3) if x != q]
    A, B = np.meshgrid(np.arange(1, n[o1] - 1), np.arange(1, n[o2] - 1), indexing="ij")
    A, B = A.reshape(-1), B.reshape(-1)
    c0, step = (n[q] - 1, -1) if hi else (0, 1)

    def at(c):
        x = [None] * 3
        x[q], x[o1], x[o2] = np.full(A.shape, c), A, B
        return L.r(*x)
    return at(c0), [at(c0 + step * m) for m in (1, 2, 3)], [at(c0 - step * m) for m in (1, 2, 3)]


@pytest.mark.parametrize("mach,aoa", [(0.5, 0.0), (0.3, math.radians(4.0))])
def test_side_rules_closed_form(mach, aoa):
    bc = (WALL, OUTFLOW, EXTRAP, FARFIELD, OUTFLOW, WALL)
    d, geo = box(bc, aoa=aoa)
    L = bounded.Layout(d)
    before = random_flow(L.n_own, mach, 7)
    rc = bounded.reference_case(d, geo, before, mach)    # applies the side conditions (mode 2)
    F = rc["FLOW"]
    pre = np.zeros(L.N + 1, dtype=F.dtype)
    pre[1:L.n_own + 1] = before
    M2 = 1.4 * mach * mach
    for side, kind in enumerate(bc):
        b0, inner, ghost = side_nodes(L, side)
        i1 = inner[0]
        if kind == WALL:
            assert all((F[c][b0] == 0).all() for c in "uvw")
            assert np.array_equal(F["p"][b0], pre["p"][i1])
            assert np.allclose(F["rho"][b0], pre["rho"][i1], rtol=4e-16, atol=0)     # same p and T
            for g, i in zip(ghost, inner):
                for c in PRIM:
                    sign = -1.0 if c in "uvw" else 1.0
                    assert np.array_equal(F[c][g], sign * pre[c][i]), (side, c)
        elif kind == OUTFLOW:
            for c in ("u", "v", "w", "p"):
                assert np.array_equal(F[c][b0], pre[c][i1]), (side, c)
            assert np.allclose(F["rho"][b0], pre["rho"][i1], rtol=4e-16, atol=0)
            for g in ghost:
                for c in PRIM:
                    assert np.array_equal(F[c][g], F[c][b0])
        elif kind == FARFIELD:
            want = {"rho": M2 * (1.0 / M2), "u": math.cos(aoa), "v": math.sin(aoa), "w": 0.0, "p": 1.0 / M2}
            for g in [b0] + ghost:
                for c in PRIM:
                    assert np.allclose(F[c][g], want[c], rtol=1e-15, atol=1e-15), (side, c)
        elif kind == EXTRAP:
            for c in PRIM:
                assert np.array_equal(F[c][b0], pre[c][i1])
            for g, i in zip(ghost, inner):
                for c in PRIM:
                    assert np.array_equal(F[c][g], pre[c][i])
    # owned nodes off the sides are untouched
    own = L.r(*np.meshgrid(np.arange(1, d.ni - 1), np.arange(1, d.nj - 1), np.arange(1, d.nk - 1), indexing="ij"))
    for c in PRIM:
        assert np.array_equal(F[c][own.reshape(-1)], pre[c][own.reshape(-1)])


def test_stage_vectors_of_ghosts():
    """At stages 0 and 1 the ghosts' next stage vector is det * Q of their side-condition state."""
    d, geo = box((WALL, OUTFLOW, EXTRAP, FARFIELD, OUTFLOW, WALL))
    L = bounded.Layout(d)
    mach = 0.4
    rc = bounded.reference_case(d, geo, random_flow(L.n_own, mach, 3), mach)
    F = rc["FLOW"]
    U = np.zeros((L.N + 1, 5))
    bounded.apply_bcs(rc, F, U, 0)
    for side in range(6):
        _, _, ghost = side_nodes(L, side)
        g = np.concatenate(ghost)
        rho, u, v, w, p = (F[c][g] for c in PRIM)
        q = np.stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)], axis=1)
        assert np.allclose(U[g], rc["det"][g][:, None] * q, rtol=1e-14, atol=0), side


def test_wake_cut_mirror():
    dsc, g = cfdx.cgrid_desc(48, 16, 4, naca=(0.02, 0.4, 0.12), n_wake=12, aoa_deg=2.0)
    xyz, geo = cfdx.cgrid_geometry(dsc, g)
    L = bounded.Layout(dsc)
    mach = 0.5
    rc = bounded.reference_case(dsc, geo, random_flow(L.n_own, mach, 11), mach)
    F = rc["FLOW"]
    ni, nj, nk = L.n
    U = np.random.default_rng(5).uniform(1, 2, (L.N + 1, 5))
    U0 = U.copy() + 1.0
    Uc, U0c = U.copy(), U0.copy()
    bounded.apply_bcs(rc, F, Uc, 0, U0=U0c)
    det = rc["det"]
    for k in range(1, nk - 1):
        for i in range(1, dsc.cut_i0):
            im = ni - 1 - i
            for m in (1, 2, 3):
                dst, src = L.r(i, -m, k), L.r(im, m, k)
                for c in PRIM:
                    assert F[c][dst] == F[c][src]
                assert np.allclose(Uc[dst], U[src] * det[dst] / det[src], rtol=1e-15)
                assert np.allclose(U0c[dst], U0[src] * det[dst] / det[src], rtol=1e-15)
            dst, src = L.r(i, 0, k), L.r(im, 0, k)
            for c in PRIM:
                assert F[c][dst] == F[c][src]
        # wall columns of the section: the wall rule
        for i in range(dsc.cut_i0 + 1, dsc.cut_i1):
            b0 = L.r(i, 0, k)
            assert F["u"][b0] == 0 and F["v"][b0] == 0 and F["w"][b0] == 0
