"""Host geometry of libcfdx (CPU): neighbour tables, node metrics, half-point metrics,
compact-scheme Jacobians and initial states, against closed forms."""
import math

import numpy as np
import pytest

import cfdx


def test_periodic_table_walks():
    ni, nj, nk = 5, 6, 7
    G = np.zeros(ni * nj * nk + 1, dtype=cfdx.GRID_DT)
    cfdx._check(cfdx._lib.cfdx_geom_periodic_table(ni, nj, nk, None, cfdx._ptr(G)))
    nid = lambda i, j, k: 1 + (i % ni) + ni * ((j % nj) + nj * (k % nk))
    for (i, j, k) in [(0, 0, 0), (4, 5, 6), (2, 3, 1)]:
        g = G[nid(i, j, k)]
        assert (g["N1"], g["N3"]) == (nid(i + 1, j, k), nid(i - 1, j, k))
        assert (g["N0"], g["N2"]) == (nid(i, j + 1, k), nid(i, j - 1, k))
        assert (g["N4"], g["N5"]) == (nid(i, j, k + 1), nid(i, j, k - 1))
        assert g["loc"] == 0
    assert G[0]["N1"] == 0 and G[0]["loc"] == 0


def test_cartesian_metrics():
    case = cfdx.periodic_case(8, 10, 12, grid="box", length=(4.0, 5.0, 6.0))
    h = np.array([0.5, 0.5, 0.5])
    assert np.allclose(case["det"][1:], h.prod(), rtol=0, atol=1e-15)
    m = case["metric"][1:, :9]
    assert np.allclose(m[:, [0, 4, 8]], 1 / h, atol=1e-14) and np.abs(m[:, [1, 2, 3, 5, 6, 7]]).max() == 0
    # half-point metrics of a constant field are the field; deter likewise
    assert np.allclose(case["metric"][1:, 9:].reshape(-1, 6, 9)[:, :, 0], 2.0, atol=1e-14)
    assert np.allclose(case["deter"][1:], 0.125, atol=1e-15)


def test_half_point_interpolation_exact_for_quintics():
    """Main.cpp:2368-2435: the 6-point midpoint formula is exact for degree <= 5."""
    ni, nj, nk = 16, 4, 4
    G = np.zeros(ni * nj * nk + 1, dtype=cfdx.GRID_DT)
    cfdx._check(cfdx._lib.cfdx_geom_periodic_table(ni, nj, nk, None, cfdx._ptr(G)))
    rng = np.random.default_rng(3)
    coef = rng.normal(size=6)
    poly = lambda x: sum(c * x ** p for p, c in enumerate(coef))
    n = ni * nj * nk
    metric = np.zeros((n + 1, 63))
    det = np.ones(n + 1)
    idx = np.arange(n)
    i = idx % ni
    metric[1:, 0] = poly(i.astype(float))
    det[1:] = poly(i.astype(float))
    deter = np.zeros((n + 1, 6))
    cfdx._check(cfdx._lib.cfdx_geom_half_metrics(cfdx._ptr(G), n, cfdx._ptr(metric), cfdx._ptr(det), cfdx._ptr(deter)))
    inner = (i >= 3) & (i <= ni - 4)   # stencils that do not wrap
    want_p = poly(i[inner] + 0.5)
    want_m = poly(i[inner] - 0.5)
    sc = np.abs(want_p).max()
    assert np.abs(metric[1:, 9][inner] - want_p).max() < 1e-12 * sc        # zeta_x at ip
    assert np.abs(metric[1:, 18][inner] - want_m).max() < 1e-12 * sc       # zeta_x at im
    assert np.abs(deter[1:, 0][inner] - want_p).max() < 1e-12 * sc
    assert np.abs(deter[1:, 1][inner] - want_m).max() < 1e-12 * sc


@pytest.mark.parametrize("n", [16, 32])
def test_compact_jacobian_matches_analytic(n):
    L = np.array([2 * math.pi] * 3)
    amp, ph = 0.4 * 2 * math.pi / 16, np.array([0.3, 1.1, 2.0])
    xyz = np.zeros((n ** 3, 3))
    cfdx._check(cfdx._lib.cfdx_geom_coords_sinusoid(n, n, n, cfdx._ptr(L), amp, cfdx._ptr(ph), cfdx._ptr(xyz)))
    ja = np.zeros((n ** 3, 9))
    jc = np.zeros((n ** 3, 9))
    cfdx._check(cfdx._lib.cfdx_geom_jacobian_sinusoid(n, n, n, cfdx._ptr(L), amp, cfdx._ptr(ph), cfdx._ptr(ja)))
    cfdx._check(cfdx._lib.cfdx_geom_jacobian_compact_periodic(n, n, n, cfdx._ptr(L), cfdx._ptr(xyz), cfdx._ptr(jc)))
    err = np.abs(ja - jc).max() / np.abs(ja).max()
    assert err < (2e-5 if n == 16 else 4e-7), err


def test_compact_converges_sixth_order():
    errs = []
    for n in (16, 32):
        L = np.array([2 * math.pi] * 3)
        amp, ph = 0.3, np.array([0.3, 1.1, 2.0])
        xyz = np.zeros((n ** 3, 3))
        cfdx._check(cfdx._lib.cfdx_geom_coords_sinusoid(n, n, n, cfdx._ptr(L), amp, cfdx._ptr(ph), cfdx._ptr(xyz)))
        ja = np.zeros((n ** 3, 9)); jc = np.zeros((n ** 3, 9))
        cfdx._check(cfdx._lib.cfdx_geom_jacobian_sinusoid(n, n, n, cfdx._ptr(L), amp, cfdx._ptr(ph), cfdx._ptr(ja)))
        cfdx._check(cfdx._lib.cfdx_geom_jacobian_compact_periodic(n, n, n, cfdx._ptr(L), cfdx._ptr(xyz), cfdx._ptr(jc)))
        errs.append(np.abs(ja - jc).max() / (2 * math.pi / n))
    assert errs[0] / errs[1] > 40, errs   # 2^6 = 64


def test_node_metrics_invert_jacobian():
    case = cfdx.periodic_case(8, 8, 8, grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.3, phases=(0.1, 0.2, 0.3))
    J = case["jac"].reshape(-1, 3, 3)            # rows x, y, z; columns zeta, eta, xi
    M = case["metric"][1:, :9].reshape(-1, 3, 3)  # rows zeta, eta, xi; columns x, y, z
    assert np.abs(np.einsum("nij,njk->nik", M, J) - np.eye(3)).max() < 1e-13
    assert np.allclose(case["det"][1:], np.linalg.det(J), rtol=1e-13)


def test_initial_states():
    case = cfdx.periodic_case(8, 8, 8, grid="box", init="freestream", mach=0.4)
    f = case["FLOW"][1:]
    assert np.allclose(f["u"], 1) and np.allclose(f["p"], 1 / (1.4 * 0.16)) and np.allclose(f["rho"], 1)
    tgv = cfdx.periodic_case(16, 16, 16, grid="box", length=(2 * math.pi,) * 3, init="tgv", mach=0.1)
    f = tgv["FLOW"][1:]
    assert abs(f["u"].mean()) < 1e-12 and abs(f["w"]).max() == 0 and np.allclose(f["t"], 1.0)
    vortex = cfdx.periodic_case(32, 32, 4, grid="box", init="vortex", mach=0.5, beta=5.0)
    f = vortex["FLOW"][1:]
    # isentropic: p / rho^gamma constant
    s = f["p"] / f["rho"] ** 1.4
    assert np.ptp(s) < 1e-12 * s.mean()
