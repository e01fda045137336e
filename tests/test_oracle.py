"""The oracle (C restatement of U_Q.cl / Viscous.cl / Weno.cl + the Solver.cpp stage
sequence) pinned three ways, all on the CPU:
  1. against the golden fixtures written from the reference's own kernels (bitwise);
  2. against those kernels live (oracle/_ref, when built) on more cases (bitwise);
  3. against what the mathematics fixes: quadratic exactness and 5th-order convergence of
     the WCNS interpolation, its ENO behaviour at a jump, L R = I and L A R = diag(lambda)
     for the flux Jacobian A, free-stream preservation, and discrete conservation.
"""
import math
import os

import numpy as np
import pytest

import cases
import pyoracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixture(name):
    z = np.load(os.path.join(GOLD, name))
    n = z["in_det"].shape[0]
    G = np.ascontiguousarray(z["in_Grid_P"]).view(pyoracle.GRID_DT).reshape(n).copy()
    flow = np.zeros(n, dtype=pyoracle.FLOW_DT)
    for i, f in enumerate(pyoracle.FLOW_FIELDS):
        flow[f] = z["in_FLOW"][:, i]
    case = dict(Grid_P=G, det=z["in_det"].copy(), metric=z["in_metric"].copy(), deter=z["in_deter"].copy(), FLOW=flow)
    return case, z


def _as_matrix(v):
    return cases.flow_matrix(v, pyoracle.FLOW_FIELDS) if getattr(v, "dtype", None) is not None and v.dtype.names else v


def _roe_used(a):
    return a[:, [i for i in range(36) if i % 6 != 3]]


def test_golden_periodic(orc):
    case, z = load_fixture("periodic_tgv_7x8x9.npz")
    out = orc.stage_outputs(case, float(z["mach"]), float(z["dt"]), pyoracle.reference_params())
    for k, v in out.items():
        got, want = _as_matrix(v)[1:], z[f"out_{k}"][1:]
        if k == "ROE":
            got, want = _roe_used(got), _roe_used(want)
        assert np.array_equal(got, want), k
    orc.steps(case, 2, float(z["mach"]), float(z["dt"]), pyoracle.reference_params())
    assert np.array_equal(_as_matrix(case["FLOW"])[1:], z["steps2_FLOW"][1:])


@pytest.mark.parametrize("name,min_weno", [("wallbox_3x4x5.npz", 0), ("channel_7x8x9.npz", 300)])
def test_golden_walls(orc, name, min_weno):
    case, z = load_fixture(name)
    out = orc.stage_outputs(case, float(z["mach"]), float(z["dt"]), pyoracle.reference_params())
    n_core = int(z["n_core"])
    core = np.zeros(len(case["det"]), dtype=bool)
    core[1:n_core + 1] = True
    vmask, wmask = z["mask_viscous"] & core, z["mask_weno"] & z["mask_viscous"] & core
    assert vmask.sum() > 0 and wmask.sum() >= min_weno
    for k, v in out.items():
        rows = core if k in ("U_C", "Q_C", "ROE") else vmask if k in ("F", "E", "G", "Fv", "Ev", "Gv", "mu") else wmask
        got, want = _as_matrix(v)[rows], z[f"out_{k}"][rows]
        if k == "ROE":
            got, want = _roe_used(got), _roe_used(want)
        assert np.array_equal(got, want), k


@pytest.mark.parametrize("kind,init,seed", [("box", "vortex", 1), ("sinusoid", "tgv", 2), ("sinusoid", "vortex", 3)])
def test_reference_kernels_live(orc, ref, kind, init, seed):
    case = cases.periodic(9, 7, 8, kind=kind, init=init, mach=0.4, noise=0.08, seed=seed)
    prm = pyoracle.reference_params()
    a = orc.stage_outputs(case, 0.4, 0.02, prm)
    b = ref.stage_outputs(case, 0.4, 0.02, prm)
    for k in a:
        x, y = _as_matrix(a[k])[1:], _as_matrix(b[k])[1:]
        if k == "ROE":
            x, y = _roe_used(x), _roe_used(y)
        assert np.array_equal(x, y), k
    c1 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in case.items()}
    c2 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in case.items()}
    orc.steps(c1, 2, 0.4, 0.02, prm)
    ref.steps(c2, 2, 0.4, 0.02, prm)
    assert np.array_equal(_as_matrix(c1["FLOW"])[1:], _as_matrix(c2["FLOW"])[1:])


@pytest.mark.parametrize("walls", [(True, True, True), (False, False, True), (True, False, False)])
def test_reference_kernels_walls_live(orc, ref, walls):
    case = cases.wall_box(5, 6, 7, seed=11, walls=walls)
    prm = pyoracle.reference_params()
    a = orc.stage_outputs(case, 0.3, 0.01, prm)
    b = ref.stage_outputs(case, 0.3, 0.01, prm)
    core = np.zeros(len(case["det"]), dtype=bool)
    core[1:case["n_core"] + 1] = True
    vm = cases.defined_mask(case, False) & core
    wm = cases.defined_mask(case, True) & vm
    for k in a:
        rows = core if k in ("U_C", "Q_C", "ROE") else vm if k in ("F", "E", "G", "Fv", "Ev", "Gv", "mu") else wm
        x, y = _as_matrix(a[k])[rows], _as_matrix(b[k])[rows]
        if k == "ROE":
            x, y = _roe_used(x), _roe_used(y)
        assert np.array_equal(x, y), k


# ---------------------------------------------------------------------------
# mathematics
# ---------------------------------------------------------------------------
def _wcns(orc, W, eps=1e-7):
    import ctypes as C
    L = orc.lib
    L.orc_wcns_pair.argtypes = [C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    w = (C.c_double * 6)(*W)
    p, n = C.c_double(), C.c_double()
    L.orc_wcns_pair(w, eps, C.byref(p), C.byref(n))
    return p.value, n.value


def test_wcns_exact_for_quadratics(orc):
    rng = np.random.default_rng(5)
    for _ in range(20):
        a, b, c = rng.normal(size=3)
        f = lambda x: a + b * x + c * x * x
        W = [f(x) for x in range(-2, 4)]
        p, n = _wcns(orc, W)
        assert abs(p - f(0.5)) < 1e-13 * (1 + abs(f(0.5)))
        assert abs(n - f(0.5)) < 1e-13 * (1 + abs(f(0.5)))


def test_wcns_fifth_order(orc):
    """Smooth data: error at the face falls like h^5 (linear weights 1/16, 5/8, 5/16 and mirror)."""
    errs = []
    for h in (0.1, 0.05, 0.025):
        W = [math.sin(1.0 + h * x) for x in range(-2, 4)]
        exact = math.sin(1.0 + 0.5 * h)
        p, n = _wcns(orc, W, eps=1e-40)  # eps negligible against the indicators
        errs.append(max(abs(p - exact), abs(n - exact)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert r1 > 20 and r2 > 20, errs   # 2^5 = 32; a wrong linear weight gives 2^3 = 8


def test_wcns_eno_at_jump(orc):
    """Unit jump between the left node and the right node: each side picks its smooth stencil."""
    p, n = _wcns(orc, [0, 0, 0, 1, 1, 1])
    assert abs(p - 1.0) < 1e-6 and abs(n) < 1e-6


def _face_state(rho, u, v, w, p):
    h = (3.5 * p + 0.5 * rho * (u * u + v * v + w * w)) / rho
    return np.array([u, v, w, 0.0, h, rho])


def _eigen(orc, roe, k):
    import ctypes as C
    orc.lib.orc_face_eigen.argtypes = [C.POINTER(C.c_double)] * 4
    r = (C.c_double * 6)(*roe)
    kk = (C.c_double * 3)(*k)
    L, R = (C.c_double * 25)(), (C.c_double * 25)()
    orc.lib.orc_face_eigen(r, kk, L, R)
    return np.array(L).reshape(5, 5), np.array(R).reshape(5, 5)


def _normal_flux(Q, n):
    rho, m = Q[0], Q[1:4]
    vel = m / rho
    p = 0.4 * (Q[4] - 0.5 * rho * vel @ vel)
    un = vel @ n
    return np.array([rho * un, *(m * un + p * n), (Q[4] + p) * un])


def test_eigenvectors_diagonalise_flux_jacobian(orc):
    rng = np.random.default_rng(9)
    for _ in range(10):
        rho, p = rng.uniform(0.5, 2.0), rng.uniform(0.5, 3.0)
        u, v, w = rng.normal(scale=0.5, size=3)
        k = rng.normal(size=3)
        L, R = _eigen(orc, _face_state(rho, u, v, w, p), k)
        assert np.abs(L @ R - np.eye(5)).max() < 1e-12
        n = k / np.linalg.norm(k)
        Q = np.array([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])
        A = np.zeros((5, 5))
        for j in range(5):
            dq = np.zeros(5)
            dq[j] = 1e-6 * max(1.0, abs(Q[j]))
            A[:, j] = (_normal_flux(Q + dq, n) - _normal_flux(Q - dq, n)) / (2 * dq[j])
        lam = L @ A @ R
        c = math.sqrt(1.4 * p / rho)
        th = n @ np.array([u, v, w])
        want = np.diag([th, th, th, th + c, th - c])
        assert np.abs(lam - want).max() < 1e-6 * (1 + c)


def test_freestream_preserved(orc):
    """Uniform flow on a Cartesian periodic grid: RHS zero to rounding."""
    import cfdx
    case = cfdx.periodic_case(8, 8, 8, grid="box", init="freestream", mach=0.5)
    L = orc.rhs(case, pyoracle.reference_params())[1:]
    Q = orc.stage_outputs(case, 0.5, 0.01, pyoracle.reference_params())["F"][1:]
    assert np.abs(L).max() <= 1e-13 * np.abs(Q).max()


def test_discrete_conservation(orc):
    """Periodic grid: flux differences telescope, so the node sum of L vanishes to rounding."""
    case = cases.periodic(9, 8, 7, noise=0.1)
    L = orc.rhs(case, pyoracle.reference_params())[1:]
    assert (np.abs(L.sum(axis=0)) <= 1e-12 * np.abs(L).sum(axis=0)).all()


def test_smagorinsky_closed_form_on_linear_shear(orc):
    """The Smagorinsky option on u = g y (v = w = 0, uniform p and rho) on a Cartesian box of
    spacing h with Re = inf: the 4th-order differences are exact away from the periodic jump, S_xy
    = g / 2, |S| = g, and the SGS stress is -2 rho (C_s h)^2 |S| S_xy. The det-scaled contravariant
    eta flux of x-momentum is then det / h * tau_xy = -rho C_s^2 h^4 g^2, the eta flux of energy
    (the stress's work u tau_xy) -rho C_s^2 h^4 g^2 u, and the xi flux of energy 0 (no
    temperature gradient); C_s = 0 gives no viscous flux at all."""
    import cfdx
    n, h, g, cs = 8, 0.5, 0.3, 0.17
    case = cfdx.periodic_case(n, 2 * n, n, grid="box", length=(n * h, 2 * n * h, n * h), init="freestream", mach=0.3)
    y = case["xyz"][:, 1] if case["xyz"].ndim == 2 else case["xyz"][1]
    F = case["FLOW"]
    rho = float(F["rho"][1])
    F["u"][1:] = g * y.reshape(-1)[: F.shape[0] - 1]
    F["v"][1:] = 0.0
    F["w"][1:] = 0.0
    N = case["Grid_P"].shape[0]
    w = orc.new_work(N)
    for c_s, sgs in ((cs, 1), (0.0, 1)):
        prm = pyoracle.reference_params(mach=0.3, reyl=math.inf, sgs=sgs, c_s=c_s)
        orc.uq(N, w["U0"], w["Q_C"], case["Grid_P"], case["det"], F, w["ROE"])
        orc.viscous(N, w["U0"], w["F"], w["E"], w["G"], w["Fv"], w["Ev"], w["Gv"], case["Grid_P"], case["metric"],
                    case["det"], F, prm)
        # interior rows: the +-2 stencil along y stays clear of the periodic jump at j = 0 / 2n - 1
        j = (np.arange(1, N) - 1) // n % (2 * n)
        inner = 1 + np.nonzero((j >= 2) & (j <= 2 * n - 3))[0]
        want = -rho * c_s ** 2 * h ** 4 * g ** 2
        scale = rho * cs ** 2 * h ** 4 * g ** 2
        assert np.allclose(w["Ev"][inner, 1], want, rtol=0, atol=1e-12 * scale), (c_s, w["Ev"][inner, 1][:4], want)
        assert np.allclose(w["Ev"][inner, 4], want * F["u"][inner], rtol=0, atol=1e-12 * scale * g * 2 * n * h)
        # the SGS heat flux multiplies the temperature gradient, zero up to rounding
        assert np.abs(w["Fv"][inner, 4]).max() <= 1e-11 * scale * g * 2 * n * h


def test_smagorinsky_heat_flux_closed_form(orc):
    """The Smagorinsky option's SGS heat flux: u = g y with p linear in x (rho uniform, so the
    internal energy e = p / (0.4 rho) has the gradient b along x) and Re = inf: the xi flux of
    energy is det / h * (-1.4 / Pr) rho (C_s h)^2 |S| b = -(1.4 / Pr) rho C_s^2 h^4 g b."""
    import cfdx
    n, h, g, cs, pr = 8, 0.5, 0.3, 0.17, 0.72
    case = cfdx.periodic_case(2 * n, 2 * n, n, grid="box", length=(2 * n * h, 2 * n * h, n * h), init="freestream",
                              mach=0.3)
    xyz = case["xyz"].reshape(-1, 3) if case["xyz"].shape[-1] == 3 else case["xyz"].reshape(3, -1).T
    F = case["FLOW"]
    rho, p0 = float(F["rho"][1]), float(F["p"][1])
    m = F.shape[0] - 1
    F["u"][1:] = g * xyz[:m, 1]
    F["p"][1:] = p0 * (1.0 + 0.05 * xyz[:m, 0])
    b = p0 * 0.05 / (0.4 * rho)
    N = case["Grid_P"].shape[0]
    w = orc.new_work(N)
    prm = pyoracle.reference_params(mach=0.3, reyl=math.inf, pr=pr, sgs=1, c_s=cs)
    orc.uq(N, w["U0"], w["Q_C"], case["Grid_P"], case["det"], F, w["ROE"])
    orc.viscous(N, w["U0"], w["F"], w["E"], w["G"], w["Fv"], w["Ev"], w["Gv"], case["Grid_P"], case["metric"],
                case["det"], F, prm)
    idx = np.arange(1, N) - 1
    i, j = idx % (2 * n), idx // (2 * n) % (2 * n)
    inner = 1 + np.nonzero((i >= 2) & (i <= 2 * n - 3) & (j >= 2) & (j <= 2 * n - 3))[0]
    want = -(1.4 / pr) * rho * cs ** 2 * h ** 4 * g * b
    assert np.allclose(w["Fv"][inner, 4], want, rtol=1e-10, atol=0), (w["Fv"][inner, 4][:4], want)
