"""Test inputs in the reference layouts (test infrastructure).

periodic(): the product's periodic geometry (cfdx.periodic_case).
wall_box(): a non-periodic box with one boundary layer (loc 1..26) and three ghost
layers (loc 100) around it, numbered like the reference's subdomains (interior and
boundary nodes first, then ghosts, Main.cpp:1426), neighbours beyond the lattice
pointing at the dummy node 0. It exercises the one-sided loc branches of
Viscous.cl:285-435 and Weno.cl:2732-2856.
"""
import numpy as np

import cfdx


def periodic(ni, nj, nk, kind="sinusoid", init="tgv", mach=0.3, noise=0.05, seed=20261202):
    if kind == "sinusoid":
        h = 2 * np.pi / max(ni, nj, nk)
        return cfdx.periodic_case(ni, nj, nk, grid="sinusoid", length=(2 * np.pi,) * 3, amp=0.3 * h,
                                  phases=(0.3, 1.1, 2.0), init=init, mach=mach, noise=noise, seed=seed)
    return cfdx.periodic_case(ni, nj, nk, grid="box", length=(10.0, 10.0, 10.0), init=init, mach=mach, noise=noise,
                              center=(5.0, 5.0), seed=seed)


def _loc_code(on):
    """loc of a node from which faces of the box it lies on (0 = interior)."""
    nfaces = sum(on)
    if nfaces == 0:
        return 0
    if nfaces == 1:
        return 1 + on.index(True)               # faces 1..6
    if nfaces == 2:
        return 15 + (hash(tuple(on)) % 12)      # edges 15..26
    return 7 + (hash(tuple(on)) % 8)            # corners 7..14


def wall_box(ni, nj, nk, mach=0.3, noise=0.05, seed=7):
    """ni x nj x nk boundary+interior nodes, 3 ghost layers around them."""
    H = 3
    Ni, Nj, Nk = ni + 2 * H, nj + 2 * H, nk + 2 * H
    lat = np.zeros((Ni, Nj, Nk), dtype=np.int64)
    order_core, order_ghost = [], []
    for k in range(Nk):
        for j in range(Nj):
            for i in range(Ni):
                core = H <= i < H + ni and H <= j < H + nj and H <= k < H + nk
                (order_core if core else order_ghost).append((i, j, k))
    ids = {}
    for t, ijk in enumerate(order_core + order_ghost):
        ids[ijk] = t + 1
    n_core, n = len(order_core), len(ids)
    G = np.zeros(n + 1, dtype=cfdx.GRID_DT)
    xyz = np.zeros((n + 1, 3))
    hgrid = 2 * np.pi / max(ni, nj, nk)
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 3)
    a = 0.3 * hgrid

    def X(s1, s2, s3):
        return (s1 + a * np.sin(s2 + ph[0]) * np.sin(s3 + ph[0]), s2 + a * np.sin(s3 + ph[1]) * np.sin(s1 + ph[1]),
                s3 + a * np.sin(s1 + ph[2]) * np.sin(s2 + ph[2]))

    jac = np.zeros((n + 1, 9))
    for (i, j, k), g in ids.items():
        s = np.array([i, j, k], dtype=float) * hgrid
        xyz[g] = X(*s)
        # analytic Jacobian d(x,y,z)/d(i,j,k)
        J = np.zeros((3, 3))
        s1, s2, s3 = s
        J[0] = [1, a * np.cos(s2 + ph[0]) * np.sin(s3 + ph[0]), a * np.sin(s2 + ph[0]) * np.cos(s3 + ph[0])]
        J[1] = [a * np.sin(s3 + ph[1]) * np.cos(s1 + ph[1]), 1, a * np.cos(s3 + ph[1]) * np.sin(s1 + ph[1])]
        J[2] = [a * np.cos(s1 + ph[2]) * np.sin(s2 + ph[2]), a * np.sin(s1 + ph[2]) * np.cos(s2 + ph[2]), 1]
        jac[g] = (J * hgrid).reshape(-1)
        nb = lambda di, dj, dk: ids.get((i + di, j + dj, k + dk), 0)
        G[g]["N0"], G[g]["N1"], G[g]["N2"] = nb(0, 1, 0), nb(1, 0, 0), nb(0, -1, 0)
        G[g]["N3"], G[g]["N4"], G[g]["N5"] = nb(-1, 0, 0), nb(0, 0, 1), nb(0, 0, -1)
        core = H <= i < H + ni and H <= j < H + nj and H <= k < H + nk
        if core:
            on = [i == H, i == H + ni - 1, j == H, j == H + nj - 1, k == H, k == H + nk - 1]
            G[g]["loc"] = _loc_code(on)
        else:
            G[g]["loc"] = 100
        G[g]["x"], G[g]["y"], G[g]["z"] = xyz[g]
    jac[0] = np.eye(3).reshape(-1)
    det = np.ones(n + 1)
    metric = np.zeros((n + 1, 63))
    d_node = np.zeros(n + 1)
    m_node = np.zeros((n + 1, 63))
    cfdx._check(cfdx._lib.cfdx_geom_node_metrics(n + 1, cfdx._ptr(jac), cfdx._ptr(d_node), cfdx._ptr(m_node)))
    det[:] = d_node
    metric[:, :9] = m_node[:, :9]
    deter = np.ones((n + 1, 6))
    cfdx._check(cfdx._lib.cfdx_geom_half_metrics(cfdx._ptr(G), n, cfdx._ptr(metric), cfdx._ptr(det), cfdx._ptr(deter)))
    flow = np.zeros(n + 1, dtype=cfdx.FLOW_DT)
    cfdx._check(cfdx._lib.cfdx_geom_init_tgv(n, cfdx._ptr(np.ascontiguousarray(xyz[1:])), mach, cfdx._ptr(flow)))
    flow[0] = flow[1]
    fl = flow[1:]
    x, y, z = xyz[1:, 0], xyz[1:, 1], xyz[1:, 2]
    for f in ("rho", "u", "v", "w", "p"):
        kx, ky, kz = rng.integers(1, 3, 3)
        p3 = rng.uniform(0, 2 * np.pi, 3)
        base = fl[f].mean() if f in ("rho", "p") else 1.0
        fl[f] += noise * base * np.sin(kx * x + p3[0]) * np.cos(ky * y + p3[1]) * np.sin(kz * z + p3[2])
    fl["t"] = 1.4 * mach * mach * fl["p"] / fl["rho"]
    fl["e"] = fl["p"] / (0.4 * fl["rho"])
    fl["a"] = np.sqrt(1.4 * fl["p"] / fl["rho"])
    return dict(Grid_P=G, det=det, metric=metric, deter=deter, FLOW=flow, n_core=n_core, n=n)


def defined_mask(case, weno_form):
    """Nodes where every direction takes one of the loc branches (value defined in the reference)."""
    G = case["Grid_P"]
    loc = G["loc"]
    n1 = len(G)
    ok = np.zeros(n1, dtype=bool)
    plus = {0: "N1", 1: "N0", 2: "N4"}
    minus = {0: "N3", 1: "N2", 2: "N5"}
    for g in range(1, n1):
        good = True
        for d in range(3):
            p1 = G[g][plus[d]]; p2 = G[p1][plus[d]]
            m1 = G[g][minus[d]]; m2 = G[m1][minus[d]]
            LI, LP1, LP2, LM1, LM2 = loc[g], loc[p1], loc[p2], loc[m1], loc[m2]
            hit = (0 < LI <= 52 and LP1 == 100) or (0 < LI <= 52 and LM1 == 100) or \
                  (LI == 0 and 0 < LP1 <= 52 and LP2 == 100) or \
                  (LI == 0 and LM1 > 0 and (LM1 <= 52 or not weno_form) and LM2 == 100) or \
                  (LI == 0 and 100 not in (LM2, LM1, LP1, LP2)) or \
                  ((not weno_form) and LI > 0 and 100 not in (LM2, LM1, LP1, LP2))
            good &= bool(hit)
        ok[g] = good
    return ok


def flow_matrix(flow, fields=("rho", "u", "v", "w", "p")):
    return np.stack([flow[f] for f in fields], axis=1)


def rel_err(a, b):
    """max |a-b| / max |b| per column (component)."""
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    if a.ndim == 1:
        a, b = a[:, None], b[:, None]
    num = np.abs(a - b).max(axis=0)
    den = np.abs(b).max(axis=0)
    den[den == 0] = 1.0
    return num / den
