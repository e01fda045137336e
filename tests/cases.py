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


def wall_box(ni, nj, nk, mach=0.3, noise=0.05, seed=7, walls=(True, True, True)):
    """ni x nj x nk boundary+interior nodes; along each walled direction a boundary layer
    (loc 1..26 on the first / last plane) and 3 ghost layers (loc 100) beyond it; the other
    directions wrap periodically."""
    H = 3
    ext = [H if w else 0 for w in walls]
    dims = (ni, nj, nk)
    Ni, Nj, Nk = (dims[d] + 2 * ext[d] for d in range(3))
    order_core, order_ghost = [], []
    for k in range(Nk):
        for j in range(Nj):
            for i in range(Ni):
                core = all(ext[d] <= (i, j, k)[d] < ext[d] + dims[d] for d in range(3))
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
    per = [2 * np.pi / hgrid / dims[d] if not walls[d] else None for d in range(3)]

    def X(s1, s2, s3):
        return (s1 + a * np.sin(s2 + ph[0]) * np.sin(s3 + ph[0]), s2 + a * np.sin(s3 + ph[1]) * np.sin(s1 + ph[1]),
                s3 + a * np.sin(s1 + ph[2]) * np.sin(s2 + ph[2]))

    def nb(i, j, k, d, o):
        q = [i, j, k]
        q[d] += o
        if not walls[d]:
            q[d] %= dims[d]
        return ids.get(tuple(q), 0)

    jac = np.zeros((n + 1, 9))
    for (i, j, k), g in ids.items():
        # periodic directions use a map period equal to the block length (2 pi / h nodes * h)
        s = np.array([i, j, k], dtype=float) * np.array([2 * np.pi / dims[d] if not walls[d] else hgrid for d in range(3)])
        xyz[g] = X(*s)
        hs = np.array([2 * np.pi / dims[d] if not walls[d] else hgrid for d in range(3)])
        s1, s2, s3 = s
        J = np.zeros((3, 3))
        J[0] = [1, a * np.cos(s2 + ph[0]) * np.sin(s3 + ph[0]), a * np.sin(s2 + ph[0]) * np.cos(s3 + ph[0])]
        J[1] = [a * np.sin(s3 + ph[1]) * np.cos(s1 + ph[1]), 1, a * np.cos(s3 + ph[1]) * np.sin(s1 + ph[1])]
        J[2] = [a * np.cos(s1 + ph[2]) * np.sin(s2 + ph[2]), a * np.sin(s1 + ph[2]) * np.cos(s2 + ph[2]), 1]
        jac[g] = (J * hs[None, :]).reshape(-1)
        G[g]["N0"], G[g]["N1"], G[g]["N2"] = nb(i, j, k, 1, 1), nb(i, j, k, 0, 1), nb(i, j, k, 1, -1)
        G[g]["N3"], G[g]["N4"], G[g]["N5"] = nb(i, j, k, 0, -1), nb(i, j, k, 2, 1), nb(i, j, k, 2, -1)
        core = all(ext[d] <= (i, j, k)[d] < ext[d] + dims[d] for d in range(3))
        if core:
            on = []
            for d in range(3):
                c = (i, j, k)[d]
                on += [walls[d] and c == ext[d], walls[d] and c == ext[d] + dims[d] - 1]
            G[g]["loc"] = _loc_code(on)
        else:
            G[g]["loc"] = 100
        G[g]["x"], G[g]["y"], G[g]["z"] = xyz[g]
    del per
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


def _branch_offsets(LI, LM2, LM1, LP1, LP2, weno_form):
    """Stencil offsets of the loc branch Viscous.cl:285-435 / Weno.cl:2732-2856 selects
    (later matches overwrite earlier ones), or None when no branch matches."""
    used = None
    if 0 < LI <= 52 and LP1 == 100:
        used = (0, -1, -2, -3)
    if 0 < LI <= 52 and LM1 == 100:
        used = (0, 1, 2, 3)
    if LI == 0 and 0 < LP1 <= 52 and LP2 == 100:
        used = (1, 0, -1, -2)
    if LI == 0 and LM1 > 0 and (LM1 <= 52 or not weno_form) and LM2 == 100:
        used = (-1, 0, 1, 2)
    if 100 not in (LM2, LM1, LP1, LP2) and (LI == 0 or ((not weno_form) and LI > 0)):
        used = (1, -1, 2, -2)
    return used


def defined_mask(case, weno_form):
    """Nodes whose value the reference defines: every direction takes one of the loc branches
    and, for the Weno form, every node that branch reads has a defined viscous flux."""
    G = case["Grid_P"]
    loc = G["loc"]
    n1 = len(G)
    plus = {0: "N1", 1: "N0", 2: "N4"}
    minus = {0: "N3", 1: "N2", 2: "N5"}

    def walk(g, d):
        idx = {0: g}
        p = m = g
        for s in (1, 2, 3):
            p, m = G[p][plus[d]], G[m][minus[d]]
            idx[s], idx[-s] = p, m
        return idx

    def own(g, wf):
        out = []
        for d in range(3):
            idx = walk(g, d)
            used = _branch_offsets(loc[g], loc[idx[-2]], loc[idx[-1]], loc[idx[1]], loc[idx[2]], wf)
            if used is None:
                return None
            out.append([idx[o] for o in used])
        return out

    visc = np.zeros(n1, dtype=bool)
    for g in range(1, n1):
        visc[g] = own(g, False) is not None
    if not weno_form:
        return visc
    ok = np.zeros(n1, dtype=bool)
    for g in range(1, n1):
        nodes = own(g, True)
        ok[g] = nodes is not None and all(visc[x] for row in nodes for x in row)
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
