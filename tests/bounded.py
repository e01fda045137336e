"""Oracle-side driver for bounded blocks (test infrastructure).

Builds the reference-numbered arrays of a bounded block from its padded geometry (owned nodes
first, then the ghost and other stored nodes; `loc` 0 interior, 1 side node, 100 ghost), applies
the side conditions to FLOW between stages (the face rules of Flow_Initialisation.cpp:91-136,
:384-418, :640-700, :952-988 and the C-grid wake cut), and runs the reference stage sequence
(Solver.cpp:356-541 with the boundary fill of :657) through the oracle kernels.

Contract for what the reference leaves undefined: the stage vectors U1 / U2 of ghost nodes are
det * Q of their side-condition state; across a wake cut the ghost rows copy the nodes across
(FLOW, and the stage vectors' Q with the ghost's own det) and row 0 of the lower half copies the
upper half the same way; the viscous fluxes of the cut ghost rows are those across with the i and
j components reversed.
"""
import numpy as np

import cfdx

WALL, FARFIELD, OUTFLOW, EXTRAP, WALL_CUT = 1, 2, 3, 4, 5
PRIM = ("rho", "u", "v", "w", "p")


class Layout:
    def __init__(self, desc):
        self.desc = desc
        dims, off = cfdx.padded(desc)
        self.dims, self.off = dims, off
        self.n = (desc.ni, desc.nj, desc.nk)
        self.N = int(np.prod(dims))
        self.periodic = [bool((desc.periodic >> q) & 1) for q in range(3)]
        own = cfdx.owned_index(desc)
        rest = np.setdiff1d(np.arange(self.N), own)
        self.order = np.concatenate([own, rest])          # reference id r = position + 1
        self.rid = np.empty(self.N, dtype=np.int64)
        self.rid[self.order] = np.arange(1, self.N + 1)
        self.n_own = own.size
        self.cut = (not self.periodic[1]) and desc.bc[2] == WALL_CUT

    def g(self, i, j, k):
        """stored index of owned coordinates (arrays allowed)"""
        d, o = self.dims, self.off
        return (np.asarray(i) + o[0]) + d[0] * ((np.asarray(j) + o[1]) + d[1] * (np.asarray(k) + o[2]))

    def r(self, i, j, k):
        return self.rid[self.g(i, j, k)]

    def wall_col(self, i):
        return (i >= self.desc.cut_i0) & (i <= self.desc.cut_i1)


def loc_codes(L: Layout) -> np.ndarray:
    """Per stored node: 0 interior, 1 side node, 100 ghost; cut ghosts take the code across."""
    c = np.full(L.N, 100, dtype=np.int32)
    ni, nj, nk = L.n
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    side = np.zeros(i.shape, dtype=bool)
    for q, x in enumerate((i, j, k)):
        if L.periodic[q]:
            continue
        side |= x == L.n[q] - 1
        lo = x == 0
        if q == 1 and L.cut:
            lo &= L.wall_col(i)
        side |= lo
    c[L.g(i, j, k)] = np.where(side, 1, 0)
    if L.cut:
        kk, ii = np.meshgrid(np.arange(nk), np.arange(ni), indexing="ij")
        m_ = ~L.wall_col(ii)
        for m in (1, 2, 3):
            c[L.g(ii[m_], -m, kk[m_])] = c[L.g(ni - 1 - ii[m_], m, kk[m_])]
    return c


def reference_case(desc, geo: np.ndarray, flow_owned: np.ndarray, mach: float) -> dict:
    L = Layout(desc)
    N = L.N
    geo = geo.reshape(22, N)
    det_s, m_s, kf_s, df_s = geo[0], geo[1:10], geo[10:19], geo[19:22]
    dims = L.dims
    strides = (1, dims[0], dims[0] * dims[1])
    coords = np.stack(np.unravel_index(np.arange(N), (dims[2], dims[1], dims[0]))[::-1], axis=1)  # (I, J, K)

    def neighbour(q, o):
        x = coords[:, q] + o
        if L.periodic[q]:
            x = x % dims[q]
            ok = np.ones(N, dtype=bool)
        else:
            ok = (x >= 0) & (x < dims[q])
        nb = np.arange(N) + o * strides[q]
        if L.periodic[q]:
            nb = np.arange(N) + (x - coords[:, q]) * strides[q]
        return np.where(ok, nb, -1)

    G = np.zeros(N + 1, dtype=cfdx.GRID_DT)
    code = loc_codes(L)
    rid = L.rid
    for name, q, o in (("N1", 0, 1), ("N3", 0, -1), ("N0", 1, 1), ("N2", 1, -1), ("N4", 2, 1), ("N5", 2, -1)):
        nb = neighbour(q, o)
        G[name][rid] = np.where(nb >= 0, rid[np.maximum(nb, 0)], 0)
    G["loc"][rid] = code
    det = np.ones(N + 1)
    det[rid] = det_s
    metric = np.zeros((N + 1, 63))
    metric[0, [0, 4, 8]] = 1.0
    metric[0, 9:] = np.tile([1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0], 6)
    for c in range(9):
        metric[rid, c] = m_s[c]
    deter = np.ones((N + 1, 6))
    for d in range(3):
        minus = neighbour(d, -1)
        for c in range(3):
            metric[rid, 9 + 9 * (2 * d) + 3 * d + c] = kf_s[3 * d + c]
            metric[rid, 9 + 9 * (2 * d + 1) + 3 * d + c] = np.where(minus >= 0, kf_s[3 * d + c][np.maximum(minus, 0)], 0.0)
        deter[rid, 2 * d] = df_s[d]
        deter[rid, 2 * d + 1] = np.where(minus >= 0, df_s[d][np.maximum(minus, 0)], 0.0)
    FLOW = np.zeros(N + 1, dtype=cfdx.FLOW_DT)
    fs = cfdx.freestream(N + 1, mach)
    for f in ("rho", "u", "v", "w", "p", "t", "e", "a"):
        FLOW[f] = fs[f]
    FLOW[1:L.n_own + 1] = flow_owned
    rc = dict(layout=L, Grid_P=G, det=det, metric=metric, deter=deter, FLOW=FLOW, mach=mach)
    rc["mirror_pairs"] = mirror_pairs(L)
    apply_bcs(rc, FLOW, None, 2)
    return rc


def mirror_pairs(L: Layout):
    """(dst, src) reference ids: cut ghost rows <- rows across, lower row 0 <- upper row 0."""
    if not L.cut:
        return np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64)
    ni, nj, nk = L.n
    kk, ii = np.meshgrid(np.arange(nk), np.arange(ni), indexing="ij")
    m_ = ~L.wall_col(ii)
    ic, kc = ii[m_], kk[m_]
    dst, src = [], []
    for m in (1, 2, 3):
        dst.append(L.r(ic, -m, kc)); src.append(L.r(ni - 1 - ic, m, kc))
    low = ic < ni - 1 - ic
    dst.append(L.r(ic[low], 0, kc[low])); src.append(L.r(ni - 1 - ic[low], 0, kc[low]))
    return np.concatenate(dst), np.concatenate(src)


def prim_to_u(FLOW, det, ids):
    rho, u, v, w, p = (FLOW[f][ids] for f in PRIM)
    ke = 0.5 * rho * (u * u + v * v + w * w)
    q = np.stack([rho, rho * u, rho * v, rho * w, p * 2.5 + ke], axis=1)
    return det[ids][:, None] * q


def apply_bcs(rc, FLOW, Unext, mode, U0=None):
    """Side conditions in the order i-, i+, j-, j+, k-, k+ (all face nodes of a side at once).
    Unext (mode 0 / 1): stage vector of the next stage, set at ghosts; cut copies also carry U0."""
    L, M = rc["layout"], rc["mach"]
    M2 = 1.4 * M * M
    a = L.desc.aoa
    ni, nj, nk = L.n
    for side in range(6):
        q, hi = side // 2, side % 2
        if L.periodic[q]:
            continue
        o1, o2 = (1, 2) if q == 0 else ((0, 2) if q == 1 else (0, 1))
        B_, A_ = np.meshgrid(np.arange(L.n[o2]), np.arange(L.n[o1]), indexing="ij")
        A_, B_ = A_.reshape(-1), B_.reshape(-1)
        n = L.n[q]
        c0, step = (n - 1, -1) if hi else (0, 1)

        def at(c, A=A_, B=B_):
            x = [None, None, None]
            x[q], x[o1], x[o2] = c, A, B
            return L.r(*[np.broadcast_to(v, A.shape) for v in x])

        kind = np.full(A_.shape, L.desc.bc[side])
        if L.desc.bc[side] == WALL_CUT:
            kind = np.where(L.wall_col(A_), WALL, WALL_CUT)
        cut = kind == WALL_CUT
        if cut.any():
            dst, src = [], []
            Ac, Bc = A_[cut], B_[cut]
            for m in (1, 2, 3):
                dst.append(at(-m, Ac, Bc)); src.append(L.r(ni - 1 - Ac, m, Bc))
            low = Ac < ni - 1 - Ac
            dst.append(L.r(Ac[low], 0, Bc[low])); src.append(L.r(ni - 1 - Ac[low], 0, Bc[low]))
            dst, src = np.concatenate(dst), np.concatenate(src)
            FLOW[dst] = FLOW[src]
            ratio = (rc["det"][dst] / rc["det"][src])[:, None]
            for arr in (Unext, U0):
                if arr is not None:
                    arr[dst] = arr[src] * ratio
        for kd in (WALL, FARFIELD, OUTFLOW, EXTRAP):
            sel = kind == kd
            if not sel.any():
                continue
            A, B = A_[sel], B_[sel]
            b0 = at(c0, A, B)
            n1 = at(c0 + step, A, B)
            inner = [at(c0 + step * m, A, B) for m in (1, 2, 3)]
            ghost = [at(c0 - step * m, A, B) for m in (1, 2, 3)]
            rho1, p1 = FLOW["rho"][n1].copy(), FLOW["p"][n1].copy()
            vel1 = [FLOW[f][n1].copy() for f in ("u", "v", "w")]
            innerv = [{f: FLOW[f][g].copy() for f in PRIM} for g in inner]
            if kd == WALL:
                t1 = M2 * (p1 / rho1)
                w0 = {"rho": M2 * (p1 / t1), "u": 0.0 * p1, "v": 0.0 * p1, "w": 0.0 * p1, "p": p1}
                gw = [{"rho": iv["rho"], "u": -iv["u"], "v": -iv["v"], "w": -iv["w"], "p": iv["p"]} for iv in innerv]
            elif kd == FARFIELD:
                p = 1.0 / M2
                w0 = {"rho": np.full(b0.shape, M2 * (p / 1.0)), "u": np.full(b0.shape, np.cos(a)),
                      "v": np.full(b0.shape, np.sin(a)), "w": np.zeros(b0.shape), "p": np.full(b0.shape, p)}
                gw = [w0] * 3
            elif kd == OUTFLOW:
                t1 = M2 * (p1 / rho1)
                w0 = {"rho": M2 * (p1 / t1), "u": vel1[0], "v": vel1[1], "w": vel1[2], "p": p1}
                gw = [w0] * 3
            else:
                w0 = {"rho": rho1, "u": vel1[0], "v": vel1[1], "w": vel1[2], "p": p1}
                gw = innerv
            for f in PRIM:
                FLOW[f][b0] = w0[f]
            for g, wv in zip(ghost, gw):
                for f in PRIM:
                    FLOW[f][g] = wv[f]
            if Unext is not None and mode < 2:
                gids = np.concatenate(ghost)
                Unext[gids] = prim_to_u(FLOW, rc["det"], gids)


def oracle_steps(orc, rc, nsteps, mach, dt, prm):
    """nsteps RK3 steps of the bounded case; returns the owned FLOW."""
    L = rc["layout"]
    n = L.N + 1
    FLOW = rc["FLOW"].copy()
    w = orc.new_work(n)
    G, det, metric, deter = rc["Grid_P"], rc["det"], rc["metric"], rc["deter"]
    dst, src = rc["mirror_pairs"]
    ghost_rows = None
    if L.cut:
        ni, nj, nk = L.n
        kk, ii = np.meshgrid(np.arange(nk), np.arange(ni), indexing="ij")
        m_ = ~L.wall_col(ii)
        ghost_rows = [(L.r(ii[m_], -m, kk[m_]), L.r(ni - 1 - ii[m_], m, kk[m_])) for m in (1, 2, 3)]
    for _ in range(nsteps):
        for j in range(3):
            orc.uq(n, w["U0"] if j == 0 else w["F"], w["Q_C"], G, det, FLOW, w["ROE"])
            if j == 0 and dst.size:
                w["U0"][dst] = w["U0"][src] * (det[dst] / det[src])[:, None]
            Ust = (w["U0"], w["U1"], w["U2"])[j]
            orc.viscous(n, Ust, w["F"], w["E"], w["G"], w["Fv"], w["Ev"], w["Gv"], G, metric, det, FLOW, prm)
            if ghost_rows:
                for gd, gs in ghost_rows:
                    w["Fv"][gd] = -w["Fv"][gs]
                    w["Ev"][gd] = -w["Ev"][gs]
                    w["Gv"][gd] = w["Gv"][gs]
            orc.weno(L.n_own + 1, j, mach, dt, deter, det, w["U0"], w["U1"], w["U2"], G, w["Q_C"], metric, w["ROE"],
                     w["Fv"], w["Ev"], w["Gv"], w["F"], w["E"], w["G"], FLOW, prm)
            apply_bcs(rc, FLOW, (w["U1"], w["U2"], None)[j], j, U0=w["U0"])
    return FLOW[1:L.n_own + 1].copy()
