"""cfdx — MI355X-native per-stage RHS + RK3 update of the high-order curvilinear
CFD solver, behind the reference's kernel interface (U_Q / Viscous / Weno).

Host-side mirror of the reference's Solver.cpp (stage sequence and argument
layout) over the C ABI of libcfdx.so (include/cfdx.h). Device memory, copies,
streams and events go through that library too; every computation runs in its
HIP kernels, and importing this package fails loudly when the library is missing.

Do not import torch in a process that uses this package on the GPU: torch's
wheel ships its own HIP runtime under another file name, and two HIP runtimes
in one process do not share the device.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcfdx.so")

# ---------------------------------------------------------------------------
# packed layouts (functions.h:49-112)
# ---------------------------------------------------------------------------
FLOW_FIELDS = ["u", "v", "w", "p", "t", "rho", "mu", "a", "e"]
FLOW_DT = np.dtype([(f, "<f8") for f in FLOW_FIELDS])                       # FLOW_VARIABLES, 72 B
GRID_DT = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")] +
                   [(f"N{i}", "<i4") for i in range(6)] + [("ID", "<i4"), ("loc", "<i4")])  # GRID, 56 B
VEC5_N, ROE_N, METRIC_N, DETER_N = 5, 36, 63, 6                            # doubles per record
assert FLOW_DT.itemsize == 72 and GRID_DT.itemsize == 56

FP64, MIXED = 0, 1
E_INVALID, E_HIP, E_NOMEM, E_UNSUPPORTED, E_NONFINITE, E_CAPACITY = -1, -2, -3, -4, -5, -6


class Params(C.Structure):
    """cfdx_params: the constants the reference hard-codes, plus precision / viscous switches."""
    _fields_ = [("mach", C.c_double), ("reyl", C.c_double), ("pr", C.c_double), ("free_t", C.c_double),
                ("c_mts", C.c_double), ("c_t", C.c_double), ("eps", C.c_double),
                ("precision", C.c_int32), ("viscous", C.c_int32), ("sgs", C.c_int32), ("pad_", C.c_int32),
                ("c_s", C.c_double)]


BC_WALL, BC_FARFIELD, BC_OUTFLOW, BC_EXTRAP, BC_WALL_CUT = 1, 2, 3, 4, 5


class BlockDesc(C.Structure):
    _fields_ = [("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32), ("periodic", C.c_int32),
                ("precision", C.c_int32), ("bc", C.c_int32 * 6), ("cut_i0", C.c_int32), ("cut_i1", C.c_int32),
                ("aoa", C.c_double)]


class CGrid(C.Structure):
    _fields_ = [("m", C.c_double), ("p", C.c_double), ("t", C.c_double), ("wake", C.c_double),
                ("radius", C.c_double), ("wall", C.c_double), ("span", C.c_double), ("n_wake", C.c_int32),
                ("pad_", C.c_int32)]


INIT_KINDS = {"freestream": 0, "vortex": 1, "tgv": 2}


class CaseDesc(C.Structure):
    _fields_ = [("length", C.c_double * 3), ("amp", C.c_double), ("phase", C.c_double * 3), ("init", C.c_int32),
                ("pad_", C.c_int32), ("mach", C.c_double), ("beta", C.c_double), ("center", C.c_double * 2)]


class Tiles(C.Structure):
    """cfdx_tiles: (face i, face j, face k, node) variant indices and the node-tile strip (rows of
    node tiles per strip walked plane by plane; 0: whole planes)."""
    _fields_ = [("face", C.c_int32 * 3), ("node", C.c_int32), ("strip", C.c_int32)]

    def as_tuple(self) -> tuple:
        return (self.face[0], self.face[1], self.face[2], self.node, self.strip)

    @classmethod
    def of(cls, t) -> "Tiles":
        return cls((C.c_int32 * 3)(t[0], t[1], t[2]), t[3], t[4] if len(t) > 4 else 0)


class CfdxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cfdx error {code}: {msg}")
        self.code = code


# ---------------------------------------------------------------------------
# library
# ---------------------------------------------------------------------------
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with build.py (hipcc, gfx950) before importing cfdx")

_lib = C.CDLL(LIB_PATH)
_vp, _i64, _i32, _d = C.c_void_p, C.c_int64, C.c_int32, C.c_double

EXPORTS = {
    # name: (restype, argtypes)
    "cfdx_last_error": (C.c_char_p, []),
    "cfdx_params_default": (None, [C.POINTER(Params)]),
    "cfdx_ctx_create": (C.c_int, [C.c_int, _vp, C.POINTER(_vp)]),
    "cfdx_ctx_destroy": (C.c_int, [_vp]),
    "cfdx_ctx_set_stream": (C.c_int, [_vp, _vp]),
    "cfdx_ctx_set_overlap": (C.c_int, [_vp, C.c_int]),
    "cfdx_ctx_stream": (_vp, [_vp]),
    "cfdx_sync": (C.c_int, [_vp]),
    "cfdx_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "cfdx_mem_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "cfdx_pool_trim": (C.c_int, [_vp]),
    "cfdx_device_alloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "cfdx_device_free": (C.c_int, [_vp, _vp]),
    "cfdx_copy_to_device": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "cfdx_copy_to_host": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "cfdx_device_copy": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "cfdx_event_record": (C.c_int, [_vp, C.c_int]),
    "cfdx_event_elapsed": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(_d)]),
    "cfdx_uq": (C.c_int, [_vp] * 7 + [_i64]),
    "cfdx_viscous": (C.c_int, [_vp] * 12 + [_i64, C.POINTER(Params)]),
    "cfdx_weno": (C.c_int, [_vp, C.c_int, _d, _d] + [_vp] * 16 + [_i64, C.POINTER(Params)]),
    "cfdx_block_bytes": (_i64, [C.POINTER(BlockDesc)]),
    "cfdx_block_create": (C.c_int, [_vp, C.POINTER(BlockDesc), C.POINTER(_vp)]),
    "cfdx_block_destroy": (C.c_int, [_vp]),
    "cfdx_block_set_geometry": (C.c_int, [_vp, _vp, _vp, _vp]),
    "cfdx_block_set_geometry_soa": (C.c_int, [_vp, _vp]),
    "cfdx_block_padded": (C.c_int, [C.POINTER(BlockDesc), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "cfdx_geom_cgrid_desc": (C.c_int, [_i32, _i32, _i32, _i32, C.POINTER(CGrid), _d, C.POINTER(BlockDesc)]),
    "cfdx_geom_cgrid_coords": (C.c_int, [C.POINTER(BlockDesc), C.POINTER(CGrid), _vp]),
    "cfdx_geom_box_coords": (C.c_int, [C.POINTER(BlockDesc), _vp, _d, _vp, _vp]),
    "cfdx_geom_block_metrics": (C.c_int, [C.POINTER(BlockDesc), _vp, _vp, _vp]),
    "cfdx_block_set_flow": (C.c_int, [_vp, _vp]),
    "cfdx_block_get_flow": (C.c_int, [_vp, _vp]),
    "cfdx_block_init_case": (C.c_int, [_vp, C.POINTER(CaseDesc)]),
    "cfdx_block_cfl_dt": (C.c_int, [_vp, _d, C.POINTER(_d)]),
    "cfdx_block_check_finite": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "cfdx_block_configure": (C.c_int, [_vp, _d, _d, C.POINTER(Params)]),
    "cfdx_block_step": (C.c_int, [_vp, _d, _d, C.c_int, C.POINTER(Params)]),
    "cfdx_batch_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "cfdx_batch_destroy": (C.c_int, [_vp]),
    "cfdx_batch_add": (C.c_int, [_vp, _vp]),
    "cfdx_batch_remove": (C.c_int, [_vp, _vp]),
    "cfdx_batch_size": (C.c_int, [_vp]),
    "cfdx_batch_set_grid": (C.c_int, [_vp, C.c_uint32]),
    "cfdx_batch_step": (C.c_int, [_vp, C.c_int, C.c_int]),
    "cfdx_block_rhs": (C.c_int, [_vp, C.POINTER(Params), _vp]),
    "cfdx_block_rhs_noalloc": (C.c_int, [_vp, C.POINTER(Params), _vp]),
    "cfdx_block_set_timing": (C.c_int, [_vp, C.c_int]),
    "cfdx_block_timing": (C.c_int, [_vp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d), C.POINTER(C.c_int)]),
    "cfdx_block_timing_split": (C.c_int, [_vp, C.POINTER(_d)]),
    "cfdx_tile_variants": (C.c_int, [C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "cfdx_block_set_tiles": (C.c_int, [_vp, C.POINTER(Tiles)]),
    "cfdx_block_get_tiles": (C.c_int, [_vp, C.POINTER(Tiles)]),
    "cfdx_batch_set_tiles": (C.c_int, [_vp, C.POINTER(Tiles)]),
    "cfdx_op_count": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(_i64)]),
    "cfdx_math_eval": (C.c_int, [_vp, C.c_int, _vp, _vp, _i64]),
    "cfdx_block_set_monitor": (C.c_int, [_vp, C.c_int]),
    "cfdx_block_set_check": (C.c_int, [_vp, C.c_int]),
    "cfdx_block_residual": (C.c_int, [_vp, C.POINTER(_d)]),
    "cfdx_write_restart": (C.c_int, [C.c_char_p, _i32, _i64, _vp, _i64, _i32]),
    "cfdx_restart_latest": (C.c_int, [C.c_char_p, _i64, C.POINTER(_i32), C.POINTER(_i64)]),
    "cfdx_read_restart": (C.c_int, [C.c_char_p, _i32, _i64, _vp, C.POINTER(_i64)]),
    "cfdx_write_tecplot": (C.c_int, [C.c_char_p, _i32, _i32, _i32, _vp, _vp, _vp, _d]),
    "cfdx_geom_periodic_table": (C.c_int, [_i32, _i32, _i32, _vp, _vp]),
    "cfdx_geom_coords_box": (C.c_int, [_i32, _i32, _i32, _vp, _vp]),
    "cfdx_geom_coords_sinusoid": (C.c_int, [_i32, _i32, _i32, _vp, _d, _vp, _vp]),
    "cfdx_geom_jacobian_sinusoid": (C.c_int, [_i32, _i32, _i32, _vp, _d, _vp, _vp]),
    "cfdx_geom_jacobian_compact_periodic": (C.c_int, [_i32, _i32, _i32, _vp, _vp, _vp]),
    "cfdx_geom_node_metrics": (C.c_int, [_i64, _vp, _vp, _vp]),
    "cfdx_geom_half_metrics": (C.c_int, [_vp, _i64, _vp, _vp, _vp]),
    "cfdx_geom_init_freestream": (C.c_int, [_i64, _d, _vp]),
    "cfdx_geom_init_vortex": (C.c_int, [_i64, _vp, _d, _d, _d, _d, _vp]),
    "cfdx_geom_init_tgv": (C.c_int, [_i64, _vp, _d, _vp]),
}

for _name, (_res, _args) in EXPORTS.items():
    _f = getattr(_lib, _name)   # AttributeError here means the library lacks a declared symbol
    _f.restype = _res
    _f.argtypes = _args


def _check(rc: int) -> None:
    if rc != 0:
        raise CfdxError(rc, _lib.cfdx_last_error().decode())


def _ptr(a) -> C.c_void_p:
    """Address of a numpy array (host) or a DeviceArray (device)."""
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, np.ndarray):
        return C.c_void_p(a.ctypes.data)
    return C.c_void_p(a.ptr)


def tile_variants(kind: int) -> list[tuple[int, int]]:
    """Compiled tile shapes of a kernel class: 0 = face sweep along i (lines across j, faces
    along i), 1 = face sweeps along j / k (lines across i, faces along the sweep), 2 = node
    kernels (along i, along j)."""
    n = C.c_int32(0)
    _check(_lib.cfdx_tile_variants(kind, C.byref(n), None))
    shapes = (C.c_int32 * (2 * n.value))()
    _check(_lib.cfdx_tile_variants(kind, C.byref(n), shapes))
    return [(shapes[2 * v], shapes[2 * v + 1]) for v in range(n.value)]


def device_count() -> int:
    n = C.c_int(0)
    rc = _lib.cfdx_device_count(C.byref(n))
    return n.value if rc == 0 else 0


SGS_MTS, SGS_SMAGORINSKY = 0, 1


def default_params(**kw) -> Params:
    """cfdx_params_default with the given fields replaced (names of cfdx_params; sgs="smagorinsky"
    or "mts" is accepted for the SGS switch)."""
    p = Params()
    _lib.cfdx_params_default(C.byref(p))
    names = {f[0] for f in Params._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"default_params: no parameter {k!r}")
        if k == "sgs" and isinstance(v, str):
            v = {"mts": SGS_MTS, "smagorinsky": SGS_SMAGORINSKY}[v]
        setattr(p, k, v)
    return p


def euler_params(**kw) -> Params:
    """Inviscid form: infinite Re, no SGS model, viscous kernels skipped."""
    return default_params(reyl=math.inf, c_mts=0.0, viscous=0, **kw)


# ---------------------------------------------------------------------------
# host geometry -> reference-numbered case (node 0 = dummy)
# ---------------------------------------------------------------------------
def periodic_case(ni: int, nj: int, nk: int, grid: str = "box", length=(10.0, 10.0, 10.0), amp: float = 0.0,
                  phases=(0.0, 0.0, 0.0), init: str = "freestream", mach: float = 0.5, beta: float = 5.0,
                  center=(5.0, 5.0), metrics: str = "analytic", seed: int = 20261202, noise: float = 0.0) -> dict:
    """Builds Grid_P, det, metric (node + half point), deter and FLOW for a periodic block.

    grid: 'box' (Cartesian) or 'sinusoid' (seeded curvilinear map, amplitude amp).
    init: 'freestream', 'vortex' (isentropic, strength beta) or 'tgv' (Taylor-Green).
    metrics: 'analytic' (exact Jacobian of the map) or 'compact' (6th-order compact scheme).
    noise: optional seeded smooth perturbation amplitude added to rho, u, v, w, p (tests).
    """
    n = ni * nj * nk
    L = np.ascontiguousarray(length, dtype=np.float64)
    ph = np.ascontiguousarray(phases, dtype=np.float64)
    xyz = np.zeros((n, 3))
    if grid == "box":
        _check(_lib.cfdx_geom_coords_box(ni, nj, nk, _ptr(L), _ptr(xyz)))
        a = 0.0
    elif grid == "sinusoid":
        _check(_lib.cfdx_geom_coords_sinusoid(ni, nj, nk, _ptr(L), amp, _ptr(ph), _ptr(xyz)))
        a = amp
    else:
        raise ValueError(grid)
    jac = np.zeros((n, 9))
    if metrics == "analytic":
        _check(_lib.cfdx_geom_jacobian_sinusoid(ni, nj, nk, _ptr(L), a, _ptr(ph), _ptr(jac)))
    elif metrics == "compact":
        _check(_lib.cfdx_geom_jacobian_compact_periodic(ni, nj, nk, _ptr(L), _ptr(xyz), _ptr(jac)))
    else:
        raise ValueError(metrics)
    det = np.ones(n + 1)
    metric = np.zeros((n + 1, METRIC_N))
    metric[0, [0, 4, 8]] = 1.0  # dummy node: identity metrics
    metric[0, 9:] = np.tile([1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0], 6)
    d_node = np.zeros(n)
    m_node = np.zeros((n, METRIC_N))
    _check(_lib.cfdx_geom_node_metrics(n, _ptr(jac), _ptr(d_node), _ptr(m_node)))
    det[1:] = d_node
    metric[1:, :9] = m_node[:, :9]
    grid_p = np.zeros(n + 1, dtype=GRID_DT)
    _check(_lib.cfdx_geom_periodic_table(ni, nj, nk, _ptr(xyz), _ptr(grid_p)))
    deter = np.ones((n + 1, DETER_N))
    _check(_lib.cfdx_geom_half_metrics(_ptr(grid_p), n, _ptr(metric), _ptr(det), _ptr(deter)))
    flow = np.zeros(n + 1, dtype=FLOW_DT)
    if init == "freestream":
        _check(_lib.cfdx_geom_init_freestream(n, mach, _ptr(flow)))
    elif init == "vortex":
        _check(_lib.cfdx_geom_init_vortex(n, _ptr(xyz), mach, beta, center[0], center[1], _ptr(flow)))
    elif init == "tgv":
        _check(_lib.cfdx_geom_init_tgv(n, _ptr(xyz), mach, _ptr(flow)))
    else:
        raise ValueError(init)
    if noise:
        rng = np.random.default_rng(seed)
        x, y, z = (xyz[:, 0] * 2 * np.pi / L[0], xyz[:, 1] * 2 * np.pi / L[1], xyz[:, 2] * 2 * np.pi / L[2])
        for f, scale in (("rho", 1.0), ("u", 1.0), ("v", 1.0), ("w", 1.0), ("p", 1.0)):
            kx, ky, kz = rng.integers(1, 3, size=3)
            phs = rng.uniform(0, 2 * np.pi, size=3)
            base = np.abs(flow[f][1:]).mean() + 1.0 if f in ("rho", "p") else 1.0
            flow[f][1:] += noise * base * scale * (np.sin(kx * x + phs[0]) * np.cos(ky * y + phs[1]) * np.sin(kz * z + phs[2]))
        fl = flow[1:]
        fl["t"] = 1.4 * mach * mach * fl["p"] / fl["rho"]
        fl["e"] = fl["p"] / (0.4 * fl["rho"])
        fl["a"] = np.sqrt(1.4 * fl["p"] / fl["rho"])
    return dict(ni=ni, nj=nj, nk=nk, n=n, xyz=xyz, jac=jac, Grid_P=grid_p, det=det, metric=metric, deter=deter,
                FLOW=flow)


# ---------------------------------------------------------------------------
# bounded blocks: padded layout, C-grid and box geometry
# ---------------------------------------------------------------------------
def padded(desc: BlockDesc) -> tuple[tuple, tuple]:
    """(stored sizes, owned offsets) of a descriptor: 3 ghost layers per side of a bounded direction."""
    dims, off = (C.c_int32 * 3)(), (C.c_int32 * 3)()
    _check(_lib.cfdx_block_padded(C.byref(desc), dims, off))
    return tuple(dims), tuple(off)


def cgrid_desc(ni: int, nj: int, nk: int, precision: int = FP64, naca=(0.0, 0.4, 0.12), wake: float = 20.0,
               radius: float = 20.0, wall: float = 1e-4, span: float = 0.25, n_wake: int | None = None,
               aoa_deg: float = 0.0) -> tuple[BlockDesc, CGrid]:
    """Descriptor and parameters of a NACA 4-digit C-grid block (naca = (m, p, t) as chord fractions)."""
    g = CGrid(naca[0], naca[1], naca[2], wake, radius, wall, span, n_wake if n_wake else ni // 4, 0)
    d = BlockDesc()
    _check(_lib.cfdx_geom_cgrid_desc(ni, nj, nk, precision, C.byref(g), math.radians(aoa_deg), C.byref(d)))
    return d, g


def cgrid_geometry(desc: BlockDesc, g: CGrid, full: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """(xyz of the stored nodes, the 22 geometry arrays) of a C-grid block. The grid is a spanwise
    extrusion, so every k plane has the same metrics: they are computed on a 3-plane block of the
    same spanwise spacing and repeated (full=True computes every plane)."""
    dims, _ = padded(desc)
    N = int(np.prod(dims))
    if full or desc.nk <= 3:
        xyz = np.zeros((N, 3))
        _check(_lib.cfdx_geom_cgrid_coords(C.byref(desc), C.byref(g), _ptr(xyz)))
        L = np.array([0.0, 0.0, g.span])
        geo = np.zeros(22 * N)
        _check(_lib.cfdx_geom_block_metrics(C.byref(desc), _ptr(xyz), _ptr(L), _ptr(geo)))
        return xyz, geo
    d3 = BlockDesc.from_buffer_copy(desc)
    d3.nk = 3
    g3 = CGrid.from_buffer_copy(g)
    g3.span = 3.0 * g.span / desc.nk
    x3, geo3 = cgrid_geometry(d3, g3, full=True)
    plane = dims[0] * dims[1]
    xyz = np.tile(x3[:plane], (desc.nk, 1))
    xyz[:, 2] = np.repeat(np.arange(desc.nk) * (g.span / desc.nk), plane)
    geo = np.tile(geo3.reshape(22, 3, plane)[:, :1, :], (1, desc.nk, 1)).reshape(-1)
    return xyz, geo


def box_geometry(desc: BlockDesc, length=(1.0, 1.0, 1.0), amp: float = 0.0, phases=(0.0, 0.0, 0.0)):
    """(xyz, geometry arrays) of a sinusoidally mapped box over the stored nodes of a descriptor."""
    dims, _ = padded(desc)
    N = int(np.prod(dims))
    L = np.ascontiguousarray(length, dtype=np.float64)
    ph = np.ascontiguousarray(phases, dtype=np.float64)
    xyz = np.zeros((N, 3))
    _check(_lib.cfdx_geom_box_coords(C.byref(desc), _ptr(L), amp, _ptr(ph), _ptr(xyz)))
    geo = np.zeros(22 * N)
    _check(_lib.cfdx_geom_block_metrics(C.byref(desc), _ptr(xyz), _ptr(L), _ptr(geo)))
    return xyz, geo


def owned_index(desc: BlockDesc) -> np.ndarray:
    """Stored index of every owned node, in owned order."""
    dims, off = padded(desc)
    k, j, i = np.meshgrid(np.arange(desc.nk), np.arange(desc.nj), np.arange(desc.ni), indexing="ij")
    return ((i + off[0]) + dims[0] * ((j + off[1]) + dims[1] * (k + off[2]))).reshape(-1)


def freestream(n: int, mach: float, aoa_deg: float = 0.0) -> np.ndarray:
    """Uniform free stream u = cos(aoa), v = sin(aoa), p = 1 / (1.4 M^2), T = 1 (Flow_Initialisation.cpp:35-48)."""
    f = np.zeros(n, dtype=FLOW_DT)
    a = math.radians(aoa_deg)
    p = 1.0 / (1.4 * mach * mach)
    f["u"], f["v"], f["w"], f["p"], f["t"] = math.cos(a), math.sin(a), 0.0, p, 1.0
    f["rho"] = (1.4 * mach * mach) * (p / 1.0)
    f["e"] = p / (0.4 * f["rho"])
    f["a"] = np.sqrt(1.4 * p / f["rho"])
    return f


VISC_RADIUS = 1.372 * 1.372


def cfl_dt(case: dict, cfl: float, params: "Params | None" = None) -> float:
    """dt = cfl / max_nodes (r_c + r_v) on the initial state (computed once), the formula of
    cfdx_block_cfl_dt: r_c = sum_d (|u . grad xi_d| + c |grad xi_d|); with viscous params,
    r_v = 1.372^2 max(4/3, 1.4/Pr) mu / (rho Re) (sum_d |grad xi_d|)^2 (mu by Sutherland's law at
    the node temperature 1.4 M^2 0.4 e with the params' Mach, T_inf). The reference itself runs a
    fixed del_t (Main.cpp:47); this step is the build's choice for its workloads."""
    # the owned nodes: a bounded case (tests/bounded.py) numbers them first, before its ghosts
    n = case["layout"].n_own if "layout" in case else len(case["FLOW"]) - 1
    fl = case["FLOW"][1:n + 1]
    m = case["metric"][1:n + 1, :9].reshape(-1, 3, 3)
    vel = np.stack([fl["u"], fl["v"], fl["w"]], axis=1)
    c = np.sqrt(1.4 * fl["p"] / fl["rho"])
    s = np.zeros(len(fl))
    sm = np.zeros(len(fl))
    for d in range(3):
        mm = np.linalg.norm(m[:, d], axis=1)
        s += np.abs(np.einsum("ni,ni->n", vel, m[:, d])) + c * mm
        sm += mm
    if params is not None and params.viscous:
        T = (1.4 * params.mach * params.mach * 0.4) * (fl["p"] / (0.4 * fl["rho"]))
        suth = 110.4 / params.free_t
        mu = (T * np.sqrt(T)) * ((1.0 + suth) / (T + suth))
        s += VISC_RADIUS * max(4.0 / 3.0, 1.4 / params.pr) * (mu * (1.0 / params.reyl) / fl["rho"]) * sm * sm
    return float(cfl / s.max())


# ---------------------------------------------------------------------------
# context + drop-in kernels on device buffers
# ---------------------------------------------------------------------------
class Ctx:
    """A device + stream (owned by the context unless a hipStream_t address is given)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        h = _vp()
        _check(_lib.cfdx_ctx_create(device, _vp(stream) if stream else None, C.byref(h)))
        self.h = h
        self.device = device

    def sync(self) -> None:
        _check(_lib.cfdx_sync(self.h))

    def close(self) -> None:
        if self.h:
            _check(_lib.cfdx_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mem_info(self) -> tuple[int, int]:
        """(free, total) bytes of the device (hipMemGetInfo)."""
        f, t = _i64(0), _i64(0)
        _check(_lib.cfdx_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def set_overlap(self, on: bool) -> None:
        """Mixed-precision stages run the viscous kernel beside the face sweeps on a second stream
        (cfdx_ctx_set_overlap; on by default)."""
        _check(_lib.cfdx_ctx_set_overlap(self.h, 1 if on else 0))

    def pool_trim(self) -> None:
        """Frees the block memory kept for reuse on this device (cfdx_pool_trim)."""
        _check(_lib.cfdx_pool_trim(self.h))

    def op_count(self, params=None) -> dict:
        """Static operation count of the scheme's arithmetic per face / node (cfdx_op_count):
        {group: {add, mul, div, sqrt, rcp, cbrt, total}} for groups face, cons, node_flux, viscous."""
        out = (_i64 * 24)()
        _check(_lib.cfdx_op_count(self.h, C.byref(params) if params is not None else None, out))
        res = {}
        for g, name in enumerate(("face", "cons", "node_flux", "viscous")):
            d = dict(zip(("add", "mul", "div", "sqrt", "rcp", "cbrt"), out[6 * g:6 * g + 6]))
            d["total"] = sum(d.values())
            res[name] = d
        return res

    MATH_FNS = ("pinv", "psqrt", "prsq", "prcbrt")

    def math_eval(self, fn: str, x, precision: str = "fp64"):
        """The kernels' positive-operand helper `fn` (pinv 1/x, psqrt sqrt x, prsq 1/sqrt x,
        prcbrt x^(-1/3)) evaluated on the device in the plain (fp64) or mixed (fp32) form
        (cfdx_math_eval): a float64 array of the results."""
        if fn not in self.MATH_FNS or precision not in ("fp64", "fp32"):
            raise ValueError(f"math_eval: unknown helper {fn!r} / precision {precision!r}")
        code = self.MATH_FNS.index(fn) + (4 if precision == "fp32" else 0)
        xa = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(xa)
        _check(_lib.cfdx_math_eval(self.h, code, xa.ctypes.data, y.ctypes.data, xa.size))
        return y

    def record(self, slot: int) -> None:
        _check(_lib.cfdx_event_record(self.h, slot))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = _d()
        _check(_lib.cfdx_event_elapsed(self.h, a, b, C.byref(ms)))
        return ms.value

    # device buffers
    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.nbytes, a.dtype, a.shape)
        _check(_lib.cfdx_copy_to_device(self.h, _vp(d.ptr), _ptr(a), a.nbytes))
        return d

    def copy(self, dst: "DeviceArray", src: "DeviceArray", nbytes: int | None = None) -> None:
        _check(_lib.cfdx_device_copy(self.h, _vp(dst.ptr), _vp(src.ptr), nbytes or src.nbytes))

    def zeros(self, shape, dtype=np.float64) -> "DeviceArray":
        return self.to_device(np.zeros(shape, dtype=dtype))

    # reference argument order (Solver.cpp:277-314); buffers are DeviceArray
    def uq(self, U_C, Q_C, Grid_P, det, FLOW, ROE, n_launch: int) -> None:
        _check(_lib.cfdx_uq(self.h, _ptr(U_C), _ptr(Q_C), _ptr(Grid_P), _ptr(det), _ptr(FLOW), _ptr(ROE), n_launch))

    def viscous(self, U_C, F, E, G, Fv, Ev, Gv, Grid_P, metric, det, FLOW, n_launch: int, params: Params) -> None:
        _check(_lib.cfdx_viscous(self.h, _ptr(U_C), _ptr(F), _ptr(E), _ptr(G), _ptr(Fv), _ptr(Ev), _ptr(Gv),
                                 _ptr(Grid_P), _ptr(metric), _ptr(det), _ptr(FLOW), n_launch, C.byref(params)))

    def weno(self, j, mach, dt, deter, det, U0, U1, U2, Grid_P, Q_C, metric, ROE, Fv, Ev, Gv, F, E, G, FLOW,
             n_launch: int, params: Params) -> None:
        _check(_lib.cfdx_weno(self.h, j, mach, dt, _ptr(deter), _ptr(det), _ptr(U0), _ptr(U1), _ptr(U2),
                              _ptr(Grid_P), _ptr(Q_C), _ptr(metric), _ptr(ROE), _ptr(Fv), _ptr(Ev), _ptr(Gv),
                              _ptr(F), _ptr(E), _ptr(G), _ptr(FLOW), n_launch, C.byref(params)))


class DeviceArray:
    """hipMalloc'd buffer with the numpy dtype / shape of its host image."""

    def __init__(self, ctx: Ctx, nbytes: int, dtype, shape):
        p = _vp()
        _check(_lib.cfdx_device_alloc(ctx.h, nbytes, C.byref(p)))
        self.ctx, self.ptr, self.nbytes, self.dtype, self.shape = ctx, p.value, nbytes, np.dtype(dtype), tuple(shape)

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(_lib.cfdx_copy_to_host(self.ctx.h, _ptr(out), _vp(self.ptr), self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.h:
            _check(_lib.cfdx_device_free(self.ctx.h, _vp(self.ptr)))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DropInSolver:
    """Solver.cpp:356-541 on device: per stage U_Q -> Viscous -> Weno through the drop-in
    kernels, with the USE_GPU==1 data contract (RK base U0 from U_Q at j = 0, raw stage
    vectors U1 / U2 into the viscous kernel at j = 1, 2). Buffers stay in HBM between
    stages; periodic single subdomain (no halo exchange / boundary fill)."""

    def __init__(self, case: dict, ctx: Ctx | None = None):
        self.ctx = c = ctx or Ctx()
        self.n = n = case["Grid_P"].shape[0]
        self.case = case
        self.Grid_P = c.to_device(case["Grid_P"])
        self.det = c.to_device(case["det"])
        self.metric = c.to_device(case["metric"])
        self.deter = c.to_device(case["deter"])
        self.FLOW = c.to_device(case["FLOW"])
        self.U0, self.U1, self.U2, self.Q_C, self.Uscratch = (c.zeros((n, 5)) for _ in range(5))
        self.F, self.E, self.G, self.Fv, self.Ev, self.Gv = (c.zeros((n, 5)) for _ in range(6))
        self.ROE = c.zeros((n, ROE_N))

    def step(self, nsteps: int, mach: float, dt: float, params: Params) -> None:
        n = self.n
        c = self.ctx
        for _ in range(nsteps):
            for j in range(3):
                c.uq(self.U0 if j == 0 else self.Uscratch, self.Q_C, self.Grid_P, self.det, self.FLOW, self.ROE, n)
                uv = (self.U0, self.U1, self.U2)[j]
                c.viscous(uv, self.F, self.E, self.G, self.Fv, self.Ev, self.Gv, self.Grid_P, self.metric, self.det,
                          self.FLOW, n, params)
                c.weno(j, mach, dt, self.deter, self.det, self.U0, self.U1, self.U2, self.Grid_P, self.Q_C,
                       self.metric, self.ROE, self.Fv, self.Ev, self.Gv, self.F, self.E, self.G, self.FLOW, n, params)

    def flow(self) -> np.ndarray:
        return self.FLOW.numpy()


# ---------------------------------------------------------------------------
# structured block (SoA, directional sweeps)
# ---------------------------------------------------------------------------
class Block:
    """One ni x nj x nk block resident in HBM (owned nodes in i-fastest order). Fully periodic by
    default; a BlockDesc (e.g. from cgrid_desc) gives bounded directions with side conditions."""

    def __init__(self, ni: int, nj: int, nk: int, precision: int = FP64, ctx: Ctx | None = None,
                 desc: BlockDesc | None = None):
        self.ctx = ctx or Ctx()
        self.precision = precision
        if desc is None:
            desc = BlockDesc(ni, nj, nk, 7, precision)
        else:
            desc = BlockDesc.from_buffer_copy(desc)
            desc.precision = precision
            if (desc.ni, desc.nj, desc.nk) != (ni, nj, nk):
                raise ValueError(f"descriptor {desc.ni}x{desc.nj}x{desc.nk} does not match the block {ni}x{nj}x{nk}")
        self.desc = desc
        h = _vp()
        _check(_lib.cfdx_block_create(self.ctx.h, C.byref(self.desc), C.byref(h)))
        self.h = h
        self.n = ni * nj * nk
        self.dt = 0.0       # del_t of the last configure / step

    @classmethod
    def of(cls, desc: BlockDesc, ctx: Ctx | None = None) -> "Block":
        return cls(desc.ni, desc.nj, desc.nk, precision=desc.precision, ctx=ctx, desc=desc)

    @staticmethod
    def bytes_for(ni: int, nj: int, nk: int, precision: int = FP64, desc: BlockDesc | None = None) -> int:
        d = desc if desc is not None else BlockDesc(ni, nj, nk, 7, precision)
        if desc is not None:
            d = BlockDesc.from_buffer_copy(desc)
            d.precision = precision
        n = int(_lib.cfdx_block_bytes(C.byref(d)))
        if n < 0:
            raise CfdxError(n, _lib.cfdx_last_error().decode())
        return n

    def set_geometry_soa(self, geo: np.ndarray) -> None:
        dims, _ = padded(self.desc)
        geo = np.ascontiguousarray(geo, dtype=np.float64)
        if geo.size != 22 * int(np.prod(dims)):
            raise ValueError(f"geometry holds {geo.size} values, the block needs 22 x {int(np.prod(dims))}")
        _check(_lib.cfdx_block_set_geometry_soa(self.h, _ptr(geo)))

    def close(self) -> None:
        if self.h:
            _check(_lib.cfdx_block_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_case(self, case: dict) -> None:
        """Geometry and state of a reference-numbered case (node 0 dropped)."""
        det = np.ascontiguousarray(case["det"][1:])
        metric = np.ascontiguousarray(case["metric"][1:])
        deter = np.ascontiguousarray(case["deter"][1:])
        _check(_lib.cfdx_block_set_geometry(self.h, _ptr(det), _ptr(metric), _ptr(deter)))
        self.set_flow(case["FLOW"][1:])

    def init_case(self, grid: str = "box", length=(10.0, 10.0, 10.0), amp: float = 0.0, phases=(0.0, 0.0, 0.0),
                  init: str = "freestream", mach: float = 0.5, beta: float = 5.0, center=(5.0, 5.0)) -> None:
        """Geometry (exact metrics of the map) and initial state built on the device; the same
        case as periodic_case(..., metrics='analytic') without the host arrays. Asynchronous."""
        if grid not in ("box", "sinusoid"):
            raise ValueError(grid)
        c = CaseDesc((C.c_double * 3)(*length), amp if grid == "sinusoid" else 0.0, (C.c_double * 3)(*phases),
                     INIT_KINDS[init], 0, mach, beta, (C.c_double * 2)(*center))
        _check(_lib.cfdx_block_init_case(self.h, C.byref(c)))

    def cfl_dt(self, cfl: float) -> float:
        """cfl / max spectral radius of the current state (synchronous)."""
        dt = _d(0.0)
        _check(_lib.cfdx_block_cfl_dt(self.h, cfl, C.byref(dt)))
        return dt.value

    def check_finite(self) -> bool:
        """True if every value of the current state (RK base) is finite (synchronous)."""
        ok = C.c_int32(0)
        _check(_lib.cfdx_block_check_finite(self.h, C.byref(ok)))
        return bool(ok.value)

    def set_flow(self, flow: np.ndarray) -> None:
        flow = np.ascontiguousarray(flow)
        if flow.dtype != FLOW_DT or flow.shape[0] != self.n:
            raise ValueError(f"flow must be a FLOW_DT array of {self.n} owned nodes")
        _check(_lib.cfdx_block_set_flow(self.h, _ptr(flow)))

    def get_flow(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=FLOW_DT)
        _check(_lib.cfdx_block_get_flow(self.h, _ptr(out)))
        return out

    def configure(self, mach: float, dt: float, params: Params) -> None:
        _check(_lib.cfdx_block_configure(self.h, mach, dt, C.byref(params)))
        self.dt = dt

    def step(self, nsteps: int, mach: float, dt: float, params: Params) -> None:
        _check(_lib.cfdx_block_step(self.h, mach, dt, nsteps, C.byref(params)))
        self.dt = dt

    def rhs(self, params: Params) -> np.ndarray:
        out = np.zeros((5, self.n))
        _check(_lib.cfdx_block_rhs(self.h, C.byref(params), _ptr(out)))
        return out

    def rhs_noalloc(self, params: Params, out: "DeviceArray") -> None:
        """The right-hand side into the device array `out` (5 x owned nodes doubles) through the
        entry point that allocates nothing; asynchronous on the context's stream."""
        if out.nbytes < 5 * self.n * 8:
            raise ValueError(f"rhs_noalloc: out holds {out.nbytes} bytes, {5 * self.n * 8} needed")
        _check(_lib.cfdx_block_rhs_noalloc(self.h, C.byref(params), _vp(out.ptr)))

    def set_timing(self, on: bool) -> None:
        _check(_lib.cfdx_block_set_timing(self.h, int(on)))

    def set_check(self, on: bool = True) -> None:
        """Non-finite check after every step() call (CfdxError E_NONFINITE when the state breaks)."""
        _check(_lib.cfdx_block_set_check(self.h, int(on)))

    def set_monitor(self, on: bool = True) -> None:
        _check(_lib.cfdx_block_set_monitor(self.h, int(on)))

    def residual(self) -> tuple[float, float, float]:
        """max |du|, |dv|, |de| of the last step (monitor on)."""
        out = (_d * 3)()
        _check(_lib.cfdx_block_residual(self.h, out))
        return out[0], out[1], out[2]

    def timing_split(self) -> list[float]:
        """ms of the timed steps: viscous, face i, face j, face k, update."""
        ms = (_d * 5)()
        _check(_lib.cfdx_block_timing_split(self.h, ms))
        return list(ms)

    def set_tiles(self, t) -> None:
        """t = (face_i, face_j, face_k, node) variant indices."""
        _check(_lib.cfdx_block_set_tiles(self.h, C.byref(Tiles.of(t))))

    def get_tiles(self) -> tuple:
        t = Tiles()
        _check(_lib.cfdx_block_get_tiles(self.h, C.byref(t)))
        return t.as_tuple()

    def timing(self) -> dict:
        a, b, c, n = _d(), _d(), _d(), C.c_int()
        _check(_lib.cfdx_block_timing(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(n)))
        return dict(visc_ms=a.value, face_ms=b.value, update_ms=c.value, launches=n.value)


class Batch:
    """Blocks of different sizes / run parameters stepped together (one launch per kernel class)."""

    def __init__(self, precision: int = FP64, ctx: Ctx | None = None):
        self.ctx = ctx or Ctx()
        h = _vp()
        _check(_lib.cfdx_batch_create(self.ctx.h, precision, C.byref(h)))
        self.h = h
        self.blocks: list[Block] = []

    def add(self, b: "Block") -> None:
        _check(_lib.cfdx_batch_add(self.h, b.h))
        self.blocks.append(b)

    def remove(self, b: "Block") -> None:
        _check(_lib.cfdx_batch_remove(self.h, b.h))
        self.blocks.remove(b)

    def __len__(self) -> int:
        return int(_lib.cfdx_batch_size(self.h))

    def set_grid(self, cap: int) -> None:
        _check(_lib.cfdx_batch_set_grid(self.h, cap))

    def step(self, nsteps: int, graph: bool = True) -> None:
        _check(_lib.cfdx_batch_step(self.h, nsteps, int(graph)))

    def set_tiles(self, t) -> None:
        _check(_lib.cfdx_batch_set_tiles(self.h, C.byref(Tiles.of(t))))

    def close(self) -> None:
        if self.h:
            _check(_lib.cfdx_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# on-disk formats
# ---------------------------------------------------------------------------
def write_restart(directory: str, slot: int, flow: np.ndarray, iteration: int, digits: int = 17) -> None:
    flow = np.ascontiguousarray(flow)
    if flow.dtype != FLOW_DT:
        raise ValueError("flow must be a FLOW_DT array")
    _check(_lib.cfdx_write_restart(directory.encode(), slot, flow.shape[0], _ptr(flow), iteration, digits))


def restart_latest(directory: str, n: int) -> tuple[int, int]:
    """(slot, iteration) of the newer complete restart file; (0, -1) when none."""
    slot, it = _i32(0), _i64(0)
    _check(_lib.cfdx_restart_latest(directory.encode(), n, C.byref(slot), C.byref(it)))
    return slot.value, it.value


def read_restart(directory: str, slot: int, n: int) -> tuple[np.ndarray, int]:
    flow = np.zeros(n, dtype=FLOW_DT)
    it = _i64(0)
    _check(_lib.cfdx_read_restart(directory.encode(), slot, n, _ptr(flow), C.byref(it)))
    return flow, it.value


def write_tecplot(path: str, ni: int, nj: int, nk: int, xyz: np.ndarray, flow: np.ndarray, loc=None,
                  solution_time: float = 0.0) -> None:
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    flow = np.ascontiguousarray(flow)
    locp = None if loc is None else np.ascontiguousarray(loc, dtype=np.int32)
    _check(_lib.cfdx_write_tecplot(path.encode(), ni, nj, nk, _ptr(xyz), _ptr(flow), _ptr(locp), solution_time))


def exported_symbols() -> list[str]:
    return list(EXPORTS)
