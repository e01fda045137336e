"""ORACLE — test infrastructure only.

ctypes binding of oracle/liboracle.so (the C restatement) and
oracle/_ref/libref.so (the reference kernels compiled as C). Both export the
same orc_* entry points on numpy arrays in the packed reference layouts.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libref.so")
REF_SRC = "/root/reference"

FLOW_FIELDS = ["u", "v", "w", "p", "t", "rho", "mu", "a", "e"]
FLOW_DT = np.dtype([(f, "<f8") for f in FLOW_FIELDS])
GRID_DT = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")] +
                   [(f"N{i}", "<i4") for i in range(6)] + [("ID", "<i4"), ("loc", "<i4")])
assert FLOW_DT.itemsize == 72 and GRID_DT.itemsize == 56


class Params(C.Structure):
    _fields_ = [("mach", C.c_double), ("reyl", C.c_double), ("pr", C.c_double), ("free_t", C.c_double),
                ("c_mts", C.c_double), ("c_t", C.c_double), ("eps", C.c_double), ("sgs", C.c_int32),
                ("pad_", C.c_int32), ("c_s", C.c_double)]


def reference_params(**kw) -> Params:
    """The constants hard-coded in Viscous.cl:30-33,490-491 and Weno.cl:59."""
    p = Params(1.5, 500.0, 0.72, 273.0, 0.03, 10.0, 10e-8, 0, 0, 0.17)
    names = {f[0] for f in Params._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"reference_params: no parameter {k!r}")
        setattr(p, k, v)
    return p


class Work(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ["U0", "U1", "U2", "Q_C", "F", "E", "G", "Fv", "Ev", "Gv", "ROE"]]


def build(ref: bool | None = None) -> None:
    """Builds liboracle.so, and _ref/libref.so when the reference sources are present."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref or (ref is None and os.path.isdir(REF_SRC)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Oracle:
    """One of the two CPU implementations (kind 'port' = restatement, 'reference' = reference kernels)."""

    def __init__(self, which: str = "port"):
        path = LIB_ORACLE if which == "port" else LIB_REF
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing; run pyoracle.build()")
        self.which = which
        self.lib = C.CDLL(path)
        L = self.lib
        vp = C.c_void_p
        L.orc_uq.argtypes = [C.c_int64, vp, vp, vp, vp, vp, vp]
        L.orc_viscous.argtypes = [C.c_int64] + [vp] * 11 + [C.POINTER(Params)]
        L.orc_weno.argtypes = [C.c_int64, C.c_int, C.c_double, C.c_double] + [vp] * 16 + [C.POINTER(Params)]
        L.orc_solver_steps.argtypes = [C.c_int64, C.c_int64, C.c_int, C.c_double, C.c_double, vp, vp, vp, vp, vp,
                                       C.POINTER(Work), C.POINTER(Params)]
        L.orc_kind.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]

    def set_threads(self, n: int) -> None:
        self.lib.orc_set_threads(int(n))

    # -- single kernels ----------------------------------------------------
    def uq(self, n, U_C, Q_C, Grid_P, det, FLOW, ROE):
        self.lib.orc_uq(n, _ptr(U_C), _ptr(Q_C), _ptr(Grid_P), _ptr(det), _ptr(FLOW), _ptr(ROE))

    def viscous(self, n, U_C, F, E, G, Fv, Ev, Gv, Grid_P, metric, det, FLOW, prm: Params):
        self.lib.orc_viscous(n, _ptr(U_C), _ptr(F), _ptr(E), _ptr(G), _ptr(Fv), _ptr(Ev), _ptr(Gv), _ptr(Grid_P),
                             _ptr(metric), _ptr(det), _ptr(FLOW), C.byref(prm))

    def weno(self, n, j, mach, dt, deter, det, U0, U1, U2, Grid_P, Q_C, metric, ROE, Fv, Ev, Gv, F, E, G, FLOW,
             prm: Params):
        self.lib.orc_weno(n, j, mach, dt, _ptr(deter), _ptr(det), _ptr(U0), _ptr(U1), _ptr(U2), _ptr(Grid_P),
                          _ptr(Q_C), _ptr(metric), _ptr(ROE), _ptr(Fv), _ptr(Ev), _ptr(Gv), _ptr(F), _ptr(E),
                          _ptr(G), _ptr(FLOW), C.byref(prm))

    # -- whole steps ---------------------------------------------------------
    def new_work(self, n_g: int) -> dict:
        w = {k: np.zeros((n_g, 5)) for k in ["U0", "U1", "U2", "Q_C", "F", "E", "G", "Fv", "Ev", "Gv"]}
        w["ROE"] = np.zeros((n_g, 36))
        return w

    def steps(self, case: dict, nsteps: int, mach: float, dt: float, prm: Params, work: dict | None = None):
        """Runs nsteps RK3 steps in place on case['FLOW']; returns the work arrays."""
        n = case["Grid_P"].shape[0]
        w = work if work is not None else self.new_work(n)
        ws = Work(*[C.c_void_p(w[k].ctypes.data) for k in ["U0", "U1", "U2", "Q_C", "F", "E", "G", "Fv", "Ev", "Gv", "ROE"]])
        self.lib.orc_solver_steps(n, n, nsteps, mach, dt, _ptr(case["Grid_P"]), _ptr(case["det"]),
                                  _ptr(case["metric"]), _ptr(case["deter"]), _ptr(case["FLOW"]), C.byref(ws),
                                  C.byref(prm))
        return w

    def stage_outputs(self, case: dict, mach: float, dt: float, prm: Params) -> dict:
        """Stage 0 of one step, keeping every kernel output (inputs are not modified)."""
        n = case["Grid_P"].shape[0]
        FLOW = case["FLOW"].copy()
        w = self.new_work(n)
        out = {}
        self.uq(n, w["U0"], w["Q_C"], case["Grid_P"], case["det"], FLOW, w["ROE"])
        out.update(U_C=w["U0"].copy(), Q_C=w["Q_C"].copy(), ROE=w["ROE"].copy())
        self.viscous(n, w["U0"], w["F"], w["E"], w["G"], w["Fv"], w["Ev"], w["Gv"], case["Grid_P"], case["metric"],
                     case["det"], FLOW, prm)
        out.update({k: w[k].copy() for k in ["F", "E", "G", "Fv", "Ev", "Gv"]})
        out["mu"] = FLOW["mu"].copy()
        self.weno(n, 0, mach, dt, case["deter"], case["det"], w["U0"], w["U1"], w["U2"], case["Grid_P"], w["Q_C"],
                  case["metric"], w["ROE"], w["Fv"], w["Ev"], w["Gv"], w["F"], w["E"], w["G"], FLOW, prm)
        out["U1"] = w["U1"].copy()
        out["FLOW"] = FLOW
        return out


    def rhs(self, case: dict, prm: Params) -> np.ndarray:
        """Stage-0 right-hand side L (Weno.cl:3078-3082) of case['FLOW'], shape (n_nodes, 5):
        Weno with del_t = 1 and U0 = 0 returns U1 = L exactly."""
        n = case["Grid_P"].shape[0]
        FLOW = case["FLOW"].copy()
        w = self.new_work(n)
        self.uq(n, w["U0"], w["Q_C"], case["Grid_P"], case["det"], FLOW, w["ROE"])
        self.viscous(n, w["U0"], w["F"], w["E"], w["G"], w["Fv"], w["Ev"], w["Gv"], case["Grid_P"], case["metric"],
                     case["det"], FLOW, prm)
        zero = np.zeros_like(w["U0"])
        self.weno(n, 0, 1.0, 1.0, case["deter"], case["det"], zero, w["U1"], w["U2"], case["Grid_P"], w["Q_C"],
                  case["metric"], w["ROE"], w["Fv"], w["Ev"], w["Gv"], w["F"], w["E"], w["G"], FLOW, prm)
        return w["U1"]


def available(which: str) -> bool:
    return os.path.exists(LIB_ORACLE if which == "port" else LIB_REF)
