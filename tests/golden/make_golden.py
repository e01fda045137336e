"""Writes the golden fixtures of tests/golden from the reference's own kernels.

Runs in a container that has /root/reference: oracle/Makefile compiles
U_Q.cl / Viscous.cl / Weno.cl as C into oracle/_ref/libref.so, and this script
drives those kernels (reference constants of Viscous.cl:30-33) on two small cases:
  periodic_tgv_7x8x9.npz  seeded curvilinear periodic block, noisy Taylor-Green state:
                          every stage-0 kernel output and the state after 2 RK3 steps
  wallbox_3x4x5.npz       box walled on all sides: a boundary layer (loc 1..26) and three ghost
                          layers (loc 100); every stage-0 kernel output (edge / corner branches)
  channel_7x8x9.npz       periodic in i, j, walled in k: every stage-0 kernel output, with the
                          one-sided wall stencils defined at every interior node
Inputs and outputs are stored together (npz, little-endian float64 / int32).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import conftest  # noqa: E402,F401  (loads the cfdx package)
import cases  # noqa: E402
import pyoracle  # noqa: E402

MACH, DT = 0.3, 0.01


def pack(case, out, extra=None):
    d = {f"in_{k}": case[k] for k in ("Grid_P", "det", "metric", "deter", "FLOW")}
    d["in_Grid_P"] = case["Grid_P"].view(np.uint8).reshape(len(case["Grid_P"]), -1)
    d["in_FLOW"] = cases.flow_matrix(case["FLOW"], pyoracle.FLOW_FIELDS)
    for k, v in out.items():
        d[f"out_{k}"] = cases.flow_matrix(v, pyoracle.FLOW_FIELDS) if v.dtype.names else v
    d.update(extra or {})
    return d


def main():
    pyoracle.build(ref=True)
    ref = pyoracle.Oracle("reference")
    prm = pyoracle.reference_params()
    c1 = cases.periodic(7, 8, 9)
    out = ref.stage_outputs(c1, MACH, DT, prm)
    c1s = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in c1.items()}
    ref.steps(c1s, 2, MACH, DT, prm)
    d = pack(c1, out, {"steps2_FLOW": cases.flow_matrix(c1s["FLOW"], pyoracle.FLOW_FIELDS),
                       "mach": MACH, "dt": DT, "n_sd": len(c1["Grid_P"])})
    np.savez_compressed(os.path.join(HERE, "periodic_tgv_7x8x9.npz"), **d)
    for name, dims, walls in (("wallbox_3x4x5.npz", (3, 4, 5), (True, True, True)),
                              ("channel_7x8x9.npz", (7, 8, 9), (False, False, True))):
        c2 = cases.wall_box(*dims, walls=walls)
        out = ref.stage_outputs(c2, MACH, DT, prm)
        d = pack(c2, out, {"mach": MACH, "dt": DT, "n_core": c2["n_core"],
                           "mask_viscous": cases.defined_mask(c2, False), "mask_weno": cases.defined_mask(c2, True)})
        np.savez_compressed(os.path.join(HERE, name), **d)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
