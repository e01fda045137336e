"""Throughput of the per-step RHS + SSP-RK3 update on one MI355X.

Headline workload: 256^3 periodic curvilinear Taylor-Green vortex (the 128^3 case's seeded
sinusoidal map, amplitude 0.05 of the spacing), FP64 Navier-Stokes with the reference's
mixed-time-scale LES model, Mach 0.1, Re 1600, CFL 0.5 (dt fixed from the initial state), full
RHS (characteristic WCNS + LLF + 6th-order correction + viscous/SGS fluxes) and RK3: one GPU's
share of a 512^3 run. A step = one full SSP-RK3 step over the whole grid with the state resident
in HBM; value = interior nodes x steps / device time. Prints one JSON line.

`roofline.achieved` = 848 algorithmic HBM bytes per cell-update x cell-updates/s; the counter
figures of the step (measured HBM bytes and VALU instructions per kernel, from the committed
rocprofv3 summary of the workload) name the bound; the face sweep's own figures (HIP-event timed
in this run) are a sub-field.

Legs on the same line (unless --legs drops them): the other precision on the headline grid with
per-field errors of mixed vs FP64 ("mixed_vs_fp64"), the 128^3 case in FP64 and mixed ("c2"),
the 128^3 case through the drop-in kernels ("dropin_path"), the 64-cube batched mixed graph step
("batched_c4"), the NACA 0012 C-grid in mixed and FP64 with per-field errors ("airfoil_c3"), and
the 200-request Poisson stream of NACA C-grid solves served by continuous batching vs plain FP64
one at a time ("serving": p50 / p99 latency, throughput). cpu_baseline: the C restatement of the
reference's kernels and stage sequence (pinned bitwise to the reference's own kernels) on the host
cores for the 64^3 vortex (100 steps), the 128^3 case (20) and 256^3 (2), all cores and then one
thread, each beside the GPU block path on the same case, dt and step count (rates side by side and
the per-field differences of the two final states); the reference's own kernels are timed on the
128^3 case too.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ns256|c2|c1|c3] [--precision fp64|mixed]
                    [--legs compare,c2,dropin,c4,c3,serve|none] [--cpu c1,c2,ns256|none] [--no-tune]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "high-order-accurate-cfd-solver-code-for-heterogenous-hardware-cpu-and-gpgpu-_amd")

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles per SIMD
VALU_PEAK_WAVE_INSTR_S = 256 * 4 * 2.4e9 / 4
ALG_BYTES_PER_CELL_UPDATE = 848  # minimal FP64 SoA traffic of one RK3 step (3 stages)
# Operations per node written out in the node kernels (counted from block_kernels.hip; the device
# physics they call is counted by cfdx_op_count): viscous kernel stencil differences and MTS
# similarity terms; update kernel face-flux differences, viscous divergence, mean RK stage and
# primitive recovery.
VISC_KERNEL_OPS = 121
UPDATE_KERNEL_OPS_EULER = 30 + 70.0 / 3 + 20   # 70 / 3: the mean RK stage, 5 x {2, 6, 6} over the 3 stages
UPDATE_KERNEL_OPS_VISC_DIV = 72


def alg_ops_per_cell_update(counts: dict, viscous: bool) -> float:
    """Algorithmic operations of one node over one SSP-RK3 step: per stage three faces, three node
    fluxes, the conserved state and (viscous) the viscous node fluxes, plus the update kernel's
    arithmetic; at the last stage the RK base det * Q."""
    stage = 3 * counts["face"]["total"] + 3 * counts["node_flux"]["total"] + counts["cons"]["total"]
    stage += UPDATE_KERNEL_OPS_EULER
    if viscous:
        stage += counts["viscous"]["total"] + VISC_KERNEL_OPS + UPDATE_KERNEL_OPS_VISC_DIV
    return 3 * stage + counts["cons"]["total"] + 5


def face_bytes_per_node(precision: str) -> int:
    """Algorithmic HBM bytes of one face of a face sweep: primitives 5, metric row 3, det, face
    normal 3, face det read and F-hat 5 written, all in the flux type (a mixed block reads the
    float copies of its geometry): 18 values."""
    return 18 * (8 if precision == "fp64" else 4)


def _import_pkg():
    spec = importlib.util.spec_from_file_location("cfdx", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cfdx"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(name: str):
    """Timed-block definition of a workload: (config spec, case definition, params keywords, viscous)."""
    from cfdx import workloads
    if name in ("c2", "ns256"):
        n = 128 if name == "c2" else 256
        spec = dict(workload=f"taylor_green_les_{n}^3_curvilinear", grid=[n, n, n], mach=0.1, re=1600.0,
                    sgs="mixed_time_scale(C=0.03,C_T=10)", cfl=0.5, map="sinusoidal amp=0.05h seeded phases")
        return spec, dict(ni=n, nj=n, nk=n, kw=workloads.tgv_case_kw(n)), dict(mach=0.1, reyl=1600.0), True
    if name == "c3":
        spec = dict(workload="naca0012_cgrid_384x128x32_les", grid=[384, 128, 32], mach=0.5, aoa_deg=2.0, re=1e5,
                    sgs="mixed_time_scale(C=0.03,C_T=10)", cfl=0.5,
                    geometry="C-grid, far field 20c, wake 20c, wall 1e-4c, span 0.25c periodic")
        case = dict(ni=384, nj=128, nk=32, cgrid=dict(naca=(0.0, 0.4, 0.12), wake=20.0, radius=20.0, wall=1e-4,
                                                       span=0.25, aoa_deg=2.0), kw=dict(mach=0.5))
        return spec, case, dict(mach=0.5, reyl=1e5), True
    if name == "c1":
        spec = dict(workload="isentropic_vortex_euler_64^3", grid=[64, 64, 64], mach=0.5, beta=5.0, cfl=0.5)
        return spec, dict(ni=64, nj=64, nk=64, kw=workloads.vortex_case_kw(64)), dict(), False
    raise SystemExit(f"unknown workload {name}")


def profile_rows(workload_name: str, precision: str):
    """Per-kernel rows of the committed rocprofv3 summary of this workload
    (profiles/block_kernels_<workload>_<precision>.md): one machine-readable row per kernel and
    tile variant, `| kernel<...> | launches | avg_us | hbm_bytes | valu_instr | fp64_flop | fp32_flop |`,
    hbm_bytes = FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration, valu_instr = SQ_INSTS_VALU
    (wave instructions), per launch; "-" where the counter runs did not launch that variant."""
    path = os.path.join(ROOT, "profiles", f"block_kernels_{workload_name}_{precision}.md")
    if not os.path.exists(path):
        return None, path
    rows = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0].startswith("k_"):
            try:
                rows[cells[0]] = dict(launches=int(cells[1]), avg_us=float(cells[2]),
                                      hbm_bytes=float(cells[3]), valu_instr=float(cells[4]),
                                      fp64_flop=float(cells[5]), fp32_flop=float(cells[6]))
            except ValueError:
                continue
    return rows, path


def step_kernels(cfdx, precision: str, tiles, viscous: bool = True):
    """Kernel names (with template arguments) of one RK stage of a periodic block with tile
    variants `tiles` = (face i, face j, face k, node variant indices)."""
    t = tuple(tiles) if tiles else (0, 0, 0, 0)
    S = "double" if precision == "fp64" else "float"
    names = []
    nx, ny = cfdx.tile_variants(2)[t[3]]
    if viscous:
        names.append(f"k_visc<{S},{nx},{ny}>")
    for d in range(3):
        ta, tb = cfdx.tile_variants(0 if d == 0 else 1)[t[d]]
        names.append(f"k_face<{S},{d},{ta},{tb}>")
    names.append(f"k_update<{S},{nx},{ny}>")
    return names


def counter_figures(cfdx, workload_name: str, precision: str, tiles, viscous: bool, nodes: int):
    """Counter figures of one RK stage from the committed profile: per kernel the measured HBM
    bytes and VALU wave instructions per launch (the profiled run's own tuned variants when this
    run tuned others of the same class), and the stage totals."""
    rows, path = profile_rows(workload_name, precision)
    if not rows:
        return None
    want = step_kernels(cfdx, precision, tiles, viscous)
    per, same = {}, True
    for k in want:
        r = rows.get(k)
        if r is None or r["hbm_bytes"] != r["hbm_bytes"]:
            # the counter runs used another variant of this class: take the counted one
            cls = k[:k.index("<")] + "<" + ",".join(k[k.index("<") + 1:].split(",")[:2 if k.startswith("k_face") else 1])
            k2 = next((n for n, v in rows.items() if n.startswith(cls) and v["hbm_bytes"] == v["hbm_bytes"]), None)
            if k2 is None:
                return None
            same = False
            r = rows[k2]
            k = k2
        per[k] = r
    stage_bytes = sum(r["hbm_bytes"] for r in per.values())
    stage_valu = sum(r["valu_instr"] for r in per.values())
    stage_us = sum(r["avg_us"] for r in per.values())
    return {"source": os.path.relpath(path, ROOT), "kernels": per, "same_variants": same,
            "stage_hbm_bytes": stage_bytes, "stage_valu_instr": stage_valu, "stage_profiled_us": stage_us,
            "hbm_bytes_per_cell_update": 3 * stage_bytes / nodes}


def cpu_baseline(cfdx, which=("c1", "c2", "ns256"), single_thread=("c1", "c2", "ns256"), gpu=True):
    """The reference's CPU path, timed on this host beside the GPU on the same inputs: the C
    restatement of the reference's kernels and stage sequence (oracle/liboracle.so, gcc -O3
    -fopenmp, OpenMP over nodes, the reference's AoS arrays), which runs the configured physics
    (Mach, Re, the Euler switch) and is pinned bitwise to the reference's own kernels at their
    built-in constants. Configurations: C1 (64^3 vortex, Euler, 100 steps), C2 (128^3 curvilinear
    TGV, 20 steps after one warm-up), NS (the C2 flow at 256^3, 2 steps after one warm-up), on all
    cores of the process affinity, then on one thread C1 (10 steps), C2 and NS (1 step each).
    Rates are interior nodes x steps / wall time. With gpu, the same case, dt and step count run on
    the GPU block path (direct FP64 steps, timed with HIP events) and the two final states are
    compared (per-field relative L2 / Linf of the primitives). For C2, the reference's own kernels
    compiled as C (oracle/_ref/libref.so) are also timed on the same arrays and steps: they keep the
    constants written into them (Re 500 and Mach 1.5 in the viscous terms), so their state is not
    compared, only their speed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from cfdx import workloads
    if not pyoracle.available("port"):
        return None
    orc = pyoracle.Oracle("port")
    ref = pyoracle.Oracle("reference") if pyoracle.available("reference") else None
    cores = len(os.sched_getaffinity(0))
    plan = {"c1": (64, 100, 0), "c2": (128, 20, 1), "ns256": (256, 2, 1)}
    single = {"c1": 10, "c2": 1, "ns256": 1}
    out = {}
    for name in which:
        n, steps, warm = plan[name]
        t_case = time.perf_counter()
        if name == "c1":
            case = cfdx.periodic_case(n, n, n, **workloads.vortex_case_kw(n))
            prm = pyoracle.reference_params(reyl=math.inf, c_mts=0.0)
            gprm = cfdx.euler_params()
            mach = 0.5
        else:
            case = cfdx.periodic_case(n, n, n, **workloads.tgv_case_kw(n))
            prm = pyoracle.reference_params(mach=0.1, reyl=1600.0)
            gprm = cfdx.default_params(mach=0.1, reyl=1600.0)
            mach = 0.1
        dt = cfdx.cfl_dt(case, 0.5)
        work = orc.new_work(case["Grid_P"].shape[0])
        flow0 = case["FLOW"].copy() if (ref is not None and name == "c2") else None
        t_case = time.perf_counter() - t_case
        rec = {"grid": f"{n}^3", "steps": steps, "warmup": warm, "dt": dt, "host_case_s": round(t_case, 1)}
        g = None
        if gpu:
            b = cfdx.Block(n, n, n)
            b.set_case(case)
            ctx = b.ctx
            b.step(warm, mach, dt, gprm)
            ctx.sync()
            ctx.record(0)
            b.step(steps, mach, dt, gprm)
            ctx.record(1)
            ctx.sync()
            ms = ctx.elapsed_ms(0, 1)
            g = b.get_flow()
            b.close()
            rec["gpu"] = {"value": n ** 3 * steps / (ms / 1e3), "ms_per_step": ms / steps,
                          "stepping": "direct FP64 block steps"}
        orc.set_threads(cores)
        if warm:
            orc.steps(case, warm, mach, dt, prm, work)
        t0 = time.perf_counter()
        orc.steps(case, steps, mach, dt, prm, work)
        sec = time.perf_counter() - t0
        rec.update({"threads": cores, "seconds": round(sec, 2), "value": n ** 3 * steps / sec,
                    "ms_per_step": 1e3 * sec / steps})
        if g is not None:
            rec["gpu"]["over_cpu"] = rec["gpu"]["value"] / rec["value"]
            rec["gpu"]["field_errors_vs_cpu"] = field_errors(g, case["FLOW"][1:])
            del g
        if name in single_thread:
            orc.set_threads(1)
            t0 = time.perf_counter()
            orc.steps(case, single[name], mach, dt, prm, work)
            s1 = time.perf_counter() - t0
            rec["single_thread"] = {"steps": single[name], "seconds": round(s1, 2), "value": n ** 3 * single[name] / s1,
                                    "ms_per_step": 1e3 * s1 / single[name]}
        if flow0 is not None:
            # the reference's own kernels on the same arrays and step count (their built-in constants)
            case["FLOW"][:] = flow0
            ref.set_threads(cores)
            rw = ref.new_work(case["Grid_P"].shape[0])
            if warm:
                ref.steps(case, warm, mach, dt, prm, rw)
            t0 = time.perf_counter()
            ref.steps(case, steps, mach, dt, prm, rw)
            rs = time.perf_counter() - t0
            rec["reference_kernels"] = {"threads": cores, "steps": steps, "seconds": round(rs, 2),
                                        "value": n ** 3 * steps / rs,
                                        "library": "oracle/_ref/libref.so (the reference's .cl kernels compiled as C)"}
            del rw
        out[name] = rec
        msg = f"bench: cpu {name} {rec['value']:.3e} cell-updates/s on {cores} threads"
        if "single_thread" in rec:
            msg += f", {rec['single_thread']['value']:.3e} on one"
        if "reference_kernels" in rec:
            msg += f" (reference kernels {rec['reference_kernels']['value']:.3e})"
        if "gpu" in rec:
            e = max(v["rel_linf"] for v in rec["gpu"]["field_errors_vs_cpu"].values())
            msg += f"; GPU {rec['gpu']['value']:.3e}, max rel Linf vs CPU {e:.1e}"
        print(msg, file=sys.stderr, flush=True)
        del case, work
    return {"kind": "port", "cores": cores, "configs": out,
            "library": "oracle/liboracle.so (the C restatement of the reference's kernels and stage sequence, "
                       "gcc -O3 -fopenmp, pinned bitwise to oracle/_ref/libref.so)"}


def make_block(cfdx, cdef, precision, prm, ctx=None, host_case=None):
    """The timed block of a workload with its CFL-0.5 step: a periodic case (uploaded from the host
    case when given, else set up on the device) or the C-grid of `cdef["cgrid"]` from the free
    stream."""
    mach = cdef["kw"]["mach"]
    if "cgrid" in cdef:
        cg = cdef["cgrid"]
        d, g = cfdx.cgrid_desc(cdef["ni"], cdef["nj"], cdef["nk"], precision, naca=cg["naca"], wake=cg["wake"],
                               radius=cg["radius"], wall=cg["wall"], span=cg["span"], aoa_deg=cg["aoa_deg"])
        if "geo" not in cg:
            cg["geo"] = cfdx.cgrid_geometry(d, g)[1]
        b = cfdx.Block(cdef["ni"], cdef["nj"], cdef["nk"], precision=precision, ctx=ctx, desc=d)
        b.set_geometry_soa(cg["geo"])
        b.configure(mach, 0.0, prm)
        b.set_flow(cfdx.freestream(cdef["ni"] * cdef["nj"] * cdef["nk"], mach, cg["aoa_deg"]))
    else:
        b = cfdx.Block(cdef["ni"], cdef["nj"], cdef["nk"], precision=precision, ctx=ctx)
        if host_case is not None:
            b.set_case(host_case)
        else:
            b.init_case(**cdef["kw"])
        b.configure(mach, 0.0, prm)
    dt = b.cfl_dt(0.5)
    b.configure(mach, dt, prm)
    return b, dt


def tuned_tiles(cfdx, cdef, precision, mach, prm):
    """Tile variants for this grid and precision, tuned on a scratch block of the same case (the
    tuner runs real steps), or taken from the tile cache."""
    from cfdx import tune
    scratch, dt = make_block(cfdx, cdef, precision, prm)
    res = tune.tune_block(scratch, mach, dt, prm, steps=2, cache=tune.TileCache())
    scratch.close()
    return res


def graph_steps(cfdx, block, ctx, tiles=None):
    """A one-block batch around `block` whose step(n) replays a captured RK3-step graph: the
    library's graph path for a single block (the mixed form's viscous/face overlap becomes graph
    edges instead of stream events: 3-4 % faster mixed, FP64 unchanged). The block keeps its
    state and run parameters; remove it (q.remove, q.close) to step it directly again."""
    q = cfdx.Batch(block.precision, ctx)
    q.add(block)
    if tiles:
        q.set_tiles(tuple(tiles))
    return q


def time_block(block, ctx, mach, dt, prm, steps, warmup, profile_steps, cfdx=None, tiles=None):
    """Warm-up, K timed steps bracketed by HIP events on the block's stream (replayed as a
    captured graph when cfdx is given, see graph_steps), then a separate per-kernel-class
    timing pass of direct steps."""
    q = graph_steps(cfdx, block, ctx, tiles) if cfdx is not None else None
    step = (lambda n: q.step(n, True)) if q else (lambda n: block.step(n, mach, dt, prm))
    step(warmup)
    ctx.sync()
    ctx.record(0)
    step(steps)
    ctx.record(1)
    ctx.sync()
    dev_ms = ctx.elapsed_ms(0, 1)
    if q:
        q.remove(block)
        q.close()
    block.set_timing(True)
    block.step(profile_steps, mach, dt, prm)
    tk = block.timing()
    tk["split_ms"] = block.timing_split()   # viscous, face i, face j, face k, update
    block.set_timing(False)
    return dev_ms, tk


def field_errors(a, b):
    """Per-field relative L2 / Linf of a against b (rho, u, v, w, p); velocity components are
    scaled by the norm of the whole velocity field."""
    import numpy as np
    vel = np.sqrt(sum(np.asarray(b[c]) ** 2 for c in "uvw"))
    out = {}
    for f in ("rho", "u", "v", "w", "p"):
        d = a[f] - b[f]
        ref = vel if f in "uvw" else np.abs(b[f])
        out[f] = {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(ref)), "rel_linf": float(np.abs(d).max() / ref.max())}
    return out


def dropin_path(cfdx, case, mach, dt, prm, block_value, steps: int = 3):
    """The headline case through the drop-in kernels (the reference's AoS records and its per-stage
    U_Q -> Viscous -> Weno sequence, each node computing both of its faces), on the device: the
    throughput the reference's own data layout and call sequence reach on this GPU, beside the
    structured block path. Plain FP64 (its parity with the reference is tested separately)."""
    import numpy as np
    s = cfdx.DropInSolver(case)
    ctx = s.ctx
    s.step(1, mach, dt, prm)                  # loads the kernels' code
    ctx.sync()
    s = cfdx.DropInSolver(case, ctx=ctx)
    ctx.record(2)
    s.step(steps, mach, dt, prm)
    ctx.record(3)
    ctx.sync()
    ms = ctx.elapsed_ms(2, 3)
    n = case["Grid_P"].shape[0] - 1
    finite = bool(np.isfinite(cfdx_matrix(s.flow()[1:])).all())
    value = n * steps / (ms / 1e3)
    return {"workload": "the headline case through cfdx_uq / cfdx_viscous / cfdx_weno (AoS, both faces per node)",
            "steps": steps, "value": value, "unit": "cell-updates/s", "ms_per_step": ms / steps,
            "block_path_over_dropin": block_value / value if block_value else None, "state_finite": finite}


def timed_run(cfdx, name, precision, steps, warmup, tune, profile_steps=2, ctx=None, host_case=None, keep_flow=True):
    """One workload's timed block: tuned tiles, CFL-0.5 dt from the initial state, `warmup`
    untimed and `steps` timed RK3 steps replayed from the captured step graph (device time from
    HIP events on the block's stream), then the per-kernel-class timing pass."""
    spec, cdef, pkw, viscous = workload(name)
    prm = cfdx.default_params(**pkw) if viscous else cfdx.euler_params()
    mach = cdef["kw"]["mach"]
    tiles = tuned_tiles(cfdx, cdef, precision, mach, prm) if tune else None
    block, dt = make_block(cfdx, cdef, precision, prm, ctx=ctx, host_case=host_case)
    if tiles:
        block.set_tiles(tuple(tiles["tiles"]))
    ctx = block.ctx
    dev_ms, tk = time_block(block, ctx, mach, dt, prm, steps, warmup, profile_steps, cfdx,
                            tiles["tiles"] if tiles else None)
    n = cdef["ni"] * cdef["nj"] * cdef["nk"]
    out = {"spec": spec, "cdef": cdef, "prm": prm, "mach": mach, "viscous": viscous, "dt": dt, "nodes": n,
           "tiles": tiles, "dev_ms": dev_ms, "tk": tk, "value": n * steps / (dev_ms / 1e3),
           "ms_per_step": dev_ms / steps, "finite": block.check_finite()}
    if keep_flow:
        out["flow"] = block.get_flow()
    block.close()
    return out


def per_stage_ms(tk, profile_steps):
    out = {k: v / (3 * profile_steps) for k, v in tk.items() if k.endswith("_ms") and not isinstance(v, list)}
    if "split_ms" in tk:
        names = ("visc", "face_i", "face_j", "face_k", "update")
        out["split_ms"] = {n: v / (3 * profile_steps) for n, v in zip(names, tk["split_ms"])}
    return out


def c2_leg(cfdx, steps, warmup, tune, profile_steps=2):
    """C2 (128^3 curvilinear TGV, MTS LES) in FP64 and in mixed precision from the same initial
    state and step count: throughput of both and per-field errors of mixed against FP64."""
    r = {p: timed_run(cfdx, "c2", prec, steps, warmup, tune, profile_steps)
         for p, prec in (("fp64", cfdx.FP64), ("mixed", cfdx.MIXED))}
    out = {"workload": r["fp64"]["spec"]["workload"], "steps": steps, "warmup": warmup, "dt": r["fp64"]["dt"]}
    for p in ("fp64", "mixed"):
        out[p] = {"value": r[p]["value"], "unit": "cell-updates/s", "ms_per_step": r[p]["ms_per_step"],
                  "kernel_ms_per_stage": per_stage_ms(r[p]["tk"], profile_steps),
                  "tiles": r[p]["tiles"]["tiles"] if r[p]["tiles"] else None, "state_finite": r[p]["finite"]}
    out["mixed_over_fp64"] = r["mixed"]["value"] / r["fp64"]["value"]
    out["hbm_alg_frac_fp64"] = ALG_BYTES_PER_CELL_UPDATE * r["fp64"]["value"] / (HBM_PEAK_GBS * 1e9)
    out["steps_compared"] = warmup + steps + profile_steps
    out["field_errors_mixed_vs_fp64"] = field_errors(r["mixed"]["flow"], r["fp64"]["flow"])
    return out


def batched_c4(cfdx, steps: int, warmup: int, nblocks: int = 64, grid_cap: int = 0):
    """C4: the 64 seeded periodic cubes (edge U{32,48,64,96,128}, isentropic vortex / TGV ~
    Bernoulli(1/2), Mach U(0.1,0.6), Re U(500,5000)), set up on the device; mixed precision, one
    batched launch per kernel class replayed from a hipGraph of one RK3 step, tuned tiles. Also
    times the same solves in plain FP64 one block after another."""
    from cfdx import tune, workloads
    ctx = cfdx.Ctx(0)
    specs = workloads.c4_specs(workloads.SEED, nblocks)

    def build(precision):
        blocks = []
        for sp in specs:
            b = cfdx.Block(sp.edge, sp.edge, sp.edge, precision=precision, ctx=ctx)
            b.init_case(**sp.case_kw())
            prm = cfdx.default_params(mach=sp.mach, reyl=sp.reyl)
            b.configure(sp.mach, 0.0, prm)
            dt = b.cfl_dt(0.5)
            b.configure(sp.mach, dt, prm)
            blocks.append((b, dt, prm, sp.mach))
        return blocks

    nodes = sum(sp.edge ** 3 for sp in specs)
    blocks = build(cfdx.MIXED)
    q = cfdx.Batch(cfdx.MIXED, ctx)
    for b, *_ in blocks:
        q.add(b)
    tuned = tune.tune_batch(q, ctx, key=f"batch:c4:{workloads.SEED}:{nblocks}:p{cfdx.MIXED}", steps=1,
                            cache=tune.TileCache())
    if grid_cap:
        q.set_grid(grid_cap)
    q.step(warmup, True)
    ctx.sync()
    ctx.record(0)
    q.step(steps, True)
    ctx.record(1)
    ctx.sync()
    ms = ctx.elapsed_ms(0, 1)
    finite = all(b.check_finite() for b, *_ in blocks)
    q.close()
    for b, *_ in blocks:
        b.close()
    plain = build(cfdx.FP64)
    for b, dt, prm, mach in plain:
        b.step(warmup, mach, dt, prm)
    ctx.sync()
    ctx.record(2)
    for b, dt, prm, mach in plain:
        b.step(steps, mach, dt, prm)
    ctx.record(3)
    ctx.sync()
    ms_plain = ctx.elapsed_ms(2, 3)
    for b, *_ in plain:
        b.close()
    hist = {e: sum(1 for sp in specs if sp.edge == e) for e in workloads.C4_EDGES}
    return {"workload": "64 periodic cubes, edge U{32,48,64,96,128}, vortex/TGV, Mach U(0.1,0.6), Re U(500,5000)",
            "edges": hist, "nodes": nodes, "steps": steps, "precision": "mixed", "graph": True,
            "value": nodes * steps / (ms / 1e3), "unit": "cell-updates/s", "ms_per_step": ms / steps,
            "tiles": {"face_i_j_k_node": tuned["tiles"], "grid_cap": tuned["grid_cap"], "cached": tuned["cached"]},
            # one RK3 step per grid size (0: one workgroup per tile; else a persistent grid of that many
            # workgroups striding over the tiles), timed by the tuner
            "grid_cap_step_ms": (dict(zip(map(str, tune.GRID_CAPS), tuned["times_ms"]["grid_cap"]))
                                 if tuned.get("times_ms") and "grid_cap" in tuned["times_ms"] else None),
            "plain_fp64_sequential": nodes * steps / (ms_plain / 1e3),
            "speedup_vs_plain_fp64": ms_plain / ms, "state_finite": finite}


def airfoil_c3(cfdx, steps: int):
    """NACA 0012 C-grid 384x128x32 (far field 20 chords, wall spacing 1e-4 chord, span 0.25), Mach
    0.5, AoA 2 deg, Re 1e5, MTS LES: the mixed form and the plain FP64 form from the same free
    stream for `steps` RK3 steps; throughput of both and per-field errors of mixed against FP64."""
    import numpy as np
    ctx = cfdx.Ctx(0)
    d, g = cfdx.cgrid_desc(384, 128, 32, naca=(0.0, 0.4, 0.12), wake=20.0, radius=20.0, wall=1e-4, span=0.25,
                           aoa_deg=2.0)
    t0 = time.perf_counter()
    _, geo = cfdx.cgrid_geometry(d, g)
    t_geo = time.perf_counter() - t0
    n = 384 * 128 * 32
    prm = cfdx.default_params(mach=0.5, reyl=1e5)
    flow0 = cfdx.freestream(n, 0.5, 2.0)
    # FP64 from the free stream with u perturbed by 1e-7 relative (seeded): how much the flow itself
    # amplifies a perturbation of the size of one FP32 rounding over the same steps
    rng = np.random.default_rng(20261202)
    flow_p = flow0.copy()
    flow_p["u"] *= 1 + 1e-7 * rng.standard_normal(n)
    res, dt = {}, None
    for name, prec in (("fp64", cfdx.FP64), ("mixed", cfdx.MIXED), ("fp64_perturbed", cfdx.FP64)):
        b = cfdx.Block(384, 128, 32, precision=prec, ctx=ctx, desc=cfdx.BlockDesc.from_buffer_copy(d))
        b.set_geometry_soa(geo)
        b.configure(0.5, 0.0, prm)
        b.set_flow(flow_p if name == "fp64_perturbed" else flow0)
        if dt is None:
            dt = b.cfl_dt(0.5)
        b.configure(0.5, dt, prm)
        q = graph_steps(cfdx, b, ctx)
        ctx.sync()
        ctx.record(0)
        q.step(steps, True)
        ctx.record(1)
        ctx.sync()
        ms = ctx.elapsed_ms(0, 1)
        q.remove(b)
        q.close()
        res[name] = (b.get_flow(), n * steps / (ms / 1e3), ms / steps)
        b.close()
    errs = field_errors(res["mixed"][0], res["fp64"][0])
    sens = field_errors(res["fp64_perturbed"][0], res["fp64"][0])
    finite = all(bool(np.isfinite(cfdx_matrix(res[k][0])).all()) for k in res)
    return {"workload": "NACA 0012 C-grid 384x128x32, far field 20c, wall 1e-4c, Mach 0.5, AoA 2 deg, Re 1e5, MTS LES",
            "steps": steps, "dt": dt, "geometry_s": round(t_geo, 3),
            "fp64_value": res["fp64"][1], "mixed_value": res["mixed"][1], "unit": "cell-updates/s",
            "fp64_ms_per_step": res["fp64"][2], "mixed_ms_per_step": res["mixed"][2],
            "mixed_over_fp64": res["mixed"][1] / res["fp64"][1], "field_errors_mixed_vs_fp64": errs,
            "fp64_sensitivity": {"perturbation": "u *= 1 + 1e-7 N(0,1) at t = 0", "field_errors_vs_fp64": sens},
            "state_finite": finite}


SERVE_MAX_NODES = 5_000_000


def serving(cfdx, nreq: int, rate: float, nsteps: int, seed: int = 20261202, kind: str = "airfoil",
            sjf: bool = False):
    """A Poisson request stream (the full C5 stream: 200 requests) served by continuous batching
    (mixed, arrival order; with sjf also least work first) and by plain FP64 one request at a
    time, on the same stream. Latency runs from arrival to the final state on the host."""
    from cfdx import serve
    if kind == "airfoil":
        reqs = serve.airfoil_stream(nreq, rate, seed=seed, nsteps=nsteps)
        out = {"stream": f"{nreq} requests, Poisson {rate}/s, {nsteps} RK3 steps each, NACA 4-digit C-grids "
                         "128x64x16..384x128x32, camber U(0,0.04), thickness U(0.08,0.18), AoA U(-4,10) deg, "
                         "Mach U(0.3,0.7), Re U(1e5,1e6)"}
    else:
        reqs = serve.periodic_stream(nreq, rate, seed=seed, nsteps=nsteps)
        out = {"stream": f"{nreq} requests, Poisson {rate}/s, {nsteps} RK3 steps each, grids 128x64x16..384x128x32, "
                         "periodic TGV / vortex (curvilinear map), Mach U(0.1,0.6), Re U(500,5000)"}
    runs = [("batched_mixed", cfdx.MIXED, True, "fifo"), ("plain_fp64", cfdx.FP64, False, "fifo")]
    if sjf:
        runs.insert(1, ("batched_mixed_sjf", cfdx.MIXED, True, "sjf"))
    for name, prec, batched, policy in runs:
        # batched: at most SERVE_MAX_NODES nodes in flight (enough to fill the GPU; the earlier
        # requests finish first under load); "sjf" admits the queued request with the least work
        srv = serve.Server(precision=prec, batched=batched, chunk=10, graph=True, keep_flow=True,
                           max_nodes=SERVE_MAX_NODES if batched else None, policy=policy)
        res = srv.run(reqs, clock="wall")
        rep = serve.report(reqs, res)
        rep["peak_active"] = srv.peak_active
        rep["max_nodes_in_flight"] = srv.max_nodes
        rep["peak_bytes"] = srv.peak_bytes
        # checked on the device when each request retires (Result.finite)
        rep["state_finite"] = all(r.finite for r in res.values())
        del res
        srv.close()
        out[name] = rep
        print(f"bench: serving {name}: p50 {rep['p50_s']:.3f}s p99 {rep['p99_s']:.3f}s "
              f"{rep['cell_updates_per_s']:.3e} cell-updates/s", file=sys.stderr, flush=True)
    out["throughput_ratio"] = out["batched_mixed"]["cell_updates_per_s"] / out["plain_fp64"]["cell_updates_per_s"]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ns256", choices=["ns256", "c2", "c1", "c3"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--legs", default="compare,c2,dropin,c4,c3,serve",
                    help="legs after the headline: compare (the other precision on the headline grid, per-field "
                         "errors), c2 (128^3 FP64 and mixed), dropin (C2 through the drop-in kernels), c4, c3, serve "
                         "(none: the headline only)")
    ap.add_argument("--cpu", default="c1,c2,ns256",
                    help="CPU-baseline configurations (all cores, then one thread; each beside the GPU on the same "
                         "case, dt and steps), or none")
    ap.add_argument("--no-cpu", dest="cpu", action="store_const", const="none")
    ap.add_argument("--c2-steps", type=int, default=20)
    ap.add_argument("--c3-steps", type=int, default=500)
    ap.add_argument("--c4-steps", type=int, default=5)
    ap.add_argument("--serve-kind", default="airfoil", choices=["airfoil", "periodic"])
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-steps", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-sjf", action="store_true", help="also serve the stream least-work-first")
    ap.add_argument("--no-tune", dest="tune", action="store_false", help="default tiles instead of the autotuned ones")
    ap.add_argument("--profile-steps", type=int, default=2, help="steps of the per-kernel event timing pass")
    ap.add_argument("--copy-mib", type=int, default=256, help="size of the reference copy kernel run (HBM counter scale)")
    ap.add_argument("--sgs", default="mts", choices=["mts", "smagorinsky"],
                    help="SGS model of the timed workload: the reference's mixed time scale (default) or Smagorinsky")
    ap.add_argument("--c-s", dest="c_s", type=float, default=0.17, help="Smagorinsky constant")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: one GPU per process; --gpus {args.gpus} is not supported (the periodic block is "
              "single-GPU; the multi-GPU halo exchange is a design only)", file=sys.stderr)
        return 2
    legs = {x for x in args.legs.split(",") if x and x != "none"}
    cpu_cfgs = [x for x in args.cpu.split(",") if x and x != "none"]

    cfdx = _import_pkg()
    t_setup = time.perf_counter()
    spec, cdef, pkw, viscous = workload(args.workload)
    if args.sgs == "smagorinsky" and viscous:
        # the Smagorinsky option (no reference counterpart; the reference's own model is MTS)
        pkw = dict(pkw, sgs="smagorinsky", c_s=args.c_s)
        spec["sgs"] = f"smagorinsky(C_s={args.c_s})"
    prm = cfdx.default_params(**pkw) if viscous else cfdx.euler_params()
    mach = cdef["kw"]["mach"]
    n = cdef["ni"] * cdef["nj"] * cdef["nk"]
    precision = cfdx.MIXED if args.precision == "mixed" else cfdx.FP64
    tiles = tuned_tiles(cfdx, cdef, precision, mach, prm) if args.tune else None
    block, dt = make_block(cfdx, cdef, precision, prm)
    if tiles:
        block.set_tiles(tuple(tiles["tiles"]))
    print(f"bench: {spec['workload']} n={n} dt={dt:.4e} setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr,
          flush=True)

    ctx = block.ctx
    t0 = time.perf_counter()
    dev_ms, tk = time_block(block, ctx, mach, dt, prm, args.steps, args.warmup, args.profile_steps, cfdx,
                            tiles["tiles"] if tiles else None)
    wall_s = time.perf_counter() - t0
    ms_per_step = dev_ms / args.steps
    value = n * args.steps / (dev_ms / 1e3)
    print(f"bench: {args.steps} steps device {dev_ms:.1f}ms ({wall_s:.2f}s wall incl. timing pass) "
          f"{value:.3e} cell-updates/s", file=sys.stderr, flush=True)
    finite = block.check_finite()
    flow = block.get_flow() if "compare" in legs else None
    block.close()

    # reference copy kernel of known size (the calibration of the FETCH_SIZE / WRITE_SIZE counters of a
    # profiled run); the first launch loads the code object, so time the second
    nb = args.copy_mib << 20
    a, b = ctx.zeros(nb // 8), ctx.zeros(nb // 8)
    ctx.copy(b, a)
    ctx.record(2)
    ctx.copy(b, a)
    ctx.record(3)
    copy_ms = ctx.elapsed_ms(2, 3)
    a.free()
    b.free()

    # whole-step figures: the algorithmic 848 B per cell-update, the counters' measured bytes and VALU
    # instructions per RK stage (committed profile of this workload), the static operation count
    t_tiles = tiles["tiles"] if tiles else None
    cf = counter_figures(cfdx, args.workload, args.precision, t_tiles, viscous, n)
    stage_ms = per_stage_ms(tk, args.profile_steps)
    alg_gbs = ALG_BYTES_PER_CELL_UPDATE * value / 1e9
    step_s = ms_per_step / 1e3
    counters = None
    bound = None
    if cf:
        traffic_step = 3 * cf["stage_hbm_bytes"]
        valu_step = 3 * cf["stage_valu_instr"]
        hbm_meas = traffic_step / step_s / 1e9
        valu_frac_step = valu_step / step_s / VALU_PEAK_WAVE_INSTR_S
        per_kernel = {}
        for k, r in cf["kernels"].items():
            sec = r["avg_us"] * 1e-6
            per_kernel[k] = {"avg_us_profiled": r["avg_us"], "hbm_bytes_per_launch": r["hbm_bytes"],
                             "hbm_GBs": r["hbm_bytes"] / sec / 1e9, "hbm_frac": r["hbm_bytes"] / sec / (HBM_PEAK_GBS * 1e9),
                             "valu_wave_instr_per_launch": r["valu_instr"],
                             "valu_issue_frac": r["valu_instr"] / sec / VALU_PEAK_WAVE_INSTR_S,
                             "fp64_tflops": r["fp64_flop"] / sec / 1e12, "fp32_tflops": r["fp32_flop"] / sec / 1e12}
            per_kernel[k]["bound"] = "valu" if per_kernel[k]["valu_issue_frac"] > per_kernel[k]["hbm_frac"] else "hbm"
        # the step's bound: the limiter of the kernels that take most of its time
        t_by = {}
        for k, r in cf["kernels"].items():
            t_by[per_kernel[k]["bound"]] = t_by.get(per_kernel[k]["bound"], 0.0) + r["avg_us"]
        bound = max(t_by, key=t_by.get)
        counters = {"source": cf["source"], "same_variants": cf["same_variants"],
                    "hbm_bytes_per_step": traffic_step, "hbm_bytes_per_cell_update": traffic_step / n,
                    "traffic_over_alg": traffic_step / n / ALG_BYTES_PER_CELL_UPDATE,
                    "measured_hbm_GBs": hbm_meas, "measured_hbm_frac": hbm_meas / HBM_PEAK_GBS,
                    "valu_wave_instr_per_step": valu_step, "valu_issue_frac_step": valu_frac_step,
                    "profiled_stage_us": cf["stage_profiled_us"], "kernel_time_by_bound_us": t_by,
                    "per_kernel": per_kernel}
    # the face sweep, timed live (HIP events around each face-kernel launch of the timing pass)
    face_launches = 9 * args.profile_steps
    face_ms = tk["face_ms"] / face_launches
    fbytes = face_bytes_per_node(args.precision) * n
    face = {"kernel": "k_face (i, j, k sweeps; mean per launch)", "avg_ms": face_ms,
            "alg_bytes_per_launch": fbytes, "alg_GBs": fbytes / (face_ms / 1e3) / 1e9,
            "alg_hbm_frac": fbytes / (face_ms / 1e3) / (HBM_PEAK_GBS * 1e9)}
    if counters:
        fk = [k for k in counters["per_kernel"] if k.startswith("k_face")]
        face["valu_issue_frac"] = sum(counters["per_kernel"][k]["valu_wave_instr_per_launch"] for k in fk) / len(fk) \
            / (face_ms / 1e3) / VALU_PEAK_WAVE_INSTR_S
        face["valu_wave_instr_per_face"] = sum(counters["per_kernel"][k]["valu_wave_instr_per_launch"] for k in fk) \
            / len(fk) * 64 / n
        face["fp64_tflops"] = sum(cf["kernels"][k]["fp64_flop"] for k in fk) / len(fk) / (face_ms / 1e3) / 1e12

    counts = ctx.op_count(prm)
    ops = alg_ops_per_cell_update(counts, viscous)
    peak_ops = (FP64_PEAK_TFLOPS if precision == cfdx.FP64 else FP32_PEAK_TFLOPS) * 1e12
    valu_alg = {"ops_per_cell_update": ops, "ops_per_face": counts["face"]["total"],
                "achieved_tops": ops * value / 1e12, "peak_tops": peak_ops / 1e12, "frac": ops * value / peak_ops,
                "note": "static count of the kernels' arithmetic (cfdx_op_count + the node kernels' own lines); "
                        "divisions, square roots and reciprocals count one each"}

    line = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == cfdx.FP64 else "f32 fluxes / f64 state",
        "data": "synthetic",
        "config": dict(spec, parallelism="single GPU, one block (one GPU's share of a 512^3 run split 2x2x2)"
                       if args.workload == "ns256" else "single GPU, one block", precision=args.precision,
                       stepping="captured RK3-step graph replays (one-block batch)"),
        "roofline": {"bound": bound or "valu", "achieved": alg_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg_gbs / HBM_PEAK_GBS,
                     "traffic": counters["hbm_bytes_per_step"] if counters else None,
                     "achieved_is": f"{ALG_BYTES_PER_CELL_UPDATE} B algorithmic HBM bytes per cell-update x "
                                    "cell-updates/s (whole RK3 step)",
                     "traffic_is": "measured HBM bytes per RK3 step (FETCH_SIZE + WRITE_SIZE of the step's kernels, "
                                   "copy-kernel calibrated)",
                     "counters": counters, "face_kernel": face},
        "valu_alg_whole_step": valu_alg,
        "kernel_ms_per_stage": stage_ms,
        "state_finite": finite,
        "copy_kernel": {"bytes_moved": 2 * nb, "ms": copy_ms, "GBs": 2 * nb / (copy_ms / 1e3) / 1e9},
        "tiles": ({"face_i_j_k_node": tiles["tiles"], "cached": tiles["cached"], "times_ms": tiles.get("times_ms")}
                  if tiles else None),
    }

    if "compare" in legs:
        # the other precision on the same grid, initial state and step count: throughput and per-field error
        other = cfdx.FP64 if precision == cfdx.MIXED else cfdx.MIXED
        o = timed_run(cfdx, args.workload, other, args.steps, args.warmup, args.tune, args.profile_steps, ctx=ctx)
        mixed_flow, fp64_flow = (o["flow"], flow) if other == cfdx.MIXED else (flow, o["flow"])
        line["mixed_vs_fp64"] = {
            "other_precision": "mixed" if other == cfdx.MIXED else "fp64",
            "other_value": o["value"], "other_ms_per_step": o["ms_per_step"],
            "mixed_over_fp64": (o["value"] / value) if other == cfdx.MIXED else (value / o["value"]),
            "other_kernel_ms_per_stage": per_stage_ms(o["tk"], args.profile_steps),
            "other_tiles": o["tiles"]["tiles"] if o["tiles"] else None,
            "steps_compared": args.warmup + args.steps + args.profile_steps,
            "field_errors": field_errors(mixed_flow, fp64_flow),
        }
        del o, mixed_flow, fp64_flow
        print(f"bench: other precision {line['mixed_vs_fp64']['other_value']:.3e} cell-updates/s", file=sys.stderr,
              flush=True)
    del flow
    if "c2" in legs:
        line["c2"] = c2_leg(cfdx, args.c2_steps, args.warmup, args.tune, args.profile_steps)
        print(f"bench: c2 fp64 {line['c2']['fp64']['value']:.3e} mixed {line['c2']['mixed']['value']:.3e}",
              file=sys.stderr, flush=True)
    if "dropin" in legs:
        _, c2def, c2kw, _ = workload("c2")
        case = cfdx.periodic_case(128, 128, 128, **c2def["kw"])
        c2prm = cfdx.default_params(**c2kw)
        c2dt = cfdx.cfl_dt(case, 0.5, c2prm)
        c2val = line["c2"]["fp64"]["value"] if "c2" in line else None
        line["dropin_path"] = dropin_path(cfdx, case, 0.1, c2dt, c2prm, c2val)
        del case
        print(f"bench: drop-in kernels {line['dropin_path']['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
    if "c4" in legs:
        line["batched_c4"] = batched_c4(cfdx, args.c4_steps, 1)
        print(f"bench: c4 {line['batched_c4']['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
    if "c3" in legs:
        line["airfoil_c3"] = airfoil_c3(cfdx, args.c3_steps)
        print(f"bench: c3 mixed {line['airfoil_c3']['mixed_value']:.3e} fp64 {line['airfoil_c3']['fp64_value']:.3e}",
              file=sys.stderr, flush=True)
    if "serve" in legs:
        line["serving"] = serving(cfdx, args.serve_requests, args.serve_rate, args.serve_steps, kind=args.serve_kind,
                                  sjf=args.serve_sjf)

    if cpu_cfgs:
        cb = cpu_baseline(cfdx, tuple(cpu_cfgs), tuple(cpu_cfgs))
        if cb:
            # the headline workload's all-core rate is the baseline value; every configuration in "configs"
            head = cb["configs"].get(args.workload)
            key = args.workload if head else next(iter(cb["configs"]))
            rec = cb["configs"][key]
            line["cpu_baseline"] = {
                "value": rec["value"], "unit": "cell-updates/s", "cores": rec["threads"], "kind": cb["kind"],
                "sample": f"{rec['steps']} RK3 steps of {key} ({rec['grid']}) after {rec['warmup']} warm-up, the "
                          f"reference's stage sequence over its AoS arrays, OpenMP over nodes on all "
                          f"{rec['threads']} cores of the process affinity",
                "library": cb["library"], "configs": cb["configs"]}
            line["gpu_over_cpu"] = value / rec["value"]
            gpu_rates = {args.workload: value}
            if "c2" in line:
                gpu_rates["c2"] = line["c2"]["fp64"]["value"]
            line["cpu_baseline"]["gpu_over_cpu_by_config"] = {k: gpu_rates[k] / cb["configs"][k]["value"]
                                                              for k in gpu_rates if k in cb["configs"]}
    print(json.dumps(line))
    return 0


def cfdx_matrix(flow):
    import numpy as np
    return np.stack([flow[f] for f in ("rho", "u", "v", "w", "p")], axis=1)


if __name__ == "__main__":
    sys.exit(main())
