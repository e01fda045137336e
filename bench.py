"""Throughput of the per-step RHS + SSP-RK3 update on one MI355X.

Workload (default): 128^3 periodic curvilinear Taylor-Green vortex, FP64 Navier-Stokes
with the reference's mixed-time-scale LES model, Mach 0.1, Re 1600, CFL 0.5 (dt fixed
from the initial state), full RHS (characteristic WCNS + LLF + 6th-order correction +
viscous/SGS fluxes) and RK3. A step = one full SSP-RK3 step over the whole grid with
the state resident in HBM. Prints one JSON line.

Beside the headline the same JSON line carries (unless --legs drops them): the other precision
on the same case with per-field errors of mixed vs FP64 ("mixed_vs_fp64"), the same flow at 256^3
in FP64 (one GPU's share of a 512^3 run, "one_gpu_share_256"), the 64-cube batched
mixed-precision graph step ("batched_c4"), the NACA 0012 C-grid in mixed and plain FP64 with
per-field errors ("airfoil_c3"), and a bounded sample of the Poisson stream of NACA C-grid solves
served by continuous batching vs plain FP64 one at a time ("serving": p50 / p99 latency,
throughput).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|ns256|c1|c3] [--precision fp64|mixed]
                    [--legs compare,ns256,c4,c3,serve|none] [--no-cpu] [--no-tune]

--workload c3 makes the NACA 0012 C-grid block (384x128x32, free stream start) the timed block.
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
    import numpy as np
    rng = np.random.default_rng(20261202)
    if name in ("c2", "ns256"):
        n = 128 if name == "c2" else 256
        h = 2 * math.pi / n
        phases = tuple(rng.uniform(0, 2 * math.pi, 3))
        spec = dict(workload=f"taylor_green_les_{n}^3_curvilinear", grid=[n, n, n], mach=0.1, re=1600.0,
                    sgs="mixed_time_scale(C=0.03,C_T=10)", cfl=0.5, map="sinusoidal amp=0.05h seeded phases")
        case = dict(ni=n, nj=n, nk=n, kw=dict(grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05 * h,
                                               phases=phases, init="tgv", mach=0.1))
        return spec, case, dict(mach=0.1, reyl=1600.0), True
    if name == "c3":
        spec = dict(workload="naca0012_cgrid_384x128x32_les", grid=[384, 128, 32], mach=0.5, aoa_deg=2.0, re=1e5,
                    sgs="mixed_time_scale(C=0.03,C_T=10)", cfl=0.5,
                    geometry="C-grid, far field 20c, wake 20c, wall 1e-4c, span 0.25c periodic")
        case = dict(ni=384, nj=128, nk=32, cgrid=dict(naca=(0.0, 0.4, 0.12), wake=20.0, radius=20.0, wall=1e-4,
                                                       span=0.25, aoa_deg=2.0), kw=dict(mach=0.5))
        return spec, case, dict(mach=0.5, reyl=1e5), True
    if name == "c1":
        spec = dict(workload="isentropic_vortex_euler_64^3", grid=[64, 64, 64], mach=0.5, beta=5.0, cfl=0.5)
        case = dict(ni=64, nj=64, nk=64, kw=dict(grid="box", length=(10.0,) * 3, init="vortex", mach=0.5))
        return spec, case, dict(), False
    raise SystemExit(f"unknown workload {name}")


def profile_counts(cfdx, workload_name: str, precision: str, tiles=None):
    """Per-launch HBM bytes and VALU counts of the face kernels, from the committed rocprofv3
    summary of this workload (profiles/block_kernels_<workload>_<precision>.md). The summary
    holds one machine-readable row per kernel and tile variant,
    `| kernel<S,DIR,TA,TB> | launches | avg_us | hbm_bytes | valu_instr | fp64_flop | fp32_flop |`;
    the rows of the variants this run uses (tiles = face i, j, k, node variant indices) are read."""
    path = os.path.join(ROOT, "profiles", f"block_kernels_{workload_name}_{precision}.md")
    if not os.path.exists(path):
        return None
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
    t = tuple(tiles) if tiles else (0, 0, 0, 0)
    S = "double" if precision == "fp64" else "float"
    names = []
    for d in range(3):
        ta, tb = cfdx.tile_variants(0 if d == 0 else 1)[t[d]]
        names.append(f"k_face<{S},{d},{ta},{tb}>")
    face, used = [], []
    for d, k in enumerate(names):
        if k not in rows:   # the profiled run tuned another variant of this class: take its counted row
            k = next((r for r in rows if r.startswith(f"k_face<{S},{d},")), None)
            if k is None:
                return None
        face.append(rows[k])
        used.append(k)
    mean = {k: sum(f[k] for f in face) / 3 for k in ("avg_us", "hbm_bytes", "valu_instr", "fp64_flop", "fp32_flop")}
    mean["source"] = os.path.relpath(path, ROOT)
    mean["rows"] = used
    mean["same_variants"] = used == names
    return mean


def cpu_baseline(cfdx, max_seconds: float = 15.0):
    """The reference's own kernels (oracle/_ref, gcc -O3, OpenMP) or the restatement, on a
    bounded sample of the same workload: a 64^3 block of the curvilinear TGV, whole RK3 steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    kind = "reference" if pyoracle.available("reference") else "port"
    if not pyoracle.available(kind):
        return None
    orc = pyoracle.Oracle(kind)
    cores = len(os.sched_getaffinity(0))
    orc.set_threads(cores)
    n = 64
    h = 2 * math.pi / n
    case = cfdx.periodic_case(n, n, n, grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05 * h,
                              phases=(0.3, 1.1, 2.0), init="tgv", mach=0.1)
    prm = pyoracle.reference_params(mach=0.1, reyl=1600.0)
    dt = cfdx.cfl_dt(case, 0.5)
    work = orc.new_work(case["Grid_P"].shape[0])
    t0 = time.perf_counter()
    orc.steps(case, 1, 0.1, dt, prm, work)
    t1 = time.perf_counter()
    steps = max(1, int(max_seconds / max(t1 - t0, 1e-6)) - 1)
    steps = min(steps, 100)
    t2 = time.perf_counter()
    orc.steps(case, steps, 0.1, dt, prm, work)
    t3 = time.perf_counter()
    cups = n ** 3 * steps / (t3 - t2)
    # the same block on one thread, a few steps
    orc.set_threads(1)
    t4 = time.perf_counter()
    orc.steps(case, 2, 0.1, dt, prm, work)
    t5 = time.perf_counter()
    orc.set_threads(cores)
    return {"value": cups, "unit": "cell-updates/s", "cores": cores, "kind": kind,
            "sample": f"{steps} RK3 steps of a {n}^3 curvilinear TGV block (same kernels, FP64, OpenMP over nodes)",
            "seconds": round(t3 - t2, 2),
            "single_thread": {"value": n ** 3 * 2 / (t5 - t4), "cores": 1, "sample": f"2 RK3 steps of the same {n}^3 block",
                              "seconds": round(t5 - t4, 2)}}


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
            "block_path_over_dropin": block_value / value, "state_finite": finite}


def one_gpu_share(cfdx, tune: bool, steps: int = 10, warmup: int = 2):
    """The C2 map and flow at 256^3, plain FP64: one GPU's share of a 512^3 run split 2x2x2 (the
    periodic wrap stands in for the neighbours' halo), set up on the device."""
    spec, cdef, pkw, _ = workload("ns256")
    prm = cfdx.default_params(**pkw)
    mach = cdef["kw"]["mach"]
    tiles = tuned_tiles(cfdx, cdef, cfdx.FP64, mach, prm) if tune else None
    b, dt = make_block(cfdx, cdef, cfdx.FP64, prm)
    if tiles:
        b.set_tiles(tuple(tiles["tiles"]))
    ctx = b.ctx
    b.step(warmup, mach, dt, prm)
    ctx.sync()
    ctx.record(0)
    b.step(steps, mach, dt, prm)
    ctx.record(1)
    ctx.sync()
    ms = ctx.elapsed_ms(0, 1)
    import numpy as np
    finite = bool(np.isfinite(cfdx_matrix(b.get_flow())).all())
    b.close()
    n = cdef["ni"] * cdef["nj"] * cdef["nk"]
    return {"workload": spec["workload"], "precision": "fp64", "steps": steps, "value": n * steps / (ms / 1e3),
            "unit": "cell-updates/s", "ms_per_step": ms / steps, "state_finite": finite,
            "tiles": tiles["tiles"] if tiles else None}


C4_EDGES = (32, 48, 64, 96, 128)


def batched_c4(cfdx, steps: int, warmup: int, seed: int = 20261202, nblocks: int = 64, grid_cap: int = 0):
    """64 periodic cubes, edge U{32,48,64,96,128}, isentropic vortex / TGV ~ Bernoulli(1/2),
    Mach U(0.1,0.6), Re U(500,5000), set up on the device; mixed precision, one batched launch per
    kernel class replayed from a hipGraph of one RK3 step. Also times the same solves in plain
    FP64 one block after another."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ctx = cfdx.Ctx(0)
    specs = []
    for _ in range(nblocks):
        e = int(rng.choice(C4_EDGES))
        tgv = bool(rng.integers(2))
        specs.append(dict(e=e, tgv=tgv, mach=float(rng.uniform(0.1, 0.6)), reyl=float(rng.uniform(500, 5000)),
                          phases=tuple(float(x) for x in rng.uniform(0, 2 * math.pi, 3))))

    def build(precision):
        blocks = []
        for sp in specs:
            e = sp["e"]
            L = 2 * math.pi if sp["tgv"] else 10.0
            b = cfdx.Block(e, e, e, precision=precision, ctx=ctx)
            b.init_case(grid="sinusoid", length=(L,) * 3, amp=0.05 * L / e, phases=sp["phases"],
                        init="tgv" if sp["tgv"] else "vortex", mach=sp["mach"], center=(5.0, 5.0))
            prm = cfdx.default_params(mach=sp["mach"], reyl=sp["reyl"])
            b.configure(sp["mach"], 0.0, prm)
            dt = b.cfl_dt(0.5)
            b.configure(sp["mach"], dt, prm)
            blocks.append((b, dt, prm, sp["mach"]))
        return blocks

    nodes = sum(sp["e"] ** 3 for sp in specs)
    blocks = build(cfdx.MIXED)
    q = cfdx.Batch(cfdx.MIXED, ctx)
    for b, *_ in blocks:
        q.add(b)
    from cfdx import tune
    tuned = tune.tune_batch(q, ctx, key=f"batch:c4:{seed}:{nblocks}:p{cfdx.MIXED}", steps=1, cache=tune.TileCache())
    if grid_cap:
        q.set_grid(grid_cap)
    q.step(warmup, True)
    ctx.sync()
    ctx.record(0)
    q.step(steps, True)
    ctx.record(1)
    ctx.sync()
    ms = ctx.elapsed_ms(0, 1)
    finite = all(bool(np.isfinite(cfdx_matrix(b.get_flow())).all()) for b, *_ in blocks[:4])
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
    hist = {e: sum(1 for sp in specs if sp["e"] == e) for e in C4_EDGES}
    return {"workload": "64 periodic cubes, edge U{32,48,64,96,128}, vortex/TGV, Mach U(0.1,0.6), Re U(500,5000)",
            "edges": hist, "nodes": nodes, "steps": steps, "precision": "mixed", "graph": True,
            "value": nodes * steps / (ms / 1e3), "unit": "cell-updates/s", "ms_per_step": ms / steps,
            "tiles": {"face_i_j_k_node": tuned["tiles"], "grid_cap": tuned["grid_cap"], "cached": tuned["cached"]},
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


def serving(cfdx, nreq: int, rate: float, nsteps: int, seed: int = 20261202, kind: str = "airfoil"):
    """A bounded sample of a Poisson request stream served by continuous batching (mixed; arrival
    order, and least work first) and by plain FP64 one request at a time, on the same stream."""
    import numpy as np
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
    runs = (("batched_mixed", cfdx.MIXED, True, "fifo"), ("batched_mixed_sjf", cfdx.MIXED, True, "sjf"),
            ("plain_fp64", cfdx.FP64, False, "fifo"))
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
        rep["state_finite"] = all(bool(np.isfinite(cfdx_matrix(r.flow)).all()) for r in res.values())
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
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--legs", default="compare,dropin,ns256,c4,c3,serve",
                    help="extra legs after the headline: compare (other precision, per-field errors), dropin (the "
                         "same case through the drop-in kernels), ns256 (256^3 FP64, one GPU's share of 512^3), c4, "
                         "c3, serve (none: the headline only)")
    ap.add_argument("--c3-steps", type=int, default=500)
    ap.add_argument("--serve-kind", default="airfoil", choices=["airfoil", "periodic"])
    ap.add_argument("--c4-steps", type=int, default=5)
    ap.add_argument("--no-tune", dest="tune", action="store_false", help="default tiles instead of the autotuned ones")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-steps", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--profile-steps", type=int, default=2, help="steps of the per-kernel event timing pass")
    ap.add_argument("--copy-mib", type=int, default=256, help="size of the reference copy kernel run (HBM counter scale)")
    ap.add_argument("--sgs", default="mts", choices=["mts", "smagorinsky"],
                    help="SGS model of the timed workload: the reference's mixed time scale (default) or Smagorinsky")
    ap.add_argument("--c-s", dest="c_s", type=float, default=0.17, help="Smagorinsky constant")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs one GPU per process; --gpus {args.gpus} is not supported "
              "(the periodic block is single-GPU, multi-GPU halo exchange is design-only)", file=sys.stderr)
        return 2
    legs = {x for x in args.legs.split(",") if x}

    import numpy as np
    cfdx = _import_pkg()
    spec, cdef, pkw, viscous = workload(args.workload)
    t_setup = time.perf_counter()
    # periodic workloads: the case built on the host (the same arrays the oracle would take)
    case = None if "cgrid" in cdef else cfdx.periodic_case(cdef["ni"], cdef["nj"], cdef["nk"], **cdef["kw"])
    if args.sgs == "smagorinsky" and viscous:
        # the Smagorinsky option (no reference counterpart; the reference's own model is MTS)
        pkw = dict(pkw, sgs="smagorinsky", c_s=args.c_s)
        spec["sgs"] = f"smagorinsky(C_s={args.c_s})"
    prm = cfdx.default_params(**pkw) if viscous else cfdx.euler_params()
    mach = cdef["kw"]["mach"]
    n = cdef["ni"] * cdef["nj"] * cdef["nk"]
    precision = cfdx.MIXED if args.precision == "mixed" else cfdx.FP64
    tiles = tuned_tiles(cfdx, cdef, precision, mach, prm) if args.tune else None
    block, dt = make_block(cfdx, cdef, precision, prm, host_case=case)
    if tiles:
        block.set_tiles(tuple(tiles["tiles"]))
    print(f"bench: {spec['workload']} n={n} dt={dt:.4e} setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)

    ctx = block.ctx
    t0 = time.perf_counter()
    dev_ms, tk = time_block(block, ctx, mach, dt, prm, args.steps, args.warmup, args.profile_steps, cfdx,
                            tiles["tiles"] if tiles else None)
    wall_s = time.perf_counter() - t0
    ms_per_step = dev_ms / args.steps
    value = n * args.steps / (dev_ms / 1e3)
    print(f"bench: {args.steps} steps device {dev_ms:.1f}ms ({wall_s:.2f}s wall incl. timing pass)", file=sys.stderr, flush=True)
    face_launches = 9 * args.profile_steps
    face_ms = tk["face_ms"] / face_launches
    face_bytes = face_bytes_per_node(args.precision)
    achieved = face_bytes * n / (face_ms / 1e3) / 1e9
    flow = block.get_flow()
    finite = bool(np.isfinite(cfdx_matrix(flow)).all())

    # reference copy kernel of known size (scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run);
    # the first launch loads the code object, so time the second
    nb = args.copy_mib << 20
    a, b = ctx.zeros(nb // 8), ctx.zeros(nb // 8)
    ctx.copy(b, a)
    ctx.record(2)
    ctx.copy(b, a)
    ctx.record(3)
    copy_ms = ctx.elapsed_ms(2, 3)
    a.free(); b.free()

    prof = profile_counts(cfdx, args.workload, args.precision, tiles["tiles"] if tiles else None)
    traffic = prof["hbm_bytes"] if prof else None
    valu = None
    if prof:
        sec = face_ms / 1e3
        valu = {"wave_instr_per_launch": prof["valu_instr"],
                "achieved_wave_instr_s": prof["valu_instr"] / sec,
                "peak_wave_instr_s": VALU_PEAK_WAVE_INSTR_S,
                "issue_frac": prof["valu_instr"] / sec / VALU_PEAK_WAVE_INSTR_S,
                "fp64_tflops": prof["fp64_flop"] / sec / 1e12, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                "fp32_tflops": prof["fp32_flop"] / sec / 1e12,
                "counts_from": prof["source"], "profile_rows": prof["rows"], "same_variants": prof["same_variants"],
                "profiled_kernel_us": prof["avg_us"]}

    counts = ctx.op_count(prm)
    ops = alg_ops_per_cell_update(counts, viscous)
    alg = {"ops_per_cell_update": ops, "ops_per_face": counts["face"]["total"],
           "counts": {k: v for k, v in counts.items()},
           "achieved_tops": ops * value / 1e12, "peak_tops": FP64_PEAK_TFLOPS,
           "frac": ops * value / (FP64_PEAK_TFLOPS * 1e12),
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
        "config": dict(spec, parallelism="single GPU, one block", precision=args.precision,
                       stepping="captured RK3-step graph replays (one-block batch)"),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_face (i, j, k sweeps; mean per launch)", "kernel_ms": face_ms,
                     "bytes_per_launch": face_bytes * n},
        "valu": valu,
        "hbm_alg_frac_whole_step": ALG_BYTES_PER_CELL_UPDATE * value / (HBM_PEAK_GBS * 1e9),
        "valu_alg_whole_step": alg,
        "kernel_ms_per_stage": {k: v / (3 * args.profile_steps) for k, v in tk.items() if k.endswith("_ms")},
        "state_finite": finite,
        "copy_kernel": {"bytes_moved": 2 * nb, "ms": copy_ms, "GBs": 2 * nb / (copy_ms / 1e3) / 1e9},
        "tiles": ({"face_i_j_k_node": tiles["tiles"], "cached": tiles["cached"], "times_ms": tiles.get("times_ms")}
                  if tiles else None),
    }

    if "compare" in legs:
        # the other precision on the same initial state and step count: throughput and per-field error
        other = cfdx.FP64 if precision == cfdx.MIXED else cfdx.MIXED
        otiles = tuned_tiles(cfdx, cdef, other, mach, prm) if args.tune else None
        ob, _ = make_block(cfdx, cdef, other, prm, ctx=ctx, host_case=case)
        ob.configure(mach, dt, prm)
        if otiles:
            ob.set_tiles(tuple(otiles["tiles"]))
        o_ms, o_tk = time_block(ob, ctx, mach, dt, prm, args.steps, args.warmup, args.profile_steps, cfdx,
                                otiles["tiles"] if otiles else None)
        oflow = ob.get_flow()
        ob.close()
        mixed_flow, fp64_flow = (oflow, flow) if other == cfdx.MIXED else (flow, oflow)
        o_value = n * args.steps / (o_ms / 1e3)
        line["mixed_vs_fp64"] = {
            "other_precision": "mixed" if other == cfdx.MIXED else "fp64",
            "other_value": o_value,
            "mixed_over_fp64": (o_value / value) if other == cfdx.MIXED else (value / o_value),
            "other_kernel_ms_per_stage": {k: v / (3 * args.profile_steps) for k, v in o_tk.items() if k.endswith("_ms")},
            "steps_compared": args.warmup + args.steps + args.profile_steps,
            "field_errors": field_errors(mixed_flow, fp64_flow),
        }
        print(f"bench: other precision {o_value:.3e} cell-updates/s", file=sys.stderr, flush=True)
    if "dropin" in legs and "cgrid" not in cdef:
        line["dropin_path"] = dropin_path(cfdx, case, mach, dt, prm, value)
        print(f"bench: drop-in kernels {line['dropin_path']['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
    del case
    block.close()

    if "ns256" in legs:
        line["one_gpu_share_256"] = one_gpu_share(cfdx, args.tune)
        print(f"bench: 256^3 {line['one_gpu_share_256']['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
    if "c4" in legs:
        line["batched_c4"] = batched_c4(cfdx, args.c4_steps, 1)
        print(f"bench: c4 {line['batched_c4']['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
    if "c3" in legs:
        line["airfoil_c3"] = airfoil_c3(cfdx, args.c3_steps)
        print(f"bench: c3 mixed {line['airfoil_c3']['mixed_value']:.3e} fp64 {line['airfoil_c3']['fp64_value']:.3e}",
              file=sys.stderr, flush=True)
    if "serve" in legs:
        line["serving"] = serving(cfdx, args.serve_requests, args.serve_rate, args.serve_steps, kind=args.serve_kind)

    if not args.no_cpu:
        cb = cpu_baseline(cfdx)
        if cb:
            line["cpu_baseline"] = cb
            line["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(line))
    return 0


def cfdx_matrix(flow):
    import numpy as np
    return np.stack([flow[f] for f in ("rho", "u", "v", "w", "p")], axis=1)


if __name__ == "__main__":
    sys.exit(main())
