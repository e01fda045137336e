"""Serving loop: parametric flow solves admitted as they arrive, stepped together on one GPU.

Each request is one block (grid size, map, initial state, Mach, Re, number of RK3 steps). The
server keeps the state of every admitted request resident in HBM and steps all of them in the
same batched launches (continuous batching): between chunks of steps it retires finished
requests (result copied to the host, block freed) and admits queued ones (first come, first
served, or with policy "sjf" the queued request with the least work first) while the device has
free memory (hipMemGetInfo minus a reserve, optionally capped by a
byte budget) and, optionally, while the nodes in flight stay under a cap. A request's
arithmetic does not depend on what else is in the batch, so its result is the same alone and
in the mix (the tests check bitwise equality).

Requests are periodic blocks set up on the device (TGV / isentropic vortex on a mapped box) or
NACA 4-digit C-grid solves (grid = "cgrid"), whose geometry is built on the host by worker threads
as requests arrive, overlapping the GPU work.

Clocks: "wall" takes arrival times in seconds from the start of run(); "steps" measures arrival
in RK3 steps executed by the server, which makes a schedule reproducible for tests.
"""
from __future__ import annotations

import bisect
import collections
import concurrent.futures
import contextlib
import dataclasses
import math
import threading
import time

import numpy as np

from . import (FP64, MIXED, Batch, Block, BlockDesc, Ctx, cgrid_desc, cgrid_geometry, default_params, euler_params,
               freestream)


@dataclasses.dataclass
class Request:
    rid: int
    ni: int
    nj: int
    nk: int
    nsteps: int
    arrival: float = 0.0             # seconds ("wall") or steps ("steps")
    grid: str = "sinusoid"
    length: tuple = (2 * math.pi,) * 3
    amp_frac: float = 0.05           # map amplitude as a fraction of the i spacing
    phases: tuple = (0.0, 0.0, 0.0)
    init: str = "tgv"
    mach: float = 0.1
    reyl: float = 1600.0
    viscous: bool = True
    beta: float = 5.0
    center: tuple = (5.0, 5.0)
    cfl: float = 0.5
    naca: tuple = (0.0, 0.4, 0.12)   # grid == "cgrid": (max camber, its position, thickness)
    aoa_deg: float = 0.0             # grid == "cgrid": far-field flow angle

    @property
    def nodes(self) -> int:
        return self.ni * self.nj * self.nk

    def desc(self, precision: int) -> BlockDesc | None:
        if self.grid != "cgrid":
            return None
        return cgrid_desc(self.ni, self.nj, self.nk, precision, naca=self.naca, aoa_deg=self.aoa_deg)[0]

    def bytes(self, precision: int) -> int:
        return Block.bytes_for(self.ni, self.nj, self.nk, precision, desc=self.desc(precision))

    def params(self):
        if not self.viscous:
            return euler_params()
        return default_params(mach=self.mach, reyl=self.reyl)


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    dt: float
    flow: np.ndarray | None
    finite: bool = True              # the final state holds no NaN / infinity (checked on the device)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def prepare(req: Request) -> dict | None:
    """Host-side set-up of a request (C-grid geometry and the free stream); None when the block is
    set up on the device. Runs in worker threads."""
    if req.grid != "cgrid":
        return None
    d, g = cgrid_desc(req.ni, req.nj, req.nk, FP64, naca=req.naca, aoa_deg=req.aoa_deg)
    _, geo = cgrid_geometry(d, g)
    return {"desc": d, "geo": geo, "flow": freestream(req.nodes, req.mach, req.aoa_deg)}


def make_block(req: Request, precision: int, ctx: Ctx, prep: dict | None = None,
               lock=None) -> tuple[Block, float]:
    """Creates, initialises and configures the block of one request; returns it with its CFL time
    step (computed once from the initial state). `lock`, if given, is held around each library
    call (see Server)."""
    g = lock if lock is not None else contextlib.nullcontext()
    if req.grid == "cgrid":
        prep = prep or prepare(req)
        d = BlockDesc.from_buffer_copy(prep["desc"])
        with g:
            b = Block(req.ni, req.nj, req.nk, precision=precision, ctx=ctx, desc=d)
    else:
        with g:
            b = Block(req.ni, req.nj, req.nk, precision=precision, ctx=ctx)
    try:
        if req.grid == "cgrid":
            with g:
                b.set_geometry_soa(prep["geo"])
            with g:
                b.configure(req.mach, 0.0, req.params())      # the side conditions use the block's Mach
            with g:
                b.set_flow(prep["flow"])
        else:
            amp = req.amp_frac * req.length[0] / req.ni if req.grid == "sinusoid" else 0.0
            with g:
                b.init_case(grid=req.grid, length=req.length, amp=amp, phases=req.phases, init=req.init,
                            mach=req.mach, beta=req.beta, center=req.center)
            with g:
                b.configure(req.mach, 0.0, req.params())      # the stable step uses the request's Re
        with g:
            dt = b.cfl_dt(req.cfl)
        with g:
            b.configure(req.mach, dt, req.params())
    except BaseException:
        with g:
            b.close()
        raise
    return b, dt


def run_alone(req: Request, precision: int, ctx: Ctx, keep_flow: bool = True) -> Result:
    """The request on its own (single-block steps), for equivalence checks."""
    b, dt = make_block(req, precision, ctx)
    b.step(req.nsteps, req.mach, dt, req.params())
    finite = b.check_finite()
    flow = b.get_flow() if keep_flow else None
    b.close()
    return Result(req.rid, 0.0, 0.0, 0.0, dt, flow, finite)


class _Gate:
    """Orders the set-up thread's library calls (`call`) and the chunk loop's graph captures
    (`capture`): one at a time, and a waiting capture goes before the next call, so that a capture
    waits for one call at most (a plain lock, released and taken again at once by the set-up thread
    between its calls, let a capture wait for a whole set-up)."""

    def __init__(self):
        self._cv = threading.Condition()
        self._busy = False
        self._captures = 0
        self.call = _GateSide(self, False)
        self.capture = _GateSide(self, True)

    def _enter(self, capture: bool) -> None:
        with self._cv:
            if capture:
                self._captures += 1
                while self._busy:
                    self._cv.wait()
                self._captures -= 1
            else:
                while self._busy or self._captures:
                    self._cv.wait()
            self._busy = True

    def _exit(self) -> None:
        with self._cv:
            self._busy = False
            self._cv.notify_all()


class _GateSide:
    def __init__(self, gate: _Gate, capture: bool):
        self.gate, self.capture = gate, capture

    def __enter__(self):
        self.gate._enter(self.capture)
        return self

    def __exit__(self, *exc) -> bool:
        self.gate._exit()
        return False


class Server:
    """Continuous batching on one GPU. The batched steps run on the main context's stream. With
    the wall clock, blocks are created, set up (each on a stream of its own), read back and
    released by a set-up thread while chunks of steps run, so block set-up, result copies and host
    work overlap the device work (pipelined: one chunk in flight; before a step that captures its
    graph the chunk loop waits at most for the one library call the set-up thread is making); with
    the steps clock the same happens on the calling thread and a second context. Memory of retired
    blocks is released after the next chunk completes."""

    def __init__(self, precision: int = MIXED, ctx: Ctx | None = None, reserve_bytes: int = 2 << 30,
                 budget_bytes: int | None = None, max_active: int = 256, chunk: int = 10, graph: bool = True,
                 grid_cap: int = 0, batched: bool = True, keep_flow: bool = True, prep_workers: int = 8,
                 max_nodes: int | None = None, pipelined: bool = True, policy: str = "fifo"):
        if policy not in ("fifo", "sjf"):
            raise ValueError(f"policy {policy!r}: 'fifo' or 'sjf'")
        # admission order of the queued requests: arrival ("fifo") or least work, nodes x steps,
        # first ("sjf": shorter median latency under load; a large request waits while smaller
        # ones keep arriving)
        self.policy = policy
        self.ctx = ctx or Ctx()
        # block set-up and result copies on the calling thread (same device, another stream)
        self.io = Ctx(self.ctx.device) if pipelined else self.ctx
        self.pipelined = pipelined
        self.precision = precision
        self.reserve = reserve_bytes
        self.budget = budget_bytes
        # in-flight node cap: enough nodes to fill the GPU, so that under load the earlier requests
        # finish first instead of every admitted request sharing the device until the end
        self.max_nodes = max_nodes
        self.max_active = max_active if batched else 1
        self.chunk = max(1, chunk)
        self.graph = graph
        self.keep_flow = keep_flow
        self.batch = Batch(precision, self.ctx)
        if grid_cap:
            self.batch.set_grid(grid_cap)
        self.peak_active = 0
        self.peak_bytes = 0
        self.chunks = 0
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=max(1, prep_workers))
        # one thread for the set-up, read-back and release of blocks
        self.setup = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        # taken by the set-up thread around each library call and by the chunk loop while the
        # batch captures its step graph: the HIP runtime invalidates a capture when another thread
        # allocates or creates streams during it (a capture failed that way with 12 M nodes in
        # flight), so no other library call runs while a capture is open
        self.gate = _Gate()

    def _fits(self, req: Request, used: int, nodes: int = 0, busy: bool = False, cap: int = 1,
              unallocated: int = 0) -> bool:
        """Whether `req` may be admitted: nodes (in flight and set up) stay within cap x max_nodes,
        the bytes within the budget, and the device keeps the reserve free after `unallocated`
        bytes of blocks still being set up."""
        if self.max_nodes is not None and busy and nodes + req.nodes > cap * self.max_nodes:
            return False
        need = req.bytes(self.precision)
        if self.budget is not None and used + need > self.budget:
            return False
        free, _ = self.ctx.mem_info()
        return free - unallocated - need >= self.reserve

    def run(self, requests: list[Request], clock: str = "wall") -> dict:
        if clock not in ("wall", "steps"):
            raise ValueError(clock)
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        queue: list[Request] = []          # arrived, not admitted; admission takes queue[0]
        order = (lambda r: (r.nodes * r.nsteps, r.arrival, r.rid)) if self.policy == "sjf" else None
        active: dict[int, dict] = {}      # in the batch
        setting_up: list[dict] = []        # admitted, block being set up on the set-up thread
        admitted: list[dict] = []          # set up, joining the batch at a chunk boundary
        finishing: list[dict] = []         # done, result being read back on the set-up thread
        results: dict[int, Result] = {}
        retired: list = []                 # blocks whose memory is released after the next sync
        closing: list = []                 # releases running on the set-up thread
        used = 0
        vclock = 0                         # steps executed (steps clock)
        inflight = 0                       # steps of the chunk running on the device (0: none)
        t0 = time.perf_counter()
        # wall clock: block set-up, result read-back and release run on the set-up thread beside
        # the batched chunks, so the chunk loop does not wait for the uploads; blocks are set up ahead,
        # up to max_nodes more nodes than are in flight. Steps clock: on the calling thread and the
        # set-up context, so that a schedule is reproducible.
        threaded = clock == "wall" and self.pipelined

        def now() -> float:
            return float(vclock) if clock == "steps" else time.perf_counter() - t0

        def nodes_in_flight() -> int:
            return (sum(a["req"].nodes for a in active.values()) + sum(a["req"].nodes for a in admitted)
                    + sum(a["req"].nodes for a in setting_up))

        lock = self.gate.call if threaded else contextlib.nullcontext()

        def set_up(req: Request, fut):
            # threaded: each block on a stream of its own, so that the batch it joins waits for
            # its set-up only, not for the uploads of the blocks set up after it
            prep = fut.result() if fut is not None else None
            if not threaded:
                return make_block(req, self.precision, self.io, prep)
            with lock:
                c = Ctx(self.ctx.device)
            try:
                return make_block(req, self.precision, c, prep, lock=lock)
            finally:
                with lock:
                    c.close()   # the block holds the context until it is destroyed

        def read_back(b: Block):
            # a diverged solve is reported as such (Result.finite, report()["nonfinite"]), not
            # returned as a NaN field
            with lock:
                finite = b.check_finite()
            with lock:
                flow = b.get_flow() if self.keep_flow else None
            return finite, flow, time.perf_counter() - t0

        def close(b: Block) -> None:
            with lock:
                b.close()

        def finish(a: dict, finite: bool, flow, t_done: float) -> None:
            rq = a["req"]
            results[rq.rid] = Result(rq.rid, rq.arrival, a["admitted"], t_done, a["dt"], flow, finite)
            retired.append(a["block"])

        def release(blocks: list) -> None:
            for b in blocks:
                if threaded:
                    closing.append(self.setup.submit(close, b))
                else:
                    b.close()
            blocks.clear()

        preps: dict[int, concurrent.futures.Future] = {}
        try:
            while pending or queue or active or admitted or setting_up or finishing or inflight:
                t = now()
                while pending and pending[0].arrival <= t:
                    req = pending.popleft()
                    if order is None:
                        queue.append(req)
                    else:
                        bisect.insort(queue, req, key=order)
                    if req.grid == "cgrid":
                        preps[req.rid] = self.pool.submit(prepare, req)
                # set-ups and read-backs that completed, in order
                while setting_up and setting_up[0]["fut"].done():
                    a = setting_up.pop(0)
                    a["block"], a["dt"] = a.pop("fut").result()
                    admitted.append(a)
                while finishing and finishing[0]["fut"].done():
                    a = finishing.pop(0)
                    try:
                        out = a.pop("fut").result()
                    except BaseException:
                        retired.append(a["block"])   # released by the failure path
                        raise
                    finish(a, *out)
                # admission in queue order while memory, slots and the node cap allow; the block is
                # set up on the set-up thread (wall clock), or here once its host set-up is ready
                # (steps clock; waited for when nothing else could run)
                busy = bool(active or admitted or setting_up)
                while queue and len(active) + len(admitted) + len(setting_up) < self.max_active and self._fits(
                        queue[0], used, nodes_in_flight(), busy, cap=2 if threaded else 1,
                        unallocated=sum(a["bytes"] for a in setting_up)):
                    req = queue[0]
                    fut = preps.get(req.rid)
                    if not threaded and fut is not None and not fut.done() and (busy or inflight):
                        break
                    queue.pop(0)
                    fut = preps.pop(req.rid, None)
                    nb = req.bytes(self.precision)
                    used += nb
                    a = dict(req=req, left=req.nsteps, admitted=now(), bytes=nb)
                    if threaded:
                        a["fut"] = self.setup.submit(set_up, req, fut)
                        setting_up.append(a)
                    else:
                        a["block"], a["dt"] = set_up(req, fut)
                        admitted.append(a)
                    busy = True
                if queue and not (active or admitted or setting_up or finishing or inflight):
                    if retired or not all(f.done() for f in closing):
                        # the blocks released last still hold their memory: release them, then retry
                        release(retired)
                        concurrent.futures.wait(closing)
                        continue
                    raise RuntimeError(f"request {queue[0].rid} ({queue[0].ni}x{queue[0].nj}x{queue[0].nk}) does not fit "
                                       "in the free device memory")
                # the chunk in flight completes
                done = []
                if inflight:
                    self.ctx.sync()
                    self.chunks += 1
                    vclock += inflight
                    for rid, a in active.items():
                        a["left"] -= inflight
                        if a["left"] == 0:
                            done.append(active[rid])
                    inflight = 0
                    release(retired)
                changed = bool(done)               # the batch captures its step again
                for a in done:
                    active.pop(a["req"].rid)
                    self.batch.remove(a["block"])
                # set-up blocks join in admission order while the in-flight node cap allows (the
                # first always joins an empty batch)
                while admitted and (self.max_nodes is None or not active or sum(
                        x["req"].nodes for x in active.values()) + admitted[0]["req"].nodes <= self.max_nodes):
                    a = admitted.pop(0)
                    self.batch.add(a["block"])
                    active[a["req"].rid] = a
                    changed = True
                # next chunk, asynchronous
                if active:
                    self.peak_active = max(self.peak_active, len(active))
                    self.peak_bytes = max(self.peak_bytes, used)
                    k = min(self.chunk, min(a["left"] for a in active.values()))
                    if clock == "steps" and pending:
                        k = max(1, min(k, int(math.ceil(pending[0].arrival - vclock))))
                    with self.gate.capture if threaded and changed else contextlib.nullcontext():
                        self.batch.step(k, self.graph)
                    inflight = k
                    if not self.pipelined:
                        self.ctx.sync()
                # results of the requests that finished, read back on the set-up stream during the chunk
                for a in done:
                    used -= a["bytes"]
                    if threaded:
                        a["fut"] = self.setup.submit(read_back, a["block"])
                        finishing.append(a)
                    else:
                        finite, flow, _ = read_back(a["block"])
                        finish(a, finite, flow, now())
                if not (active or inflight):
                    if setting_up and not admitted:
                        concurrent.futures.wait([setting_up[0]["fut"]])   # nothing to step until it is set up
                    elif finishing and not (admitted or setting_up):
                        concurrent.futures.wait([finishing[0]["fut"]])
                    elif pending and not (queue or admitted or setting_up or finishing):
                        release(retired)
                        if clock == "steps":   # idle until the next arrival
                            vclock = max(vclock, int(math.ceil(pending[0].arrival)))
                        else:
                            time.sleep(max(0.0, pending[0].arrival - now()))
            release(retired)
            concurrent.futures.wait(closing)
            for f in closing:
                f.result()
            return results
        except BaseException:
            self._abandon(active, admitted, retired, setting_up, finishing, closing)
            raise

    def _abandon(self, active: dict, admitted: list, retired: list, setting_up: list = (), finishing: list = (),
                 closing: list = ()) -> None:
        """After a failure inside run(): waits for the device and the set-up thread and releases
        every block the run holds, so that the server and its batch can be used again."""
        try:
            self.ctx.sync()
        except Exception:
            pass
        for a in list(active.values()):
            try:
                self.batch.remove(a["block"])
            except Exception:
                pass
        blocks = [a["block"] for a in active.values()] + [a["block"] for a in admitted] + list(retired)
        for a in setting_up:
            try:
                blocks.append(a["fut"].result()[0])
            except Exception:
                pass
        for a in finishing:
            try:
                a["fut"].result()
            except Exception:
                pass
            blocks.append(a["block"])
        concurrent.futures.wait(list(closing))
        for b in blocks:
            try:
                b.close()
            except Exception:
                pass

    def close(self) -> None:
        self.batch.close()
        self.pool.shutdown(wait=True)
        self.setup.shutdown(wait=True)
        if self.io is not self.ctx:
            self.io.close()


def report(requests: list[Request], results: dict[int, Result]) -> dict:
    """p50 / p99 latency (arrival -> result on the host) and aggregate cell-updates/s."""
    lat = np.array([results[r.rid].latency for r in requests])
    cells = float(sum(r.nodes * r.nsteps for r in requests))
    span = max(res.finished for res in results.values()) - min(r.arrival for r in requests)
    return {"requests": len(requests), "nonfinite": sum(1 for r in requests if not results[r.rid].finite),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "makespan_s": span,
            "cell_updates": cells, "cell_updates_per_s": cells / span if span > 0 else float("nan")}


def periodic_stream(n: int, rate: float, seed: int = 20261202, sizes=((128, 64, 16), (192, 96, 24), (256, 128, 32),
                                                                       (384, 128, 32)),
                    nsteps: int = 300) -> list[Request]:
    """Poisson arrivals (rate per second) of periodic TGV / vortex solves with seeded sizes,
    Mach U(0.1, 0.6) and Re U(500, 5000); the same request stream every call for a given seed."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        ni, nj, nk = sizes[rng.integers(len(sizes))]
        tgv = bool(rng.integers(2))
        mach = float(rng.uniform(0.1, 0.6))
        reyl = float(rng.uniform(500.0, 5000.0))
        phases = tuple(float(x) for x in rng.uniform(0, 2 * math.pi, 3))
        if tgv:
            out.append(Request(rid, ni, nj, nk, nsteps, t, grid="sinusoid", length=(2 * math.pi,) * 3, phases=phases,
                               init="tgv", mach=mach, reyl=reyl))
        else:
            out.append(Request(rid, ni, nj, nk, nsteps, t, grid="sinusoid", length=(10.0,) * 3, phases=phases,
                               init="vortex", mach=mach, reyl=reyl, center=(5.0, 5.0)))
    return out


C5_LADDER = ((128, 64, 16), (192, 96, 24), (256, 128, 32), (384, 128, 32))


def airfoil_stream(n: int, rate: float = 20.0, seed: int = 20261202, sizes=C5_LADDER, nsteps: int = 300) -> list[Request]:
    """Poisson arrivals of NACA 4-digit C-grid solves: camber U(0, 0.04) at 0.4 chord, thickness
    U(0.08, 0.18), AoA U(-4, 10) deg, Mach U(0.3, 0.7), Re U(1e5, 1e6), grid sizes from the ladder."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        ni, nj, nk = sizes[rng.integers(len(sizes))]
        naca = (float(rng.uniform(0.0, 0.04)), 0.4, float(rng.uniform(0.08, 0.18)))
        out.append(Request(rid, ni, nj, nk, nsteps, t, grid="cgrid", naca=naca, aoa_deg=float(rng.uniform(-4.0, 10.0)),
                           mach=float(rng.uniform(0.3, 0.7)), reyl=float(rng.uniform(1e5, 1e6))))
    return out
