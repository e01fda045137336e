"""Iteration loop of a block run, the host side of the reference's solver() (Solver.cpp:356-727 and
the master loop of Main.cpp:6080-6192): RK3 steps on the device, the residual monitor line of
every iteration, restart files written in turn every `restart_every` iterations, Tecplot point
files every `tecplot_every` iterations, and resumption from the newer complete restart file.

Iteration numbering: a restart file tagged N holds the state after N completed steps, and a
resumed run continues with step N + 1. (The reference tags the file with the number of the
iteration about to start and resumes one later.)
"""
from __future__ import annotations

import os
import sys

import numpy as np

from . import Block, read_restart, restart_latest, write_restart, write_tecplot


class Run:
    def __init__(self, block: Block, mach: float, dt: float, params, outdir: str | None = None,
                 restart_every: int = 100, tecplot_every: int | None = None, digits: int = 17, xyz=None,
                 monitor: bool = True, log=None):
        self.b = block
        self.mach, self.dt, self.params = mach, dt, params
        self.outdir = outdir
        self.restart_every = restart_every
        self.tecplot_every = tecplot_every
        self.digits = digits
        self.xyz = xyz
        self.monitor = monitor
        self.log = log if log is not None else (lambda s: print(s, file=sys.stderr))
        self.iteration = 0          # completed steps
        self.next_slot = 1
        self.history: list[tuple[int, float, float, float]] = []
        if monitor:
            block.set_monitor(True)

    def resume(self) -> int:
        """Loads the newer complete restart file of outdir, if any; returns the completed steps."""
        if not self.outdir:
            return self.iteration
        slot, it = restart_latest(self.outdir, self.b.n)
        if slot == 0:
            return self.iteration
        flow, it = read_restart(self.outdir, slot, self.b.n)
        self.b.set_flow(flow)
        self.iteration = it
        self.next_slot = 2 if slot == 1 else 1
        self.log(f"resumed from restart_file_{slot}.neu at iteration {it}")
        return it

    def _write_outputs(self) -> None:
        it = self.iteration
        if self.outdir and self.restart_every and it % self.restart_every == 0:
            write_restart(self.outdir, self.next_slot, self.b.get_flow(), it, self.digits)
            self.next_slot = 2 if self.next_slot == 1 else 1
        if self.outdir and self.tecplot_every and it % self.tecplot_every == 0 and self.xyz is not None:
            d = self.b.desc
            write_tecplot(os.path.join(self.outdir, f"nodefile_{it}.dat"), d.ni, d.nj, d.nk, self.xyz, self.b.get_flow(),
                          solution_time=it * self.dt)

    def run(self, iterations: int) -> int:
        """Steps until `iterations` steps are complete; returns the completed steps."""
        if self.outdir:
            os.makedirs(self.outdir, exist_ok=True)
        while self.iteration < iterations:
            self.b.step(1, self.mach, self.dt, self.params)
            self.iteration += 1
            if self.monitor:
                du, dv, de = self.b.residual()
                self.history.append((self.iteration, du, dv, de))
                self.log(f"iteration = {self.iteration} max_div_u = {du:.6g} max_div_v = {dv:.6g} max_div_e = {de:.6g}")
                if not (np.isfinite(du) and np.isfinite(dv) and np.isfinite(de)):
                    raise FloatingPointError(f"non-finite residual at iteration {self.iteration}")
            self._write_outputs()
        return self.iteration
