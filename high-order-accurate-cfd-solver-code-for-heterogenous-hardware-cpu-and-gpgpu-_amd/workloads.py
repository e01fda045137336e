"""The seeded workloads of the benchmark configurations, shared by the bench and the tests.

C1  64^3 isentropic vortex (strength 5, Mach 0.5 along x), Euler, Cartesian box [0, 10)^3 through
    the metric path.
C2  128^3 Taylor-Green vortex on [0, 2 pi)^3 with a seeded sinusoidal map (amplitude 0.05 of the
    spacing), Mach 0.1, Re 1600, the mixed-time-scale LES model. NS is the same flow at 256^3.
C4  64 periodic cubes, edge drawn from {32, 48, 64, 96, 128}, isentropic vortex or Taylor-Green with
    probability 1/2, Mach U(0.1, 0.6), Re U(500, 5000), seeded map phases.
Seeds follow the benchmark definition (20261202).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SEED = 20261202
C4_EDGES = (32, 48, 64, 96, 128)


def c2_phases(seed: int = SEED) -> tuple:
    """Map phases of the C2 / NS grid: the first three draws of the seeded generator."""
    return tuple(float(x) for x in np.random.default_rng(seed).uniform(0, 2 * math.pi, 3))


def tgv_case_kw(n: int, mach: float = 0.1, seed: int = SEED) -> dict:
    """periodic_case / init_case keywords of the C2 flow on an n^3 grid."""
    h = 2 * math.pi / n
    return dict(grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05 * h, phases=c2_phases(seed), init="tgv",
                mach=mach)


def vortex_case_kw(n: int = 64) -> dict:
    """periodic_case / init_case keywords of the C1 flow."""
    return dict(grid="box", length=(10.0,) * 3, init="vortex", mach=0.5)


@dataclasses.dataclass
class CubeSpec:
    """One cube of the C4 batch."""
    edge: int
    tgv: bool
    mach: float
    reyl: float
    phases: tuple

    @property
    def length(self) -> float:
        return 2 * math.pi if self.tgv else 10.0

    def case_kw(self) -> dict:
        """periodic_case / init_case keywords (map amplitude 0.05 of the spacing)."""
        L = self.length
        return dict(grid="sinusoid", length=(L,) * 3, amp=0.05 * L / self.edge, phases=self.phases,
                    init="tgv" if self.tgv else "vortex", mach=self.mach, center=(5.0, 5.0))


def c4_specs(seed: int = SEED, nblocks: int = 64) -> list[CubeSpec]:
    """The seeded cubes of C4 (the same list for a given seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nblocks):
        e = int(rng.choice(C4_EDGES))
        tgv = bool(rng.integers(2))
        out.append(CubeSpec(e, tgv, float(rng.uniform(0.1, 0.6)), float(rng.uniform(500, 5000)),
                            tuple(float(x) for x in rng.uniform(0, 2 * math.pi, 3))))
    return out
