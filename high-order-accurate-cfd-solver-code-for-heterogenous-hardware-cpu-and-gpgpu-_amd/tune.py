"""Tile autotuner for the block kernels.

Every kernel class (sweep along i, along j, along k; the viscous node kernel) has a few compiled
tile shapes (`cfdx.tile_variants`). The tuner times them on the actual grid with HIP events on
the block's stream and keeps the fastest: a coordinate search, one kernel class at a time with
the others fixed at the best found so far (the classes are separate launches, so their times
are nearly independent). For a batch the measure is the whole batched step, and the persistent
grid size is tuned too. Results are cached in a JSON table keyed by grid shape, precision, the
viscous switch and the periodic mask (`CFDX_TILE_CACHE`, default ~/.cache/cfdx/tiles.json).

Tuning runs real RK3 steps: the state of the block or batch advances by them. Tune a scratch
copy (same shape) when the state matters.
"""
from __future__ import annotations

import json
import os

from . import Batch, Block, Ctx, tile_variants

DEFAULT_CACHE = os.environ.get("CFDX_TILE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "cfdx", "tiles.json"))
GRID_CAPS = (0, 1024, 2048, 4096, 8192)
# node-tile orders tried: strips of this many tile rows walked plane by plane (0: whole planes)
STRIPS = (0, 4, 8, 16)


class TileCache:
    """A JSON table key -> tile choice, saved after every update."""

    def __init__(self, path: str | None = DEFAULT_CACHE):
        self.path = path
        self.table: dict = {}
        if path and os.path.exists(path):
            try:
                with open(path) as f:
                    self.table = json.load(f)
            except (OSError, ValueError):
                self.table = {}

    def get(self, key: str):
        return self.table.get(key)

    def put(self, key: str, value) -> None:
        self.table[key] = value
        if self.path:
            try:
                os.makedirs(os.path.dirname(self.path), exist_ok=True)
                tmp = self.path + ".tmp"
                with open(tmp, "w") as f:
                    json.dump(self.table, f, indent=1, sort_keys=True)
                os.replace(tmp, self.path)
            except OSError:
                pass


def tables_tag() -> str:
    """The compiled tile shapes, so that a cached choice is dropped when the library's tables change."""
    shapes = [tile_variants(kind) for kind in range(3)]
    return ("t" + "-".join("_".join(f"{a}x{b}" for a, b in kind) for kind in shapes) + "-s" + "_".join(map(str, STRIPS))
            + f"-tie{TIE}")


def block_key(ni: int, nj: int, nk: int, precision: int, viscous: bool, periodic: int = 7) -> str:
    """Cache key of a block: shape, precision, viscous switch, (bounded blocks) the periodic mask and
    the library's tile tables."""
    tail = "" if periodic == 7 else f":per{periodic}"
    return f"block:{ni}x{nj}x{nk}:p{precision}:v{int(bool(viscous))}{tail}:{tables_tag()}"


def _counts() -> list[int]:
    return [len(tile_variants(0)), len(tile_variants(1)), len(tile_variants(1)), len(tile_variants(2))]


def _options() -> list[tuple]:
    """Per tuned class, the values tried: variant indices of the face sweeps and node kernels, then
    the node-tile strips."""
    return [tuple(range(n)) for n in _counts()] + [STRIPS]


NAMES = ("face_i", "face_j", "face_k", "node", "strip")
# a variant replaces the first listed one only when it is faster by more than this fraction, so
# that near-equal timings give the same choice from run to run
TIE = 0.02


def _pick(opts, times):
    """The first option whose time is within TIE of the fastest."""
    best = min(times)
    return next(o for o, t in zip(opts, times) if t <= best * (1 + TIE))


def tune_block(block: Block, mach: float, dt: float, params, steps: int = 2, cache: TileCache | None = None,
               force: bool = False) -> dict:
    """Chooses and sets the tile variants of one block; returns {"tiles", "times_ms", "cached"}."""
    key = block_key(block.desc.ni, block.desc.nj, block.desc.nk, block.precision, params.viscous, block.desc.periodic)
    if cache is not None and not force:
        hit = cache.get(key)
        if hit:
            block.set_tiles(tuple(hit["tiles"]))
            return dict(hit, cached=True)
    best = list(block.get_tiles())
    times: dict[str, list[float]] = {}
    for cls, opts in enumerate(_options()):
        t_cls = []
        for v in opts:
            cfg = list(best)
            cfg[cls] = v
            block.set_tiles(tuple(cfg))
            block.step(1, mach, dt, params)           # first launch of a variant loads its code
            block.set_timing(True)
            block.step(steps, mach, dt, params)
            ms = block.timing_split()
            block.set_timing(False)
            t_cls.append(ms[1 + cls] if cls < 3 else ms[0] + ms[4])
        times[NAMES[cls]] = t_cls
        best[cls] = _pick(opts, t_cls)
    block.set_tiles(tuple(best))
    out = {"tiles": best, "times_ms": times}
    if cache is not None:
        cache.put(key, out)
    return dict(out, cached=False)


def tune_batch(batch: Batch, ctx: Ctx, key: str | None = None, steps: int = 2, cache: TileCache | None = None,
               graph: bool = True, force: bool = False) -> dict:
    """Chooses and sets the tile variants and the persistent grid size of a batch by the time of
    whole batched steps; returns {"tiles", "grid_cap", "times_ms", "cached"}."""
    if key:
        key = f"{key}:{tables_tag()}"
    if cache is not None and key and not force:
        hit = cache.get(key)
        if hit:
            batch.set_tiles(tuple(hit["tiles"]))
            batch.set_grid(hit["grid_cap"])
            return dict(hit, cached=True)

    def timed(cfg, cap) -> float:
        batch.set_tiles(tuple(cfg))
        batch.set_grid(cap)
        batch.step(1, graph)
        ctx.sync()
        ctx.record(6)
        batch.step(steps, graph)
        ctx.record(7)
        ctx.sync()
        return ctx.elapsed_ms(6, 7)

    best, cap = [0, 0, 0, 0, 0], 0
    times: dict[str, list[float]] = {}
    for cls, opts in enumerate(_options()):
        t_cls = []
        for v in opts:
            cfg = list(best)
            cfg[cls] = v
            t_cls.append(timed(cfg, cap))
        times[NAMES[cls]] = t_cls
        best[cls] = _pick(opts, t_cls)
    t_cap = [timed(best, c) for c in GRID_CAPS]
    times["grid_cap"] = t_cap
    cap = _pick(GRID_CAPS, t_cap)
    batch.set_tiles(tuple(best))
    batch.set_grid(cap)
    out = {"tiles": best, "grid_cap": cap, "times_ms": times}
    if cache is not None and key:
        cache.put(key, out)
    return dict(out, cached=False)
