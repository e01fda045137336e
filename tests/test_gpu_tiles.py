"""Tile variants and the autotuner (GPU).

Results may not depend on the tile shapes: every compiled variant of every kernel class, on a
block whose sizes are not multiples of any tile (partial tiles in all directions), and in a
batch, gives bitwise the state of the default tiles. The tuner picks valid variants and its
cache reproduces the choice.
"""
import math

import numpy as np
import pytest

import cfdx
from cfdx import tune

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p")


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


def ragged(ctx, precision, ni=19, nj=13, nk=11):
    b = cfdx.Block(ni, nj, nk, precision=precision, ctx=ctx)
    b.init_case(grid="sinusoid", length=(2 * math.pi,) * 3, amp=0.05 * 2 * math.pi / ni, phases=(0.4, 1.0, 2.2),
                init="tgv", mach=0.2)
    return b


def all_configs():
    counts = [len(cfdx.tile_variants(0)), len(cfdx.tile_variants(1)), len(cfdx.tile_variants(1)),
              len(cfdx.tile_variants(2))]
    cfgs = []
    for cls, n in enumerate(counts):
        for v in range(1, n):
            c = [0, 0, 0, 0, 0]
            c[cls] = v
            cfgs.append(tuple(c))
    # node-tile strips (rows of tiles per strip, each strip walked plane by plane), with each node
    # variant, including a strip taller than the block
    for v in range(counts[3]):
        for s in (1, 3, 64):
            cfgs.append((0, 0, 0, v, s))
    cfgs.append((counts[0] - 1, counts[1] - 1, counts[2] - 1, counts[3] - 1, 2))
    return cfgs


def test_variant_table():
    fi, fjk, node = cfdx.tile_variants(0), cfdx.tile_variants(1), cfdx.tile_variants(2)
    assert fi[0] == (4, 64) and fjk[0] == (16, 16) and node[0] == (64, 4)
    assert all(a * b == 256 for a, b in fi + fjk + node)


@pytest.mark.parametrize("precision", [cfdx.FP64, cfdx.MIXED], ids=["fp64", "mixed"])
def test_results_independent_of_tiles(ctx, precision):
    prm = cfdx.default_params(mach=0.2, reyl=900.0)
    ref = ragged(ctx, precision)
    dt = ref.cfl_dt(0.5)
    ref.step(2, 0.2, dt, prm)
    want = ref.get_flow()
    for cfg in all_configs():
        b = ragged(ctx, precision)
        b.set_tiles(cfg)
        assert b.get_tiles() == cfg
        b.step(2, 0.2, dt, prm)
        got = b.get_flow()
        for f in FIELDS:
            assert np.array_equal(got[f], want[f]), (cfg, f)
        b.close()


def test_batch_tiles_and_grid(ctx):
    prm = cfdx.default_params(mach=0.2, reyl=900.0)
    shapes = [(19, 13, 11), (8, 9, 10), (17, 5, 7)]
    alone = []
    for s in shapes:
        b = ragged(ctx, cfdx.MIXED, *s)
        dt = b.cfl_dt(0.5)
        b.step(3, 0.2, dt, prm)
        alone.append(b.get_flow())
    for cfg, cap in [((1, 2, 3, 1, 0), 0), ((2, 1, 2, 3, 2), 7), ((0, 3, 1, 2, 5), 64), ((0, 0, 0, 0, 1), 0)]:
        q = cfdx.Batch(cfdx.MIXED, ctx)
        blocks = []
        for s in shapes:
            b = ragged(ctx, cfdx.MIXED, *s)
            b.configure(0.2, b.cfl_dt(0.5), prm)
            q.add(b)
            blocks.append(b)
        q.set_tiles(cfg)
        q.set_grid(cap)
        q.step(3, True)
        for b, want in zip(blocks, alone):
            got = b.get_flow()
            for f in FIELDS:
                assert np.array_equal(got[f], want[f]), (cfg, cap, f)
        q.close()


def test_tuner_and_cache(ctx, tmp_path):
    cache = tune.TileCache(str(tmp_path / "tiles.json"))
    prm = cfdx.default_params(mach=0.2, reyl=900.0)
    b = ragged(ctx, cfdx.MIXED, 32, 24, 16)
    dt = b.cfl_dt(0.5)
    res = tune.tune_block(b, 0.2, dt, prm, steps=1, cache=cache)
    assert not res["cached"] and b.get_tiles() == tuple(res["tiles"])
    assert set(res["times_ms"]) == {"face_i", "face_j", "face_k", "node", "strip"}
    assert all(t > 0 for ts in res["times_ms"].values() for t in ts)
    again = tune.TileCache(str(tmp_path / "tiles.json"))
    b2 = ragged(ctx, cfdx.MIXED, 32, 24, 16)
    res2 = tune.tune_block(b2, 0.2, dt, prm, cache=again)
    assert res2["cached"] and b2.get_tiles() == tuple(res["tiles"])
    q = cfdx.Batch(cfdx.MIXED, ctx)
    b3 = ragged(ctx, cfdx.MIXED, 24, 16, 8)
    b3.configure(0.2, b3.cfl_dt(0.5), prm)
    q.add(b)
    q.add(b3)
    rq = tune.tune_batch(q, ctx, key="test-batch", steps=1, cache=cache)
    assert rq["grid_cap"] in tune.GRID_CAPS and len(rq["tiles"]) == 5 and rq["tiles"][4] in tune.STRIPS
    q.close()


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_overlap_gives_the_same_bits(ctx, graph):
    """Mixed-precision stages with the viscous kernel on the second stream (the default) give
    bitwise the results of the stages run in order, for a block and a captured / eager batch
    with a bounded C-grid block in it, over several steps."""
    prm = cfdx.default_params(mach=0.3, reyl=1200.0)
    d, g = cfdx.cgrid_desc(48, 16, 4, cfdx.MIXED, naca=(0.02, 0.4, 0.12), n_wake=12, aoa_deg=3.0)
    geo = cfdx.cgrid_geometry(d, g)[1]
    out = {}
    for on in (True, False):
        ctx.set_overlap(on)
        b = ragged(ctx, cfdx.MIXED, 33, 17, 12)
        dt = b.cfl_dt(0.5)
        b.step(3, 0.2, dt, prm)
        q = cfdx.Batch(cfdx.MIXED, ctx)
        b1 = ragged(ctx, cfdx.MIXED, 24, 16, 8)
        b1.configure(0.2, b1.cfl_dt(0.5), prm)
        c = cfdx.Block(48, 16, 4, precision=cfdx.MIXED, ctx=ctx, desc=d)
        c.set_geometry_soa(geo)
        # Re 1200 on this C-grid (wall spacing 1e-4 chord) at the full stable step: the viscous bound of
        # cfdx_block_cfl_dt keeps the impulsive start finite
        c.configure(0.3, 0.0, prm)
        c.set_flow(cfdx.freestream(48 * 16 * 4, 0.3, 3.0))
        c.configure(0.3, c.cfl_dt(0.5), prm)
        q.add(b1)
        q.add(c)
        q.step(4, graph)
        ctx.sync()
        out[on] = (b.get_flow(), b1.get_flow(), c.get_flow())
        q.close()
        for x in (b, b1, c):
            x.close()
    ctx.set_overlap(True)
    for a, r in zip(out[True], out[False]):
        for f in ("rho", "u", "v", "w", "p"):
            assert np.array_equal(a[f], r[f]), f
            assert np.isfinite(a[f]).all()


def test_overlap_switch_recaptures(ctx):
    """Switching the overlap after a batch captured its step: the next step captures again (the graph
    holds the fork / join of the old setting); the state stays bitwise that of the block alone."""
    prm = cfdx.default_params(mach=0.2, reyl=900.0)
    alone = ragged(ctx, cfdx.MIXED, 24, 16, 8)
    dt = alone.cfl_dt(0.5)
    alone.configure(0.2, dt, prm)
    alone.step(6, 0.2, dt, prm)
    want = alone.get_flow()
    b = ragged(ctx, cfdx.MIXED, 24, 16, 8)
    b.configure(0.2, dt, prm)
    q = cfdx.Batch(cfdx.MIXED, ctx)
    q.add(b)
    try:
        ctx.set_overlap(True)
        q.step(2, True)
        ctx.set_overlap(False)
        q.step(2, True)
        ctx.set_overlap(True)
        q.step(2, True)
        ctx.sync()
    finally:
        ctx.set_overlap(True)
    got = b.get_flow()
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f
    q.close()
    b.close()
    alone.close()


def test_remove_then_destroy_across_contexts(ctx):
    """Blocks on a second context in a batch of the first: a block removed from the batch and
    destroyed at once (its memory back in the pool while the batch's launches may still run), then
    a new block created and added, gives every block bitwise its result alone."""
    io = cfdx.Ctx(0)
    prm = cfdx.default_params(mach=0.2, reyl=900.0)

    def make(shape):
        b = ragged(io, cfdx.MIXED, *shape)
        b.configure(0.2, 0.0, prm)
        b.configure(0.2, b.cfl_dt(0.5), prm)
        return b

    shapes = [(24, 16, 8), (19, 13, 11), (20, 16, 8)]
    alone = []
    for sh, n in zip(shapes, (6, 3, 3)):
        b = make(sh)
        b.step(n, 0.2, b.dt, prm)
        alone.append(b.get_flow())
        b.close()
    q = cfdx.Batch(cfdx.MIXED, ctx)
    a, b = make(shapes[0]), make(shapes[1])
    q.add(a)
    q.add(b)
    q.step(3, True)               # asynchronous on the batch's stream
    q.remove(b)
    got_b = b.get_flow()          # waits for the batch's launches over b
    b.close()
    c = make(shapes[2])           # may take b's memory from the pool
    q.add(c)
    q.step(3, True)
    ctx.sync()
    for x, want in ((a, alone[0]), (c, alone[2])):
        got = x.get_flow()
        for f in FIELDS:
            assert np.array_equal(got[f], want[f]), f
    for f in FIELDS:
        assert np.array_equal(got_b[f], alone[1][f]), f
    q.close()
    a.close()
    c.close()
    io.close()


def test_destroyed_block_leaves_its_batches(ctx):
    """A block destroyed while a batch (captured graph) holds it leaves the batch: the batch then
    steps the remaining block, which gets its result alone."""
    prm = cfdx.default_params(mach=0.2, reyl=900.0)
    q = cfdx.Batch(cfdx.MIXED, ctx)
    a, b = ragged(ctx, cfdx.MIXED, 24, 16, 8), ragged(ctx, cfdx.MIXED, 19, 13, 11)
    for x in (a, b):
        x.configure(0.2, x.cfl_dt(0.5), prm)
        q.add(x)
    q.step(2, True)
    ctx.sync()
    b.close()                      # not removed first
    assert len(q) == 1
    q.step(2, True)
    ctx.sync()
    got = a.get_flow()
    q.close()
    alone = ragged(ctx, cfdx.MIXED, 24, 16, 8)
    dt = alone.cfl_dt(0.5)        # the step the batch's block took (computed before its configure)
    alone.configure(0.2, dt, prm)
    alone.step(4, 0.2, dt, prm)
    want = alone.get_flow()
    for f in ("rho", "u", "v", "w", "p"):
        assert np.array_equal(got[f], want[f]), f
    a.close()
    alone.close()


def test_strip_out_of_range_refused(ctx):
    b = ragged(ctx, cfdx.FP64)
    with pytest.raises(cfdx.CfdxError) as e:
        b.set_tiles((0, 0, 0, 0, 5000))
    assert e.value.code == cfdx.E_INVALID and "strip" in str(e.value)
    with pytest.raises(cfdx.CfdxError):
        b.set_tiles((0, 0, 0, 0, -1))
    assert b.get_tiles() == (0, 0, 0, 0, 4)   # the default: strips of 4 tile rows
    b.close()
