"""Tile configuration on the host (CPU): the tuner's choice rule, the cache key, the strip field of
cfdx_tiles through the ctypes binding, and the library's refusal of a strip outside its range."""
import ctypes as C

import pytest

import cfdx
from cfdx import tune


def test_pick_keeps_the_first_option_within_the_tie_band():
    assert tune._pick((0, 1, 2), [5.23, 5.77, 5.33]) == 0          # 2 % band: the first stays
    assert tune._pick((0, 1, 2), [5.40, 5.00, 5.30]) == 1          # 8 % faster: replaced
    assert tune._pick((0, 4, 8, 16), [14.6, 14.0, 13.5, 13.8]) == 8
    assert tune._pick(tune.GRID_CAPS, [1.0, 0.99, 0.985, 1.2, 1.3]) == 0


def test_cache_key_names_the_tables_strips_and_band():
    tag = tune.tables_tag()
    assert all(f"{a}x{b}" in tag for a, b in cfdx.tile_variants(0) + cfdx.tile_variants(1) + cfdx.tile_variants(2))
    assert "-s" + "_".join(map(str, tune.STRIPS)) in tag and f"tie{tune.TIE}" in tag
    assert tune.block_key(32, 24, 16, cfdx.MIXED, True).endswith(tag)


def test_tiles_strip_round_trip():
    t = cfdx.Tiles.of((1, 2, 3, 1, 8))
    assert t.as_tuple() == (1, 2, 3, 1, 8)
    assert cfdx.Tiles.of((1, 2, 3, 1)).as_tuple() == (1, 2, 3, 1, 0)   # four entries: whole planes
    assert C.sizeof(cfdx.Tiles) == 5 * 4


def test_tiles_refused_without_a_block():
    # the C ABI checks its arguments before touching a device: a null block is refused
    rc = cfdx._lib.cfdx_block_set_tiles(None, C.byref(cfdx.Tiles.of((0, 0, 0, 0, 8))))
    assert rc == cfdx.E_INVALID
