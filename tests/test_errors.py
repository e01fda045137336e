"""Error paths of the C ABI and the Python mirror that need no GPU: every refusal is a CFDX_E_*
code with a message (the library never exits), or a ValueError for a malformed host array."""
import numpy as np
import pytest

import cfdx


def desc(ni, nj, nk, periodic, bc=(0,) * 6, precision=cfdx.FP64):
    d = cfdx.BlockDesc(ni, nj, nk, periodic, precision)
    for q in range(6):
        d.bc[q] = bc[q]
    return d


@pytest.mark.parametrize("d, words", [
    (desc(2, 8, 8, 7), "dimension"),
    (desc(8, 8, 8, 9), "3-bit mask"),
    (desc(8, 8, 8, 7, precision=5), "precision"),
    (desc(8, 6, 8, 5, (0, 0, 1, 2, 0, 0)), ">= 7 nodes"),
    (desc(8, 8, 8, 5, (0, 0, 1, 0, 0, 0)), "no condition"),
    (desc(8, 8, 8, 6, (5, 5, 0, 0, 0, 0)), "j- side"),
])
def test_block_bytes_refuses_bad_descriptors(d, words):
    with pytest.raises(cfdx.CfdxError) as e:
        cfdx.Block.bytes_for(d.ni, d.nj, d.nk, d.precision, desc=d)
    assert e.value.code == cfdx.E_INVALID and words in str(e.value)


def test_block_bytes_refuses_blocks_past_the_index_range():
    """Kernels index a block's arrays with 32-bit byte offsets: 2^29 stored nodes or more are
    refused with CFDX_E_CAPACITY; one node row fewer is accepted."""
    big = desc(1024, 1024, 512, 7)                       # 2^29 nodes
    with pytest.raises(cfdx.CfdxError) as e:
        cfdx.Block.bytes_for(big.ni, big.nj, big.nk, big.precision, desc=big)
    assert e.value.code == cfdx.E_CAPACITY and "2^29" in str(e.value)
    ok = desc(1024, 1023, 512, 7)
    assert cfdx.Block.bytes_for(ok.ni, ok.nj, ok.nk, ok.precision, desc=ok) > 1023 * 1024 * 512 * 520
    # a bounded direction counts its ghost layers: 1018 owned + 6 ghosts = 1024 stored
    bnd = desc(1018, 1024, 512, 6, (3, 3, 0, 0, 0, 0))
    with pytest.raises(cfdx.CfdxError):
        cfdx.Block.bytes_for(bnd.ni, bnd.nj, bnd.nk, bnd.precision, desc=bnd)


def test_block_bytes_valid():
    d = desc(10, 9, 8, 4, (3, 2, 1, 1, 0, 0))
    n = 16 * 15 * 8
    assert cfdx.Block.bytes_for(10, 9, 8, cfdx.FP64, desc=d) >= n * (33 * 8 + 32 * 8)
    assert cfdx.Block.bytes_for(10, 9, 8, cfdx.MIXED, desc=d) < cfdx.Block.bytes_for(10, 9, 8, cfdx.FP64, desc=d)


def test_cgrid_refusals():
    with pytest.raises(cfdx.CfdxError):
        cfdx.cgrid_desc(40, 12, 4, n_wake=17)            # fewer than 8 section points
    d, g = cfdx.cgrid_desc(48, 16, 4)
    g.radius = 0.5
    with pytest.raises(cfdx.CfdxError) as e:
        cfdx.cgrid_geometry(d, g)
    assert "parameters" in str(e.value)
    g.radius, g.wall = 20.0, 5.0                          # first spacing larger than the rows allow
    with pytest.raises(cfdx.CfdxError):
        cfdx.cgrid_geometry(d, g)
    box = desc(8, 8, 8, 7)
    with pytest.raises(cfdx.CfdxError):
        cfdx.cgrid_geometry(box, cfdx.cgrid_desc(48, 16, 4)[1])   # not a C-grid descriptor


def test_tile_variant_kind():
    with pytest.raises(cfdx.CfdxError):
        cfdx.tile_variants(3)


def test_restart_errors(tmp_path):
    assert cfdx.restart_latest(str(tmp_path), 10)[0] == 0        # no complete file: slot 0
    with pytest.raises(cfdx.CfdxError):
        cfdx.read_restart(str(tmp_path), 1, 10)
    with pytest.raises(ValueError):
        cfdx.write_restart(str(tmp_path), 1, np.zeros((10, 9)), 0)
    with pytest.raises(cfdx.CfdxError):
        cfdx.write_restart(str(tmp_path / "missing" / "dir"), 1, cfdx.freestream(10, 0.5), 0)
