"""The C++ layer's EnzoInitialInclinedWave against the oracle's restatement of
the reference initializer (src/Enzo/initial/EnzoInitialInclinedWave.cpp),
on the GPU blocks it fills.

* Every field of every block, ghost zones included, equals the oracle's
  initial state bit for bit, for all eight wave types, on the reference
  decks' domain [0,3]x[0,1.5]^2 split over 2x2x1 blocks (the blocks' lower
  corners enter the positions), both wave directions.
* Evolved on the GPU from the product's own initial state, the N16 fast
  wave reproduces the reference's golden L1 error
  (input/vlct/run_MHD_linear_wave_test.py:178-188) at rtol 1e-13.
"""
import numpy as np
import pytest

import pyoracle as po
from test_gpu_parity import method_for

pytestmark = pytest.mark.gpu

PARAMS = "alpha = %r; beta = %r; amplitude = 1e-6; lambda = 1.0; positive_vel = %s; wave_type = %s"


def _ev():
    import enzoe_vlct as ev
    return ev


def _product_grid(grid, wave, positive, method):
    """DeviceGrid of the oracle grid's shape, filled by the C++ initializer."""
    ev = _ev()
    nb = grid.nblocks
    cws = [tuple(grid.block(i).cw) for i in range(nb[0] * nb[1] * nb[2])]
    dg = ev.DeviceGrid(method, nb, grid.ncells, cws, bc=("periodic",) * 3)
    params = PARAMS % (po.WAVE_ALPHA, po.WAVE_BETA, "true" if positive else "false", wave)
    for bz in range(nb[2]):
        for by in range(nb[1]):
            for bx in range(nb[0]):
                i = (bz * nb[1] + by) * nb[0] + bx
                # Block::lower (src/Cello/mesh_Block.cpp:832-880) on [0,3]x[0,1.5]^2
                ib, upper = (bx, by, bz), (3.0, 1.5, 1.5)
                lower = tuple((1.0 - 1.0 * ib[a] / nb[a]) * 0.0 + 1.0 * ib[a] / nb[a] * upper[a]
                              for a in range(3))
                ev.initial_inclined_wave(params, dg.blocks[i], lower)
    return dg


@pytest.mark.parametrize("wave,mhd", [
    ("fast", True), ("alfven", True), ("slow", True), ("mhd_entropy", True),
    ("sound", False), ("hd_entropy", False), ("hd_transv_entropy_v1", False),
    ("hd_transv_entropy_v2", False)])
@pytest.mark.parametrize("positive", [True, False])
def test_inclined_wave_initial_state_matches_the_oracle(wave, mhd, positive):
    ev = _ev()
    h = ev.host_lib()
    grid = po.linear_wave_grid(wave, 8, mhd=mhd, nblocks=(2, 2, 1), positive=positive)
    h.enzoe_host_set_fluid_props(grid.cfg.gamma, grid.cfg.density_floor,
                                 grid.cfg.pressure_floor)
    method = method_for(grid.cfg)
    dg = _product_grid(grid, wave, positive, method)
    ev.synchronize()
    for i in range(4):
        got = dg.blocks[i].download()
        ref = grid.block_arrays(i)
        for name, a in got.items():
            if name == "pressure":
                continue
            assert np.array_equal(a, ref[name]), (wave, positive, i, name)


def test_fast_wave_golden_through_the_product_initializer():
    ev = _ev()
    h = ev.host_lib()
    grid = po.linear_wave_grid("fast", 16)
    h.enzoe_host_set_fluid_props(grid.cfg.gamma, grid.cfg.density_floor,
                                 grid.cfg.pressure_floor)
    fields = po.MHD_L1_FIELDS
    init = [grid.gather(f) for f in fields]
    method = method_for(grid.cfg)
    dg = _product_grid(grid, "fast", True, method)
    ev.synchronize()
    dg.evolve(0.0, 0.5)
    arrs = dg.blocks[0].download()
    view = grid.block_arrays(0)
    for k, v in arrs.items():
        if k in view:
            view[k][...] = v
    final = [grid.gather(f) for f in fields]
    l1 = po.l1_norm(init, final, res=16)
    print("L1", repr(l1))
    assert np.isclose(l1, 1.6388526155394664e-07, rtol=1e-13, atol=7e-14)
