"""The C++ layer's EnzoInitialInclinedWave against the oracle's restatement of
the reference initializer (src/Enzo/initial/EnzoInitialInclinedWave.cpp),
on the GPU blocks it fills.

* Every field of every block, ghost zones included, equals the oracle's
  initial state bit for bit, for all eight wave types, on the reference
  decks' domain [0,3]x[0,1.5]^2 split over 2x2x1 blocks (the blocks' lower
  corners enter the positions), both wave directions.
* Evolved on the GPU from the product's own initial state, every MHD and
  HD linear wave of the reference's tests (N16 and N32) reproduces its
  golden L1 error at rtol 1e-13 / atol 7e-14, and the left- and
  right-propagating N32 sound waves agree to 6 decimal places.
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


def _product_l1(wave, res, t_stop, mhd, nblocks, positive=True):
    """The reference's linear-wave L1 error computed entirely through the
    product: the C++ initializer, the GPU cycles, the L1 over the gathered
    fields (the oracle grid only lends its gather)."""
    ev = _ev()
    h = ev.host_lib()
    grid = po.linear_wave_grid(wave, res, mhd=mhd, nblocks=nblocks, positive=positive)
    h.enzoe_host_set_fluid_props(grid.cfg.gamma, grid.cfg.density_floor,
                                 grid.cfg.pressure_floor)
    fields = po.MHD_L1_FIELDS if mhd else po.HD_L1_FIELDS
    method = method_for(grid.cfg)
    dg = _product_grid(grid, wave, positive, method)
    ev.synchronize()
    nb = nblocks[0] * nblocks[1] * nblocks[2]

    def pull():
        for i in range(nb):
            arrs = dg.blocks[i].download()
            view = grid.block_arrays(i)
            for k, v in arrs.items():
                if k in view:
                    view[k][...] = v
        return [grid.gather(f) for f in fields]
    init = pull()
    ncycles, _ = dg.evolve(0.0, t_stop)
    final = pull()
    return po.l1_norm(init, final, res=res), ncycles


# the reference's golden L1 errors (input/vlct/run_MHD_linear_wave_test.py:
# 178-188, input/vlct/run_HD_linear_wave_test.py:70-80; the HD decks split
# the grid over 2x2x2 blocks)
GOLDEN = [
    ("fast", 16, 0.5, True, 1.6388526155394664e-07),
    ("fast", 32, 0.5, True, 3.302538226654406e-08),
    ("alfven", 16, 1.0, True, 1.927245356389947e-07),
    ("alfven", 32, 1.0, True, 3.005870212811956e-08),
    ("slow", 16, 2.0, True, 2.2373810027584788e-07),
    ("slow", 32, 2.0, True, 4.43702e-08),
    ("mhd_entropy", 16, 1.0, True, 1.0021263485338544e-07),
    ("mhd_entropy", 32, 1.0, True, 2.9194839706868883e-08),
    ("sound", 16, 1.0, False, 1.3704437791196907e-07),
    ("sound", 32, 1.0, False, 2.3484946798755385e-08),
    ("hd_entropy", 16, 1.0, False, 8.736217559091042e-08),
    ("hd_entropy", 32, 1.0, False, 2.2383944940640457e-08),
    ("hd_transv_entropy_v1", 16, 1.0, False, 7.671004076321944e-08),
    ("hd_transv_entropy_v1", 32, 1.0, False, 1.6725791096617187e-08),
    ("hd_transv_entropy_v2", 16, 1.0, False, 8.640730176761958e-08),
    ("hd_transv_entropy_v2", 32, 1.0, False, 1.6960029746447077e-08),
]


@pytest.mark.parametrize("wave,res,t_stop,mhd,ref", GOLDEN)
def test_linear_wave_golden_through_the_product(wave, res, t_stop, mhd, ref):
    print("golden", wave, res, flush=True)
    l1, ncycles = _product_l1(wave, res, t_stop, mhd, (1, 1, 1) if mhd else (2, 2, 2))
    print("cycles", ncycles, "L1", repr(l1), "ref", ref)
    assert ncycles > 10
    # the reference's tolerance (input/vlct/testing_utils.py:275-292)
    assert np.isclose(l1, ref, rtol=1e-13, atol=7e-14)


def test_sound_waves_left_and_right_have_the_same_l1():
    """identical_l1_error_linwave(1, 'sound', 32, ..., prec=6)
    (input/vlct/run_HD_linear_wave_test.py:82-84): the left- and the
    right-propagating N32 sound waves agree to 6 decimal places."""
    l1_r, _ = _product_l1("sound", 32, 1.0, False, (2, 2, 2), positive=True)
    l1_l, _ = _product_l1("sound", 32, 1.0, False, (2, 2, 2), positive=False)
    print("right", repr(l1_r), "left", repr(l1_l))
    assert ("%.6e" % l1_r) == ("%.6e" % l1_l)


def _tube_grid(axis, n=256, ntrans=4, params=None, method=None):
    """The product's shock tube on the grid of po.rj2a_grid (one block)."""
    ev = _ev()
    grid = po.rj2a_grid(axis, n=n, ntrans=ntrans)
    ev.host_lib().enzoe_host_set_fluid_props(grid.cfg.gamma, grid.cfg.density_floor,
                                             grid.cfg.pressure_floor)
    method = method or method_for(grid.cfg)
    bc = ["periodic"] * 3
    bc[axis] = "outflow"
    dg = ev.DeviceGrid(method, (1, 1, 1), grid.ncells, [tuple(grid.block(0).cw)], bc=tuple(bc))
    params = params or ("setup_name = rj2a; aligned_ax = %s" % "xyz"[axis])
    ev.initial_shock_tube(params, dg.blocks[0], (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    return grid, dg


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_rj2a_initial_state_matches_the_oracle(axis):
    ev = _ev()
    grid, dg = _tube_grid(axis, n=64)
    ev.synchronize()
    got = dg.blocks[0].download()
    ref = grid.block_arrays(0)
    for name, a in got.items():
        if name != "pressure":
            assert np.array_equal(a, ref[name]), (axis, name)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_rj2a_golden_through_the_product_initializer(axis):
    """input/vlct/run_MHD_shock_tube_test.py:82-87 with the reference's table
    (tests/golden/rj2a_shock_tube_t0.2_res256.csv), every axis."""
    import os
    ev = _ev()
    print("rj2a axis", axis, flush=True)
    table = po.load_reference_table(os.path.join(os.path.dirname(__file__), "golden",
                                                 "rj2a_shock_tube_t0.2_res256.csv"))
    grid, dg = _tube_grid(axis)
    dg.evolve(0.0, 0.2)
    arrs = dg.blocks[0].download()
    view = grid.block_arrays(0)
    for k, v in arrs.items():
        if k in view:
            view[k][...] = v
    l1 = po.table_l1(grid, table, axis)
    ref = {0: 0.012523489882320429, 1: 0.012523489882320308, 2: 0.012523489882320315}[axis]
    print("L1", repr(l1), "ref", ref)
    assert np.isclose(l1, ref, rtol=1e-13, atol=7e-14)


def test_sod_flip_and_velocities_closed_form():
    """"sod" on a hydro block, flipped, with axis and transverse velocities:
    the two uniform states of EnzoInitialShockTube.cpp:183-292 split at
    x = 0.5 (cell centres below 0.5 hold the left state)."""
    ev = _ev()
    ev.host_lib().enzoe_host_set_fluid_props(1.4, 1e-30, 1e-30)
    n, g = 16, 3
    m = n + 2 * g
    db = ev.DeviceBlock(m, m, m, g, (1.0 / n,) * 3, mhd=False)
    ev.initial_shock_tube("setup_name = sod; aligned_ax = y; flip_initialize = true; "
                          "axis_velocity = 0.25; transverse_velocity = -0.5",
                          db, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    ev.synchronize()
    f = db.download()
    yc = (np.arange(m) - g + 0.5) / n
    left = (yc < 0.5)[None, :, None] * np.ones((m, m, m), bool)
    # flipped: the left state is sod's right one, vectors reversed
    assert np.all(f["density"][left] == 0.125) and np.all(f["density"][~left] == 1.0)
    # velocity_0 (along y) = flip * (v0 + axis_velocity); velocity_1 (along z)
    assert np.all(f["velocity_y"] == -0.25)
    assert np.all(f["velocity_z"] == 0.5)
    assert np.all(f["velocity_x"] == 0.0)
    e_left = 0.1 / (0.4 * 0.125) + 0.5 * (0.25 ** 2 + 0.5 ** 2)
    e_right = 1.0 / (0.4 * 1.0) + 0.5 * (0.25 ** 2 + 0.5 ** 2)
    np.testing.assert_allclose(f["total_energy"][left], e_left, rtol=1e-15)
    np.testing.assert_allclose(f["total_energy"][~left], e_right, rtol=1e-15)
