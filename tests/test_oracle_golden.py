"""End-to-end pins of the CPU oracle against the reference's own golden
values: the inclined linear-wave L1 errors
(input/vlct/run_MHD_linear_wave_test.py:178-188,
input/vlct/run_HD_linear_wave_test.py:70-80) and the RJ2a shock-tube L1 error
against the committed table (input/vlct/run_MHD_shock_tube_test.py:82-87).
Tolerance as in input/vlct/testing_utils.py:275-292 (rtol 1e-13, atol 7e-14).
"""
import os

import numpy as np
import pytest

import pyoracle as po

RTOL, ATOL = 1e-13, 7e-14
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

MHD_CASES = [
    ("fast", 16, 0.5, 1.6388526155394664e-07),
    ("alfven", 16, 1.0, 1.927245356389947e-07),
    ("slow", 16, 2.0, 2.2373810027584788e-07),
    ("mhd_entropy", 16, 1.0, 1.0021263485338544e-07),
    ("fast", 32, 0.5, 3.302538226654406e-08),
    pytest.param("alfven", 32, 1.0, 3.005870212811956e-08, marks=pytest.mark.slow),
    pytest.param("slow", 32, 2.0, 4.43702e-08, marks=pytest.mark.slow),
    pytest.param("mhd_entropy", 32, 1.0, 2.9194839706868883e-08,
                 marks=pytest.mark.slow),
]

HD_CASES = [
    ("sound", 16, 1.3704437791196907e-07),
    ("hd_entropy", 16, 8.736217559091042e-08),
    ("hd_transv_entropy_v1", 16, 7.671004076321944e-08),
    ("hd_transv_entropy_v2", 16, 8.640730176761958e-08),
    pytest.param("sound", 32, 2.3484946798755385e-08, marks=pytest.mark.slow),
    pytest.param("hd_entropy", 32, 2.2383944940640457e-08, marks=pytest.mark.slow),
    pytest.param("hd_transv_entropy_v1", 32, 1.6725791096617187e-08,
                 marks=pytest.mark.slow),
    pytest.param("hd_transv_entropy_v2", 32, 1.6960029746447077e-08,
                 marks=pytest.mark.slow),
]

RJ2A = {0: 0.012523489882320429, 1: 0.012523489882320308,
        2: 0.012523489882320315}


@pytest.mark.parametrize("wave,res,t_stop,ref", MHD_CASES)
def test_mhd_linear_wave_l1(wave, res, t_stop, ref):
    l1, ncyc = po.run_linear_wave(wave, res, t_stop, mhd=True)
    assert ncyc > 10
    assert np.isclose(l1, ref, rtol=RTOL, atol=ATOL), (l1, ref)


@pytest.mark.parametrize("wave,res,ref", HD_CASES)
def test_hd_linear_wave_l1(wave, res, ref):
    # the HD decks split the grid over 2x2x2 blocks
    l1, ncyc = po.run_linear_wave(wave, res, 1.0, mhd=False, nblocks=(2, 2, 2),
                                  nthreads=4)
    assert ncyc > 10
    assert np.isclose(l1, ref, rtol=RTOL, atol=ATOL), (l1, ref)


@pytest.mark.parametrize("axis", [0, pytest.param(1, marks=pytest.mark.slow),
                                  pytest.param(2, marks=pytest.mark.slow)])
def test_rj2a_shock_tube_l1(axis):
    table = po.load_reference_table(
        os.path.join(GOLDEN, "rj2a_shock_tube_t0.2_res256.csv"))
    grid = po.rj2a_grid(axis)
    grid.evolve(0.0, 0.2)
    l1 = po.table_l1(grid, table, axis)
    assert np.isclose(l1, RJ2A[axis], rtol=RTOL, atol=ATOL), (l1, RJ2A[axis])
    # the reference checks the per-slice L1 spread is exactly 0 for one block
    # (input/vlct/run_MHD_shock_tube_test.py:83-87): the transverse slices of
    # a 1D problem must be identical
    rho = grid.gather("density")
    ax_np = 2 - axis
    moved = np.moveaxis(rho, ax_np, 0).reshape(rho.shape[ax_np], -1)
    assert np.all(moved == moved[:, :1])
