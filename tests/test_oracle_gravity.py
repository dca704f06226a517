"""Pins of the oracle's gravity source (EnzoSourceGravity::calculate_source,
src/Enzo/hydro-mhd/toolkit/EnzoSourceGravity.cpp:17-69, applied on the stage
with index 1 by EnzoMHDIntegratorStageCommands.cpp:265-295), on the CPU.

* A uniform state at rest under a uniform acceleration a gets v = dt a
  after one step, with the specific total energy unchanged (the work term
  uses the start-of-step velocity, which is zero).
* On a random periodic MHD block the source is additive: the conserved
  momentum and energy with and without the acceleration differ by exactly
  dt rho0 a and dt rho0 (v0 . a) (the full-step fluxes come from the
  half-step state, which has no gravity), to roundoff.
* A one-stage ("euler") integration never adds it (the reference adds it on
  the stage with index 1 only).
"""
import ctypes

import numpy as np

import pyoracle as po
from _states import random_periodic_grid


def _accel(shape, rng=None, const=None):
    if const is not None:
        return [np.full(shape, c) for c in const]
    return [np.ascontiguousarray(rng.normal(size=shape)) for _ in range(3)]


def _run(grid, dt, acc):
    ptrs = (ctypes.POINTER(ctypes.c_double) * 3)(*[po.dptr(a) for a in acc])
    assert po.lib().o_vlct_compute_accel(ctypes.byref(grid.cfg), ctypes.byref(grid.block(0)),
                                         dt, ptrs) == 0


def test_uniform_state_at_rest_accelerates_as_dt_a():
    grid = random_periodic_grid(mhd=False, seed=1, n=(8, 8, 8))
    arr = grid.block_arrays(0)
    for k in ("velocity_x", "velocity_y", "velocity_z"):
        arr[k][...] = 0.0
    arr["density"][...] = 1.25
    arr["total_energy"][...] = 1.5
    a = (0.5, -0.25, 0.125)
    acc = _accel(arr["density"].shape, const=a)
    dt = 0.01
    _run(grid, dt, acc)
    s = (slice(3, -3),) * 3
    for i, k in enumerate(("velocity_x", "velocity_y", "velocity_z")):
        np.testing.assert_allclose(arr[k][s], dt * a[i], rtol=1e-15)
    np.testing.assert_allclose(arr["density"][s], 1.25, rtol=0)
    np.testing.assert_allclose(arr["total_energy"][s], 1.5, rtol=1e-15)


def test_source_is_additive_on_a_random_mhd_block():
    rng = np.random.default_rng(7)
    g0 = random_periodic_grid(mhd=True, seed=3)
    g1 = random_periodic_grid(mhd=True, seed=3)
    a0 = g0.block_arrays(0)
    rho0 = a0["density"].copy()
    v0 = [a0[k].copy() for k in ("velocity_x", "velocity_y", "velocity_z")]
    acc = _accel(rho0.shape, rng=rng)
    dt = 0.5 * po.lib().o_vlct_timestep(ctypes.byref(g0.cfg), ctypes.byref(g0.block(0)))
    po.lib().o_vlct_timestep(ctypes.byref(g1.cfg), ctypes.byref(g1.block(0)))
    assert po.lib().o_vlct_compute(ctypes.byref(g0.cfg), ctypes.byref(g0.block(0)), dt) == 0
    _run(g1, dt, acc)
    a1 = g1.block_arrays(0)
    s = (slice(3, -3),) * 3
    assert np.array_equal(a0["density"][s], a1["density"][s])
    for i, k in enumerate(("velocity_x", "velocity_y", "velocity_z")):
        m0 = a0["density"][s] * a0[k][s]
        m1 = a1["density"][s] * a1[k][s]
        expect = dt * rho0[s] * acc[i][s]
        scale = np.abs(m0).max() + np.abs(expect).max()
        assert np.abs((m1 - m0) - expect).max() <= 1e-13 * scale, k
        assert np.abs(expect).max() > 1e-4 * scale
    e0 = a0["density"][s] * a0["total_energy"][s]
    e1 = a1["density"][s] * a1["total_energy"][s]
    work = dt * rho0[s] * sum(v0[i][s] * acc[i][s] for i in range(3))
    assert np.abs((e1 - e0) - work).max() <= 1e-13 * np.abs(e0).max()
    # cell-centred and face B do not see gravity
    for k in ("bfield_x", "bfield_y", "bfield_z", "bfieldi_x"):
        assert np.array_equal(a0[k], a1[k])


def test_one_stage_integration_adds_no_gravity():
    g0 = random_periodic_grid(mhd=False, seed=5, recon=("plm",))
    g1 = random_periodic_grid(mhd=False, seed=5, recon=("plm",))
    for g in (g0, g1):
        g.cfg.nstages = 1
    acc = _accel(g0.block_arrays(0)["density"].shape, const=(1.0, 2.0, 3.0))
    assert po.lib().o_vlct_compute(ctypes.byref(g0.cfg), ctypes.byref(g0.block(0)), 0.01) == 0
    _run(g1, 0.01, acc)
    for k, v in g0.block_arrays(0).items():
        assert np.array_equal(v, g1.block_arrays(0)[k]), k
