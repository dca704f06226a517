"""Pins of the oracle's dual-energy formalism ("modern" formulation,
src/Enzo/fluid-props/EnzoDualEnergyConfig.hpp:40-80) on the CPU.

End to end: the reference's dual-energy Sod shock tube (Mach ~10 background
flow along the tube, input/vlct/run_dual_energy_shock_tube_test.py:52-90:
one golden L1 for the x, y, z and reversed-x decks, and an exactly zero
spread of the per-slice L1 norms).

Per function (closed forms on exact binary data):
* the internal-energy flux is the upwind specific internal energy
  p / ((gamma-1) rho) times the density flux, and vi_bar is the flow
  velocity, for supersonic flow either way and for a moving contact
  (HLLD, HLLE, HLLC: EnzoRiemannHLLD.hpp:404-447, EnzoRiemannHLL.hpp:
  322-344, EnzoRiemannHLLC.hpp:159-170);
* the synchronisation of EnzoPhysicsFluidProps::apply_floor_to_energy_and_
  sync (EnzoPhysicsFluidProps.cpp:158-290) takes eint from etot where the
  implied sound speed beats eta times the flow scales, keeps it otherwise,
  applies the floor and rebuilds etot;
* the internal-energy source (EnzoSourceInternalEnergy.cpp:16-92) of a
  uniform pressure and a linear interface velocity is -dt p dv/dx exactly,
  on the reference's region only;
* a uniform flow stays uniform, its internal energy untouched.
"""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden",
                      "sod_shock_tube_t0.25_res128.csv")
DE_L1 = 0.02605861216339738   # run_dual_energy_shock_tube_test.py:59
RTOL, ATOL = 1e-13, 7e-14     # input/vlct/testing_utils.py:275-292 (abs_tol)


def _de_tube_l1(axis, reverse=False):
    grid = po.sod_de_grid(axis, reverse=reverse)
    ncyc, _ = grid.evolve(0.0, 0.25, nthreads=4)
    table = po.load_reference_table(GOLDEN)
    cw = (grid.upper[axis] - grid.lower[axis]) / 508
    bkg = [0.0, 0.0, 0.0]
    bkg[axis] = -po.DE_TUBE_AXIS_VELOCITY if reverse else po.DE_TUBE_AXIS_VELOCITY
    off = -po.DE_TUBE_OFFSET if reverse else po.DE_TUBE_OFFSET
    kw = dict(offset=off, bkg_velocity=bkg, reverse=reverse)
    l1 = po.shifted_table_l1(grid.gather, table, axis, grid.lower[axis], cw, **kw)
    cols = po.shifted_table_l1(grid.gather, table, axis, grid.lower[axis], cw,
                               per_column=True, **kw)
    return l1, cols, ncyc, grid


@pytest.mark.parametrize("axis,reverse", [
    (0, False), pytest.param(1, False, marks=pytest.mark.slow),
    pytest.param(2, False, marks=pytest.mark.slow),
    pytest.param(0, True, marks=pytest.mark.slow)])
def test_dual_energy_shock_tube_golden(axis, reverse):
    l1, cols, ncyc, grid = _de_tube_l1(axis, reverse)
    assert ncyc > 1000
    assert np.isclose(l1, DE_L1, rtol=RTOL, atol=ATOL), (l1, DE_L1)
    # the --std check: every slice along the tube has the same L1 norm
    assert np.std(cols) == 0.0
    # the formalism was active both ways (the deck's eta is chosen so):
    # somewhere eint differs from etot - ekin, somewhere it equals it
    e = grid.gather("internal_energy")
    v2 = sum(np.square(grid.gather(k)) for k in ("velocity_x", "velocity_y",
                                                  "velocity_z"))
    b2 = sum(np.square(grid.gather(k)) for k in ("bfield_x", "bfield_y", "bfield_z"))
    e1 = grid.gather("total_energy") - (0.5 * v2 + 0.5 * b2 / grid.gather("density"))
    assert np.any(e == e1) and np.any(e != e1)


def test_sod_initial_state_closed_form():
    """EnzoInitialShockTube.cpp:183-292 with axis_velocity and the internal
    energy field: left of x = 0.5 (1, 1), right (0.125, 0.1)."""
    grid = po.sod_de_grid(0, n=64, nblocks=2)
    d = grid.gather("density")
    e = grid.gather("internal_energy")
    cw = 3.96875 / 64
    x = (np.arange(64) + 0.5) * cw
    left = (x < 0.5)[None, None, :]
    assert np.all(d[np.broadcast_to(left, d.shape)] == 1.0)
    assert np.all(d[np.broadcast_to(~left, d.shape)] == 0.125)
    np.testing.assert_allclose(e[np.broadcast_to(left, e.shape)], 1.0 / 0.4, rtol=1e-15)
    np.testing.assert_allclose(e[np.broadcast_to(~left, e.shape)], 0.1 / (0.4 * 0.125),
                               rtol=1e-15)
    assert np.all(grid.gather("velocity_x") == 11.875)
    etot = grid.gather("total_energy")
    np.testing.assert_allclose(etot, e + 0.5 * 11.875 ** 2, rtol=1e-15)


# ------------------------------------------------------------ Riemann outputs

def _mhd_state(rho, v, p, b=(0.0, 0.0, 0.0)):
    w = np.zeros(8)
    w[0], w[1], w[2], w[3] = rho, b[0], b[1], b[2]
    w[4], w[5], w[6], w[7] = v[0], v[1], v[2], p
    return w


def _hydro_state(rho, v, p):
    return np.array([rho, v[0], v[1], v[2], p], dtype=np.float64)


def _solve(kind, wl, wr, gamma):
    L = po.lib()
    fo = np.zeros(wl.size)
    de = np.zeros(2)
    fn = {"hlld": L.o_hlld_face_de, "hlle": L.o_hlle_mhd_face_de,
          "hllc": L.o_hllc_face_de}[kind]
    fn(po.dptr(wl), po.dptr(wr), gamma, po.dptr(fo), po.dptr(de))
    return fo, de


@pytest.mark.parametrize("kind", ["hlld", "hlle", "hllc"])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_supersonic_eint_flux_and_interface_velocity(kind, sign):
    # gamma 1.5: eint = p / (0.5 rho); left rho 2, p 3 -> eint 3; right
    # rho 4, p 2 -> eint 1; |v| = 8 is far above the sound speeds (~1.5)
    gamma = 1.5
    v = (8.0 * sign, 0.5, -0.25)
    mk = _hydro_state if kind == "hllc" else _mhd_state
    wl, wr = mk(2.0, v, 3.0), mk(4.0, v, 2.0)
    fo, de = _solve(kind, wl, wr, gamma)
    up_rho, up_eint = (2.0, 3.0) if sign > 0 else (4.0, 1.0)
    assert fo[0] == up_rho * v[0]                 # the upwind density flux
    assert de[0] == up_eint * (up_rho * v[0])     # eint carried with it
    assert de[1] == v[0]                          # the flow velocity


@pytest.mark.parametrize("kind", ["hlld", "hlle", "hllc"])
def test_subsonic_moving_contact(kind):
    """A contact (density jump, same p and v) moving at v: vi_bar = v and
    the internal-energy flux carries the upwind eint (HLLD and HLLC resolve
    the contact exactly; HLLE diffuses the density flux but its vi_bar is
    still v)."""
    gamma = 1.5
    v = (0.375, 0.0, 0.0)
    mk = _hydro_state if kind == "hllc" else _mhd_state
    wl, wr = mk(2.0, v, 3.0), mk(1.0, v, 3.0)
    fo, de = _solve(kind, wl, wr, gamma)
    np.testing.assert_allclose(de[1], v[0], rtol=1e-14)
    eint_l = 3.0 / (0.5 * 2.0)
    np.testing.assert_allclose(de[0], eint_l * fo[0], rtol=1e-15)
    if kind != "hlle":
        np.testing.assert_allclose(fo[0], 2.0 * v[0], rtol=1e-14)


# ------------------------------------------------------------ synchronisation

def _sync(cfg, cells, mhd=False):
    """cells: list of dicts rho, v(3), b(3), etot, eint on a 1x1xN block."""
    n = len(cells)
    U = [np.zeros(n) for _ in range(8)]
    eint = np.zeros(n)
    for i, c in enumerate(cells):
        U[0][i] = c["rho"]
        for a in range(3):
            U[1 + a][i] = c["v"][a]
            U[5 + a][i] = c.get("b", (0.0, 0.0, 0.0))[a]
        U[4][i] = c["etot"]
        eint[i] = c["eint"]
    ptrs = (ctypes.POINTER(ctypes.c_double) * 8)(*[po.dptr(u) for u in U])
    po.lib().o_floor_energy_and_sync(ctypes.byref(cfg), ptrs, po.dptr(eint), 1, 1, n, 0)
    return U[4], eint


def test_sync_rules():
    # gamma 1.5 -> ggm1 = 0.75 (exact in float); floor p 1/1024 -> eint
    # floor = p_floor / (0.5 rho)
    cfg = po.make_config(gamma=1.5, mhd=True, pressure_floor=2.0 ** -10,
                         dual_energy_eta=0.0)
    v = (2.0, 0.0, 0.0)   # ekin = 2
    cells = [
        # eta = 0: eint replaced by etot - ekin whenever that is positive
        dict(rho=1.0, v=v, etot=5.0, eint=7.0),
        # etot - ekin <= 0: eint kept, etot rebuilt from it
        dict(rho=1.0, v=v, etot=1.5, eint=0.25),
        # eint kept below the floor: floored, etot rebuilt
        dict(rho=1.0, v=v, etot=2.0, eint=2.0 ** -20),
    ]
    etot, eint = _sync(cfg, cells)
    assert eint.tolist() == [3.0, 0.25, 2.0 ** -9]
    assert etot.tolist() == [5.0, 2.25, 2.0 + 2.0 ** -9]

    cfg.eta = 0.5
    cells = [
        # cs2_1 = 0.75 * 3 = 2.25 > eta v2 = 2: take etot - ekin
        dict(rho=1.0, v=v, etot=5.0, eint=4.0),
        # cs2_1 = 0.75 * 2 = 1.5 < 2: keep eint, rebuild etot
        dict(rho=1.0, v=v, etot=4.0, eint=2.5),
        # cs2_1 = 0.75 * 3 > 2 but eint_1 = 3 < eint / 2: keep eint
        dict(rho=1.0, v=v, etot=5.0, eint=8.0),
        # magnetic scale: eta b2 / rho = 0.5 * 16 / 1 = 8 > 2.25: keep
        dict(rho=1.0, v=v, b=(4.0, 0.0, 0.0), etot=13.0, eint=2.5),
    ]
    etot, eint = _sync(cfg, cells)
    assert eint.tolist() == [3.0, 2.5, 8.0, 2.5]
    assert etot.tolist() == [5.0, 4.5, 10.0, 12.5]


def test_sync_without_dual_energy_is_the_energy_floor():
    cfg = po.make_config(gamma=1.5, mhd=False, pressure_floor=0.5)
    # eint floor = 0.5 / (0.5 * 2) = 0.5 at rho = 2; ekin = 0.5 v^2 = 2
    cells = [dict(rho=2.0, v=(2.0, 0.0, 0.0), etot=2.25, eint=-1.0),
             dict(rho=2.0, v=(2.0, 0.0, 0.0), etot=3.0, eint=-1.0)]
    etot, eint = _sync(cfg, cells)
    assert etot.tolist() == [2.5, 3.0]
    assert eint.tolist() == [-1.0, -1.0]      # never touched


# ------------------------------------------------------------ source term

@pytest.mark.parametrize("dim", [0, 1, 2])
def test_internal_energy_source_linear_velocity(dim):
    cfg = po.make_config(gamma=1.5, pressure_floor=2.0 ** -30, dual_energy_eta=0.001)
    m = [7, 6, 5]
    mx, my, mz = m
    p = np.full((mz, my, mx), 1.5)
    p[0, 0, 0] = 2.0 ** -40                  # outside the region; floored if read
    fshape = [mz, my, mx]
    fshape[2 - dim] -= 1
    # vbar at face f (between cells f and f+1 along dim) = 0.25 f
    f = np.arange(fshape[2 - dim], dtype=np.float64) * 0.25
    sh = [1, 1, 1]
    sh[2 - dim] = -1
    vbar = np.ascontiguousarray(np.broadcast_to(f.reshape(sh), fshape))
    de = np.full((mz, my, mx), 10.0)
    dt, cw, s = 0.5, 0.25, 1
    po.lib().o_internal_energy_source(ctypes.byref(cfg), dim, dt, cw, po.dptr(p),
                                      po.dptr(de), po.dptr(vbar), mz, my, mx, s)
    # region: [s+1, m-s-1) along dim, [s, m-s) across
    want = np.full((mz, my, mx), 10.0)
    reg = []
    for a in (2, 1, 0):
        lo = s + 1 if a == dim else s
        hi = m[a] - s - 1 if a == dim else m[a] - s
        reg.append(slice(lo, hi))
    want[tuple(reg)] = 10.0 - (dt / cw) * 1.5 * 0.25
    assert np.array_equal(de, want)


def test_uniform_flow_keeps_its_internal_energy():
    """Every face flux of a uniform state is the same, so the divergence
    and the source vanish; with rho = 1 the update is exact."""
    from _states import random_periodic_grid
    grid = random_periodic_grid(mhd=True, seed=2, n=(8, 8, 8))
    grid.cfg.dual_energy = 1
    grid.cfg.eta = 0.001
    grid.cfg.gamma = 1.5
    arr = grid.block_arrays(0)
    arr["density"][...] = 1.0
    for k, val in zip(("velocity_x", "velocity_y", "velocity_z"), (0.5, -0.25, 0.125)):
        arr[k][...] = val
    for k, val in zip(("bfield_x", "bfield_y", "bfield_z"), (0.5, 0.25, -0.5)):
        arr[k][...] = val
    for k, val in zip(("bfieldi_x", "bfieldi_y", "bfieldi_z"), (0.5, 0.25, -0.5)):
        arr[k][...] = val
    arr["internal_energy"][...] = 3.0
    ke = 0.5 * (0.25 + 0.0625 + 0.015625)
    me = 0.5 * (0.25 + 0.0625 + 0.25)
    arr["total_energy"][...] = 3.0 + ke + me
    blk = grid.block(0)
    dt = po.lib().o_vlct_timestep(ctypes.byref(grid.cfg), ctypes.byref(blk))
    assert po.lib().o_vlct_compute(ctypes.byref(grid.cfg), ctypes.byref(blk), dt) == 0
    s = (slice(3, -3),) * 3
    np.testing.assert_allclose(arr["internal_energy"][s], 3.0, rtol=1e-14)
    np.testing.assert_allclose(arr["total_energy"][s], 3.0 + ke + me, rtol=1e-14)
