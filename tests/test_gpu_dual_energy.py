"""The dual-energy formalism ("modern" formulation) of the fused gfx950 step
against the oracle and the reference's golden values.

* Random periodic MHD (HLLD, HLLE; PLM, PLM-Athena, NN) and hydro (HLLC)
  blocks with an internal energy that differs from etot - ekin, and an eta
  large enough that some cells keep it while others take it from the total
  energy: three Cello cycles (CFL dt with its synchronisation, refresh of
  every field including internal_energy, step) equal the oracle bit for
  bit, every field, ghost zones included.
* The CFL reduction synchronises the energies in place and returns the
  oracle's dt exactly.
* A device-driven graph of cycles ends in the same state as the eager
  cycles, bit for bit, in the active region (its zero-length cycles past
  t_stop refresh the ghost zones and leave the energies alone).
* The product's initializers write internal_energy as the oracle does.
* The dual-energy Sod shock tube decks (x, y, z and the reversed x deck)
  through the product initializer and the GPU step reproduce the reference's
  golden L1 (input/vlct/run_dual_energy_shock_tube_test.py:52-90, rtol
  1e-13 / atol 7e-14), every slice along the tube has the same L1 norm (the
  reference's --std check), and the final state equals the oracle's.
"""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as po
from _states import random_periodic_grid
from test_gpu_parity import device_copy, method_for

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden",
                      "sod_shock_tube_t0.25_res128.csv")
DE_L1 = 0.02605861216339738


def _ev():
    import enzoe_vlct as ev
    return ev


def _de_grid(case, seed=5, eta=2.0):
    grid = random_periodic_grid(mhd=case["mhd"], seed=seed, n=(14, 12, 10),
                                riemann=case["riemann"], recon=case["recon"],
                                theta=case.get("theta", 2.0), vamp=0.8, bamp=0.08)
    grid.cfg.dual_energy = 1
    grid.cfg.eta = eta
    arr = grid.block_arrays(0)
    rng = np.random.default_rng(seed + 100)
    v2 = sum(np.square(arr[k]) for k in ("velocity_x", "velocity_y", "velocity_z"))
    b2 = 0.0
    if case["mhd"]:
        b2 = sum(np.square(arr[k]) for k in ("bfield_x", "bfield_y", "bfield_z"))
    eint = arr["total_energy"] - 0.5 * v2 - 0.5 * b2 / arr["density"]
    arr["internal_energy"][...] = eint * (1.0 + 0.05 * rng.standard_normal(eint.shape))
    return grid


CASES = [
    dict(mhd=True, riemann="hlld", recon=("nn", "plm")),
    dict(mhd=True, riemann="hlld", recon=("nn", "plm_athena"), theta=1.5),
    dict(mhd=True, riemann="hlld", recon=("nn", "nn")),
    dict(mhd=True, riemann="hlle", recon=("nn", "plm")),
    dict(mhd=False, riemann="hllc", recon=("nn", "plm")),
    dict(mhd=False, riemann="hllc", recon=("nn", "plm_athena")),
]


def _assert_equal_grids(dg, grid):
    for i in range(len(dg.blocks)):
        got = dg.blocks[i].download()
        ref = grid.block_arrays(i)
        for name, a in got.items():
            if name in ref:
                assert np.array_equal(a, ref[name]), (i, name)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1]))
def test_dual_energy_cycles_equal_the_oracle(case):
    print("case", case, flush=True)
    grid = _de_grid(case)
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    t = 0.0
    for _ in range(3):
        dt_o = grid.cycle(t, 10.0)
        dt_g = dg.cycle(t, 10.0)
        assert dt_g == dt_o
        t += dt_o
    _ev().synchronize()
    _assert_equal_grids(dg, grid)
    # both branches of the synchronisation were taken
    a = grid.block_arrays(0)
    v2 = sum(np.square(a[k]) for k in ("velocity_x", "velocity_y", "velocity_z"))
    b2 = sum(np.square(a[k]) for k in ("bfield_x", "bfield_y", "bfield_z")) if case["mhd"] else 0
    e1 = a["total_energy"] - (0.5 * v2 + 0.5 * b2 / a["density"])
    s = (slice(3, -3),) * 3
    assert np.any(a["internal_energy"][s] == e1[s])
    assert np.any(a["internal_energy"][s] != e1[s])


def test_dual_energy_timestep_synchronises_in_place():
    case = CASES[0]
    grid = _de_grid(case, seed=9)
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    dt_g = method.timestep(dg.blocks)
    dt_o = po.lib().o_vlct_timestep(ctypes.byref(grid.cfg), ctypes.byref(grid.block(0)))
    assert dt_g == dt_o
    _ev().synchronize()
    _assert_equal_grids(dg, grid)


def _active(a, g):
    """The active region (a face-centred array's extra face included)."""
    return a[tuple(slice(g, a.shape[i] - g) for i in range(3))]


def test_dual_energy_graph_equals_eager_cycles():
    """Device-driven cycles past t_stop are zero-length: they refresh the
    ghost zones but leave the active cells -- internal and total energy
    included -- exactly as the eager cycles end."""
    ev = _ev()
    case = CASES[0]
    g1, g2 = _de_grid(case, seed=4), _de_grid(case, seed=4)
    m = method_for(g1.cfg)
    d1, d2 = device_copy(g1, m), device_copy(g2, m)
    n, _ = d1.evolve(0.0, 0.05)
    k, t = d2.evolve_graph(0.0, 0.05, ncycles=4)
    assert t == 0.05 and n > 4 and 4 * k > n
    ev.synchronize()
    for b1, b2 in zip(d1.blocks, d2.blocks):
        x, y = b1.download(), b2.download()
        for name in x:
            if name == "pressure":
                continue
            assert np.array_equal(_active(x[name], 3), _active(y[name], 3)), name


def _tube(axis, reverse=False, n=508, ntrans=4, nblocks=4):
    """The product's grid of the dual-energy tube deck, filled by the C++
    EnzoInitialShockTube; the oracle grid of the same deck beside it."""
    ev = _ev()
    grid = po.sod_de_grid(axis, reverse=reverse, n=n, ntrans=ntrans, nblocks=nblocks)
    ev.host_lib().enzoe_host_set_fluid_props(1.4, 1e-200, 1e-200)
    method = method_for(grid.cfg)
    bc = ["periodic"] * 3
    bc[axis] = "outflow"
    nb = grid.nblocks
    cws = [tuple(grid.block(i).cw) for i in range(nblocks)]
    dg = ev.DeviceGrid(method, nb, grid.ncells, cws, bc=tuple(bc))
    params = ("setup_name = sod; aligned_ax = %s; axis_velocity = 11.875; "
              "flip_initialize = %s" % ("xyz"[axis], "true" if reverse else "false"))
    for i in range(nblocks):
        ib = [0, 0, 0]
        ib[axis] = i
        # Block::lower / Block::upper (src/Cello/mesh_Block.cpp:832-880)
        lo = tuple((1.0 - ib[a] / nb[a]) * grid.lower[a] + ib[a] / nb[a] * grid.upper[a]
                   for a in range(3))
        hi = tuple((1.0 - (ib[a] + 1) / nb[a]) * grid.lower[a] +
                   (ib[a] + 1) / nb[a] * grid.upper[a] for a in range(3))
        ev.initial_shock_tube(params, dg.blocks[i], lo, hi)
    ev.synchronize()
    return grid, dg


def test_dual_energy_tube_initial_state_equals_the_oracle():
    for axis, reverse in ((0, False), (1, False), (0, True)):
        grid, dg = _tube(axis, reverse, n=64, nblocks=2)
        for i in range(2):
            got = dg.blocks[i].download()
            ref = grid.block_arrays(i)
            for name, a in got.items():
                if name != "pressure":
                    assert np.array_equal(a, ref[name]), (axis, reverse, i, name)


def test_inclined_wave_internal_energy_equals_the_oracle():
    ev = _ev()
    h = ev.host_lib()
    for wave, mhd in (("fast", True), ("sound", False)):
        grid = po.linear_wave_grid(wave, 8, mhd=mhd, nblocks=(2, 1, 1))
        grid.cfg.dual_energy = 1
        grid.cfg.eta = 0.001
        assert po.lib().o_init_inclined_wave(
            ctypes.byref(grid.cfg), ctypes.byref(grid.g), wave.encode(), po.WAVE_ALPHA,
            po.WAVE_BETA, 1e-6, 1.0, 1) == 0
        h.enzoe_host_set_fluid_props(grid.cfg.gamma, grid.cfg.density_floor,
                                     grid.cfg.pressure_floor)
        method = method_for(grid.cfg)
        cws = [tuple(grid.block(i).cw) for i in range(2)]
        dg = ev.DeviceGrid(method, (2, 1, 1), grid.ncells, cws)
        params = ("alpha = %r; beta = %r; amplitude = 1e-6; lambda = 1.0; "
                  "positive_vel = true; wave_type = %s" % (po.WAVE_ALPHA, po.WAVE_BETA, wave))
        for i in range(2):
            lower = (1.5 * i, 0.0, 0.0)
            ev.initial_inclined_wave(params, dg.blocks[i], lower)
        ev.synchronize()
        for i in range(2):
            got = dg.blocks[i].download(["internal_energy"])["internal_energy"]
            assert np.array_equal(got, grid.block_arrays(i)["internal_energy"]), (wave, i)


def test_dual_energy_tube_equals_the_oracle_after_200_cycles():
    grid, dg = _tube(0)
    t = 0.0
    for _ in range(200):
        dt_o = grid.cycle(t, 0.25, nthreads=4)
        dt_g = dg.cycle(t, 0.25)
        assert dt_g == dt_o
        t += dt_o
    _ev().synchronize()
    _assert_equal_grids(dg, grid)


@pytest.mark.parametrize("axis,reverse", [(0, False), (1, False), (2, False), (0, True)])
def test_dual_energy_tube_golden_through_the_product(axis, reverse):
    print("de tube axis", axis, "reverse", reverse, flush=True)
    grid, dg = _tube(axis, reverse)
    ncyc, _ = dg.evolve(0.0, 0.25)
    assert ncyc > 1000
    for i in range(len(dg.blocks)):
        arrs = dg.blocks[i].download()
        view = grid.block_arrays(i)
        for k, v in arrs.items():
            if k in view:
                view[k][...] = v
    table = po.load_reference_table(GOLDEN)
    cw = (grid.upper[axis] - grid.lower[axis]) / 508
    bkg = [0.0, 0.0, 0.0]
    bkg[axis] = -po.DE_TUBE_AXIS_VELOCITY if reverse else po.DE_TUBE_AXIS_VELOCITY
    kw = dict(offset=-po.DE_TUBE_OFFSET if reverse else po.DE_TUBE_OFFSET,
              bkg_velocity=bkg, reverse=reverse)
    l1 = po.shifted_table_l1(grid.gather, table, axis, grid.lower[axis], cw, **kw)
    cols = po.shifted_table_l1(grid.gather, table, axis, grid.lower[axis], cw,
                               per_column=True, **kw)
    print("cycles", ncyc, "L1", repr(l1), "ref", DE_L1)
    # input/vlct/testing_utils.py:275-292 (abs_tol) and the exact --std check
    assert np.isclose(l1, DE_L1, rtol=1e-13, atol=7e-14)
    assert np.std(cols) == 0.0


# ------------------------------------------------------------- plugin path
# EnzoMethodMHDVlct with device_fused = false: the reference's composition
# stage by stage, the dual-energy calls included (EnzoRiemann::solve with
# interface velocities, EnzoSourceInternalEnergy, the dual-energy update and
# the synchronisation before the CFL dt), each a single-operation entry.

def _host_params(case):
    return ("mhd_choice = %s; riemann_solver = %s; reconstruct_method = %s; "
            "theta_limiter = %r; courant = 0.4; device_fused = false"
            % ("constrained_transport" if case["mhd"] else "no_bfield",
               case["riemann"], case["recon"][1], case.get("theta", 2.0)))


PLUGIN_CASES = [CASES[0], CASES[1], CASES[3], CASES[4]]


@pytest.mark.parametrize("case", PLUGIN_CASES, ids=lambda c: "%s-%s-%s" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1]))
def test_dual_energy_staged_plugin_path_equals_the_oracle(case):
    from test_gpu_plugins import host_cycle
    print("staged case", case, flush=True)
    ev = _ev()
    grid = _de_grid(case)
    dg = device_copy(grid, method_for(grid.cfg))
    hm = ev.HostMethod(_host_params(case), gamma=grid.cfg.gamma,
                       dual_energy=("modern", [grid.cfg.eta]))
    t = 0.0
    for _ in range(3):
        dt_o = grid.cycle(t, 10.0)
        dt_g = host_cycle(hm, dg, t, 10.0)
        assert dt_g == dt_o
        t += dt_o
    ev.synchronize()
    _assert_equal_grids(dg, grid)


def _dev(arr):
    ev = _ev()
    buf = ev.DeviceBuffer(arr.nbytes)
    ev.to_device(buf.ptr.value, np.ascontiguousarray(arr))
    return buf


@pytest.mark.parametrize("solver,mhd", [("hlld", 1), ("hlle", 1), ("hllc", 0)])
def test_single_operation_riemann_dual_energy(solver, mhd):
    """enzoe_riemann_solve_dual_energy against the oracle's solve with the
    internal-energy flux and vi_bar, on random face states (both signs of
    the normal velocity, sub- and supersonic)."""
    ev = _ev()
    rng = np.random.default_rng(21)
    mz, my, mx, dim, s = 7, 8, 9, 2, 1
    fshape = (mz - 1, my, mx)
    nq = 8 if mhd else 5
    keys = ["density", "velocity_x", "velocity_y", "velocity_z", "pressure",
            "bfield_x", "bfield_y", "bfield_z"][:nq]
    fkeys = ["density", "velocity_x", "velocity_y", "velocity_z", "total_energy",
             "bfield_x", "bfield_y", "bfield_z"][:nq]

    def states():
        w = np.zeros((8,) + fshape)
        w[0] = rng.uniform(0.3, 2.0, fshape)
        w[1:4] = rng.normal(0, 2.0, (3,) + fshape)
        w[4] = rng.uniform(0.1, 1.5, fshape)
        if mhd:
            w[5:8] = rng.normal(0, 0.7, (3,) + fshape)
        return w
    wl, wr = states(), states()
    fn = int(np.prod(fshape))
    dl, dr = _dev(wl), _dev(wr)
    dflux, dvel = _dev(np.zeros((9,) + fshape)), _dev(np.zeros(fshape))
    pl = ev.fmap({k: dl.ptr.value + 8 * fn * q for q, k in enumerate(keys)}, fshape)
    pr = ev.fmap({k: dr.ptr.value + 8 * fn * q for q, k in enumerate(keys)}, fshape)
    fmapd = {k: dflux.ptr.value + 8 * fn * q for q, k in enumerate(fkeys)}
    fmapd["internal_energy"] = dflux.ptr.value + 8 * fn * 8
    flux = ev.fmap(fmapd, fshape)
    vel = ev.farray(dvel.ptr.value, fshape)
    L = ev.ops()
    assert L.enzoe_riemann_solve_dual_energy(solver.encode(), mhd, 5 / 3, pl, pr, flux,
                                             vel, dim, s, None) == 0
    ev.synchronize()
    got = ev.from_device(dflux.ptr.value, (9,) + fshape)
    got_v = ev.from_device(dvel.ptr.value, fshape)
    cfg = po.make_config(riemann=solver, mhd=bool(mhd))
    P = ctypes.POINTER(ctypes.c_double) * 8
    ol = [np.ascontiguousarray(wl[q]) for q in range(8)]
    orr = [np.ascontiguousarray(wr[q]) for q in range(8)]
    of = [np.zeros(fshape) for _ in range(8)]
    oe, ov = np.zeros(fshape), np.zeros(fshape)
    po.lib().o_riemann_solve_de(ctypes.byref(cfg), P(*[po.dptr(a) for a in ol]),
                                P(*[po.dptr(a) for a in orr]), P(*[po.dptr(a) for a in of]),
                                po.dptr(oe), po.dptr(ov), mz, my, mx, dim, s)
    sl = np.s_[s:fshape[0] - s, s:fshape[1] - s, s:fshape[2] - s]
    for q in range(nq):
        np.testing.assert_array_equal(got[q][sl], of[q][sl])
    np.testing.assert_array_equal(got[8][sl], oe[sl])
    np.testing.assert_array_equal(got_v[sl], ov[sl])
    assert np.any(got_v[sl] != 0.0) and np.any(got[8][sl] < 0) and np.any(got[8][sl] > 0)


@pytest.mark.parametrize("mhd,dual", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_single_operation_floor_and_sync(mhd, dual):
    """enzoe_apply_floor_to_energy_and_sync against the oracle on random
    cells: some with the internal energy replaced, some kept, some floored."""
    ev = _ev()
    rng = np.random.default_rng(33 + 2 * mhd + dual)
    shape = (6, 7, 8)
    U = np.zeros((9,) + shape)
    U[0] = rng.uniform(0.5, 2.0, shape)
    U[1:4] = rng.normal(0, 1.0, (3,) + shape)
    if mhd:
        U[5:8] = rng.normal(0, 0.5, (3,) + shape)
    ke = 0.5 * np.sum(U[1:4] ** 2, axis=0)
    me = 0.5 * np.sum(U[5:8] ** 2, axis=0) / U[0]
    U[8] = rng.uniform(-0.01, 1.0, shape)              # eint, some below the floor
    U[4] = ke + me + U[8] * rng.uniform(0.3, 1.7, shape)
    cfg = po.make_config(gamma=1.4, mhd=bool(mhd), pressure_floor=0.05,
                         dual_energy_eta=0.05 if dual else None)
    n = int(np.prod(shape))
    d = _dev(U)
    keys = ["density", "velocity_x", "velocity_y", "velocity_z", "total_energy",
            "bfield_x", "bfield_y", "bfield_z", "internal_energy"]
    use = [q for q in range(9) if (mhd or q not in (5, 6, 7)) and (dual or q != 8)]
    m = ev.fmap({keys[q]: d.ptr.value + 8 * n * q for q in use}, shape)
    assert ev.ops().enzoe_apply_floor_to_energy_and_sync(m, mhd, 1.4, 0.05, dual,
                                                         cfg.eta, 1, None) == 0
    ev.synchronize()
    got = ev.from_device(d.ptr.value, U.shape)
    ref = [U[q].copy() for q in range(8)]
    e = U[8].copy()
    P = ctypes.POINTER(ctypes.c_double) * 8
    po.lib().o_floor_energy_and_sync(ctypes.byref(cfg), P(*[po.dptr(a) for a in ref]),
                                     po.dptr(e), *shape, 1)
    np.testing.assert_array_equal(got[4], ref[4])
    if dual:
        np.testing.assert_array_equal(got[8], e)
        inner = np.s_[1:-1, 1:-1, 1:-1]
        changed = got[8][inner] != U[8][inner]
        assert np.any(changed) and np.any(~changed)


def test_dual_energy_tube_golden_through_the_staged_plugin_path():
    """The x deck of the dual-energy shock tube through the C++ layer's
    stage-by-stage composition (device_fused = false)."""
    from test_gpu_plugins import host_cycle
    ev = _ev()
    grid, dg = _tube(0)
    hm = ev.HostMethod("mhd_choice = constrained_transport; riemann_solver = hlld; "
                       "reconstruct_method = plm; theta_limiter = 2.0; courant = 0.4; "
                       "device_fused = false", gamma=1.4,
                       dual_energy=("modern", [po.DE_TUBE_ETA]))
    t, n = 0.0, 0
    while t < 0.25:
        t += host_cycle(hm, dg, t, 0.25)
        n += 1
    ev.synchronize()
    assert n > 1000
    for i in range(len(dg.blocks)):
        arrs = dg.blocks[i].download()
        view = grid.block_arrays(i)
        for k, v in arrs.items():
            if k in view:
                view[k][...] = v
    table = po.load_reference_table(GOLDEN)
    cw = 3.96875 / 508
    l1 = po.shifted_table_l1(grid.gather, table, 0, 0.0, cw, offset=po.DE_TUBE_OFFSET,
                             bkg_velocity=(po.DE_TUBE_AXIS_VELOCITY, 0.0, 0.0))
    print("staged cycles", n, "L1", repr(l1))
    assert np.isclose(l1, DE_L1, rtol=1e-13, atol=7e-14)


@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: "%s-%s-%s" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1]))
def test_dual_energy_compute_timestep_equals_compute_then_timestep(case):
    """With the dual-energy formalism the fused compute + CFL applies the
    timestep's in-place synchronisation to the cells of the final update
    (and to the ghost layers): next dt, energies, pressure and
    every other field equal compute() followed by timestep(), bit for bit."""
    print("fused de", case, flush=True)
    ev = _ev()
    grid = _de_grid(case, seed=13)
    a = device_copy(grid, method_for(grid.cfg))
    b = device_copy(grid, method_for(grid.cfg))
    dt_a = a.method.timestep(a.blocks)
    dt_b = b.method.timestep(b.blocks)
    assert dt_a == dt_b
    for _ in range(3):
        a.refresh()
        a.method.compute(a.blocks, dt_a)
        dt_a = a.method.timestep(a.blocks)
        b.refresh()
        dt_b = b.method.compute_timestep(b.blocks, dt_b)
        assert dt_b == dt_a
    ev.synchronize()
    ga, gb = a.blocks[0].download(), b.blocks[0].download()
    assert "internal_energy" in ga
    for f in ga:
        assert np.array_equal(ga[f], gb[f]), f
