"""The C++ drop-in layer (EnzoMethodMHDVlct and its plugin objects) on the
GPU against the oracle.

Two paths of the same class are checked, both to the FP64 bar of the fused
library (<= 1e-12 relative per cell, every field incl. ghosts; dt exact):
  * device_fused = true  : Method::compute is one fused launch sequence;
  * device_fused = false : Method::compute runs the reference's composition
    stage by stage (EnzoMHDIntegratorStageCommands::compute_update_stage):
    clear dU, primitives, per dim reconstruct -> CT correct -> Riemann ->
    accumulate -> upwind, then CT (centre E, edge E, face B, centre B) and
    the quantity update -- every call a single-operation C-ABI entry point
    on strided views (face-trimmed subviews of cell-shaped scratch, slices of
    4-D backing arrays), so this also covers each of those entry points.
A few single-operation calls are also compared with the oracle's restatement
of the same function directly.
"""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from _states import random_periodic_grid
from test_gpu_parity import compare_grids, device_copy, method_for

pytestmark = pytest.mark.gpu


def _ev():
    import enzoe_vlct as ev
    return ev


def params(case, fused):
    return ("mhd_choice = %s; riemann_solver = %s; reconstruct_method = %s; "
            "theta_limiter = %r; courant = 0.4; device_fused = %s"
            % ("constrained_transport" if case["mhd"] else "no_bfield",
               case["riemann"], case["recon"][1], case["theta"],
               "true" if fused else "false"))


def host_cycle(hm, dg, t, t_stop):
    """A Cello cycle driven through Method::timestep / Method::compute."""
    dt = min(hm.timestep(b) for b in dg.blocks)
    if t < t_stop < t + dt:
        dt = t_stop - t
    rem = t_stop - t
    dt = dt if dt < rem else rem
    dg.refresh()
    hm.compute(dg.blocks, dt)
    return dt


CASES = [
    dict(mhd=True, riemann="hlld", recon=("nn", "plm"), theta=2.0),
    dict(mhd=True, riemann="hlle", recon=("nn", "plm_athena"), theta=1.5),
    dict(mhd=True, riemann="hlld", recon=("nn", "nn"), theta=1.5),
    dict(mhd=False, riemann="hllc", recon=("nn", "plm"), theta=1.5),
]


@pytest.mark.parametrize("fused", [False, True], ids=["staged", "fused"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1]))
def test_host_method_matches_oracle(case, fused):
    print("case", case, "fused", fused, flush=True)
    ev = _ev()
    grid = random_periodic_grid(mhd=case["mhd"], seed=5, n=(13, 11, 9),
                                riemann=case["riemann"], recon=case["recon"],
                                theta=case["theta"], vamp=0.6, bamp=0.08)
    dg = device_copy(grid, method_for(grid.cfg))
    hm = ev.HostMethod(params(case, fused), gamma=grid.cfg.gamma)
    assert hm.name == "mhd_vlct" and hm.courant == 0.4
    t = 0.0
    for _ in range(3):
        dt_o = grid.cycle(t, 10.0)
        dt_g = host_cycle(hm, dg, t, 10.0)
        assert dt_g == dt_o
        t += dt_o
    ev.synchronize()
    print("max rel diff per field:", compare_grids(dg, grid))


def test_host_method_batched_blocks():
    """compute_blocks: several blocks in one batched launch sequence; each
    block gets its compute_done."""
    ev = _ev()
    cfg = po.make_config(riemann="hlld", recon=("nn", "plm"), theta=2.0,
                         courant=0.4)
    grid = po.linear_wave_grid("fast", 8, mhd=True, nblocks=(2, 2, 2))
    dg = device_copy(grid, method_for(grid.cfg))
    hm = ev.HostMethod(params(dict(mhd=True, riemann="hlld", recon=("nn", "plm"),
                                   theta=2.0), True), gamma=cfg.gamma)
    t = 0.0
    for _ in range(2):
        dt_o = grid.cycle(t, 10.0)
        assert host_cycle(hm, dg, t, 10.0) == dt_o
        t += dt_o
    ev.synchronize()
    compare_grids(dg, grid)


def _dev(arr):
    ev = _ev()
    buf = ev.DeviceBuffer(arr.nbytes)
    ev.to_device(buf.ptr.value, np.ascontiguousarray(arr))
    return buf


def test_single_operation_reconstruct_and_riemann_on_backing_arrays():
    """PLM reconstruction into cell-shaped 4-D scratch read through
    face-trimmed views, then HLLD on those views, against the oracle's
    reconstruct / Riemann restatements on dense face arrays."""
    ev = _ev()
    rng = np.random.default_rng(7)
    mz, my, mx = 9, 8, 10
    dim, stale = 1, 1
    keys = ["density", "velocity_x", "velocity_y", "velocity_z", "pressure",
            "bfield_x", "bfield_y", "bfield_z"]
    w = np.empty((8, mz, my, mx))
    w[0] = rng.uniform(0.5, 2.0, (mz, my, mx))
    w[1:4] = rng.normal(0, 0.5, (3, mz, my, mx))
    w[4] = rng.uniform(0.3, 1.5, (mz, my, mx))
    w[5:8] = rng.normal(0, 0.7, (3, mz, my, mx))
    n = mz * my * mx
    dprim, dl, dr = _dev(w), _dev(np.zeros_like(w)), _dev(np.zeros_like(w))
    fshape = (mz, my - 1, mx)
    dflux = _dev(np.zeros((8,) + fshape))
    cell_stride = (my * mx, mx, 1)
    prim = ev.fmap({k: dprim.ptr.value + 8 * n * q for q, k in enumerate(keys)},
                   (mz, my, mx))
    # face-trimmed views of the cell-shaped scratch (cell strides)
    pl = ev.fmap({k: dl.ptr.value + 8 * n * q for q, k in enumerate(keys)}, fshape,
                 cell_stride)
    pr = ev.fmap({k: dr.ptr.value + 8 * n * q for q, k in enumerate(keys)}, fshape,
                 cell_stride)
    fn = int(np.prod(fshape))
    flux_keys = ["density", "velocity_x", "velocity_y", "velocity_z", "total_energy",
                 "bfield_x", "bfield_y", "bfield_z"]
    flux = ev.fmap({k: dflux.ptr.value + 8 * fn * q for q, k in enumerate(flux_keys)},
                   fshape)
    L = ev.ops()
    assert L.enzoe_reconstruct_interface(b"plm", 2.0, 1e-200, 1e-200, prim, pl, pr,
                                         dim, stale, None) == 0
    cs = stale + 1
    assert L.enzoe_riemann_solve(b"hlld", 1, 5 / 3, pl, pr, flux, dim, cs, None) == 0
    ev.synchronize()
    got_l = ev.from_device(dl.ptr.value, w.shape)[:, :, :my - 1, :]
    got_f = ev.from_device(dflux.ptr.value, (8,) + fshape)
    # oracle on dense face arrays
    ol = [np.zeros(fshape) for _ in range(8)]
    orr = [np.zeros(fshape) for _ in range(8)]
    for q in range(8):
        po.lib().o_reconstruct(1, 2.0, po.dptr(np.ascontiguousarray(w[q])),
                               po.dptr(ol[q]), po.dptr(orr[q]), mz, my, mx, dim, stale,
                               int(q in (0, 4)), 1e-200)
    sl = np.s_[cs:fshape[0] - cs, cs:fshape[1] - cs, cs:fshape[2] - cs]
    for q in range(8):
        np.testing.assert_array_equal(got_l[q][sl], ol[q][sl])
    cfg = po.make_config(riemann="hlld", mhd=True)
    of = [np.zeros(fshape) for _ in range(8)]
    P = ctypes.POINTER(ctypes.c_double) * 8
    po.lib().o_riemann_solve(ctypes.byref(cfg), P(*[po.dptr(a) for a in ol]),
                             P(*[po.dptr(a) for a in orr]), P(*[po.dptr(a) for a in of]),
                             mz, my, mx, dim, cs)
    for q in range(8):
        np.testing.assert_array_equal(got_f[q][sl], of[q][sl])


def test_passive_scalars_are_advected_with_the_density():
    """Passively advected scalars through the staged plugin path
    (solve_passive_advection, update_passive_scalars_, mass fractions in
    primitive_from_integration): a scalar holding a constant mass fraction c
    keeps it to roundoff, a zero scalar stays exactly zero, and a random
    scalar's total mass is conserved on the periodic block; the flow fields
    equal the oracle's (the scalars do not feed back)."""
    ev = _ev()
    case = dict(mhd=True, riemann="hlld", recon=("nn", "plm"), theta=2.0)
    grid = random_periodic_grid(mhd=True, seed=17, n=(12, 11, 10), vamp=0.5,
                                bamp=0.06)
    dg = device_copy(grid, method_for(grid.cfg))
    dg.refresh()
    grid.refresh()
    hm = ev.HostMethod(params(case, False), gamma=grid.cfg.gamma)
    rho = dg.blocks[0].download(["density"])["density"]
    c = 0.3125
    rng = np.random.default_rng(4)
    # the random scalar is periodic: its ghost zones wrap like the fields
    g = 3
    act = rng.uniform(0.1, 1.0, size=tuple(s - 2 * g for s in rho.shape))
    rnd = np.pad(act, g, mode="wrap") * rho
    fields = {"color_a": c * rho, "color_zero": np.zeros_like(rho), "color_rnd": rnd}
    bufs = {k: _dev(v) for k, v in fields.items()}
    ptrs = {k: b.ptr.value for k, b in bufs.items()}
    dt = hm.timestep(dg.blocks[0])
    dt_o = po.lib().o_vlct_timestep(ctypes.byref(grid.cfg), ctypes.byref(grid.block(0)))
    assert dt == dt_o
    hm.compute_passive(dg.blocks[0], ptrs, dt)
    assert po.lib().o_vlct_compute(ctypes.byref(grid.cfg), ctypes.byref(grid.block(0)), dt) == 0
    ev.synchronize()
    compare_grids(dg, grid)
    s = (slice(g, -g),) * 3
    new_rho = dg.blocks[0].download(["density"])["density"]
    a = ev.from_device(ptrs["color_a"], rho.shape)
    np.testing.assert_allclose(a[s] / new_rho[s], c, rtol=1e-13, atol=0)
    z = ev.from_device(ptrs["color_zero"], rho.shape)
    assert not np.any(z[s])
    r = ev.from_device(ptrs["color_rnd"], rho.shape)
    m0, m1 = float(rnd[s].sum()), float(r[s].sum())
    assert abs(m1 - m0) <= 1e-13 * m0, (m0, m1)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_passive_advection_sound_wave_golden(axis):
    """The reference's passive-advection sound wave
    (input/vlct/run_passive_advect_sound_test.py, decks
    input/vlct/passive_advect_sound_wave/method_vlct_{x,y,z}_passive_sound.in):
    a 1e-6 sound wave along one axis of a 16x4x4 periodic block with HLLD
    + CT, and a "red" scalar holding 10% of the density plus a 1e-7
    perturbation a quarter wave apart, evolved to t = 1 through the staged
    plugin path; the L1 error norm over density, velocity, total energy,
    bfield and red is the reference's 6.918011605252798e-08 (rtol 1e-13,
    atol 7e-14, input/vlct/testing_utils.py:275-292)."""
    ev = _ev()
    print("passive sound wave axis", axis, flush=True)
    g = 3
    n = [4, 4, 4]
    n[axis] = 16
    upper = [0.25, 0.25, 0.25]
    upper[axis] = 1.0
    cw = tuple(upper[a] / n[a] for a in range(3))
    method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0,
                           courant=0.4)
    dg = ev.DeviceGrid(method, (1, 1, 1), tuple(n), [cw], bc=("periodic",) * 3)
    m = [n[a] + 2 * g for a in range(3)]
    # the "value" initializer at every cell centre, ghosts included
    c = [(np.arange(m[a]) - g + 0.5) * cw[a] for a in range(3)]
    z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    s = (x, y, z)[axis]
    sn = np.sin(2.0 * np.pi * s)
    rho = 1.0 + 1.e-6 * sn
    state = {"density": rho, "total_energy": 1.5 * (0.6 + 1.e-6 * sn) / (1.0 + 1.e-6 * sn)}
    for a, name in enumerate(("velocity_x", "velocity_y", "velocity_z")):
        state[name] = -1.e-6 * sn if a == axis else np.zeros_like(rho)
    for name in ("bfield_x", "bfield_y", "bfield_z"):
        state[name] = np.zeros_like(rho)
    state["bfieldi_x"] = np.zeros((m[2], m[1], m[0] + 1))
    state["bfieldi_y"] = np.zeros((m[2], m[1] + 1, m[0]))
    state["bfieldi_z"] = np.zeros((m[2] + 1, m[1], m[0]))
    red = (0.1 + 1.e-7 * np.sin(2.0 * np.pi * (s - 0.25))) * (1.0 + 1.e-6 * sn)
    dg.upload_from([state])
    red_buf = _dev(red)
    hm = ev.HostMethod("mhd_choice = constrained_transport; riemann_solver = hlld; "
                       "reconstruct_method = plm; theta_limiter = 2.0; courant = 0.4; "
                       "device_fused = false")
    names = ["density", "velocity_x", "velocity_y", "velocity_z", "total_energy",
             "bfield_x", "bfield_y", "bfield_z"]
    act = tuple(slice(g, g + n[a]) for a in (2, 1, 0))
    init = [state[k][act].copy() for k in names] + [red[act].copy()]
    t, t_stop, cycles = 0.0, 1.0, 0
    while t < t_stop and cycles < 10000:
        # Cello's cycle: dt from the current state, then refresh, then compute
        dt = hm.timestep(dg.blocks[0])
        if t + dt > t_stop:
            dt = t_stop - t
        dg.refresh()
        r = ev.from_device(red_buf.ptr.value, red.shape)
        r = np.pad(r[act], g, mode="wrap")        # periodic ghost zones of red
        ev.to_device(red_buf.ptr.value, np.ascontiguousarray(r))
        hm.compute_passive(dg.blocks[0], {"red": red_buf.ptr.value}, dt)
        t += dt
        cycles += 1
    ev.synchronize()
    fin = dg.blocks[0].download(names)
    final = [fin[k][act] for k in names] + [ev.from_device(red_buf.ptr.value, red.shape)[act]]
    l1 = po.l1_norm(init, final)
    print("cycles", cycles, "L1", repr(l1))
    assert np.isclose(l1, 6.918011605252798e-08, rtol=1e-13, atol=7e-14)


@pytest.mark.parametrize("mhd", [True, False])
def test_gravity_source_fused_and_staged_match_the_oracle(mhd):
    """Acceleration fields on the block: the fused step
    (enzoe_vlct_compute_gravity) and the C++ layer's staged plugin path
    (EnzoSourceGravity through enzoe_source_gravity) both equal the oracle's
    step with the gravity source bit for bit."""
    ev = _ev()
    print("gravity mhd", mhd, flush=True)
    case = dict(mhd=mhd, riemann="hlld" if mhd else "hllc", recon=("nn", "plm"), theta=2.0)
    rng = np.random.default_rng(21)
    grids = [random_periodic_grid(mhd=mhd, seed=8, n=(12, 10, 8)) for _ in range(3)]
    shape = grids[0].block_arrays(0)["density"].shape
    acc = [np.ascontiguousarray(rng.normal(size=shape)) for _ in range(3)]
    acc_dev = [_dev(a) for a in acc]
    acc_ptrs = [b.ptr.value for b in acc_dev]
    dt = 0.5 * po.lib().o_vlct_timestep(ctypes.byref(grids[0].cfg),
                                        ctypes.byref(grids[0].block(0)))
    ptrs = (ctypes.POINTER(ctypes.c_double) * 3)(*[po.dptr(a) for a in acc])
    assert po.lib().o_vlct_compute_accel(ctypes.byref(grids[0].cfg),
                                         ctypes.byref(grids[0].block(0)), dt, ptrs) == 0
    # fused
    method = method_for(grids[1].cfg)
    dg = device_copy(grids[1], method)
    ev.compute_gravity(method, dg.blocks, dt, [acc_ptrs])
    # staged (C++ plugin composition)
    dg2 = device_copy(grids[2], method)
    hm = ev.HostMethod(params(case, False), gamma=grids[2].cfg.gamma)
    hm.compute_accel(dg2.blocks[0], acc_ptrs, dt)
    ev.synchronize()
    for d in (dg, dg2):
        got = d.blocks[0].download()
        ref = grids[0].block_arrays(0)
        for k, v in got.items():
            if k != "pressure":
                assert np.array_equal(v, ref[k]), (mhd, k)


def test_staged_path_on_two_host_threads():
    """The staged plugin path (device_fused = false: every call a
    single-operation entry point, the CFL reduction by
    enzoe_integrator_timestep) driven from two host threads at once, each
    with its own Method and grid, equals the same cycles run one after the
    other, bit for bit (dt included)."""
    import threading
    ev = _ev()
    case = dict(mhd=True, riemann="hlld", recon=("nn", "plm"), theta=2.0)
    grids = [random_periodic_grid(mhd=True, seed=s, n=(13, 11, 9), vamp=0.6, bamp=0.08)
             for s in (41, 42)]

    def run(grid, out, i):
        dg = device_copy(grid, method_for(grid.cfg))
        hm = ev.HostMethod(params(case, False), gamma=grid.cfg.gamma)
        t, dts = 0.0, []
        for _ in range(3):
            dt = host_cycle(hm, dg, t, 10.0)
            dts.append(dt)
            t += dt
        ev.synchronize()
        out[i] = (dts, dg.blocks[0].download())

    seq = [None, None]
    for i, g in enumerate(grids):
        run(g, seq, i)
    par = [None, None]
    threads = [threading.Thread(target=run, args=(g, par, i)) for i, g in enumerate(grids)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    for i in range(2):
        assert par[i] is not None
        assert par[i][0] == seq[i][0], i
        for f in seq[i][1]:
            assert np.array_equal(seq[i][1][f], par[i][1][f]), (i, f)
