"""Parity of the gfx950 library (through the C ABI) with the CPU oracle.

FP64 bar: <= 1e-12 relative per cell against the oracle (which restates the
reference operation by operation), on every cell of every field including
the ghost zones the method writes; the end-to-end runs must also reproduce
the reference's golden L1 values (rtol 1e-13, atol 7e-14,
input/vlct/testing_utils.py:275-292).
"""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as po
from _states import random_periodic_grid

pytestmark = pytest.mark.gpu

REL = 1e-12
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _ev():
    import enzoe_vlct as ev
    return ev


def method_for(cfg_o, precision=0, courant=None):
    ev = _ev()
    rec = {0: "nn", 1: "plm", 2: "plm_athena"}[cfg_o.recon[1]]
    rs = {0: "hlle", 1: "hllc", 2: "hlld"}[cfg_o.riemann]
    return ev.VlctMethod(mhd=bool(cfg_o.mhd), riemann=rs, reconstruct=rec,
                         theta=cfg_o.theta_limiter,
                         courant=cfg_o.courant if courant is None else courant,
                         gamma=cfg_o.gamma, density_floor=cfg_o.density_floor,
                         pressure_floor=cfg_o.pressure_floor,
                         precision=precision,
                         dual_energy_eta=cfg_o.eta if cfg_o.dual_energy else None)


def device_copy(grid_o, method):
    ev = _ev()
    nb = grid_o.nblocks[0] * grid_o.nblocks[1] * grid_o.nblocks[2]
    cws = [tuple(grid_o.block(i).cw) for i in range(nb)]
    bcs = tuple("periodic" if grid_o.g.bc[a] == 0 else "outflow" for a in range(3))
    dg = ev.DeviceGrid(method, grid_o.nblocks, grid_o.ncells, cws, bc=bcs)
    dg.upload_from([grid_o.block_arrays(i) for i in range(nb)])
    return dg


def max_rel(a, b):
    scale = np.maximum(np.abs(b), 1e-300)
    d = np.abs(a - b)
    return float(np.max(d / np.maximum(scale, np.abs(b).max() * 1e-300)))


def compare_grids(dg, grid_o, rel=REL, fields=None):
    out = {}
    nb = len(dg.blocks)
    for i in range(nb):
        got = dg.blocks[i].download()
        ref = grid_o.block_arrays(i)
        for name in (fields or got):
            if name not in ref:
                continue
            a, b = got[name], ref[name]
            # relative to the field's magnitude scale (per cell)
            denom = np.maximum(np.abs(b), 1e-12 * max(1.0, float(np.abs(b).max())))
            r = float(np.max(np.abs(a - b) / denom))
            out[name] = max(out.get(name, 0.0), r)
    bad = {k: v for k, v in out.items() if not v <= rel}
    assert not bad, "relative differences above %g: %s" % (rel, bad)
    return out


CASES = [
    dict(mhd=True, riemann="hlld", recon=("nn", "plm"), theta=2.0),
    dict(mhd=True, riemann="hlld", recon=("nn", "plm"), theta=1.5),
    dict(mhd=True, riemann="hlld", recon=("nn", "plm_athena"), theta=1.5),
    dict(mhd=True, riemann="hlle", recon=("nn", "plm"), theta=2.0),
    dict(mhd=True, riemann="hlld", recon=("nn", "nn"), theta=1.5),
    dict(mhd=False, riemann="hllc", recon=("nn", "plm"), theta=2.0),
    dict(mhd=False, riemann="hllc", recon=("nn", "plm_athena"), theta=2.0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s-%g" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1], c["theta"]))
def test_cycles_match_oracle(case):
    print("case", case, flush=True)
    grid = random_periodic_grid(mhd=case["mhd"], seed=11, n=(14, 12, 10),
                                riemann=case["riemann"], recon=case["recon"],
                                theta=case["theta"], vamp=0.6, bamp=0.08)
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    t = 0.0
    for _ in range(3):
        dt_o = grid.cycle(t, 10.0)
        dt_g = dg.cycle(t, 10.0)
        assert dt_g == pytest.approx(dt_o, rel=1e-14, abs=0)
        t += dt_o
    _ev().synchronize()
    diffs = compare_grids(dg, grid)
    print("max rel diff per field:", diffs)


NAN_CASES = [CASES[0], CASES[2], CASES[3], CASES[5]]


@pytest.mark.parametrize("case", NAN_CASES, ids=lambda c: "%s-%s-%s" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1]))
def test_nan_cells_spread_as_in_oracle(case):
    """A step from a state holding NaN in a few cells (one velocity, one
    density, one energy value, one centred B value): the NaNs spread to the
    same cells and every finite value is bit for bit the oracle's, so the
    comparisons of the limiter, the floors and the solvers treat NaN as the
    reference's do (the PLM limiter's minimum of three magnitudes among
    them, which uses the hardware minimum)."""
    print("case", case, flush=True)
    grid = random_periodic_grid(mhd=case["mhd"], seed=21, n=(14, 12, 10),
                                riemann=case["riemann"], recon=case["recon"],
                                theta=case["theta"], vamp=0.6, bamp=0.08)
    arr = grid.block_arrays(0)
    arr["velocity_x"][6, 7, 5] = np.nan
    arr["density"][9, 4, 12] = np.nan
    arr["total_energy"][5, 10, 15] = np.nan
    if case["mhd"]:
        arr["bfield_y"][11, 8, 9] = np.nan
    grid.refresh()
    dg = device_copy(grid, method_for(grid.cfg))
    dt = 2e-3
    assert po.lib().o_vlct_compute(ctypes.byref(grid.cfg), ctypes.byref(grid.block(0)), dt) == 0
    dg.method.compute(dg.blocks, dt)
    _ev().synchronize()
    got = dg.blocks[0].download()
    ref = grid.block_arrays(0)
    spread = 0
    for k, a in got.items():
        if k not in ref or k == "pressure":
            continue
        spread += int(np.isnan(ref[k]).sum())
        np.testing.assert_array_equal(a, ref[k], err_msg=k)
    print("NaN values after the step", spread)
    assert spread > 4


def test_timestep_writes_pressure_and_matches_oracle():
    grid = random_periodic_grid(mhd=True, seed=3, n=(10, 9, 8))
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    dt_g = method.timestep(dg.blocks)
    dt_o = po.lib().o_vlct_timestep(ctypes.byref(grid.cfg), ctypes.byref(grid.block(0)))
    assert dt_g == dt_o
    p = dg.blocks[0].download(["pressure"])["pressure"]
    np.testing.assert_allclose(p, grid.block_arrays(0)["pressure"], rtol=REL, atol=0)


def _wave(wave, res, t_stop, mhd=True, nblocks=(1, 1, 1)):
    grid = po.linear_wave_grid(wave, res, mhd=mhd, nblocks=nblocks)
    fields = po.MHD_L1_FIELDS if mhd else po.HD_L1_FIELDS
    init = [grid.gather(f) for f in fields]
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    n_g, log_g = dg.evolve(0.0, t_stop)
    n_o, log_o = grid.evolve(0.0, t_stop)
    assert n_g == n_o
    np.testing.assert_allclose(log_g, log_o, rtol=1e-13)
    # copy the device state back into the oracle grid to use its gather
    nb = nblocks[0] * nblocks[1] * nblocks[2]
    ref_final = [grid.gather(f) for f in fields]
    diffs = compare_grids(dg, grid)
    for i in range(nb):
        arrs = dg.blocks[i].download()
        view = grid.block_arrays(i)
        for k, v in arrs.items():
            if k in view:
                view[k][...] = v
    final = [grid.gather(f) for f in fields]
    for a, b in zip(final, ref_final):
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-14)
    return po.l1_norm(init, final, res=res), diffs


@pytest.mark.parametrize("wave,res,t_stop,ref", [
    ("fast", 16, 0.5, 1.6388526155394664e-07),
    ("alfven", 16, 1.0, 1.927245356389947e-07),
    ("mhd_entropy", 16, 1.0, 1.0021263485338544e-07),
])
def test_mhd_linear_wave_golden_on_gpu(wave, res, t_stop, ref):
    print("wave", wave, res, flush=True)
    l1, diffs = _wave(wave, res, t_stop)
    print("L1", repr(l1), "ref", ref, "per-field max rel", diffs)
    assert np.isclose(l1, ref, rtol=1e-13, atol=7e-14)


def test_hd_sound_wave_multiblock_golden_on_gpu():
    l1, diffs = _wave("sound", 16, 1.0, mhd=False, nblocks=(2, 2, 2))
    print("L1", repr(l1), diffs)
    assert np.isclose(l1, 1.3704437791196907e-07, rtol=1e-13, atol=7e-14)


def test_rj2a_shock_tube_golden_on_gpu():
    table = po.load_reference_table(
        os.path.join(GOLDEN, "rj2a_shock_tube_t0.2_res256.csv"))
    grid = po.rj2a_grid(0)
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    dg.evolve(0.0, 0.2)
    arrs = dg.blocks[0].download()
    view = grid.block_arrays(0)
    for k, v in arrs.items():
        if k in view:
            view[k][...] = v
    l1 = po.table_l1(grid, table, 0)
    print("rj2a L1", repr(l1))
    assert np.isclose(l1, 0.012523489882320429, rtol=1e-13, atol=7e-14)


def test_div_b_stays_at_roundoff_on_gpu():
    grid = random_periodic_grid(mhd=True, seed=5, n=(16, 16, 16), ext=(1, 1, 1),
                                vamp=0.8, bamp=0.1)
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    t = 0.0
    for _ in range(6):
        t += dg.cycle(t, 10.0)
    a = dg.blocks[0].download()
    dx = 1.0 / 16
    div = ((a["bfieldi_x"][:, :, 1:] - a["bfieldi_x"][:, :, :-1]) / dx +
           (a["bfieldi_y"][:, 1:, :] - a["bfieldi_y"][:, :-1, :]) / dx +
           (a["bfieldi_z"][1:, :, :] - a["bfieldi_z"][:-1, :, :]) / dx)[3:-3, 3:-3, 3:-3]
    print("max |div B| * dx =", np.abs(div).max() * dx)
    assert np.abs(div).max() * dx < 1e-13


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s-%g" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1], c["theta"]))
def test_mixed_precision_close_to_fp64(case):
    """FP32 reconstruction + Riemann, FP64 accumulation/CT/dt: the L1 error
    against the FP64 result stays below 1e-5 relative, for every solver and
    reconstruction."""
    print("case", case, flush=True)
    grid = random_periodic_grid(mhd=case["mhd"], seed=9, n=(16, 16, 16), ext=(1, 1, 1),
                                riemann=case["riemann"], recon=case["recon"],
                                theta=case["theta"], vamp=0.3, bamp=0.05)
    m64 = method_for(grid.cfg)
    m32 = method_for(grid.cfg, precision=1)
    d64 = device_copy(grid, m64)
    d32 = device_copy(grid, m32)
    t = 0.0
    for _ in range(5):
        dt = m64.timestep(d64.blocks)
        d64.refresh(); d32.refresh()
        m64.compute(d64.blocks, dt)
        m32.compute(d32.blocks, dt)
        t += dt
    a, b = d64.blocks[0].download(), d32.blocks[0].download()
    for f in ("density", "velocity_x", "total_energy") + (("bfield_y",) if case["mhd"] else ()):
        x, y = a[f][3:-3, 3:-3, 3:-3], b[f][3:-3, 3:-3, 3:-3]
        rel = np.abs(x - y).sum() / np.abs(x).sum()
        print(f, "L1 rel", rel)
        assert rel < 1e-5


def test_orszag_tang_256_full_size_matches_oracle():
    """The benchmark workload at its full size (one 256^3 block): two
    cycles on the GPU equal the oracle's cell for cell, mass is conserved to
    roundoff and the discrete div B of the face field stays at roundoff."""
    import problems
    ev = _ev()
    n = 256
    cfg = po.make_config(riemann="hlld", recon=("nn", "plm"), theta=2.0, courant=0.4)
    blk, cw = problems.orszag_tang(n)
    grid = po.Grid(cfg, (1, 1, 1), (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    view = grid.block_arrays(0)
    for k, v in blk.items():
        view[k][...] = v
    grid.refresh()
    dg = device_copy(grid, method_for(cfg))
    s = (slice(3, -3),) * 3
    mass0 = float(view["density"][s].sum())
    t = 0.0
    for _ in range(2):
        dt_o = grid.cycle(t, 10.0, nthreads=16)
        assert dg.cycle(t, 10.0) == dt_o
        t += dt_o
    ev.synchronize()
    compare_grids(dg, grid)
    got = dg.blocks[0].download()
    assert abs(float(got["density"][s].sum()) - mass0) <= 1e-12 * mass0
    bx, by, bz = got["bfieldi_x"], got["bfieldi_y"], got["bfieldi_z"]
    div = ((bx[s[0], s[1], 4:-3] - bx[s[0], s[1], 3:-4]) +
           (by[s[0], 4:-3, s[2]] - by[s[0], 3:-4, s[2]]) +
           (bz[4:-3, s[1], s[2]] - bz[3:-4, s[1], s[2]])) / cw[0]
    bmax = float(np.abs(bx).max())
    assert float(np.abs(div).max()) <= 1e-13 * bmax / cw[0]


def test_mixed_block_shapes_in_one_call():
    """A compute call over blocks of different shapes (runs of equal shape
    each get their own launches) equals the oracle block by block; an empty
    list is rejected without touching the device."""
    ev = _ev()
    grids = [random_periodic_grid(mhd=True, seed=40 + i, n=n, vamp=0.4)
             for i, n in enumerate([(9, 8, 7), (9, 8, 7), (12, 7, 9)])]
    method = method_for(grids[0].cfg)
    dgs = [device_copy(g, method) for g in grids]
    blocks = [d.blocks[0] for d in dgs]
    for d in dgs:
        d.refresh()
    dt = min(method.timestep([b]) for b in blocks)
    method.compute(blocks, dt)
    for g in grids:
        g.refresh()
        po.lib().o_vlct_timestep(ctypes.byref(g.cfg), ctypes.byref(g.block(0)))  # pressure
        assert po.lib().o_vlct_compute(ctypes.byref(g.cfg), ctypes.byref(g.block(0)), dt) == 0
    ev.synchronize()
    for d, g in zip(dgs, grids):
        compare_grids(d, g)
    arr = (ev.Block * 1)()
    assert ev.lib().enzoe_vlct_compute(method.handle, arr, 0, dt, None) == 2


@pytest.mark.parametrize("mhd", [True, False], ids=["mhd", "hd"])
@pytest.mark.parametrize("mask", [15, 0], ids=["tiled", "plain"])
def test_tiled_cell_order_matches_oracle(mhd, mask, monkeypatch):
    """Blocks with 134 x 134 planes (more than an L2 holds) are walked in
    slabs of 32 rows marched through z when a kernel's launch-order bit is
    set (the last slab of each box is shorter): every stage kernel in the
    tiled order (mask 15), and all in the plain order (mask 0), give the
    oracle's cycles bit for bit."""
    print("tiled order", mhd, mask, flush=True)
    monkeypatch.setenv("ENZOE_XCD_REMAP", str(mask))
    grid = random_periodic_grid(mhd=mhd, seed=5, n=(128, 128, 6), ext=(1.0, 1.0, 0.05),
                                vamp=0.6, bamp=0.08)
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    t = 0.0
    for _ in range(2):
        dt_o = grid.cycle(t, 10.0)
        assert dg.cycle(t, 10.0) == dt_o
        t += dt_o
    _ev().synchronize()
    compare_grids(dg, grid, rel=0.0)


FUSED_CASES = [
    dict(mhd=True, riemann="hlld", recon=("nn", "plm"), precision=0, n=(14, 12, 10), mask=None),
    dict(mhd=True, riemann="hlle", recon=("nn", "nn"), precision=0, n=(14, 12, 10), mask=None),
    dict(mhd=False, riemann="hllc", recon=("nn", "plm"), precision=0, n=(14, 12, 10), mask=None),
    dict(mhd=True, riemann="hlld", recon=("nn", "plm"), precision=1, n=(14, 12, 10), mask=None),
    dict(mhd=True, riemann="hlld", recon=("nn", "plm"), precision=0, n=(128, 128, 6), mask=15),
]


@pytest.mark.parametrize("case", FUSED_CASES, ids=lambda c: "%s-%s-%s-%s-%s" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1],
    "mixed" if c["precision"] else "fp64", "x".join(map(str, c["n"]))))
def test_compute_timestep_equals_compute_then_timestep(case, monkeypatch):
    """enzoe_vlct_compute_timestep (the final cell update, launched over the
    whole block, also takes the CFL minimum and writes the pressure of every
    cell) gives the next dt and every field, pressure included, bit for
    bit as compute() followed by timestep(), over several cycles (NN + NN
    moves the final update box; one case runs every stage kernel in the
    tiled order)."""
    print("fused", case, flush=True)
    ev = _ev()
    if case["mask"] is not None:
        monkeypatch.setenv("ENZOE_XCD_REMAP", str(case["mask"]))
    grid = random_periodic_grid(mhd=case["mhd"], seed=7, n=case["n"], riemann=case["riemann"],
                                recon=case["recon"], vamp=0.6, bamp=0.08,
                                ext=(1.2, 1.0, 0.8) if case["n"][0] < 100 else (1.0, 1.0, 0.05))
    a = device_copy(grid, method_for(grid.cfg, precision=case["precision"]))
    b = device_copy(grid, method_for(grid.cfg, precision=case["precision"]))
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
    for f in ga:
        assert np.array_equal(ga[f], gb[f]), f


def test_compute_timestep_on_blocks_of_two_shapes():
    """The fused CFL over a call with blocks of two shapes (two runs of the
    stage kernels, one minimum) equals compute() then timestep()."""
    ev = _ev()
    g1 = random_periodic_grid(mhd=True, seed=21, n=(12, 10, 8))
    g2 = random_periodic_grid(mhd=True, seed=22, n=(10, 12, 14), vamp=0.9)
    m = method_for(g1.cfg)
    da, db = device_copy(g1, m), device_copy(g2, m)
    ea, eb = device_copy(g1, m), device_copy(g2, m)
    blocks_d = da.blocks + db.blocks
    blocks_e = ea.blocks + eb.blocks
    dt = m.timestep(blocks_d)
    m.compute(blocks_d, dt)
    ref = m.timestep(blocks_d)
    got = m.compute_timestep(blocks_e, dt)
    ev.synchronize()
    assert got == ref
    for x, y in zip(blocks_d, blocks_e):
        gx, gy = x.download(), y.download()
        for f in gx:
            assert np.array_equal(gx[f], gy[f]), f


def test_two_handles_on_two_host_threads():
    """One method handle per host thread (the reference's one Method per
    PE): two threads stepping their own grids at the same time give, bit for
    bit, what the same cycles give one after the other."""
    import threading
    ev = _ev()
    grids = [random_periodic_grid(mhd=True, seed=s, n=(16, 14, 12)) for s in (31, 32)]

    def run(grid, out, i):
        dg = device_copy(grid, method_for(grid.cfg))
        t = 0.0
        for _ in range(3):
            t += dg.cycle(t, 10.0)
        ev.synchronize()
        out[i] = dg.blocks[0].download()

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
        for f in seq[i]:
            assert np.array_equal(seq[i][f], par[i][f]), (i, f)
