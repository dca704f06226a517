"""The persistent stage launch (launch structure 5, k_stage: the flux, edge,
face and cell phases of a stage in one launch with a device-wide barrier
between them) against the separate launches on the GPU, bit for bit: every
field, the pressure field and the next dt, for HLLD MHD (PLM, PLM Athena,
nearest neighbour) and HLLC hydro, FP64, mixed and contracted FP64 (the
last two contract their flux arithmetic, which the compiler may fuse
differently in the two kernels: mixed within 1e-6 in relative L1, the
contracted FP64 within 1e-12 per active cell); a
batch whose phases hold more workgroups than the launch (each workgroup
takes several), with per-block dt and a zero-length block; a captured
graph of cycles; the sub-cycled levels driven on the device and replayed
from a graph.  The barrier reports no expired wait after any of them."""
import numpy as np
import pytest

from _states import random_periodic_grid
from test_gpu_parity import device_copy, method_for

pytestmark = pytest.mark.gpu

PERSISTENT = 5


def _ev():
    import enzoe_vlct as ev
    return ev


def _equal(blocks_a, blocks_b):
    for ba, bb in zip(blocks_a, blocks_b):
        a, b = ba.download(), bb.download()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def _close(blocks_a, blocks_b, rel, g=3):
    """Every cell-centred field within rel per active cell, relative to the
    field's largest magnitude (the contracted build may fuse its flux
    arithmetic differently in the persistent kernel than in k_flux; the
    ghost zones hold the stale partial updates, which the next refresh
    replaces)."""
    worst = 0.0
    act = (slice(g, -g),) * 3
    for ba, bb in zip(blocks_a, blocks_b):
        a, b = ba.download(), bb.download()
        for k in a:
            if k.startswith("bfieldi_"):
                continue
            scale = max(float(np.abs(b[k][act]).max()), 1e-300)
            d = float(np.abs(a[k][act] - b[k][act]).max()) / scale
            assert d <= rel, (k, d)
            worst = max(worst, d)
    return worst


CASES = [
    dict(mhd=True, riemann="hlld", recon=("nn", "plm"), theta=2.0),
    dict(mhd=True, riemann="hlld", recon=("nn", "plm_athena"), theta=1.5),
    dict(mhd=True, riemann="hlld", recon=("nn", "nn"), theta=1.5),
    dict(mhd=False, riemann="hllc", recon=("nn", "plm"), theta=2.0),
    dict(mhd=False, riemann="hllc", recon=("nn", "nn"), theta=2.0),
]


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["fp64", "mixed", "fp64_fma"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1]))
def test_persistent_stage_equals_separate_launches(case, precision):
    print("case", case, "precision", precision, flush=True)
    ev = _ev()
    grid = random_periodic_grid(mhd=case["mhd"], seed=41, n=(19, 13, 11),
                                riemann=case["riemann"], recon=case["recon"],
                                theta=case["theta"], vamp=0.5)
    runs = []
    for v in (PERSISTENT, 0):
        m = method_for(grid.cfg, precision=precision)
        m.set_fused_ct(v)
        runs.append((m, device_copy(grid, m)))
    (mp, gp), (ms, gs) = runs
    assert mp.launch_choice((25, 19, 17))[1] == PERSISTENT
    dtp = mp.timestep(gp.blocks)
    dts = ms.timestep(gs.blocks)
    assert dtp == dts
    for _ in range(3):
        gp.refresh()
        gs.refresh()
        dtp = mp.compute_timestep(gp.blocks, dtp)
        dts = ms.compute_timestep(gs.blocks, dts)
        if precision == 2:
            assert dtp == pytest.approx(dts, rel=1e-13, abs=0)
        elif precision == 1:
            assert dtp == pytest.approx(dts, rel=1e-6, abs=0)
        else:
            assert dtp == dts
    ev.synchronize()
    mp.check_persistent()
    if precision == 2:
        print("largest difference (relative to the field's scale)",
              _close(gp.blocks, gs.blocks, 1e-12))
    elif precision == 1:
        _, mag = mp.state_l1(gp.blocks, gs.blocks)
        print("relative L1 by vector magnitude", mag)
        assert max(mag.values()) <= 1e-6, mag
    else:
        _equal(gp.blocks, gs.blocks)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_persistent_stage_batch_larger_than_the_launch(precision):
    """24 blocks of 26^3 (ghost 3): about 1700 workgroups per phase, more
    than the device holds at once, so every workgroup of the launch takes
    several; per-block dt, one of them 0 (the block stays as it is)."""
    import problems
    ev = _ev()
    n, nb = 26, 24
    arrays = [problems.random_mhd_block(900 + i, n=n, width=1.0 / n)[0] for i in range(nb)]
    dts = [1e-3 * (1.0 + 0.01 * i) for i in range(nb)]
    dts[5] = 0.0
    out = []
    for v in (PERSISTENT, 0):
        m = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0,
                          courant=0.4, precision=precision)
        m.set_fused_ct(v)
        blocks = []
        for a in arrays:
            b = ev.DeviceBlock(n + 6, n + 6, n + 6, 3, (1.0 / n,) * 3, mhd=True)
            b.upload(a)
            blocks.append(b)
        for _ in range(2):
            m.compute_dt(blocks, dts)
        ev.synchronize()
        if v == PERSISTENT:
            m.check_persistent()
        out.append((m, blocks))
    if precision == 0:
        _equal(out[0][1], out[1][1])
    else:
        _, mag = out[0][0].state_l1(out[0][1], out[1][1])
        assert max(mag.values()) <= 1e-6, mag
    # the zero-length block is unchanged
    got = out[0][1][5].download()
    for k, v in arrays[5].items():
        if k in got:
            np.testing.assert_array_equal(got[k], v, err_msg=k)


def test_persistent_stage_in_a_captured_graph():
    """A graph of 2 cycles with the persistent launches, replayed 3 times,
    equals 6 eager cycles of the separate launches after every replay."""
    ev = _ev()
    grid = random_periodic_grid(mhd=True, seed=12, n=(11, 10, 9), vamp=0.4)
    mp = method_for(grid.cfg)
    mp.set_fused_ct(PERSISTENT)
    ms = method_for(grid.cfg)
    ms.set_fused_ct(0)
    dev = device_copy(grid, mp)
    eager = device_copy(grid, ms)
    t_stop = 10.0
    g = dev.graph(0.0, t_stop, 2)
    t = 0.0
    for rep in range(3):
        g.launch()
        for _ in range(2):
            t += eager.cycle(t, t_stop)
        ev.synchronize()
        assert g.time() == t
        a, b = dev.blocks[0].download(), eager.blocks[0].download()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg="replay %d %s" % (rep, k))
    g.destroy()
    mp.check_persistent()


def test_persistent_stage_levels_on_device_and_in_a_graph():
    """Sub-cycled levels (3 levels, the finest mixing two block shapes)
    stepped on the device with the persistent launches equal the host-driven
    steps of the separate launches, eagerly and replayed from a graph."""
    from test_gpu_batched import _host_coarse_step, _levels_blocks
    import problems
    ev = _ev()
    spec = [(2, 12), (3, 12), (4, 10)]
    mp = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0, courant=0.4)
    mp.set_fused_ct(PERSISTENT)
    ms = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0, courant=0.4)
    ms.set_fused_ct(0)
    host = _levels_blocks(ev, ms, spec, 40)
    dev = _levels_blocks(ev, mp, spec, 40)
    for lv in (host, dev):
        for i in range(2):
            arr = problems.random_mhd_block(77 + i, n=12, width=1.0 / 48)[0]
            b = ev.DeviceBlock(18, 18, 18, 3, (1.0 / 48,) * 3, mhd=True)
            b.upload(arr)
            lv[2].append(b)
    levels = ev.Levels(mp, dev)
    dt_h = _host_coarse_step(ms, host)
    levels.step()
    t, dt_d = levels.state()
    assert dt_d == dt_h and t == dt_h

    def same():
        for a_lv, b_lv in zip(host, dev):
            _equal(a_lv, b_lv)
    same()
    levels.capture(2)
    t_h = dt_h
    for rep in range(2):
        levels.replay()
        for _ in range(2):
            t_h += _host_coarse_step(ms, host)
        ev.synchronize()
        t, _ = levels.state()
        assert t == t_h, rep
        same()
    levels.destroy()
    mp.check_persistent()


def test_autotuner_may_choose_the_persistent_stage():
    """The autotuner times the persistent launch among the structures; the
    cycles after it equal the oracle's whatever it chose."""
    ev = _ev()
    grid = random_periodic_grid(mhd=True, seed=31, n=(20, 14, 12), vamp=0.5)
    m = method_for(grid.cfg)
    dg = device_copy(grid, m)
    m.autotune(dg.blocks, 1e-3)
    b = dg.blocks[0]
    choice = m.launch_choice((b.mx, b.my, b.mz))
    print("launch choice", choice)
    assert choice[1] in (0, 1, 2, 3, 4, PERSISTENT)
    t = 0.0
    for _ in range(2):
        dt_o = grid.cycle(t, 10.0)
        assert dg.cycle(t, 10.0) == dt_o
        t += dt_o
    ev.synchronize()
    m.check_persistent()
    from test_gpu_parity import compare_grids
    compare_grids(dg, grid)


def test_persistent_stage_many_launches_in_a_row():
    """40 persistent launches back to back (20 steps of 16 blocks of 26^3,
    more workgroups per phase than the launch holds) equal the separate
    launches bit for bit; the barrier's generation word carries across the
    launches and no wait runs out."""
    import problems
    ev = _ev()
    n, nb = 26, 16
    arrays = [problems.random_mhd_block(1300 + i, n=n, width=1.0 / n)[0] for i in range(nb)]
    dts = [7e-4] * nb
    out = []
    for v in (PERSISTENT, 0):
        m = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0,
                          courant=0.4)
        m.set_fused_ct(v)
        blocks = []
        for a in arrays:
            b = ev.DeviceBlock(n + 6, n + 6, n + 6, 3, (1.0 / n,) * 3, mhd=True)
            b.upload(a)
            blocks.append(b)
        for _ in range(20):
            m.compute_dt(blocks, dts)
        ev.synchronize()
        m.check_persistent()
        out.append((m, blocks))
    _equal(out[0][1], out[1][1])
