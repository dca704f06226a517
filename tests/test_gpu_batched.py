"""Batched and sub-cycled launches, and the (1,1,1) fast-wave workload, on
the GPU against the oracle.

* Blocks of three refinement levels advanced together with one dt per
  block (enzoe_vlct_compute_dt, the sub-cycled launch lists of the batched
  workload) equal the oracle's step of each block with its own dt, bit for
  bit (the reference has no sub-cycling, so parity is per block per launch).
* The fast magnetosonic eigenmode along (1,1,1) matches the oracle cycle by
  cycle, and its error after one period converges at second order in FP64
  (the mixed-precision result is reported beside it).
"""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from test_gpu_parity import compare_grids, device_copy, method_for

pytestmark = pytest.mark.gpu


def _ev():
    import enzoe_vlct as ev
    return ev


def _oracle_block(cfg, arrays, n, width):
    g = po.Grid(cfg, (1, 1, 1), (n, n, n), (0.0, 0.0, 0.0), (n * width,) * 3)
    view = g.block_arrays(0)
    for k, v in arrays.items():
        if k in view:
            view[k][...] = v
    return g


def test_subcycled_levels_match_oracle_per_block():
    ev = _ev()
    import workloads
    import problems
    n, counts = 12, (2, 2, 3)
    cfg = po.make_config(riemann="hlld", recon=("nn", "plm"), theta=2.0, courant=0.4)
    method = method_for(cfg)
    lv = workloads.SubcycledLevels(method, counts=counts, n=n, distinct=3, seed=99)
    assert lv.launch_sizes() == [7, 3, 5, 3]
    # the oracle copy of every block (same seeds as SubcycledLevels)
    oracle = []
    for lev, cnt in enumerate(counts):
        width = 1.0 / (n * 2 ** lev)
        states = [problems.random_mhd_block(99 + 1000 * lev + s, n=n, width=width)[0]
                  for s in range(min(3, cnt))]
        oracle.append([_oracle_block(cfg, states[i % len(states)], n, width)
                       for i in range(cnt)])
    for _ in range(2):  # two coarse steps
        cfl = [min(po.lib().o_vlct_timestep(ctypes.byref(cfg), ctypes.byref(g.block(0)))
                   for g in lev_blocks) for lev_blocks in oracle]
        dt0_o = min(2 ** lev * d for lev, d in enumerate(cfl))
        for act in lv.active:
            for lev in act:
                for g in oracle[lev]:
                    assert po.lib().o_vlct_compute(ctypes.byref(cfg), ctypes.byref(g.block(0)),
                                                   dt0_o / 2 ** lev) == 0
        dt0_g = lv.coarse_step()
        assert dt0_g == dt0_o
    ev.synchronize()
    worst = 0.0
    for lev in range(len(counts)):
        for db, g in zip(lv.levels[lev], oracle[lev]):
            got = db.download()
            ref = g.block_arrays(0)
            for name, a in got.items():
                if name == "pressure":
                    continue
                b = ref[name]
                denom = np.maximum(np.abs(b), 1e-12 * max(1.0, float(np.abs(b).max())))
                worst = max(worst, float(np.max(np.abs(a - b) / denom)))
    assert worst <= 1e-12, worst
    print("max rel diff", worst)


def test_per_block_dt_from_device_memory():
    """dt_on_device: the per-block dt array read from device memory gives
    the same result as the host array."""
    ev = _ev()
    from _states import random_periodic_grid
    grid = random_periodic_grid(mhd=True, seed=8, n=(10, 9, 8))
    method = method_for(grid.cfg)
    a = device_copy(grid, method)
    b = device_copy(grid, method)
    dt = method.timestep(a.blocks)
    assert method.timestep(b.blocks) == dt  # also writes b's pressure
    method.compute_dt(a.blocks, [dt])
    buf = ev.DeviceBuffer(8)
    ev.to_device(buf.ptr.value, np.array([dt]))
    arr = (ev.Block * 1)(b.blocks[0].desc)
    d = ctypes.cast(buf.ptr, ctypes.POINTER(ctypes.c_double))
    assert ev.lib().enzoe_vlct_compute_dt(method.handle, arr, 1, d, 1, None) == 0
    ev.synchronize()
    x, y = a.blocks[0].download(), b.blocks[0].download()
    for k in x:
        np.testing.assert_array_equal(x[k], y[k])


def _fast_wave_grid(n, method, amplitude=1e-6):
    ev = _ev()
    import problems
    blk, cw, cf, period = problems.fast_wave_111(n, amplitude=amplitude)
    dg = ev.DeviceGrid(method, (1, 1, 1), (n, n, n), [cw])
    dg.upload_from([blk])
    dg.refresh()
    return dg, blk, period


def test_fast_wave_111_matches_oracle():
    ev = _ev()
    import problems
    n = 16
    cfg = po.make_config(riemann="hlld", recon=("nn", "plm"), theta=2.0, courant=0.4)
    blk, cw, cf, period = problems.fast_wave_111(n)
    grid = po.Grid(cfg, (1, 1, 1), (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    view = grid.block_arrays(0)
    for k, v in blk.items():
        view[k][...] = v
    dg = device_copy(grid, method_for(cfg))
    t = 0.0
    for _ in range(4):
        dt_o = grid.cycle(t, period)
        assert dg.cycle(t, period) == dt_o
        t += dt_o
    ev.synchronize()
    compare_grids(dg, grid)


def test_fast_wave_111_convergence():
    """L1 error of density after one period (against the initial state) at
    16^3 and 32^3: second order in FP64; mixed precision is stated."""
    ev = _ev()
    errs = {}
    for prec in (0, 1):
        for n in (16, 32):
            method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm",
                                   theta=2.0, courant=0.4, precision=prec)
            dg, blk, period = _fast_wave_grid(n, method)
            dg.evolve(0.0, period)
            fin = dg.blocks[0].download(["density"])["density"]
            s = (slice(3, -3),) * 3
            errs[(prec, n)] = float(np.abs(fin[s] - blk["density"][s]).mean())
    order64 = np.log2(errs[(0, 16)] / errs[(0, 32)])
    order32 = np.log2(errs[(1, 16)] / errs[(1, 32)])
    print("L1(density) fp64:", errs[(0, 16)], errs[(0, 32)], "order", order64)
    print("L1(density) mixed:", errs[(1, 16)], errs[(1, 32)], "order", order32)
    assert order64 > 1.8


def test_fast_wave_111_convergence_mixed_at_amplitude_1e3():
    """The same wave at amplitude 1e-3: the FP32 flux roundoff (~1e-7 of the
    background) is far below the truncation error, so the mixed path
    converges at second order as the FP64 one does, and its L1 errors stay
    within 1% of FP64's.  (At amplitude 1e-6 the roundoff sets the mixed
    error floor, the test above.)"""
    ev = _ev()
    errs = {}
    for prec in (0, 1):
        for n in (16, 32):
            method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm",
                                   theta=2.0, courant=0.4, precision=prec)
            dg, blk, period = _fast_wave_grid(n, method, amplitude=1e-3)
            dg.evolve(0.0, period)
            fin = dg.blocks[0].download(["density"])["density"]
            s = (slice(3, -3),) * 3
            errs[(prec, n)] = float(np.abs(fin[s] - blk["density"][s]).mean())
    order64 = np.log2(errs[(0, 16)] / errs[(0, 32)])
    order32 = np.log2(errs[(1, 16)] / errs[(1, 32)])
    print("amplitude 1e-3 L1(density) fp64:", errs[(0, 16)], errs[(0, 32)], "order", order64)
    print("amplitude 1e-3 L1(density) mixed:", errs[(1, 16)], errs[(1, 32)], "order", order32)
    assert order64 > 1.8 and order32 > 1.8
    for n in (16, 32):
        assert abs(errs[(1, n)] - errs[(0, n)]) <= 0.01 * errs[(0, n)], n


def test_graph_replays_match_eager_cycles():
    """A captured graph of 2 cycles, replayed 3 times with eager work and
    synchronisations between the replays, equals 6 eager cycles after every
    replay; past t_stop the cycles are zero-length and change nothing."""
    ev = _ev()
    from _states import random_periodic_grid
    grid = random_periodic_grid(mhd=True, seed=12, n=(11, 10, 9), vamp=0.4)
    method = method_for(grid.cfg)
    eager = device_copy(grid, method)
    dev = device_copy(grid, method)
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
    # a short stop: the graph clips the last cycle and then idles
    dev2 = device_copy(grid, method)
    eager2 = device_copy(grid, method)
    t_stop = 0.6 * t
    n_rep, t_end = dev2.evolve_graph(0.0, t_stop, ncycles=4)
    n_cyc, _ = eager2.evolve(0.0, t_stop)
    ev.synchronize()
    assert t_end == t_stop
    a, b = dev2.blocks[0].download(), eager2.blocks[0].download()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])


def test_device_cycles_without_a_graph_match_eager_cycles():
    """enzoe_vlct_run_cycles (the cycles enqueued directly, device dt with
    the next cycle's CFL minimum taken by the final cell update, no host
    synchronisation) equals host-driven eager cycles bit for bit, the last
    cycle clipped to t_stop and the cycles after it zero-length; its
    pressure field is the final state's, what a CFL pass writes."""
    ev = _ev()
    from _states import random_periodic_grid
    grid = random_periodic_grid(mhd=True, seed=13, n=(11, 10, 9), vamp=0.4)
    method = method_for(grid.cfg)
    eager = device_copy(grid, method)
    dev = device_copy(grid, method)
    n_cyc, log = eager.evolve(0.0, 10.0, max_cycles=5)
    t_stop = sum(log) - 0.3 * log[-1]
    eager2 = device_copy(grid, method)
    n2, _ = eager2.evolve(0.0, t_stop)
    ncalls, t_end = dev.evolve_device(0.0, t_stop, ncycles=3)
    eager2.method.timestep(eager2.blocks)  # the pressure of the final state
    ev.synchronize()
    assert t_end == t_stop and n2 == 5 and ncalls == 2
    a, b = dev.blocks[0].download(), eager2.blocks[0].download()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_graph_clipped_cycle_that_rounds_past_t_stop():
    """A stop time whose clipped cycle ends, in floating point, just above
    t_stop: t0 + (t_stop - t0) > t_stop for this pair (found by search;
    t0 < t_stop / 2, so the difference is not exact).  The first cycle is
    clipped (the CFL dt of the block, ~0.01, exceeds t_stop - t0), the
    device time then exceeds t_stop, and the graph's remaining cycles must be
    zero-length (dt 0, not the negative t_stop - t) so that the fields equal
    the single clipped eager cycle bit for bit."""
    ev = _ev()
    from _states import random_periodic_grid
    t0, t_stop = 0.0006548562668191901, 0.0049961035292957505
    assert t0 + (t_stop - t0) > t_stop
    grid = random_periodic_grid(mhd=True, seed=1, n=(12, 10, 8))
    method = method_for(grid.cfg)
    eager = device_copy(grid, method)
    dev = device_copy(grid, method)
    dt_cfl = method.timestep(eager.blocks)
    assert dt_cfl > t_stop - t0
    n_cyc, log = eager.evolve(t0, t_stop)
    assert n_cyc == 1 and log[0] == t_stop - t0
    g = dev.graph(t0, t_stop, 3)
    for rep in range(2):
        g.launch()
        ev.synchronize()
        assert g.time() == t0 + (t_stop - t0)
        a, b = dev.blocks[0].download(), eager.blocks[0].download()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg="replay %d %s" % (rep, k))
    g.destroy()


def test_autotuned_launch_order_keeps_results():
    """The autotuner only reorders workgroups: after it has chosen a mask
    for the shape (on a copy, leaving the blocks alone), cycles still equal
    the oracle's."""
    ev = _ev()
    from _states import random_periodic_grid
    grid = random_periodic_grid(mhd=True, seed=31, n=(16, 14, 12), vamp=0.5)
    method = method_for(grid.cfg)
    dg = device_copy(grid, method)
    before = dg.blocks[0].download()
    mask = method.autotune(dg.blocks, 1e-3)
    assert 0 <= mask < 16
    after = dg.blocks[0].download()
    for k in before:
        np.testing.assert_array_equal(before[k], after[k])
    t = 0.0
    for _ in range(2):
        dt_o = grid.cycle(t, 10.0)
        assert dg.cycle(t, 10.0) == dt_o
        t += dt_o
    ev.synchronize()
    compare_grids(dg, grid)


def _levels_blocks(ev, method, spec, seed):
    """Blocks of three levels: spec[l] = (count, n) (block edge n)."""
    import problems
    levels = []
    for lev, (cnt, n) in enumerate(spec):
        width = 1.0 / (n * 2 ** lev)
        blks = []
        for i in range(cnt):
            arr = problems.random_mhd_block(seed + 100 * lev + i, n=n, width=width)[0]
            b = ev.DeviceBlock(n + 6, n + 6, n + 6, 3, (width,) * 3, mhd=True)
            b.upload(arr)
            blks.append(b)
        levels.append(blks)
    return levels


def _host_coarse_step(method, levels):
    """One coarse step driven from the host: per-level CFL dt, then each
    substep's batched compute with per-block dt (workloads.SubcycledLevels)."""
    L = len(levels)
    cfl = [method.timestep(blks) for blks in levels]
    dt0 = min(2 ** lev * d for lev, d in enumerate(cfl))
    for s in range(2 ** (L - 1)):
        act = [lev for lev in range(L) if s % (2 ** (L - 1 - lev)) == 0]
        blocks = [b for lev in act for b in levels[lev]]
        dts = [dt0 / 2 ** lev for lev in act for _ in levels[lev]]
        method.compute_dt(blocks, dts)
    return dt0


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_levels_on_device_equal_host_driven_steps(precision):
    """enzoe_vlct_levels: a coarse step with the dt on the device equals the
    host-driven step bit for bit (fields, pressure, dt_0); a captured graph
    of 2 coarse steps replayed twice equals 4 host-driven steps.  The finest
    level mixes two block shapes, so its substeps launch two runs."""
    ev = _ev()
    spec = [(2, 12), (3, 12), (4, 10)]
    m = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0, courant=0.4,
                      precision=precision)
    host = _levels_blocks(ev, m, spec, 40)
    dev = _levels_blocks(ev, m, spec, 40)
    # the finest level: two blocks of 10^3 and two of 12^3 (copies of level 1's)
    import problems
    for lv in (host, dev):
        for i in range(2):
            arr = problems.random_mhd_block(77 + i, n=12, width=1.0 / 48)[0]
            b = ev.DeviceBlock(18, 18, 18, 3, (1.0 / 48,) * 3, mhd=True)
            b.upload(arr)
            lv[2].append(b)
    levels = ev.Levels(m, dev)
    dt_h = _host_coarse_step(m, host)
    levels.step()
    t, dt_d = levels.state()
    assert dt_d == dt_h and t == dt_h

    def same():
        for a_lv, b_lv in zip(host, dev):
            for a, b in zip(a_lv, b_lv):
                x, y = a.download(), b.download()
                for k in x:
                    np.testing.assert_array_equal(x[k], y[k], err_msg=k)
    same()
    levels.capture(2)
    t_h = dt_h
    for rep in range(2):
        levels.replay()
        for _ in range(2):
            t_h += _host_coarse_step(m, host)
        ev.synchronize()
        t, _ = levels.state()
        assert t == t_h, rep
        same()
    levels.destroy()
