"""The contracted FP64 step (ENZOE_PREC_DOUBLE_FMA): the FP64 kernels with
multiply-adds fused into FMAs.  Not bit-identical to the reference, so it is
held to the FP64 bar of <= 1e-12 relative per cell against the oracle, over
the active cells and faces (the ghost zones hold the stale partial updates
of the stage region, values the next refresh replaces), with the relative
difference of a cell defined as |a - b| / |b| for the scalar fields and,
for a velocity or B component, |a - b| over the cell's vector magnitude
|b_vec| (a face B value: over the larger of its own magnitude and the
adjacent cells' |B|): a component that crosses zero has no scale of its own.
Checked on the benchmark's Orszag-Tang run for its 20 steps and at the
benchmark's full size against the reference-parity FP64 build (which equals
the oracle bit for bit), plus div B at roundoff and conserved totals; and
on every solver and reconstruction, with and without the dual-energy
formalism, over seeded rough random states, held there to 1e-11.
"""
import ctypes
import gc
import os

import numpy as np
import pytest

import pyoracle as po
from _states import random_periodic_grid
from test_gpu_dual_energy import CASES as DE_CASES, _de_grid
from test_gpu_parity import CASES, device_copy, method_for

pytestmark = pytest.mark.gpu

REL = 1e-12
# the seeded random states below are rough (independent values in
# neighbouring cells, shocks everywhere): rounding differences grow faster
# there than on any of the benchmark problems, so those cases are held to
# 1e-11 and print their figure
REL_ROUGH = 1e-11
DT_REL = 1e-13
THREADS = min(16, os.cpu_count() or 1)
FMA = 2  # ENZOE_PREC_DOUBLE_FMA


def _ev():
    import enzoe_vlct as ev
    return ev


VEC = {"velocity": ("velocity_x", "velocity_y", "velocity_z"),
       "bfield": ("bfield_x", "bfield_y", "bfield_z")}


def compare_active(dg, grid_o, rel=REL, g=3):
    """Largest per-cell relative difference of every field over the active
    cells (and faces), as the module docstring defines it; asserts <= rel."""
    out = {}
    for ib in range(len(dg.blocks)):
        got = dg.blocks[ib].download()
        ref = grid_o.block_arrays(ib)
        act = (slice(g, -g),) * 3
        mag = {v: np.sqrt(sum(np.square(ref[c]) for c in comps))
               for v, comps in VEC.items() if comps[0] in ref and comps[0] in got}
        for name, a in got.items():
            if name not in ref:
                continue
            b = ref[name]
            if name.startswith("bfieldi_"):
                ax = "xyz".index(name[-1])
                m = mag["bfield"]
                lo = [slice(g, -g)] * 3
                hi = [slice(g, -g)] * 3
                lo[2 - ax] = slice(g - 1, -g)      # cell below each active face
                hi[2 - ax] = slice(g, m.shape[2 - ax] - g + 1)
                fs = [slice(g, -g)] * 3
                fs[2 - ax] = slice(g, b.shape[2 - ax] - g)
                den = np.maximum(np.abs(b[tuple(fs)]),
                                 np.maximum(m[tuple(lo)], m[tuple(hi)]))
                d = np.abs(a[tuple(fs)] - b[tuple(fs)])
            else:
                vec = next((v for v, comps in VEC.items() if name in comps), None)
                den = mag[vec][act] if vec else np.abs(b[act])
                d = np.abs(a[act] - b[act])
            r = np.zeros(d.shape)
            nz = ~(d == 0)  # NaN differences included
            with np.errstate(divide="ignore", invalid="ignore"):
                r[nz] = d[nz] / den[nz]
            v = float(np.max(r))  # NaN if any difference is NaN
            if name not in out or v != v or v > out[name]:
                out[name] = v
    bad = {k: v for k, v in out.items() if not v <= rel}
    assert not bad, "relative differences above %g: %s" % (rel, bad)
    return out


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s-%g" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1], c["theta"]))
def test_contracted_cycles_within_fp64_bar_of_oracle(case):
    print("case", case, flush=True)
    grid = random_periodic_grid(mhd=case["mhd"], seed=11, n=(14, 12, 10),
                                riemann=case["riemann"], recon=case["recon"],
                                theta=case["theta"], vamp=0.6, bamp=0.08)
    dg = device_copy(grid, method_for(grid.cfg, precision=FMA))
    t = 0.0
    for _ in range(3):
        dt_o = grid.cycle(t, 10.0)
        dt_g = dg.cycle(t, 10.0)
        assert dt_g == pytest.approx(dt_o, rel=DT_REL, abs=0)
        t += dt_o
    _ev().synchronize()
    print("max rel diff per field:", compare_active(dg, grid, rel=REL_ROUGH))


@pytest.mark.parametrize("case", DE_CASES, ids=lambda c: "%s-%s-%s" % (
    "mhd" if c["mhd"] else "hd", c["riemann"], c["recon"][1]))
def test_contracted_dual_energy_within_fp64_bar_of_oracle(case):
    """The dual-energy formalism (eta = 2, so both branches of the
    synchronisation are taken)."""
    print("case", case, flush=True)
    grid = _de_grid(case)
    dg = device_copy(grid, method_for(grid.cfg, precision=FMA))
    t = 0.0
    for _ in range(3):
        dt_o = grid.cycle(t, 10.0)
        dt_g = dg.cycle(t, 10.0)
        assert dt_g == pytest.approx(dt_o, rel=DT_REL, abs=0)
        t += dt_o
    _ev().synchronize()
    print("max rel diff per field:", compare_active(dg, grid, rel=REL_ROUGH))


def test_contracted_orszag_tang_20_steps_within_fp64_bar_of_oracle():
    """The benchmark's Orszag-Tang run (PLM theta = 2, HLLD, CFL 0.4) for
    its 20 steps at 128^3, driven as the benchmark drives it: every step's
    dt and the final state within the FP64 bar of the oracle's cycles."""
    import problems
    ev = _ev()
    n = 128
    cfg = po.make_config(riemann="hlld", recon=("nn", "plm"), theta=2.0, courant=0.4)
    blk, cw = problems.orszag_tang(n)
    grid = po.Grid(cfg, (1, 1, 1), (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    view = grid.block_arrays(0)
    for k, v in blk.items():
        view[k][...] = v
    grid.refresh()
    dg = device_copy(grid, method_for(cfg, precision=FMA))
    m = dg.method
    dt = m.timestep(dg.blocks)
    t = 0.0
    for step in range(20):
        dt_o = grid.cycle(t, 1e9, nthreads=THREADS)
        assert dt == pytest.approx(dt_o, rel=DT_REL, abs=0), step
        dg.refresh()
        dt = m.compute_timestep(dg.blocks, dt)
        t += dt_o
    ev.synchronize()
    # the oracle's pressure field holds the last cycle's start state until
    # its CFL pass runs on the final state (the device's fused one did)
    dt_o = po.lib().o_vlct_timestep(ctypes.byref(cfg), ctypes.byref(grid.block(0)))
    assert dt == pytest.approx(dt_o, rel=DT_REL, abs=0)
    print("max rel diff per field:", compare_active(dg, grid))


def test_contracted_orszag_tang_256_full_size_against_fp64_build():
    """One 256^3 block for 20 steps in both builds from the same state: the
    largest per-cell relative difference of every field <= 1e-12, the
    relative L1 far below it, div B at roundoff and the conserved totals of
    the contracted run equal to the FP64 run's to roundoff."""
    import problems
    ev = _ev()
    n = 256
    blk, cw = problems.orszag_tang(n)
    runs = {}
    for prec in (0, FMA):
        m = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0, courant=0.4,
                          precision=prec)
        dg = ev.DeviceGrid(m, (1, 1, 1), (n, n, n), [cw])
        dg.upload_from([blk])
        dg.refresh()
        dt = m.timestep(dg.blocks)
        for _ in range(20):
            dg.refresh()
            dt = m.compute_timestep(dg.blocks, dt)
        ev.synchronize()
        runs[prec] = (m, dg, dt)
    m64, g64, dt64 = runs[0]
    mc, gc_, dtc = runs[FMA]
    assert dtc == pytest.approx(dt64, rel=DT_REL, abs=0)
    per, mag, mx = mc.state_l1(gc_.blocks, g64.blocks, max_rel=True)
    print("relative L1 by vector magnitude", mag)
    print("largest per-cell relative difference", mx)
    assert all(v <= REL for v in mx.values()), mx
    assert all(v <= REL for v in mag.values()), mag
    d64 = m64.diagnostics(g64.blocks)
    dc = mc.diagnostics(gc_.blocks)
    print("diagnostics fp64", d64, "contracted", dc)
    assert dc["nonfinite"] == 0
    assert dc["max_abs_divb"] <= 1e-13 * dc["max_abs_bface"], dc
    assert abs(dc["mass"] - d64["mass"]) <= 1e-12 * abs(d64["mass"])
    assert abs(dc["energy"] - d64["energy"]) <= 1e-12 * abs(d64["energy"])
    del runs, g64, gc_, m64, mc
    gc.collect()
