"""Flux corrections at refinement boundaries on the GPU against the oracle.

* enzoe_vlct_compute_store_fluxes over the five blocks of a two-level
  hydro composite (one batched launch) equals the oracle's step with
  save_fluxes_for_corrections_ bit for bit: fields and FluxData.
* enzoe_flux_data_send_to_coarse and enzoe_flux_correct equal the oracle
  bit for bit on random data, every axis, face and child quadrant.
* The composite driven through the device is conservative across the
  refinement boundary (the identity of the oracle test) and equals the
  oracle's composite bit for bit; the C++ Method list ["mhd_vlct",
  "flux_correct"] (fused and plugin-staged paths) gives the same fields.
* The mixed-precision step stores fluxes within FP32 accuracy of FP64.
"""
import ctypes

import numpy as np
import pytest

import problems
import pyoracle as po
from test_gpu_parity import method_for
from test_oracle_fluxcorr import (CONS, NAMES, SLOTS, _hydro_cfg, _oracle_block,
                                  composite_balance, run_composite_oracle)

pytestmark = pytest.mark.gpu


def _ev():
    import enzoe_vlct as ev
    return ev


def _device_block(arrays, n, g, dx):
    ev = _ev()
    m = n + 2 * g
    db = ev.DeviceBlock(m, m, m, g, (dx, dx, dx), mhd=False)
    db.upload({k: v for k, v in arrays.items() if k in db.shapes})
    return db


def _device_composite(n=16, g=4, precision=0):
    ev = _ev()
    cfg = _hydro_cfg()
    method = method_for(cfg, precision=precision)
    comp = problems.two_level_composite(n, g)
    blocks = [_device_block(arr, n, g, dx) for (_, _, _, dx, arr) in comp]
    fds = [ev.FluxData((n, n, n)) for _ in blocks]
    return cfg, method, comp, blocks, fds


def test_store_send_and_correct_match_the_oracle_composite():
    ev = _ev()
    n, g = 16, 4
    print("composite n=%d g=%d" % (n, g), flush=True)
    comp_o, grids_o, fds_o, before_o, after_o, dt_o = run_composite_oracle(n, g)
    cfg, method, comp, blocks, fds = _device_composite(n, g)
    dt = method.timestep(blocks)
    assert dt == dt_o
    ev.compute_store_fluxes(method, blocks, dt, fds)
    for (kind, child, _, _, _), fd in zip(comp[1:], fds[1:]):
        ev.flux_send_to_coarse(fd, fds[0], 0, 0, child)
    ev.flux_correct(blocks[0], fds[0], [0, 1, 0, 0, 0, 0])
    ev.synchronize()
    for i, (db, gr, fd, fd_o) in enumerate(zip(blocks, grids_o, fds, fds_o)):
        got = db.download()
        ref = gr.block_arrays(0)
        for name in NAMES:
            assert np.array_equal(got[name], ref[name]), (i, name)
        assert np.array_equal(fd.download(), fd_o), i
    bal = composite_balance(comp_o, fds_o, before_o, after_o, n)
    for name in NAMES:
        err, scale = bal[name]
        assert abs(err) <= 2e-14 * scale, name


def test_batched_send_and_correct_match_the_oracle_composite():
    """Three copies of the composite in one batched step, one batched send
    (12 links) and one batched correction (3 coarse blocks)."""
    ev = _ev()
    n, g, copies = 16, 4, 3
    print("batched composites x%d" % copies, flush=True)
    _, grids_o, fds_o, _, _, dt_o = run_composite_oracle(n, g)
    cfg = _hydro_cfg()
    method = method_for(cfg)
    comp = problems.two_level_composite(n, g)
    blocks, fds = [], []
    for _ in range(copies):
        blocks += [_device_block(arr, n, g, dx) for (_, _, _, dx, arr) in comp]
        fds += [ev.FluxData((n, n, n)) for _ in comp]
    ev.compute_store_fluxes(method, blocks, dt_o, fds)
    links, entries = [], []
    for c in range(copies):
        base = 5 * c
        for j, (_, child, _, _, _) in enumerate(comp[1:], start=1):
            links.append((fds[base + j], fds[base], 0, 0, child))
        entries.append((blocks[base], fds[base], [0, 1, 0, 0, 0, 0]))
    ev.flux_send_to_coarse_batch(method, links)
    ev.flux_correct_blocks(method, entries)
    ev.synchronize()
    for c in range(copies):
        for j in range(5):
            got = blocks[5 * c + j].download()
            ref = grids_o[j].block_arrays(0)
            for name in NAMES:
                assert np.array_equal(got[name], ref[name]), (c, j, name)
            assert np.array_equal(fds[5 * c + j].download(), fds_o[j]), (c, j)


def test_send_to_coarse_and_flux_correct_on_random_data():
    ev = _ev()
    rng = np.random.default_rng(5)
    n = (8, 6, 4)
    nf = 5
    size = po.lib().o_flux_data_size(nf, *n)
    for axis in range(3):
        for face in range(2):
            for child in [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]:
                fine = rng.standard_normal(size)
                coarse = rng.standard_normal(size)
                fd_f, fd_c = ev.FluxData(n), ev.FluxData(n)
                fd_f.upload(fine)
                fd_c.upload(coarse)
                ev.flux_send_to_coarse(fd_f, fd_c, axis, face, child)
                po.lib().o_flux_send_to_coarse(po.dptr(fine), po.dptr(coarse), nf, *n, axis,
                                               face, po.ints(child))
                ev.synchronize()
                assert np.array_equal(fd_c.download(), coarse), (axis, face, child)
    g = 2
    m = [n[a] + 2 * g for a in range(3)]
    for pattern in ([1, 0, 0, 0, 0, 0], [0, 1, 1, 0, 0, 1], [1, 1, 1, 1, 1, 1],
                    [0, 0, 0, 0, 0, 0]):
        fields = {k: rng.uniform(0.5, 2.0, (m[2], m[1], m[0])) for k in NAMES}
        fd = rng.standard_normal(size) * 0.01
        db = ev.DeviceBlock(m[0], m[1], m[2], g, (0.1, 0.1, 0.1), mhd=False)
        db.upload(fields)
        fdd = ev.FluxData(n)
        fdd.upload(fd)
        ev.flux_correct(db, fdd, pattern)
        arrs = [fields[k].copy() for k in NAMES]
        ptrs = (ctypes.POINTER(ctypes.c_double) * 5)(*[po.dptr(a) for a in arrs])
        assert po.lib().o_flux_correct(ptrs, 5, 0, po.ints(CONS), po.dptr(fd), *n, g,
                                       po.ints(pattern)) == 0
        ev.synchronize()
        got = db.download()
        for name, a in zip(NAMES, arrs):
            assert np.array_equal(got[name], a), (pattern, name)


@pytest.mark.parametrize("fused", ["true", "false"])
def test_method_list_with_flux_correct_matches_the_oracle(fused):
    ev = _ev()
    n, g = 16, 4
    print("host Method list, device_fused=%s" % fused, flush=True)
    comp_o, grids_o, _, _, _, dt_o = run_composite_oracle(n, g)
    cfg = _hydro_cfg()
    h = ev.host_lib()
    h.enzoe_host_set_fluid_props(cfg.gamma, cfg.density_floor, cfg.pressure_floor)
    comp = problems.two_level_composite(n, g)
    blocks = [_device_block(arr, n, g, dx) for (_, _, _, dx, arr) in comp]
    params = ("mhd_choice = no_bfield; riemann_solver = hllc; reconstruct_method = plm; "
              "theta_limiter = 2.0; courant = 0.4; device_fused = %s" % fused)
    amr = ev.HostAmr(params, "min_digits = [density, 15.0]", blocks,
                     [0] + [1] * 4)
    amr.set_face(0, 0, 1, 1)
    for i, (_, child, _, _, _) in enumerate(comp[1:], start=1):
        amr.set_face(i, 0, 0, 0, coarse_neighbor=0, child=child)
    digits, ok = amr.cycle(dt_o)
    ev.synchronize()
    for i, (db, gr) in enumerate(zip(blocks, grids_o)):
        got = db.download()
        ref = gr.block_arrays(0)
        for name in NAMES:
            assert np.array_equal(got[name], ref[name]), (fused, i, name)
    # first cycle: the sums are their own reference
    assert ok and all(d == float("inf") for d in digits), digits


def test_mixed_precision_stores_fluxes_within_fp32_accuracy():
    ev = _ev()
    n, g = 16, 4
    _, m64, _, b64, f64 = _device_composite(n, g, precision=0)
    _, m32, _, b32, f32 = _device_composite(n, g, precision=1)
    dt = m64.timestep(b64)
    ev.compute_store_fluxes(m64, b64, dt, f64)
    ev.compute_store_fluxes(m32, b32, dt, f32)
    ev.synchronize()
    worst = 0.0
    for a, b in zip(f64, f32):
        x, y = a.download(), b.download()
        scale = np.abs(x).max()
        worst = max(worst, float(np.abs(x - y).max() / scale))
    print("mixed vs fp64 FluxData, max |diff| / max |flux| = %.3e" % worst)
    assert worst < 1e-5
