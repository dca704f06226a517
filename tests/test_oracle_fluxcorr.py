"""Pins of the oracle's flux corrections at refinement boundaries
(oracle/fluxcorr_oracle.c), on the CPU.

* FluxData layout: the single-array order of FluxData::allocate
  (src/Cello/data_FluxData.cpp:57-87) walked independently.
* FaceFluxes::coarsen (data_FaceFluxes.cpp:128-250): closed forms for
  constant and linear fine fluxes on every axis and child quadrant.
* save_fluxes_for_corrections_ (EnzoMethodMHDVlct.cpp:250-330): a uniform
  flow saves dt/dx times the exact Euler flux on every face; on a smooth
  block the saved fluxes telescope with the change of the conserved totals
  of the active region (which pins the face indices g-1 and m-g-1).
* flux_correct_ (problem_MethodFluxCorrect.cpp:425-560): closed form on a
  small block (signs, faces, corners, conservative fields through rho).
* End to end: a coarse block and the four fine blocks across its high-x
  face change their conserved totals by exactly the fluxes through the
  outer faces once the coarse block is corrected, and not before.
"""
import ctypes

import numpy as np
import pytest

import problems
import pyoracle as po

NAMES = ("density", "total_energy", "velocity_x", "velocity_y", "velocity_z")
SLOTS = [0, 4, 1, 2, 3]
CONS = [0, 1, 1, 1, 1]


def face_shape(axis, n):
    nx, ny, nz = n
    return {0: (nz, ny), 1: (nz, nx), 2: (ny, nx)}[axis]


def fd_view(fd, i_f, axis, face, nbr, n):
    off = po.lib().o_flux_data_offset(i_f, axis, face, nbr, *n)
    shp = face_shape(axis, n)
    return fd[off:off + shp[0] * shp[1]].reshape(shp)


def test_flux_data_layout_follows_the_allocation_loop():
    n = (4, 6, 8)
    nf = 3
    L = po.lib()
    pos = 0
    for i_f in range(nf):
        for axis in range(3):
            m = int(np.prod(face_shape(axis, n)))
            for face in range(2):
                assert L.o_flux_data_offset(i_f, axis, face, 0, *n) == pos
                pos += m
                assert L.o_flux_data_offset(i_f, axis, face, 1, *n) == pos
                pos += m
    assert L.o_flux_data_size(nf, *n) == pos == nf * 4 * (6 * 8 + 4 * 8 + 4 * 6)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_coarsen_closed_forms(axis):
    n = (8, 6, 4)
    shp = face_shape(axis, n)
    hv, hu = shp[0] // 2, shp[1] // 2
    L = po.lib()
    for child in [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]:
        a0 = 1 if axis == 0 else 0
        a1 = 1 if axis == 2 else 2
        cu, cv = child[a0], child[a1]
        # constant fine flux: each coarse face of the quadrant gets half
        fine = np.full(shp, 3.0)
        out = np.full(shp, 7.0)
        L.o_face_fluxes_coarsen(po.dptr(fine), po.dptr(out), axis, *n, *child)
        expect = np.zeros(shp)
        expect[cv * hv:(cv + 1) * hv, cu * hu:(cu + 1) * hu] = 1.5
        assert np.array_equal(out, expect)
        # linear fine flux f(u, v) = u + 100 v: 0.5 * mean of the 4 fine faces
        v, u = np.meshgrid(np.arange(shp[0]), np.arange(shp[1]), indexing="ij")
        fine = (u + 100.0 * v).astype(np.float64)
        L.o_face_fluxes_coarsen(po.dptr(fine), po.dptr(out), axis, *n, *child)
        expect = np.zeros(shp)
        vv, uu = np.meshgrid(np.arange(hv), np.arange(hu), indexing="ij")
        expect[cv * hv:(cv + 1) * hv, cu * hu:(cu + 1) * hu] = \
            0.5 * ((2 * uu + 0.5) + 100.0 * (2 * vv + 0.5))
        assert np.array_equal(out, expect)


def _hydro_cfg():
    return po.make_config(riemann="hllc", recon=("nn", "plm"), theta=2.0, mhd=False,
                          courant=0.4, density_floor=1e-30, pressure_floor=1e-30)


def _oracle_block(cfg, arrays, n, g, lower, dx):
    up = tuple(lower[a] + n * dx for a in range(3))
    grid = po.Grid(cfg, (1, 1, 1), (n, n, n), lower, up, ghost=g)
    view = grid.block_arrays(0)
    for k, v in arrays.items():
        if k in view:
            view[k][...] = v
    return grid


def _compute_fc(cfg, grid, dt, n, g):
    fd = np.zeros(po.lib().o_flux_data_size(len(SLOTS), n, n, n))
    rc = po.lib().o_vlct_compute_fc(ctypes.byref(cfg), ctypes.byref(grid.block(0)), dt,
                                    g, len(SLOTS), po.ints(SLOTS), po.dptr(fd))
    assert rc == 0
    return fd


def test_saved_fluxes_of_a_uniform_flow_are_the_euler_flux():
    cfg = _hydro_cfg()
    n, g, gamma = 8, 4, 5.0 / 3.0
    m = n + 2 * g
    rho, v, p = 1.3, (0.4, -0.2, 0.1), 0.9
    etot = p / ((gamma - 1.0) * rho) + 0.5 * sum(c * c for c in v)
    arrays = {"density": np.full((m,) * 3, rho), "total_energy": np.full((m,) * 3, etot)}
    for a, c in enumerate("xyz"):
        arrays["velocity_" + c] = np.full((m,) * 3, v[a])
    dx = 1.0 / n
    grid = _oracle_block(cfg, arrays, n, g, (0.0, 0.0, 0.0), dx)
    dt = 0.01
    fd = _compute_fc(cfg, grid, dt, n, g)
    E = rho * etot
    for axis in range(3):
        vn = v[axis]
        mom = [rho * vn * v[b] + (p if b == axis else 0.0) for b in range(3)]
        exact = {"density": rho * vn, "total_energy": (E + p) * vn,
                 "velocity_x": mom[0], "velocity_y": mom[1], "velocity_z": mom[2]}
        for i_f, name in enumerate(NAMES):
            for face in range(2):
                blk = fd_view(fd, i_f, axis, face, 0, (n, n, n))
                nb = fd_view(fd, i_f, axis, face, 1, (n, n, n))
                np.testing.assert_allclose(blk, dt / dx * exact[name], rtol=1e-13,
                                           atol=1e-16)
                assert not nb.any()
    view = grid.block_arrays(0)
    np.testing.assert_allclose(view["density"], rho, rtol=1e-14)


def _totals(view, g, dV):
    sl = (slice(g, -g),) * 3
    rho = view["density"][sl]
    out = {"density": rho.sum() * dV, "total_energy": (rho * view["total_energy"][sl]).sum() * dV}
    for c in "xyz":
        out["velocity_" + c] = (rho * view["velocity_" + c][sl]).sum() * dV
    return out


def _abs_totals(view, g, dV):
    """Sums of |conserved density| * dV: the scale of the roundoff."""
    sl = (slice(g, -g),) * 3
    rho = view["density"][sl]
    out = {"density": np.abs(rho).sum() * dV,
           "total_energy": np.abs(rho * view["total_energy"][sl]).sum() * dV}
    for c in "xyz":
        out["velocity_" + c] = np.abs(rho * view["velocity_" + c][sl]).sum() * dV
    return out


def _face_sum(fd, i_f, axis, face, nbr, n):
    return float(fd_view(fd, i_f, axis, face, nbr, n).sum())


def test_saved_fluxes_telescope_with_the_active_totals():
    cfg = _hydro_cfg()
    n, g = 16, 4
    # not a whole period of the state, so the faces carry a net flux
    dx, lower = 0.6 / n, (0.1, 0.2, 0.3)
    arrays = problems.smooth_hydro_block(lower, dx, n, g)
    grid = _oracle_block(cfg, arrays, n, g, lower, dx)
    dt = 0.5 * po.lib().o_vlct_timestep(ctypes.byref(cfg), ctypes.byref(grid.block(0)))
    dV = dx ** 3
    before = _totals(grid.block_arrays(0), g, dV)
    size = _abs_totals(grid.block_arrays(0), g, dV)
    fd = _compute_fc(cfg, grid, dt, n, g)
    after = _totals(grid.block_arrays(0), g, dV)
    for i_f, name in enumerate(NAMES):
        through = sum((1.0 if face == 0 else -1.0) * _face_sum(fd, i_f, axis, face, 0, (n,) * 3)
                      for axis in range(3) for face in range(2)) * dV
        scale = size[name]
        assert abs((after[name] - before[name]) - through) <= 1e-13 * scale, name
        assert abs(through) > 1e-6 * scale  # the faces carry a net flux


def test_flux_correct_closed_form():
    n, g = 4, 1
    m = n + 2 * g
    fields = {k: np.full((m, m, m), val) for k, val in
              zip(NAMES, (2.0, 3.0, 0.5, -0.25, 0.125))}
    fd = np.zeros(po.lib().o_flux_data_size(5, n, n, n))
    # block 0.3 / neighbour 0.1 on every face of every field
    for i_f in range(5):
        for axis in range(3):
            for face in range(2):
                fd_view(fd, i_f, axis, face, 0, (n,) * 3)[...] = 0.3 + 0.01 * i_f
                fd_view(fd, i_f, axis, face, 1, (n,) * 3)[...] = 0.1
    correct = [0, 1, 0, 0, 1, 0]          # x high and z low have finer neighbours
    cons = [0, 1, 1, 1, 0]                # velocity_z corrected directly
    arrs = [fields[k] for k in NAMES]
    ptrs = (ctypes.POINTER(ctypes.c_double) * 5)(*[po.dptr(a) for a in arrs])
    assert po.lib().o_flux_correct(ptrs, 5, 0, po.ints(cons), po.dptr(fd), n, n, n, g,
                                   po.ints(correct)) == 0
    A = slice(g, g + n)
    xs, zs = np.zeros((n, n, n)), np.zeros((n, n, n))  # (z, y, x) active
    xs[:, :, n - 1] = 1.0                  # x high: +(b - n)
    zs[0, :, :] = 1.0                      # z low:  -(b - n)
    d = [0.3 + 0.01 * i_f - 0.1 for i_f in range(5)]
    rho_new = 2.0 + d[0] * xs - d[0] * zs
    np.testing.assert_allclose(fields["density"][A, A, A], rho_new, rtol=0, atol=1e-15)
    touched = (xs + zs) > 0
    for i_f, (name, base) in enumerate(zip(NAMES, (2.0, 3.0, 0.5, -0.25, 0.125))):
        if i_f == 0:
            continue
        corr = d[i_f] * xs - d[i_f] * zs
        if cons[i_f]:
            expect = np.where(touched, (2.0 * base + corr) / rho_new, base)
        else:
            expect = base + corr
        np.testing.assert_allclose(fields[name][A, A, A], expect, rtol=1e-15, atol=1e-15)
    # ghost zones are left alone
    assert np.all(fields["density"][0] == 2.0) and np.all(fields["density"][:, :, -1] == 2.0)
    # a conservative field needs density
    assert po.lib().o_flux_correct(ptrs, 5, -1, po.ints(cons), po.dptr(fd), n, n, n, g,
                                   po.ints(correct)) == -1


def run_composite_oracle(n=16, g=4, correct=True):
    """The two-level composite through the oracle: one step with a common
    dt, fine fluxes sent to the coarse block, coarse block corrected.
    Returns (blocks, flux data, totals before, totals after)."""
    cfg = _hydro_cfg()
    comp = problems.two_level_composite(n, g)
    grids = [_oracle_block(cfg, arr, n, g, lower, dx) for (_, _, lower, dx, arr) in comp]
    dt = min(po.lib().o_vlct_timestep(ctypes.byref(cfg), ctypes.byref(gr.block(0)))
             for gr in grids)
    before = [_totals(gr.block_arrays(0), g, c[3] ** 3) for gr, c in zip(grids, comp)]
    for b, gr, c in zip(before, grids, comp):
        b["scale"] = _abs_totals(gr.block_arrays(0), g, c[3] ** 3)
    fds = [_compute_fc(cfg, gr, dt, n, g) for gr in grids]
    for (kind, child, _, _, _), fd in zip(comp[1:], fds[1:]):
        po.lib().o_flux_send_to_coarse(po.dptr(fd), po.dptr(fds[0]), 5, n, n, n, 0, 0,
                                       po.ints(child))
    if correct:
        view = grids[0].block_arrays(0)
        ptrs = (ctypes.POINTER(ctypes.c_double) * 5)(*[po.dptr(view[k]) for k in NAMES])
        assert po.lib().o_flux_correct(ptrs, 5, 0, po.ints(CONS), po.dptr(fds[0]), n, n,
                                       n, g, po.ints([0, 1, 0, 0, 0, 0])) == 0
    after = [_totals(gr.block_arrays(0), g, c[3] ** 3) for gr, c in zip(grids, comp)]
    return comp, grids, fds, before, after, dt


def composite_balance(comp, fds, before, after, n):
    """Per field: (change of the composite totals) - (fluxes through the
    outer faces), and the scale it is measured against."""
    out = {}
    for i_f, name in enumerate(NAMES):
        change = sum(a[name] for a in after) - sum(b[name] for b in before)
        scale = sum(b["scale"][name] for b in before)
        through = 0.0
        for (kind, _, _, dx, _), fd in zip(comp, fds):
            for axis in range(3):
                for face in range(2):
                    interface = (axis == 0 and ((kind == "coarse" and face == 1) or
                                                (kind == "fine" and face == 0)))
                    if interface:
                        continue
                    s = (1.0 if face == 0 else -1.0)
                    through += s * _face_sum(fd, i_f, axis, face, 0, (n,) * 3) * dx ** 3
        out[name] = (change - through, scale)
    return out


def test_two_level_composite_is_conservative_after_correction():
    n = 16
    comp, _, fds, before, after, _ = run_composite_oracle(n, correct=True)
    bal = composite_balance(comp, fds, before, after, n)
    comp0, _, fds0, before0, after0, _ = run_composite_oracle(n, correct=False)
    bal0 = composite_balance(comp0, fds0, before0, after0, n)
    for name in NAMES:
        err, scale = bal[name]
        err0, _ = bal0[name]
        print(name, "corrected %.3e  uncorrected %.3e" % (err / scale, err0 / scale))
        assert abs(err) <= 2e-14 * scale, name
        assert abs(err0) >= 1e-9 * scale, name
