"""Per-function pins of the CPU oracle: closed forms, invariants and
brute-force cases chosen so that a plausible slip (dropped term, wrong sign,
swapped operand) fails at least one of them."""
import ctypes

import numpy as np
import pytest

import pyoracle as po

GAMMA = 5.0 / 3.0


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- limiters

def _minmod2(a, b):
    if a * b <= 0:
        return 0.0
    return np.sign(a) * min(abs(a), abs(b))


def test_limiter_theta1_is_minmod_and_theta2_is_mc():
    rng = np.random.default_rng(1)
    L = po.lib()
    vals = list(rng.normal(size=(400, 3))) + [
        (0, 0, 0), (0, 1, 1), (1, 1, 0), (1, 0, 1), (-1, 0, 2), (2, 1, 0)]
    for vm1, v, vp1 in vals:
        dl, dr = v - vm1, vp1 - v
        # theta = 1: minmod(dl, dr)
        assert L.o_limiter_enzo(vm1, v, vp1, 1.0) == pytest.approx(
            _minmod2(dl, dr), abs=1e-15)
        # theta = 2: monotonized central, minmod(2dl, 2dr, (dl+dr)/2)
        mc = _minmod2(2 * dl, 2 * dr)
        if mc != 0:
            c = 0.5 * (dl + dr)
            mc = np.sign(mc) * min(abs(mc), abs(c))
        assert L.o_limiter_enzo(vm1, v, vp1, 2.0) == pytest.approx(mc, abs=1e-15)


def test_athena_limiter_is_van_leer_harmonic_mean():
    rng = np.random.default_rng(2)
    L = po.lib()
    for vm1, v, vp1 in rng.normal(size=(400, 3)):
        a, b = v - vm1, vp1 - v
        vl = (a * abs(b) + abs(a) * b) / (abs(a) + abs(b))  # van Leer (1974)
        assert L.o_limiter_athena(vm1, v, vp1) == pytest.approx(vl, rel=1e-13,
                                                                abs=1e-15)
    assert L.o_limiter_athena(1.0, 2.0, 2.0) == 0.0  # flat side -> 0


# ---------------------------------------------------------- Riemann solvers

def _mhd_phys_flux(w, gamma):
    """Ideal-MHD flux along i from primitives in LUT order
    (rho, Bi, Bj, Bk, vi, vj, vk, p), textbook form."""
    rho, bi, bj, bk, vi, vj, vk, p = w
    b2 = bi * bi + bj * bj + bk * bk
    v2 = vi * vi + vj * vj + vk * vk
    E = p / (gamma - 1) + 0.5 * rho * v2 + 0.5 * b2
    pt = p + 0.5 * b2
    vb = vi * bi + vj * bj + vk * bk
    return np.array([rho * vi, 0.0, bj * vi - bi * vj, bk * vi - bi * vk,
                     rho * vi * vi + pt - bi * bi, rho * vi * vj - bi * bj,
                     rho * vi * vk - bi * bk, (E + pt) * vi - vb * bi])


def _hd_phys_flux(w, gamma):
    rho, vi, vj, vk, p = w
    E = p / (gamma - 1) + 0.5 * rho * (vi * vi + vj * vj + vk * vk)
    return np.array([rho * vi, rho * vi * vi + p, rho * vi * vj,
                     rho * vi * vk, (E + p) * vi])


def _face(fn, wl, wr, n):
    out = np.zeros(n)
    fn(po.dptr(_d(wl)), po.dptr(_d(wr)), GAMMA, po.dptr(out))
    return out


def _rand_mhd_state(rng):
    return np.array([rng.uniform(0.5, 2.0), rng.normal(), rng.normal(),
                     rng.normal(), rng.normal(), rng.normal(), rng.normal(),
                     rng.uniform(0.3, 2.0)])


@pytest.mark.parametrize("solver", ["hlld", "hlle"])
def test_mhd_riemann_consistency(solver):
    """F(U, U) equals the physical flux (any consistent Riemann solver)."""
    rng = np.random.default_rng(3)
    fn = po.lib().o_hlld_face if solver == "hlld" else po.lib().o_hlle_mhd_face
    for _ in range(300):
        w = _rand_mhd_state(rng)
        f = _face(fn, w, w, 8)
        np.testing.assert_allclose(f, _mhd_phys_flux(w, GAMMA), rtol=1e-12,
                                   atol=1e-12)


def test_hllc_consistency():
    rng = np.random.default_rng(4)
    for _ in range(300):
        w = np.array([rng.uniform(0.5, 2), rng.normal(), rng.normal(),
                      rng.normal(), rng.uniform(0.3, 2)])
        f = _face(po.lib().o_hllc_face, w, w, 5)
        np.testing.assert_allclose(f, _hd_phys_flux(w, GAMMA), rtol=1e-12,
                                   atol=1e-12)


@pytest.mark.parametrize("solver", ["hlld", "hlle"])
def test_mhd_riemann_supersonic_upwinding(solver):
    """All waves move right (left) -> the flux is exactly F(U_L) (F(U_R))."""
    rng = np.random.default_rng(5)
    fn = po.lib().o_hlld_face if solver == "hlld" else po.lib().o_hlle_mhd_face
    for sign in (+1, -1):
        for _ in range(50):
            wl, wr = _rand_mhd_state(rng), _rand_mhd_state(rng)
            wr[1] = wl[1]  # continuous normal B
            wl[4] = sign * 40.0 + 0.1 * rng.normal()
            wr[4] = sign * 40.0 + 0.1 * rng.normal()
            f = _face(fn, wl, wr, 8)
            up = wl if sign > 0 else wr
            np.testing.assert_allclose(f, _mhd_phys_flux(up, GAMMA),
                                       rtol=1e-12, atol=1e-11)


def test_hlld_mirror_symmetry():
    """Reflecting the problem (x -> -x) negates the odd flux components."""
    rng = np.random.default_rng(6)
    fn = po.lib().o_hlld_face
    for _ in range(200):
        wl, wr = _rand_mhd_state(rng), _rand_mhd_state(rng)
        wr[1] = wl[1]
        f = _face(fn, wl, wr, 8)
        # mirror: swap sides, flip v_i and B_i
        ml, mr = wr.copy(), wl.copy()
        ml[4] *= -1; mr[4] *= -1; ml[1] *= -1; mr[1] *= -1
        g = _face(fn, ml, mr, 8)
        # rho, Bj, Bk, vi-momentum fluxes change sign?  mass flux flips sign,
        # normal momentum flux is even, transverse momentum / B fluxes flip,
        # energy flux flips.
        expect = np.array([-f[0], 0.0, -f[2], -f[3], f[4], -f[5], -f[6], -f[7]])
        np.testing.assert_allclose(g, expect, rtol=1e-9, atol=1e-10)


def test_hlld_reduces_to_hydro_contact_for_zero_field():
    """With B = 0 the HLLD fan collapses to HLLC-like structure: a stationary
    contact (equal pressure, zero velocity, density jump) carries no flux
    except the pressure term."""
    wl = np.array([2.0, 0, 0, 0, 0, 0, 0, 1.0])
    wr = np.array([0.5, 0, 0, 0, 0, 0, 0, 1.0])
    f = _face(po.lib().o_hlld_face, wl, wr, 8)
    np.testing.assert_allclose(f, [0, 0, 0, 0, 1.0, 0, 0, 0], atol=1e-14)
    g = _face(po.lib().o_hllc_face, np.array([2.0, 0, 0, 0, 1.0]),
              np.array([0.5, 0, 0, 0, 1.0]), 5)
    np.testing.assert_allclose(g, [0, 1.0, 0, 0, 0], atol=1e-14)


# ------------------------------------------------------------- CT invariants

def _div_face_b(bfx, bfy, bfz, cw, g):
    """Discrete divergence of the face field over the active cells."""
    dx, dy, dz = cw
    div = ((bfx[:, :, 1:] - bfx[:, :, :-1]) / dx +
           (bfy[:, 1:, :] - bfy[:, :-1, :]) / dy +
           (bfz[1:, :, :] - bfz[:-1, :, :]) / dz)
    return div[g:-g, g:-g, g:-g]


def test_update_bfield_preserves_div_b_for_arbitrary_emfs():
    """The CT face update is a discrete curl: it leaves div(B) unchanged for
    any edge EMFs (telescoping identity)."""
    rng = np.random.default_rng(7)
    mz, my, mx = 9, 10, 11
    cw = np.array([0.3, 0.5, 0.7])
    bf = [rng.normal(size=(mz, my, mx + 1)), rng.normal(size=(mz, my + 1, mx)),
          rng.normal(size=(mz + 1, my, mx))]
    edge = [_d(rng.normal(size=(mz - 1, my - 1, mx))),
            _d(rng.normal(size=(mz - 1, my, mx - 1))),
            _d(rng.normal(size=(mz, my - 1, mx - 1)))]
    eptr = (ctypes.POINTER(ctypes.c_double) * 3)(*[po.dptr(e) for e in edge])
    out = [b.copy() for b in bf]
    for d in range(3):
        po.lib().o_update_bfield(po.dptr(cw), d, eptr, po.dptr(bf[d]),
                                 po.dptr(out[d]), 0.37, mz, my, mx, 0)
    d0 = _div_face_b(*bf, cw, 1)
    d1 = _div_face_b(*out, cw, 1)
    np.testing.assert_allclose(d1, d0, atol=1e-12)
    assert not np.allclose(out[0], bf[0])  # the update did something


def test_center_bfield_is_face_average():
    rng = np.random.default_rng(8)
    mz, my, mx = 5, 6, 7
    bfy = _d(rng.normal(size=(mz, my + 1, mx)))
    bc = np.zeros((mz, my, mx))
    po.lib().o_compute_center_bfield(1, po.dptr(bc), po.dptr(bfy), mz, my, mx, 0)
    np.testing.assert_array_equal(bc, 0.5 * (bfy[:, :-1, :] + bfy[:, 1:, :]))


# ------------------------------------------------ whole-method invariants

def _conserved_totals(grid):
    rho = grid.gather("density")
    out = [rho.sum()]
    for f in ("velocity_x", "velocity_y", "velocity_z"):
        out.append((rho * grid.gather(f)).sum())
    out.append((rho * grid.gather("total_energy")).sum())
    return np.array(out)


from _states import random_periodic_grid as _random_periodic_grid  # noqa: E402


@pytest.mark.parametrize("mhd", [True, False])
def test_conservation_and_div_b_over_periodic_steps(mhd):
    grid = _random_periodic_grid(mhd=mhd)
    tot0 = _conserved_totals(grid)
    t = 0.0
    for _ in range(4):
        t += grid.cycle(t, 1.0)
    tot1 = _conserved_totals(grid)
    # mass, momentum and energy of a periodic box change only by roundoff
    np.testing.assert_allclose(tot1, tot0, rtol=0, atol=1e-12 * np.abs(tot0).max())
    if mhd:
        arr = grid.block_arrays(0)
        cw = list(grid.block(0).cw)
        div = _div_face_b(arr["bfieldi_x"], arr["bfieldi_y"], arr["bfieldi_z"],
                          cw, 3)
        assert np.abs(div).max() < 1e-13


@pytest.mark.parametrize("mhd", [True, False])
def test_timestep_closed_form_uniform_state(mhd):
    cfg = po.make_config(riemann="hlld" if mhd else "hllc", mhd=mhd, courant=0.4)
    grid = po.Grid(cfg, (1, 1, 1), (8, 6, 4), (0, 0, 0), (1.0, 0.6, 0.8))
    arr = grid.block_arrays(0)
    rho, p, v, B = 1.3, 0.7, (0.2, -0.5, 0.1), (0.3, 0.4, -0.2)
    arr["density"][...] = rho
    for q, f in enumerate(("velocity_x", "velocity_y", "velocity_z")):
        arr[f][...] = v[q]
    b2 = 0.0
    if mhd:
        for q, f in enumerate(("bfield_x", "bfield_y", "bfield_z")):
            arr[f][...] = B[q]
        b2 = sum(b * b for b in B)
    arr["total_energy"][...] = (p / (GAMMA - 1) + 0.5 * rho * sum(x * x for x in v)
                                + 0.5 * b2) / rho
    dt = po.lib().o_vlct_timestep(ctypes.byref(cfg), ctypes.byref(grid.block(0)))
    cf = np.sqrt(GAMMA * p / rho + b2 / rho)
    cw = (1.0 / 8, 0.6 / 6, 0.8 / 4)
    expect = 0.4 * min(cw[a] / (abs(v[a]) + cf) for a in range(3))
    assert dt == pytest.approx(expect, rel=1e-13)
    np.testing.assert_allclose(arr["pressure"], p, rtol=1e-13)


def test_timestep_includes_ghost_cells():
    """The CFL minimum runs over every cell, ghost zones included
    (EnzoMHDIntegratorStageCommands.cpp:336-337, 361-362)."""
    cfg = po.make_config()
    grid = _random_periodic_grid(mhd=True)
    b = grid.block(0)
    dt_all = po.lib().o_vlct_timestep(ctypes.byref(cfg), ctypes.byref(b))
    arr = grid.block_arrays(0)
    # a corner ghost cell only; keep its pressure by adding the kinetic energy
    v_old = arr["velocity_x"][0, 0, 0]
    arr["velocity_x"][0, 0, 0] = 50.0
    arr["total_energy"][0, 0, 0] += 0.5 * (50.0 ** 2 - v_old ** 2)
    dt_ghost = po.lib().o_vlct_timestep(ctypes.byref(cfg), ctypes.byref(b))
    assert dt_ghost < 0.5 * dt_all
