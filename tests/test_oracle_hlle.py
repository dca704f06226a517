"""Pins of the oracle's MHD HLLE solver ("hlle": the HLL flux with the
Einfeldt wavespeeds, the Roe-averaged fast speed of Stone et al. 2008,
ApJS 178, 137, appendix B, eqs. B17-B18;
src/Enzo/hydro-mhd/riemann/EnzoRiemannHLL.hpp:44-172, 238-320).

The reference ships no HLLE golden values, so the expected fluxes here are
computed from the mathematics, written independently of the oracle:
textbook ideal-MHD fluxes and fast speeds, Roe averages (sqrt(rho)-weighted
velocity and enthalpy, rho~ = sqrt(rho_l rho_r), the cross-weighted B~ of
Cargo & Gallice that makes the Roe matrix exact), the Roe fast speed
c~f^2 = (A + sqrt(A^2 - 4 a~^2 B~x^2/rho~)) / 2 with
A = a~^2 + (B~x^2 + (gamma' - (gamma'-1) Y) B~t^2) / rho~,
a~^2 = gamma' (H~ - v~^2/2 - B~^2/rho~) - (gamma'-1) X, gamma' = gamma - 1,
Y = (rho_l + rho_r) / (2 rho~), and the HLL average
F = (b+ F_l - b- F_r + b+ b- (U_r - U_l)) / (b+ - b-) with
b+ = max(v~ + c~f, v_r + cf_r, 0), b- = min(v~ - c~f, v_l - cf_l, 0).

The reference's X term is (dB_j^2 + dB_k^2) / 2 / (sqrt(rho_l) + sqrt(rho_r))
(EnzoRiemannHLL.hpp:154-156), where Stone+08 and Athena divide by the
square of that sum.  The cases are chosen where every such form agrees:
equal transverse B (X = 0), gamma = 2 (gamma' - 1 = 0: X and the Y term
drop out), and sqrt(rho_l) + sqrt(rho_r) = 1 (both denominators are 1).
Each case asserts that the Roe speed, not a one-sided fast speed, bounds
the fan on the side it tests, so the Roe average is what the flux measures.
A swapped B~ weighting, a dropped 1/2 in Y or in X, or a wrong sign of X
changes these fluxes well beyond the tolerance.
"""
import math

import numpy as np
import pytest

import pyoracle as po

# oracle state order (EnzoRiemannImpl.hpp:40-56): density, bfield_i, _j, _k,
# velocity_i, _j, _k, pressure
RHO, BI, BJ, BK, VI, VJ, VK, P = range(8)


def _state(rho, v, b, p):
    w = np.zeros(8)
    w[RHO], w[P] = rho, p
    w[VI], w[VJ], w[VK] = v
    w[BI], w[BJ], w[BK] = b
    return w


def _cons(w, gamma):
    v = np.array([w[VI], w[VJ], w[VK]])
    b = np.array([w[BI], w[BJ], w[BK]])
    e = w[P] / (gamma - 1) + 0.5 * w[RHO] * v.dot(v) + 0.5 * b.dot(b)
    u = np.zeros(8)
    u[RHO] = w[RHO]
    u[VI], u[VJ], u[VK] = w[RHO] * v
    u[BI], u[BJ], u[BK] = b
    u[P] = e
    return u


def _flux(w, gamma):
    """Ideal-MHD flux along i (the normal-B entry is 0, as the solvers set
    it)."""
    v = np.array([w[VI], w[VJ], w[VK]])
    b = np.array([w[BI], w[BJ], w[BK]])
    rho, p = w[RHO], w[P]
    pt = p + 0.5 * b.dot(b)
    e = _cons(w, gamma)[P]
    f = np.zeros(8)
    f[RHO] = rho * v[0]
    mom = rho * v[0] * v - b[0] * b
    mom[0] += pt
    f[VI], f[VJ], f[VK] = mom
    f[BJ] = b[1] * v[0] - b[0] * v[1]
    f[BK] = b[2] * v[0] - b[0] * v[2]
    f[P] = (e + pt) * v[0] - b[0] * v.dot(b)
    return f


def _fast(rho, a2, bx2_over_rho, b2_over_rho):
    s = a2 + b2_over_rho
    return math.sqrt(0.5 * (s + math.sqrt(s * s - 4.0 * a2 * bx2_over_rho)))


def _roe_fast_speed(wl, wr, gamma):
    """Stone+08 B17-B18 (Athena's esys_roe_adb_mhd) with X in the form
    every variant shares at these inputs; returns (v~_i, c~f)."""
    sl, sr = math.sqrt(wl[RHO]), math.sqrt(wr[RHO])
    rho_t = sl * sr
    v_t = (sl * np.array([wl[VI], wl[VJ], wl[VK]]) +
           sr * np.array([wr[VI], wr[VJ], wr[VK]])) / (sl + sr)
    ul, ur = _cons(wl, gamma), _cons(wr, gamma)
    hl = (ul[P] + wl[P] + 0.5 * (wl[BI] ** 2 + wl[BJ] ** 2 + wl[BK] ** 2)) / wl[RHO]
    hr = (ur[P] + wr[P] + 0.5 * (wr[BI] ** 2 + wr[BJ] ** 2 + wr[BK] ** 2)) / wr[RHO]
    h_t = (sl * hl + sr * hr) / (sl + sr)
    # Roe B: each side weighted by the other side's sqrt(rho)
    bt = np.array([(sr * wl[BJ] + sl * wr[BJ]) / (sl + sr),
                   (sr * wl[BK] + sl * wr[BK]) / (sl + sr)])
    bx = wl[BI]
    gp = gamma - 1.0
    dbt2 = (wl[BJ] - wr[BJ]) ** 2 + (wl[BK] - wr[BK]) ** 2
    X = 0.5 * dbt2 / (sl + sr)   # (sl + sr) == 1 wherever dbt2 * (gp - 1) != 0
    Y = 0.5 * (wl[RHO] + wr[RHO]) / rho_t
    b2 = bx * bx + bt.dot(bt)
    a2 = gp * (h_t - 0.5 * v_t.dot(v_t) - b2 / rho_t) - (gp - 1.0) * X
    va2 = (bx * bx + (gp - (gp - 1.0) * Y) * bt.dot(bt)) / rho_t
    return v_t[0], _fast(rho_t, a2, bx * bx / rho_t, va2)


def _hlle_expected(wl, wr, gamma):
    vroe, croe = _roe_fast_speed(wl, wr, gamma)
    cl = _fast(wl[RHO], gamma * wl[P] / wl[RHO], wl[BI] ** 2 / wl[RHO],
               (wl[BI] ** 2 + wl[BJ] ** 2 + wl[BK] ** 2) / wl[RHO])
    cr = _fast(wr[RHO], gamma * wr[P] / wr[RHO], wr[BI] ** 2 / wr[RHO],
               (wr[BI] ** 2 + wr[BJ] ** 2 + wr[BK] ** 2) / wr[RHO])
    bp = max(vroe + croe, wr[VI] + cr, 0.0)
    bm = min(vroe - croe, wl[VI] - cl, 0.0)
    fl, fr = _flux(wl, gamma), _flux(wr, gamma)
    ul, ur = _cons(wl, gamma), _cons(wr, gamma)
    f = (bp * fl - bm * fr + bp * bm * (ur - ul)) / (bp - bm)
    f[BI] = 0.0
    roe_sides = (vroe + croe > wr[VI] + cr, vroe - croe < wl[VI] - cl)
    return f, roe_sides


def _oracle_flux(wl, wr, gamma):
    out = np.zeros(8)
    po.lib().o_hlle_mhd_face(po.dptr(np.ascontiguousarray(wl)),
                             po.dptr(np.ascontiguousarray(wr)), gamma, po.dptr(out))
    return out


def _check(wl, wr, gamma, need_roe):
    exp, roe_sides = _hlle_expected(wl, wr, gamma)
    assert any(roe_sides[s] for s in need_roe), "case does not reach the Roe speed"
    got = _oracle_flux(wl, wr, gamma)
    scale = max(1.0, float(np.abs(exp).max()))
    np.testing.assert_allclose(got, exp, rtol=1e-13, atol=1e-14 * scale)


GAMMA = 5.0 / 3.0


def test_converging_flow_closed_form():
    """Equal rho, p, B and v_i = +-u: the Roe state is at rest with
    H~ = H, so a~^2 = gamma p / rho + (gamma-1) u^2 / 2 (X = 0, Y = 1) and
    the momentum flux is p_t - B_x^2 + rho u^2 + c~f rho u; both sides of the
    fan take the Roe speed."""
    for (u, b) in [(0.3, (0.8, 0.5, -0.2)), (0.05, (1.2, 0.0, 0.7)),
                   (0.6, (0.0, 0.4, 0.3))]:
        rho, p = 1.3, 0.9
        wl = _state(rho, (u, 0.1, -0.2), b, p)
        wr = _state(rho, (-u, 0.1, -0.2), b, p)
        a2 = GAMMA * p / rho + (GAMMA - 1.0) * u * u / 2.0
        bb = b[0] ** 2 + b[1] ** 2 + b[2] ** 2
        c = _fast(rho, a2, b[0] ** 2 / rho, bb / rho)
        got = _oracle_flux(wl, wr, GAMMA)
        mom = p + 0.5 * bb - b[0] ** 2 + rho * u * u + c * rho * u
        np.testing.assert_allclose(got[VI], mom, rtol=1e-14)
        # mass: the jump of rho v is -2 rho u, the average flux 0
        np.testing.assert_allclose(got[RHO], 0.0, atol=1e-15)
        _check(wl, wr, GAMMA, (0, 1))


def test_density_jump_uniform_b():
    """A contact-like density jump with uniform B (X = 0): the sqrt(rho)
    weights of v~ and H~ and the Y term of the Alfven part of c~f."""
    for rl, rr, b in [(1.0, 0.2, (0.4, 1.1, 0.6)), (0.3, 2.5, (1.0, 0.7, -0.9)),
                      (4.0, 1.0, (0.1, 1.5, 0.0))]:
        wl = _state(rl, (0.05, 0.2, 0.0), b, 1.0)
        wr = _state(rr, (-0.02, -0.1, 0.3), b, 1.0)
        _check(wl, wr, GAMMA, (0, 1))


def test_cross_weighted_roe_bfield_gamma_2():
    """gamma = 2 (Brio & Wu's value): X and Y drop out and c~f is the fast
    speed of the Roe state with B~_t = (sqrt(rho_r) B_t,l + sqrt(rho_l)
    B_t,r) / (sqrt(rho_l) + sqrt(rho_r)); unequal densities and transverse
    fields separate this from the rho-weighted mean."""
    g2 = 2.0
    for rl, rr, bl, br in [(1.0, 0.125, (0.75, 1.0, 0.2), (0.75, -1.0, 0.5)),
                           (0.2, 1.7, (0.3, 0.1, 1.2), (0.3, 1.4, -0.3)),
                           (3.0, 0.5, (1.0, -0.6, 0.0), (1.0, 0.9, 0.8))]:
        wl = _state(rl, (0.1, 0.0, 0.2), bl, 1.0)
        wr = _state(rr, (0.0, -0.3, 0.0), br, 0.4)
        _check(wl, wr, g2, (0, 1))


def test_transverse_b_jump_unit_root_sum():
    """rho_l = rho_r = 1/4 (sqrt(rho_l) + sqrt(rho_r) = 1): the X term with
    its 1/2 and its sign at gamma = 5/3, where gamma' - 1 = -1/3."""
    for bl, br in [((0.5, 1.0, 0.0), (0.5, -0.8, 0.4)),
                   ((0.0, 0.3, -1.1), (0.0, 1.2, 0.6)),
                   ((1.1, 0.0, 0.9), (1.1, 0.7, -0.9))]:
        # a converging normal flow lifts a~^2 so the Roe speed bounds the fan
        wl = _state(0.25, (0.9, 0.1, 0.0), bl, 0.6)
        wr = _state(0.25, (-0.8, 0.0, -0.1), br, 0.6)
        _check(wl, wr, GAMMA, (0, 1))


@pytest.mark.parametrize("gamma", [GAMMA, 1.4, 2.0])
def test_roe_speed_bounds_fast_speeds_of_equal_states(gamma):
    """Equal states: the Roe average is the state itself, so c~f equals the
    state's fast speed and the flux is the physical flux."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        w = _state(rng.uniform(0.3, 3.0), rng.normal(size=3), rng.normal(size=3),
                   rng.uniform(0.2, 2.0))
        _, croe = _roe_fast_speed(w, w, gamma)
        cf = _fast(w[RHO], gamma * w[P] / w[RHO], w[BI] ** 2 / w[RHO],
                   (w[BI] ** 2 + w[BJ] ** 2 + w[BK] ** 2) / w[RHO])
        assert abs(croe - cf) <= 1e-12 * cf
        f = _oracle_flux(w, w, gamma)
        exp = _flux(w, gamma)
        np.testing.assert_allclose(f, exp, rtol=1e-12, atol=1e-12)
