"""Pins of the oracle's EnzoProlong interpolation (oracle/prolong_oracle.c:
EnzoProlong::apply_, src/Enzo/mesh/EnzoProlong.cpp:112-190, and the
interpolate.F routines tscint3d "3A", interp3d "2A", int_lin3d "2B", rank 3,
refinement 2) on the CPU, by closed forms:

* a constant coarse field gives the constant (2A and 2B exactly);
* 2A and 2B reproduce a coarse field linear in x, y and z exactly at the
  fine centres (their slopes are the corner-average differences, and the
  limiter leaves a linear field alone: s = (d1+d2+d3)/d0 = 1, frac = 1);
* every method conserves: the 8 children of a coarse cell sum to 8 times
  its value (to roundoff);
* 3A ends at the children's mean along each axis (its second bound makes
  every fraction 1), so each child equals its parent to roundoff;
* 2A flattens the children of a coarse extremum (min-limited slopes);
* 2B with positivity limits slopes to 0.8 of the parent value;
* sizes outside the caller's contract (nc = nf/2 + 2) are refused.
Planted slips (a sign of a corner slope, the 1/4 child offsets, the
conservation correction) each break one of these."""
import numpy as np
import pytest

import pyoracle as po


def _i3(v):
    return po.ints(list(v))


def _run(method, coarse, nf3, positive=0, oc3=(1, 2, 1), of3=(3, 1, 2), fill=-7.0):
    mc3 = [coarse.shape[2], coarse.shape[1], coarse.shape[0]]
    nc3 = [n // 2 + 2 for n in nf3]
    mf3 = [nf3[a] + of3[a] + 2 for a in range(3)]
    vf = np.full((mf3[2], mf3[1], mf3[0]), fill)
    rc = po.lib().o_enzo_prolong(method, positive, po.dptr(vf), _i3(mf3), _i3(of3), _i3(nf3),
                                 po.dptr(np.ascontiguousarray(coarse)), _i3(mc3), _i3(oc3),
                                 _i3(nc3))
    region = tuple(slice(of3[a], of3[a] + nf3[a]) for a in (2, 1, 0))
    return rc, vf, region


def _coarse(mc3, f):
    z, y, x = np.meshgrid(*(np.arange(mc3[a], dtype=float) + 0.5 for a in (2, 1, 0)),
                          indexing="ij")
    return f(x, y, z)


def _fine_centres(oc3, nf3):
    # centre of fine cell i of the region, in coarse index units: the
    # region's first child belongs to coarse cell oc + 1
    c = [oc3[a] + 1 + (np.arange(nf3[a]) + 0.5) / 2 for a in range(3)]
    return np.meshgrid(c[2], c[1], c[0], indexing="ij")


def _children_sums(fine, nf3):
    f = fine.reshape(nf3[2] // 2, 2, nf3[1] // 2, 2, nf3[0] // 2, 2)
    return f.sum(axis=(1, 3, 5))


NF = (8, 6, 4)
MC = (10, 8, 7)
OC = (1, 2, 1)


@pytest.mark.parametrize("method", [0, 1, 2])
def test_constant_field(method):
    coarse = np.full((MC[2], MC[1], MC[0]), 1.75)
    rc, vf, region = _run(method, coarse, NF)
    assert rc == 0
    if method == 0:
        np.testing.assert_allclose(vf[region], 1.75, rtol=4e-16)
    else:
        assert np.all(vf[region] == 1.75)
    mask = np.ones(vf.shape, bool)
    mask[region] = False
    assert np.all(vf[mask] == -7.0)


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("grad", [(1.0, 0.0, 0.0), (0.0, -0.5, 0.0), (0.0, 0.0, 2.0),
                                  (1.0, 0.5, 0.25), (-0.5, 1.0, 0.75)])
def test_linear_field_reproduced(method, grad):
    a, b, c = grad
    f = lambda x, y, z: 3.0 + a * x + b * y + c * z
    coarse = _coarse(MC, f)
    rc, vf, region = _run(method, coarse, NF, oc3=OC)
    assert rc == 0
    z, y, x = _fine_centres(OC, NF)
    np.testing.assert_array_equal(vf[region], f(x, y, z))


@pytest.mark.parametrize("method", [0, 1, 2])
def test_conservation_random(method):
    rng = np.random.default_rng(3 + method)
    coarse = 2.0 + rng.random((MC[2], MC[1], MC[0]))
    rc, vf, region = _run(method, coarse, NF, oc3=OC)
    assert rc == 0
    sums = _children_sums(vf[region], NF)
    parents = coarse[OC[2] + 1:OC[2] + 1 + NF[2] // 2, OC[1] + 1:OC[1] + 1 + NF[1] // 2,
                     OC[0] + 1:OC[0] + 1 + NF[0] // 2]
    np.testing.assert_allclose(sums, 8.0 * parents, rtol=2e-15)


def test_3a_children_equal_their_parent():
    rng = np.random.default_rng(11)
    coarse = 1.0 + rng.random((MC[2], MC[1], MC[0]))
    rc, vf, region = _run(0, coarse, NF, oc3=OC)
    assert rc == 0
    parents = coarse[OC[2] + 1:OC[2] + 1 + NF[2] // 2, OC[1] + 1:OC[1] + 1 + NF[1] // 2,
                     OC[0] + 1:OC[0] + 1 + NF[0] // 2]
    expect = np.repeat(np.repeat(np.repeat(parents, 2, 0), 2, 1), 2, 2)
    np.testing.assert_allclose(vf[region], expect, rtol=1e-15)


def test_2a_flattens_an_extremum():
    coarse = np.ones((MC[2], MC[1], MC[0]))
    # a peak at the coarse cell holding the region's first children
    peak = (OC[2] + 1, OC[1] + 1, OC[0] + 1)
    coarse[peak] = 5.0
    rc, vf, region = _run(1, coarse, NF, oc3=OC)
    assert rc == 0
    kids = vf[region][0:2, 0:2, 0:2]
    assert np.all(kids == 5.0)


def test_2b_positivity_limits_slopes():
    """A step from 0.02 to 10 across the high-x face of a coarse cell: the
    unlimited corner slopes (2.495 along three diagonals) take a child of
    the 0.02 cell below zero; with positivity each is cut to 0.8 fbar and
    the children are 0.02 -+ 0.008."""
    coarse = np.full((MC[2], MC[1], MC[0]), 0.02)
    cx = OC[0] + 1
    coarse[:, :, cx + 1:] = 10.0
    cell = (slice(0, 2), slice(0, 2), slice(0, 2))
    rc, vf, region = _run(2, coarse, NF, positive=0, oc3=OC)
    assert rc == 0
    assert np.any(vf[region][cell] < 0.0)
    rc, vf, region = _run(2, coarse, NF, positive=1, oc3=OC)
    assert rc == 0
    kids = vf[region][cell]
    assert np.all(kids > 0.0)
    np.testing.assert_allclose(np.sort(kids.ravel()), [0.012] * 4 + [0.028] * 4, rtol=1e-14)


@pytest.mark.parametrize("nf,nc_delta", [((8, 6, 4), 1), ((7, 6, 4), 0), ((8, 6, 4), -1)])
def test_sizes_outside_the_contract_are_refused(nf, nc_delta):
    coarse = np.ones((MC[2], MC[1], MC[0]))
    mc3 = list(MC)
    nc3 = [n // 2 + 2 for n in nf]
    nc3[0] += nc_delta
    of3 = (1, 1, 1)
    mf3 = [nf[a] + 3 for a in range(3)]
    vf = np.zeros((mf3[2], mf3[1], mf3[0]))
    rc = po.lib().o_enzo_prolong(1, 0, po.dptr(vf), _i3(mf3), _i3(of3), _i3(nf),
                                 po.dptr(coarse), _i3(mc3), _i3(OC), _i3(nc3))
    if nc_delta == 0 and nf[0] % 2 == 0:
        assert rc == 0
    else:
        assert rc == -1
