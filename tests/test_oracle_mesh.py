"""Pins of the oracle's coarse/fine transfer operators (oracle/mesh_oracle.c)
on the CPU: ProlongLinear (src/Cello/problem_ProlongLinear.cpp:85-370) and
RestrictLinear (src/Cello/problem_RestrictLinear.cpp:60-198).

Closed forms on dyadic data (every product and sum exact):
* a linear field sampled on the coarse cells is reproduced exactly at the
  fine cell centres, with coarse ghost cells available (weights 1/4, 3/4)
  and without them (one-sided 5/4, -1/4 at the region's ends), in rank 1,
  2 and 3, outside the region nothing changes, accumulate adds;
* restriction of a linear fine field is the field at the coarse centres.
"""
import numpy as np
import pytest

import pyoracle as po


def lin(x, y, z):
    return 1.0 + 0.5 * x - 0.25 * y + 2.0 * z


def _i3(v):
    return po.ints(list(v))


def _coarse_field(mc3, rank):
    """Linear field at coarse centres k + 0.5 (array (mz, my, mx))."""
    z, y, x = np.meshgrid(*(np.arange(mc3[a]) + 0.5 if a < rank else np.zeros(1)
                            for a in (2, 1, 0)), indexing="ij")
    return np.ascontiguousarray(lin(x, y, z))


@pytest.mark.parametrize("rank", [1, 2, 3])
@pytest.mark.parametrize("ghosts", [True, False])
def test_prolong_linear_reproduces_linear_fields(rank, ghosts):
    mc3 = [10, 8, 7][:rank] + [1] * (3 - rank)
    oc3 = [1, 2, 1][:rank] + [0] * (3 - rank)
    nc3 = [6, 5, 4][:rank] + [1] * (3 - rank)
    nf3 = [2 * (c - 2) if ghosts else 2 * c for c in nc3[:rank]] + [1] * (3 - rank)
    mf3 = [nf3[a] + 3 for a in range(rank)] + [1] * (3 - rank)
    of3 = [2, 1, 3][:rank] + [0] * (3 - rank)
    vc = _coarse_field(mc3, rank)
    vf = np.full((mf3[2], mf3[1], mf3[0]), -7.0)
    assert po.lib().o_prolong_linear(po.dptr(vf), _i3(mf3), _i3(of3), _i3(nf3),
                                     po.dptr(vc), _i3(mc3), _i3(oc3), _i3(nc3), 0) == 0
    # fine centre of fine cell i along an axis, in coarse index units
    start = [oc3[a] + (1 if ghosts else 0) for a in range(3)]
    centres = [start[a] + (np.arange(nf3[a]) + 0.5) / 2 if a < rank else np.zeros(1)
               for a in range(3)]
    z, y, x = np.meshgrid(centres[2], centres[1], centres[0], indexing="ij")
    region = tuple(slice(of3[a], of3[a] + nf3[a]) for a in (2, 1, 0))
    assert np.array_equal(vf[region], lin(x, y, z))
    mask = np.ones(vf.shape, bool)
    mask[region] = False
    assert np.all(vf[mask] == -7.0)
    # accumulate adds the same values
    before = vf.copy()
    assert po.lib().o_prolong_linear(po.dptr(vf), _i3(mf3), _i3(of3), _i3(nf3),
                                     po.dptr(vc), _i3(mc3), _i3(oc3), _i3(nc3), 1) == 0
    assert np.array_equal(vf[region], 2 * before[region])
    assert np.array_equal(vf[mask], before[mask])


def test_prolong_linear_rejects_sizes_the_reference_asserts():
    mc3, oc3, nc3 = [8, 8, 8], [1, 1, 1], [4, 4, 4]
    mf3, of3, nf3 = [12, 12, 12], [0, 0, 0], [6, 8, 8]
    vc = np.zeros(512)
    vf = np.zeros(12 ** 3)
    assert po.lib().o_prolong_linear(po.dptr(vf), _i3(mf3), _i3(of3), _i3(nf3),
                                     po.dptr(vc), _i3(mc3), _i3(oc3), _i3(nc3), 0) == -1


@pytest.mark.parametrize("rank", [1, 2, 3])
def test_restrict_linear_is_the_mean_of_the_children(rank):
    nd_f = [12, 10, 8][:rank] + [1] * (3 - rank)
    im_f = [2, 0, 2][:rank] + [0] * (3 - rank)
    n_c = [4, 5, 3][:rank] + [1] * (3 - rank)
    n_f = [2 * c for c in n_c[:rank]] + [1] * (3 - rank)
    nd_c = [7, 8, 5][:rank] + [1] * (3 - rank)
    im_c = [1, 2, 1][:rank] + [0] * (3 - rank)
    # fine centres (i + 0.5) / 2 in coarse units measured from fine index 0
    c = [(np.arange(nd_f[a]) + 0.5) / 2 if a < rank else np.zeros(1) for a in range(3)]
    z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    vf = np.ascontiguousarray(lin(x, y, z))
    vc = np.full((nd_c[2], nd_c[1], nd_c[0]), 3.0)
    assert po.lib().o_restrict_linear(po.dptr(vc), _i3(nd_c), _i3(im_c), _i3(n_c),
                                      po.dptr(vf), _i3(nd_f), _i3(im_f), _i3(n_f), 0) == 0
    cc = [im_f[a] / 2 + np.arange(n_c[a]) + 0.5 if a < rank else np.zeros(1)
          for a in range(3)]
    z, y, x = np.meshgrid(cc[2], cc[1], cc[0], indexing="ij")
    region = tuple(slice(im_c[a], im_c[a] + n_c[a]) for a in (2, 1, 0))
    assert np.array_equal(vc[region], lin(x, y, z))
    mask = np.ones(vc.shape, bool)
    mask[region] = False
    assert np.all(vc[mask] == 3.0)
    assert po.lib().o_restrict_linear(po.dptr(vc), _i3(nd_c), _i3(im_c), _i3(n_c),
                                      po.dptr(vf), _i3(nd_f), _i3(im_f), _i3(n_f), 1) == 0
    assert np.array_equal(vc[region], 2 * lin(x, y, z))
