"""Coarse/fine transfer operators on the GPU against the oracle.

* enzoe_prolong_linear / enzoe_restrict_linear equal the oracle's
  ProlongLinear / RestrictLinear bit for bit on random data (rank 1, 2, 3;
  with and without coarse ghost cells; accumulate).
* The C++ layer's EnzoProlong with use_linear is ProlongLinear on the coarse
  region without its padding; EnzoRestrict writes nothing.
* enzoe_enzo_prolong (EnzoProlong's interpolate(): "3A", "2A", "2B" with and
  without positivity) equals the oracle bit for bit on random data, inside
  a larger fine array (cells outside the region untouched), and through the
  C++ layer's EnzoProlong(type, positive, use_linear = false); a NaN coarse
  value gives 2A's error flag and writes nothing, as the oracle.
* The ghost zones on both sides of the refinement face of the two-level
  composite, filled from device data by prolongation and restriction,
  equal a linear field exactly (dyadic data): the regions line up.
"""
import numpy as np
import pytest

import problems
import pyoracle as po

pytestmark = pytest.mark.gpu


def _ev():
    import enzoe_vlct as ev
    return ev


def _dev(arr):
    ev = _ev()
    buf = ev.DeviceBuffer(arr.nbytes)
    ev.to_device(buf.ptr.value, np.ascontiguousarray(arr))
    return buf


def _cases():
    for rank in (1, 2, 3):
        for ghosts in (True, False):
            mc3 = [10, 8, 7][:rank] + [1] * (3 - rank)
            oc3 = [1, 2, 1][:rank] + [0] * (3 - rank)
            nc3 = [6, 5, 4][:rank] + [1] * (3 - rank)
            nf3 = [2 * (c - 2) if ghosts else 2 * c for c in nc3[:rank]] + [1] * (3 - rank)
            mf3 = [nf3[a] + 3 for a in range(rank)] + [1] * (3 - rank)
            of3 = [2, 1, 3][:rank] + [0] * (3 - rank)
            yield rank, ghosts, mc3, oc3, nc3, mf3, of3, nf3


def test_prolong_and_restrict_match_the_oracle_on_random_data():
    ev = _ev()
    rng = np.random.default_rng(11)
    for rank, ghosts, mc3, oc3, nc3, mf3, of3, nf3 in _cases():
        for acc in (0, 1):
            print("prolong rank %d ghosts %s accumulate %d" % (rank, ghosts, acc), flush=True)
            vc = rng.standard_normal((mc3[2], mc3[1], mc3[0]))
            vf = rng.standard_normal((mf3[2], mf3[1], mf3[0]))
            dc, df = _dev(vc), _dev(vf)
            ev.prolong_linear(df.ptr.value, mf3, of3, nf3, dc.ptr.value, mc3, oc3, nc3,
                              accumulate=acc)
            ref = vf.copy()
            assert po.lib().o_prolong_linear(po.dptr(ref), po.ints(mf3), po.ints(of3),
                                             po.ints(nf3), po.dptr(vc), po.ints(mc3),
                                             po.ints(oc3), po.ints(nc3), acc) == 0
            ev.synchronize()
            assert np.array_equal(ev.from_device(df.ptr.value, ref.shape), ref)
    for rank in (1, 2, 3):
        nd_f = [12, 10, 8][:rank] + [1] * (3 - rank)
        im_f = [2, 0, 2][:rank] + [0] * (3 - rank)
        n_c = [4, 5, 3][:rank] + [1] * (3 - rank)
        n_f = [2 * c for c in n_c[:rank]] + [1] * (3 - rank)
        nd_c = [7, 8, 5][:rank] + [1] * (3 - rank)
        im_c = [1, 2, 1][:rank] + [0] * (3 - rank)
        for acc in (0, 1):
            vf = rng.standard_normal((nd_f[2], nd_f[1], nd_f[0]))
            vc = rng.standard_normal((nd_c[2], nd_c[1], nd_c[0]))
            dc, df = _dev(vc), _dev(vf)
            ev.restrict_linear(dc.ptr.value, nd_c, im_c, n_c, df.ptr.value, nd_f, im_f, n_f,
                               accumulate=acc)
            ref = vc.copy()
            assert po.lib().o_restrict_linear(po.dptr(ref), po.ints(nd_c), po.ints(im_c),
                                              po.ints(n_c), po.dptr(vf), po.ints(nd_f),
                                              po.ints(im_f), po.ints(n_f), acc) == 0
            ev.synchronize()
            assert np.array_equal(ev.from_device(dc.ptr.value, ref.shape), ref), (rank, acc)


def test_host_enzo_prolong_and_restrict():
    ev = _ev()
    H = ev.host_lib()
    rng = np.random.default_rng(3)
    mc3, mf3 = [10, 10, 10], [8, 8, 8]
    vc = rng.standard_normal((10, 10, 10))
    vf = np.zeros((8, 8, 8))
    # EnzoProlong with use_linear: the coarse region carries one cell of
    # padding per side; ProlongLinear sees it without (EnzoProlong.cpp:62-78)
    oc_pad, nc_pad = [1, 1, 1], [8, 8, 8]
    dc, df1, df2 = _dev(vc), _dev(vf), _dev(vf)
    I3 = po.ints
    assert H.enzoe_host_prolong_apply(b"enzo", 1, df1.ptr.value, I3(mf3), I3([0, 0, 0]),
                                      I3([8, 8, 8]), dc.ptr.value, I3(mc3), I3(oc_pad),
                                      I3(nc_pad), 0, None) == 0
    assert H.enzoe_host_prolong_apply(b"linear", 0, df2.ptr.value, I3(mf3), I3([0, 0, 0]),
                                      I3([8, 8, 8]), dc.ptr.value, I3(mc3), I3([2, 2, 2]),
                                      I3([6, 6, 6]), 0, None) == 0
    ev.synchronize()
    a = ev.from_device(df1.ptr.value, (8, 8, 8))
    b = ev.from_device(df2.ptr.value, (8, 8, 8))
    assert np.array_equal(a, b) and np.any(a != 0.0)
    # EnzoRestrict writes nothing; RestrictLinear returns the bytes written
    dcr = _dev(np.full((10, 10, 10), 5.0))
    assert H.enzoe_host_restrict_apply(b"enzo", dcr.ptr.value, I3(mc3), I3([1, 1, 1]),
                                       I3([4, 4, 4]), df1.ptr.value, I3(mf3), I3([0, 0, 0]),
                                       I3([8, 8, 8]), 0, None) == 0
    ev.synchronize()
    assert np.all(ev.from_device(dcr.ptr.value, (10, 10, 10)) == 5.0)
    assert H.enzoe_host_restrict_apply(b"linear", dcr.ptr.value, I3(mc3), I3([1, 1, 1]),
                                       I3([4, 4, 4]), df1.ptr.value, I3(mf3), I3([0, 0, 0]),
                                       I3([8, 8, 8]), 0, None) == 8 * 64


@pytest.mark.parametrize("batched", [False, True])
def test_composite_interface_ghosts_reproduce_a_linear_field(batched):
    ev = _ev()
    import workloads
    n, g = 16, 4
    m = n + 2 * g
    comp = problems.two_level_composite(n, g)

    def lin(lower, dx):
        c = [lower[a] + (np.arange(m) - g + 0.5) * dx for a in range(3)]
        z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
        return 1.0 + 0.5 * x - 0.25 * y + 2.0 * z

    blocks = {}
    for kind, child, lower, dx, _ in comp:
        db = ev.DeviceBlock(m, m, m, g, (dx, dx, dx), mhd=False)
        f = lin(lower, dx)
        poisoned = f.copy()
        if kind == "coarse":
            # the high-x ghost slab over the fine blocks' y-z extent
            poisoned[g:g + n, g:g + n, g + n:] = np.nan
        else:
            poisoned[:, :, :g] = np.nan           # low-x ghost slab
        db.upload({"density": poisoned})
        blocks[(kind, child)] = (db, f)
    coarse = blocks[("coarse", None)][0]
    fines = {(c[1], c[2]): blocks[("fine", c)][0] for (k, c) in blocks if k == "fine"}
    if batched:
        method = ev.VlctMethod(mhd=False, riemann="hllc")
        restricts, prolongs = workloads.composite_interface_transfers(
            coarse, fines, n, g, ["density"])
        ev.transfer_batch(method, restricts)
        ev.transfer_batch(method, prolongs)
    else:
        workloads.composite_interface_ghosts(coarse, fines, n, g, ["density"])
    ev.synchronize()
    for (kind, child), (db, f) in blocks.items():
        got = db.download(["density"])["density"]
        if kind == "fine":
            assert np.array_equal(got, f), child
        else:
            # the slab over the fine blocks' y-z extent is restricted, the
            # rest of the array is untouched
            assert np.array_equal(got, f)


@pytest.mark.parametrize("method,positive", [("3A", 0), ("2A", 0), ("2B", 0), ("2B", 1)])
def test_enzo_prolong_equals_the_oracle(method, positive):
    ev = _ev()
    H = ev.host_lib()
    rng = np.random.default_rng(7 + 3 * positive + {"3A": 0, "2A": 1, "2B": 2}[method])
    mc3, oc3, nf3, of3 = [13, 11, 9], [2, 1, 1], [16, 12, 10], [3, 2, 1]
    nc3 = [n // 2 + 2 for n in nf3]
    mf3 = [nf3[a] + of3[a] + 2 for a in range(3)]
    # a positive field with structure on every scale (a smooth part, steps,
    # noise), so the limiters take every branch
    z, y, x = np.meshgrid(np.arange(mc3[2]), np.arange(mc3[1]), np.arange(mc3[0]),
                          indexing="ij")
    vc = (2.0 + np.sin(0.7 * x) * np.cos(0.5 * y) + 0.3 * z + (x > 6) * 1.5 +
          0.4 * rng.standard_normal((mc3[2], mc3[1], mc3[0])))
    vc = np.ascontiguousarray(np.where(rng.random(vc.shape) < 0.1, 0.05, vc))
    fill = np.full((mf3[2], mf3[1], mf3[0]), -3.0)
    mcode = {"3A": 0, "2A": 1, "2B": 2}[method]
    ref = fill.copy()
    assert po.lib().o_enzo_prolong(mcode, positive, po.dptr(ref), po.ints(mf3), po.ints(of3),
                                   po.ints(nf3), po.dptr(vc), po.ints(mc3), po.ints(oc3),
                                   po.ints(nc3)) == 0
    dc, df, dh = _dev(vc), _dev(fill), _dev(fill)
    assert ev.enzo_prolong(method, df.ptr.value, mf3, of3, nf3, dc.ptr.value, mc3, oc3, nc3,
                           positive=positive) == 0
    assert H.enzoe_host_enzo_prolong_apply(method.encode(), positive, 0, dh.ptr.value,
                                           po.ints(mf3), po.ints(of3), po.ints(nf3),
                                           dc.ptr.value, po.ints(mc3), po.ints(oc3),
                                           po.ints(nc3), 0, None) == 0
    ev.synchronize()
    got = ev.from_device(df.ptr.value, fill.shape)
    host = ev.from_device(dh.ptr.value, fill.shape)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(host, ref)


def test_enzo_prolong_2a_nan_sets_the_error_flag_and_writes_nothing():
    ev = _ev()
    mc3, oc3, nf3, of3 = [10, 8, 8], [1, 1, 1], [8, 6, 6], [0, 0, 0]
    nc3 = [n // 2 + 2 for n in nf3]
    mf3 = list(nf3)
    vc = np.full((mc3[2], mc3[1], mc3[0]), 2.0)
    vc[2, 2, 3] = np.nan
    fill = np.full((mf3[2], mf3[1], mf3[0]), -3.0)
    ref = fill.copy()
    assert po.lib().o_enzo_prolong(1, 0, po.dptr(ref), po.ints(mf3), po.ints(of3),
                                   po.ints(nf3), po.dptr(vc), po.ints(mc3), po.ints(oc3),
                                   po.ints(nc3)) == -2
    dc, df = _dev(vc), _dev(fill)
    assert ev.enzo_prolong("2A", df.ptr.value, mf3, of3, nf3, dc.ptr.value, mc3, oc3,
                           nc3) == 1
    ev.synchronize()
    np.testing.assert_array_equal(ev.from_device(df.ptr.value, fill.shape), ref)
