"""enzoe_vlct_compute_noalloc (the compute entry point that allocates
nothing: its descriptor table and stage scratch live in a fixed device
array of the library) against enzoe_vlct_compute, through the C ABI: two
calls in a row on the same workspace, the second with a smaller input, each
bitwise equal to the allocating entry point on a copy of the same blocks;
FP64 and mixed precision; sizes beyond the workspace are rejected before
anything is launched."""
import ctypes

import numpy as np
import pytest

from _states import random_periodic_grid
from test_gpu_parity import device_copy, method_for

pytestmark = pytest.mark.gpu

ENZOE_ERR_SIZE = 6


def _ev():
    import enzoe_vlct as ev
    return ev


def _noalloc(ev, m, blocks, dt):
    arr = (ev.Block * len(blocks))(*[b.desc for b in blocks])
    return ev.lib().enzoe_vlct_compute_noalloc(m.handle, arr, len(blocks),
                                               ctypes.c_double(dt), None)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_noalloc_equals_compute_twice_second_smaller(precision):
    ev = _ev()
    cap = ev.lib().enzoe_vlct_noalloc_bytes()
    assert cap > 0
    for seed, n, nb in ((3, (24, 20, 18), 2), (4, (12, 10, 9), 1)):
        grids = [random_periodic_grid(mhd=True, seed=seed + 10 * i, n=n, vamp=0.4)
                 for i in range(nb)]
        m = method_for(grids[0].cfg, precision=precision)
        ref = [device_copy(g, m) for g in grids]
        got = [device_copy(g, m) for g in grids]
        rb = [b for d in ref for b in d.blocks]
        gb = [b for d in got for b in d.blocks]
        dt = m.timestep(rb)
        assert m.timestep(gb) == dt
        for d in ref + got:
            d.refresh()
        m.compute(rb, dt)
        assert _noalloc(ev, m, gb, dt) == 0
        ev.synchronize()
        for x, y in zip(rb, gb):
            a, b = x.download(), y.download()
            for k in a:
                np.testing.assert_array_equal(a[k], b[k], err_msg="%s %s" % (n, k))


def test_noalloc_rejects_blocks_beyond_the_workspace():
    """A real 64^3 block (its stage scratch is larger than the workspace) is
    refused with ENZOE_ERR_SIZE and left untouched."""
    ev = _ev()
    cap = ev.lib().enzoe_vlct_noalloc_bytes()
    n = 16
    while (n + 6) ** 3 * 8 * 38 <= cap:
        n *= 2
    blk = ev.DeviceBlock(n + 6, n + 6, n + 6, 3, (1.0 / n,) * 3, mhd=True)
    m = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0, courant=0.4)
    rc = _noalloc(ev, m, [blk], 1e-3)
    assert rc == ENZOE_ERR_SIZE
    ev.synchronize()
    for k, a in blk.download().items():
        assert not a.any(), k
