"""Drivers of the multi-block workloads on the device (host side).

* `SubcycledLevels`: independent 32^3 blocks on three refinement levels
  advanced with 2:1 sub-cycling; every substep is one batched launch
  sequence over the blocks of the levels that start together, each block
  with its own level's dt (enzoe_vlct_compute_dt).
* `uniform_grid`: a periodic grid of equal blocks with the device refresh
  (decaying turbulence, Kelvin-Helmholtz).
* `relative_l1`: per-field sum|a-b| / sum|b| between two grids, block by
  block on the host.
"""
import ctypes
import sys

import numpy as np

import enzoe_vlct as ev
import problems


class SubcycledLevels:
    """Blocks of refinement levels 0..L-1 (cell width halving per level),
    counts[l] blocks on level l, advanced by one coarse step = 2^(L-1)
    substeps of the finest level; level l takes a step at every 2^(L-1-l)-th
    substep with dt_l = dt_0 / 2^l, dt_0 = min_l 2^l dt_cfl(l)."""

    def __init__(self, method, counts=(512, 1536, 3000), n=32, distinct=64,
                 seed=2024, ghost=3):
        self.method = method
        self.counts = tuple(counts)
        self.n = n
        self.levels = []
        L = len(counts)
        for lev, cnt in enumerate(counts):
            width = 1.0 / (n * 2 ** lev)
            states = [problems.random_mhd_block(seed + 1000 * lev + s, n=n,
                                                width=width)[0]
                      for s in range(min(distinct, cnt))]
            blocks = []
            for i in range(cnt):
                b = ev.DeviceBlock(n + 2 * ghost, n + 2 * ghost, n + 2 * ghost, ghost,
                                   (width, width, width), mhd=True)
                b.upload(states[i % len(states)])
                blocks.append(b)
            self.levels.append(blocks)
            print("level %d: %d blocks of %d^3 on the device" % (lev, cnt, n),
                  file=sys.stderr, flush=True)
        self.nsub = 2 ** (L - 1)
        # launch lists: levels active at each substep
        self.active = [[lev for lev in range(L) if s % (2 ** (L - 1 - lev)) == 0]
                       for s in range(self.nsub)]
        self.arrays = []
        for act in self.active:
            blks = [b for lev in act for b in self.levels[lev]]
            self.arrays.append(((ev.Block * len(blks))(*[b.desc for b in blks]),
                                [(lev, len(self.levels[lev])) for lev in act]))

    def zones_per_coarse_step(self):
        return self.n ** 3 * sum(c * 2 ** lev for lev, c in enumerate(self.counts))

    def launch_sizes(self):
        return [sum(len(self.levels[lev]) for lev in act) for act in self.active]

    def device_levels(self):
        """The same levels as an enzoe_vlct_levels object: coarse steps with
        the dt on the device, eagerly or replayed from a HIP graph."""
        return ev.Levels(self.method, self.levels)

    def coarse_step(self):
        m = self.method
        cfl = [m.timestep(blks) for blks in self.levels]
        dt0 = min(2 ** lev * d for lev, d in enumerate(cfl))
        for arr, parts in self.arrays:
            dts = []
            for lev, cnt in parts:
                dts += [dt0 / 2 ** lev] * cnt
            m.compute_dt(None, dts, block_array=arr)
        return dt0


def uniform_grid(method, nblocks, n, builder, ghost=3):
    """A periodic grid of equal blocks filled block by block by
    builder(ib) -> (arrays, cell widths)."""
    cw = None
    arrays = {}
    nbx, nby, nbz = nblocks
    order = [(bx, by, bz) for bz in range(nbz) for by in range(nby) for bx in range(nbx)]
    blk0, cw = builder(order[0])
    grid = ev.DeviceGrid(method, nblocks, (n, n, n), cw, ghost=ghost)
    for i, ib in enumerate(order):
        arrs = blk0 if i == 0 else builder(ib)[0]
        grid.blocks[i].upload(arrs)
        if (i + 1) % 64 == 0:
            print("initialised %d of %d blocks" % (i + 1, len(order)),
                  file=sys.stderr, flush=True)
    grid.refresh()
    ev.synchronize()
    return grid


def copy_grid_state(src, dst):
    """Device-to-device copy of every field of every block."""
    for a, b in zip(src.blocks, dst.blocks):
        for name in a.shapes:
            nbytes = int(np.prod(a.shapes[name])) * 8
            ev.hip_check(ev.hip().hipMemcpy(ctypes.c_void_p(b.ptr(name)),
                                            ctypes.c_void_p(a.ptr(name)),
                                            ctypes.c_size_t(nbytes), 3),
                         "hipMemcpy D2D")


def relative_l1(grid_a, grid_b, names, ghost=3):
    """Two L1 measures of a against b over the active cells:
    per field, sum |a - b| / sum |b|; and per field normalised by the
    magnitude of its vector (sum |a - b| / sum |b_vec| for the velocity and
    B components, as the field itself for scalars) -- the first is not
    meaningful for a component that is zero in b (e.g. v_z of a 2D slab)."""
    num = dict.fromkeys(names, 0.0)
    den = dict.fromkeys(names, 0.0)
    vec = {"velocity": 0.0, "bfield": 0.0}
    for ba, bb in zip(grid_a.blocks, grid_b.blocks):
        da, db = ba.download(names), bb.download(names)
        s = (slice(ghost, -ghost),) * 3
        for nm in names:
            x, y = da[nm][s], db[nm][s]
            num[nm] += float(np.abs(x - y).sum())
            den[nm] += float(np.abs(y).sum())
        for g in vec:
            comps = [g + "_" + c for c in "xyz"]
            if all(c in db for c in comps):
                vec[g] += float(np.sqrt(sum(db[c][s] ** 2 for c in comps)).sum())
    per_field = {nm: (num[nm] / den[nm] if den[nm] > 0 else num[nm]) for nm in names}
    by_magnitude = {}
    for nm in names:
        g = nm.rsplit("_", 1)[0]
        d = vec[g] if g in vec and vec[g] > 0 else den[nm]
        by_magnitude[nm] = num[nm] / d if d > 0 else num[nm]
    return per_field, by_magnitude


# ----------------------------------- ghost zones across the refinement face

def composite_interface_transfers(coarse, fines, n, g, names):
    """The transfers of composite_interface_ghosts as two lists (the
    restrictions, then the prolongations that read them) for
    enzoe_vlct_transfer_batch."""
    m = n + 2 * g
    R, P = ev.TRANSFER_RESTRICT, ev.TRANSFER_PROLONG
    restricts, prolongs = [], []
    for (cy, cz), fb in fines.items():
        for name in names:
            restricts.append((R, coarse.ptr(name), (m, m, m),
                              (g + n, g + cy * n // 2, g + cz * n // 2), (g, n // 2, n // 2),
                              fb.ptr(name), (m, m, m), (g, g, g), (2 * g, n, n), 0))
            prolongs.append((P, fb.ptr(name), (m, m, m), (0, 0, 0), (g, m, m),
                             coarse.ptr(name), (m, m, m),
                             (g + n - g // 2 - 1, g + cy * n // 2 - g // 2 - 1,
                              g + cz * n // 2 - g // 2 - 1),
                             (g // 2 + 2, m // 2 + 2, m // 2 + 2), 0))
    return restricts, prolongs


def composite_interface_ghosts(coarse, fines, n, g, names, stream=None):
    """Fills the ghost zones on both sides of the refinement face of the
    two-level composite (problems.two_level_composite: coarse block on
    [0,1]^3, fine children (0, cy, cz) across its high-x face), from device
    data: each fine block's low-x ghost slab (g fine cells deep, its whole
    y-z extent with ghosts) by linear prolongation of the coarse block (with
    one coarse cell of stencil around it), and the part of the coarse
    block's high-x ghost slab that each fine block covers by restriction of
    its active cells.  fines: {(cy, cz): DeviceBlock}.  The restrictions
    go first: the prolongation stencil of the fine cells next to the face
    reaches one coarse cell across it, into the restricted slab (the coarse
    ghost cells beyond the fine blocks' y-z extent come from the coarse
    block's other neighbours and are read as they are)."""
    m = n + 2 * g
    for (cy, cz), fb in fines.items():
        for name in names:
            ev.restrict_linear(coarse.ptr(name), (m, m, m),
                               (g + n, g + cy * n // 2, g + cz * n // 2),
                               (g, n // 2, n // 2), fb.ptr(name), (m, m, m),
                               (g, g, g), (2 * g, n, n), stream=stream)
    for (cy, cz), fb in fines.items():
        for name in names:
            # coarse cells under the fine ghost slab, plus one cell of stencil
            oc = (g + n - g // 2 - 1, g + cy * n // 2 - g // 2 - 1,
                  g + cz * n // 2 - g // 2 - 1)
            nc = (g // 2 + 2, m // 2 + 2, m // 2 + 2)
            ev.prolong_linear(fb.ptr(name), (m, m, m), (0, 0, 0), (g, m, m),
                              coarse.ptr(name), (m, m, m), oc, nc, stream=stream)
