"""Seeded initial states shared by the oracle and GPU parity tests."""
import numpy as np

import pyoracle as po

GAMMA = 5.0 / 3.0


def fill_periodic_solenoidal_b(arr, n, ext, rng, g, amp=0.05,
                               mean=(0.7, 0.4, 0.1)):
    """Face field = mean + discrete curl of a random periodic vector
    potential, so div(B) = 0 to roundoff and the two boundary faces of the
    periodic block agree."""
    nx, ny, nz = n
    dx, dy, dz = (ext[0] / nx, ext[1] / ny, ext[2] / nz)
    Ax, Ay, Az = (amp * rng.normal(size=(nz, ny, nx)) for _ in range(3))
    p = lambda a, ax: np.roll(a, -1, axis=ax)  # value at index + 1
    bx = mean[0] + ((p(Az, 1) - Az) / dy - (p(Ay, 0) - Ay) / dz)
    by = mean[1] + ((p(Ax, 0) - Ax) / dz - (p(Az, 2) - Az) / dx)
    bz = mean[2] + ((p(Ay, 2) - Ay) / dx - (p(Ax, 1) - Ax) / dy)
    s = slice(g, -g)
    arr["bfieldi_x"][s, s, g:g + nx] = bx
    arr["bfieldi_x"][s, s, g + nx] = bx[:, :, 0]
    arr["bfieldi_y"][s, g:g + ny, s] = by
    arr["bfieldi_y"][s, g + ny, s] = by[:, 0, :]
    arr["bfieldi_z"][g:g + nz, s, s] = bz
    arr["bfieldi_z"][g + nz, s, s] = bz[0, :, :]


def center_b(arr):
    arr["bfield_x"][...] = 0.5 * (arr["bfieldi_x"][:, :, 1:] + arr["bfieldi_x"][:, :, :-1])
    arr["bfield_y"][...] = 0.5 * (arr["bfieldi_y"][:, 1:, :] + arr["bfieldi_y"][:, :-1, :])
    arr["bfield_z"][...] = 0.5 * (arr["bfieldi_z"][1:, :, :] + arr["bfieldi_z"][:-1, :, :])


def random_periodic_grid(mhd=True, seed=0, n=(12, 10, 8), ext=(1.2, 1.0, 0.8),
                         riemann=None, recon=("nn", "plm"), theta=2.0,
                         vamp=0.3, bamp=0.05):
    """A one-block periodic grid holding a seeded random smooth-ish state."""
    cfg = po.make_config(riemann=riemann or ("hlld" if mhd else "hllc"),
                         mhd=mhd, recon=recon, theta=theta)
    grid = po.Grid(cfg, (1, 1, 1), n, (0, 0, 0), ext)
    rng = np.random.default_rng(seed)
    arr = grid.block_arrays(0)
    mz, my, mx = arr["density"].shape
    arr["density"][...] = 1.0 + 0.3 * rng.random((mz, my, mx))
    for f in ("velocity_x", "velocity_y", "velocity_z"):
        arr[f][...] = vamp * rng.normal(size=(mz, my, mx))
    if mhd:
        fill_periodic_solenoidal_b(arr, n, ext, rng, 3, amp=bamp)
    grid.refresh()
    if mhd:
        center_b(arr)
    p = 0.6 + 0.2 * rng.random((mz, my, mx))
    v2 = sum(arr[f] ** 2 for f in ("velocity_x", "velocity_y", "velocity_z"))
    b2 = 0.0
    if mhd:
        b2 = sum(arr[f] ** 2 for f in ("bfield_x", "bfield_y", "bfield_z"))
    arr["total_energy"][...] = (p / (GAMMA - 1) + 0.5 * arr["density"] * v2 +
                                0.5 * b2) / arr["density"]
    grid.refresh()
    return grid
