"""Initial conditions of the benchmark workloads (numpy, host side).

Each builder returns, per block, a dict of ghost-extended numpy arrays in
the Cello field layout (density, velocity_x..z, total_energy, bfield_x..z,
bfieldi_x..z); ghost zones are filled afterwards by the device refresh.
Magnetic fields are built as the discrete curl of a vector potential on
cell edges, so the face field is divergence free to roundoff (the recipe of
src/Enzo/initial/EnzoInitialBCenter.cpp:43-127).
"""
import numpy as np


def _empty_block(n, g, mhd=True):
    nx, ny, nz = n
    mx, my, mz = nx + 2 * g, ny + 2 * g, nz + 2 * g
    out = {f: np.zeros((mz, my, mx)) for f in
           ("density", "velocity_x", "velocity_y", "velocity_z", "total_energy")}
    if mhd:
        for c in "xyz":
            out["bfield_" + c] = np.zeros((mz, my, mx))
        out["bfieldi_x"] = np.zeros((mz, my, mx + 1))
        out["bfieldi_y"] = np.zeros((mz, my + 1, mx))
        out["bfieldi_z"] = np.zeros((mz + 1, my, mx))
    return out


def _set_active(dst, src, g):
    sl = tuple(slice(g, g + s) for s in src.shape)
    dst[sl] = src


def _faces_from_az(Az, dx, dy):
    """B = curl(0, 0, Az) on faces from Az on z-edges (x_{i-1/2}, y_{j-1/2}),
    periodic in x and y."""
    bx = (np.roll(Az, -1, axis=1) - Az) / dy      # face (x_{i-1/2}, y_j)
    by = -(np.roll(Az, -1, axis=2) - Az) / dx     # face (x_i, y_{j-1/2})
    return bx, by


def orszag_tang(n=256, gamma=5.0 / 3.0, seed=1234, vz_perturb=1e-3, g=3):
    """3D Orszag-Tang vortex on [0,1]^3, one periodic block of n^3 cells:
    rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2pi y, sin 2pi x, dvz),
    B = (-sin 2pi y, sin 4pi x, 0)/sqrt(4 pi) from
    A_z = (cos 4pi x/(4pi) + cos 2pi y/(2pi))/sqrt(4pi)
    (input/vlct/orszag-tang/orszag-tang.in:45-48), with a seeded uniform
    +-1e-3 perturbation of v_z.  Returns (block arrays, cell widths)."""
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    xf = np.arange(n) * dx
    pi = np.pi
    blk = _empty_block((n, n, n), g)
    rho0 = 25.0 / (36.0 * pi)
    p0 = 5.0 / (12.0 * pi)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    vx = -np.sin(2 * pi * Y)
    vy = np.sin(2 * pi * X)
    rng = np.random.default_rng(seed)
    vz = vz_perturb * (2.0 * rng.random((n, n, n)) - 1.0)
    del Z
    # vector potential on z-edges (x_{i-1/2}, y_{j-1/2}); uniform along z
    YF, XF = np.meshgrid(xf, xf, indexing="ij")
    az2 = (np.cos(4 * pi * XF) / (4 * pi) + np.cos(2 * pi * YF) / (2 * pi)) / np.sqrt(4 * pi)
    Az = np.broadcast_to(az2, (n, n, n))
    bxf, byf = _faces_from_az(np.ascontiguousarray(Az), dx, dx)
    bzf = np.zeros((n, n, n))
    s = slice(g, g + n)
    blk["bfieldi_x"][s, s, g:g + n] = bxf
    blk["bfieldi_x"][s, s, g + n] = bxf[:, :, 0]
    blk["bfieldi_y"][s, g:g + n, s] = byf
    blk["bfieldi_y"][s, g + n, s] = byf[:, 0, :]
    blk["bfieldi_z"][g:g + n, s, s] = bzf
    blk["bfieldi_z"][g + n, s, s] = bzf[0, :, :]
    bx = 0.5 * (bxf + np.roll(bxf, -1, axis=2))
    by = 0.5 * (byf + np.roll(byf, -1, axis=1))
    bz = np.zeros((n, n, n))
    rho = np.full((n, n, n), rho0)
    etot = (p0 / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) +
            0.5 * (bx * bx + by * by + bz * bz)) / rho
    for name, arr in (("density", rho), ("velocity_x", vx), ("velocity_y", vy),
                      ("velocity_z", vz), ("total_energy", etot),
                      ("bfield_x", bx), ("bfield_y", by), ("bfield_z", bz)):
        _set_active(blk[name], arr, g)
    return blk, (dx, dx, dx)
