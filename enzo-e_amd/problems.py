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


# ------------------------------------------------------------ helpers

def _edge_curl(Ax, Ay, Az, dx, dy, dz):
    """Face fields from a vector potential on cell edges of a ghost-extended
    block (arrays of shape (mz+1, my+1, mx+1)): A_x on x-edges
    (x_i, y_{j-1/2}, z_{k-1/2}), A_y on (x_{i-1/2}, y_j, z_{k-1/2}), A_z on
    (x_{i-1/2}, y_{j-1/2}, z_k), index (k, j, i).  Returns bfieldi_x (mz,
    my, mx+1), bfieldi_y (mz, my+1, mx), bfieldi_z (mz+1, my, mx); the
    discrete divergence of the result is zero to roundoff."""
    bx = ((Az[:-1, 1:, :] - Az[:-1, :-1, :]) / dy
          - (Ay[1:, :-1, :] - Ay[:-1, :-1, :]) / dz)
    by = ((Ax[1:, :, :-1] - Ax[:-1, :, :-1]) / dz
          - (Az[:-1, :, 1:] - Az[:-1, :, :-1]) / dx)
    bz = ((Ay[:, :-1, 1:] - Ay[:, :-1, :-1]) / dx
          - (Ax[:, 1:, :-1] - Ax[:, :-1, :-1]) / dy)
    return bx, by, bz


def _fill_state(blk, rho, v, p, bfi, gamma):
    """Conserved-form integration quantities (specific velocity and total
    energy, cell-centred B as face averages) from primitive arrays on the
    whole ghost-extended block."""
    blk["density"][...] = rho
    for a, c in enumerate("xyz"):
        blk["velocity_" + c][...] = v[a]
    ke = 0.5 * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
    if bfi is not None:
        blk["bfieldi_x"][...], blk["bfieldi_y"][...], blk["bfieldi_z"][...] = bfi
        bx = 0.5 * (bfi[0][:, :, 1:] + bfi[0][:, :, :-1])
        by = 0.5 * (bfi[1][:, 1:, :] + bfi[1][:, :-1, :])
        bz = 0.5 * (bfi[2][1:, :, :] + bfi[2][:-1, :, :])
        blk["bfield_x"][...], blk["bfield_y"][...], blk["bfield_z"][...] = bx, by, bz
        me = 0.5 * (bx * bx + by * by + bz * bz)
    else:
        me = 0.0
    blk["total_energy"][...] = (p / (gamma - 1.0) + rho * ke + me) / rho


# --------------------------------------------- fast wave along (1, 1, 1)

def fast_wave_111(n=32, amplitude=1e-6, gamma=5.0 / 3.0, g=3):
    """MHD fast magnetosonic eigenmode travelling along (1,1,1) of the unit
    cube (wavelength 1/sqrt(3), periodic on one n^3 block): background
    rho = 1, p = 1/gamma, v = 0, B = (1, sqrt(2), 0.5) in the wave frame
    (B_parallel = 1).  The eigenvector is the right eigenvector of the 1D
    ideal-MHD Jacobian in primitive variables for the fastest right-moving
    wave (numpy.linalg.eig); the B perturbation is the curl of an edge
    vector potential so div B = 0 to roundoff.  Returns (block, widths,
    fast speed, wave period)."""
    rho0, p0 = 1.0, 1.0 / gamma
    b_par, b_t1, b_t2 = 1.0, np.sqrt(2.0), 0.5
    # primitive 1D MHD along the wave normal: (rho, vn, vt1, vt2, bt1, bt2, p)
    A = np.zeros((7, 7))
    A[0, 1] = rho0
    A[1, 4] = b_t1 / rho0
    A[1, 5] = b_t2 / rho0
    A[1, 6] = 1.0 / rho0
    A[2, 4] = -b_par / rho0
    A[3, 5] = -b_par / rho0
    A[4, 1] = b_t1
    A[4, 2] = -b_par
    A[5, 1] = b_t2
    A[5, 3] = -b_par
    A[6, 1] = gamma * p0
    lam, vec = np.linalg.eig(A)
    i = int(np.argmax(lam.real))
    cf = float(lam[i].real)
    r = vec[:, i].real
    r = r / r[0]  # unit density perturbation
    # wave frame -> lab frame: normal n, tangents t1, t2
    nrm = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    t1 = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    t2 = np.cross(nrm, t1)
    kw = 2.0 * np.pi * np.sqrt(3.0)  # |k| for phase 2 pi (x + y + z)
    dx = 1.0 / n
    m = n + 2 * g
    blk = _empty_block((n, n, n), g)
    xc = (np.arange(m) - g + 0.5) * dx
    xce = (np.arange(m + 1) - g + 0.5) * dx  # cell centres, one extra entry
    xf = (np.arange(m + 1) - g) * dx
    Zc, Yc, Xc = np.meshgrid(xc, xc, xc, indexing="ij")
    phase = 2.0 * np.pi * (Xc + Yc + Zc)
    s = amplitude * np.cos(phase)
    rho = rho0 + r[0] * s
    vel = [(r[1] * nrm[a] + r[2] * t1[a] + r[3] * t2[a]) * s for a in range(3)]
    p = p0 + r[6] * s
    B0 = b_par * nrm + b_t1 * t1 + b_t2 * t2
    # dB = s (bt1 t1 + bt2 t2) = curl A with A = (amp/k) sin(phase) (db x n):
    # curl(f(phase) c) = f' k (n x c) and n x (db x n) = db for db . n = 0
    db_dir = r[4] * t1 + r[5] * t2
    cdir = np.cross(db_dir, nrm)

    def edge(coords):
        Z, Y, X = np.meshgrid(*coords, indexing="ij")
        return (amplitude / kw) * np.sin(2.0 * np.pi * (X + Y + Z))

    # all three on the (m+1)^3 edge lattice of _edge_curl
    Ax = cdir[0] * edge((xf, xf, xce))
    Ay = cdir[1] * edge((xf, xce, xf))
    Az = cdir[2] * edge((xce, xf, xf))
    bx, by, bz = _edge_curl(Ax, Ay, Az, dx, dx, dx)
    bfi = (bx + B0[0], by + B0[1], bz + B0[2])
    _fill_state(blk, rho, vel, p, bfi, gamma)
    wavelength = 1.0 / np.sqrt(3.0)
    return blk, (dx, dx, dx), cf, wavelength / cf


# ------------------------------------------ random smooth blocks (batched)

def _smooth_field(rng, shape_cells, n_modes=6, kmax=3):
    """A smooth random field of unit variance on cell centres of a block
    (sum of random low-wavenumber Fourier modes over the block's span)."""
    mz, my, mx = shape_cells
    z = np.arange(mz) / mz
    y = np.arange(my) / my
    x = np.arange(mx) / mx
    f = np.zeros(shape_cells)
    for _ in range(n_modes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        if not k.any():
            k[0] = 1
        ph = rng.uniform(0, 2 * np.pi)
        amp = rng.normal()
        f += amp * np.cos(2 * np.pi * (k[2] * z[:, None, None] + k[1] * y[None, :, None]
                                       + k[0] * x[None, None, :]) + ph)
    sd = f.std()
    return f / sd if sd > 0 else f


def random_mhd_block(seed, n=32, g=3, gamma=5.0 / 3.0, mach=2.0, beta=1.0,
                     sigma_lnrho=0.5, width=1.0 / 32):
    """One self-contained block of the batched workload: rho lognormal
    (sigma 0.5), v Gaussian with rms Mach `mach` (sound speed of p = 1,
    rho = 1), B = mean field along x plus the curl of a smooth random edge
    potential, scaled so that beta = 2 p / <B^2> = `beta`.  Every field is
    evaluated on the ghost-extended block, so the block is a complete
    Method input without a refresh."""
    rng = np.random.default_rng(seed)
    m = n + 2 * g
    shp = (m, m, m)
    rho = np.exp(sigma_lnrho * _smooth_field(rng, shp))
    p = np.ones(shp)
    cs = np.sqrt(gamma)
    vel = [mach * cs / np.sqrt(3.0) * _smooth_field(rng, shp) for _ in range(3)]
    A = [_smooth_field(rng, (m + 1, m + 1, m + 1)) for _ in range(3)]
    bx, by, bz = _edge_curl(A[0], A[1], A[2], width, width, width)
    b2 = (np.mean(bx ** 2) + np.mean(by ** 2) + np.mean(bz ** 2))
    target = 2.0 / beta  # <B^2> for p = 1
    # half the magnetic energy in the mean field, half in the fluctuation
    s = np.sqrt(0.5 * target / b2)
    b0 = np.sqrt(0.5 * target)
    bfi = (bx * s + b0, by * s, bz * s)
    blk = _empty_block((n, n, n), g)
    _fill_state(blk, rho, vel, p, bfi, gamma)
    return blk, (width, width, width)


# ------------------------------------------------- decaying turbulence

def turbulence_spectrum(seed=42, kmin=1, kmax=8, slope=-2.0):
    """Random-phase solenoidal velocity modes: integer wave vectors with
    kmin <= |k| <= kmax, complex amplitudes orthogonal to k with |v_k| ~
    |k|^slope.  Returns (k vectors (M,3), complex amplitudes (M,3))."""
    rng = np.random.default_rng(seed)
    r = np.arange(-kmax, kmax + 1)
    K = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    kk = np.sqrt((K ** 2).sum(axis=1))
    K = K[(kk >= kmin) & (kk <= kmax)]
    kk = np.sqrt((K ** 2).sum(axis=1))
    c = rng.normal(size=(len(K), 3)) + 1j * rng.normal(size=(len(K), 3))
    khat = K / kk[:, None]
    c -= (c * khat).sum(axis=1)[:, None] * khat  # solenoidal
    c *= (kk ** slope)[:, None]
    return K.astype(float), c


def _eval_modes(K, C, z, y, x):
    """Real part of sum_k C_k exp(2 pi i k.x) on the grid z x y x: the modes
    go into a dense (kz, ky, kx) cube and are summed one axis at a time."""
    kmax = int(np.abs(K).max())
    w = 2 * kmax + 1
    cube = np.zeros((w, w, w), dtype=complex)
    ki = K.astype(int) + kmax
    np.add.at(cube, (ki[:, 2], ki[:, 1], ki[:, 0]), C)
    r = np.arange(-kmax, kmax + 1)
    ez = np.exp(2j * np.pi * np.outer(r, z))  # (kz, nz)
    ey = np.exp(2j * np.pi * np.outer(r, y))
    ex = np.exp(2j * np.pi * np.outer(r, x))
    t = np.tensordot(ez, cube, axes=([0], [0]))     # (nz, ky, kx)
    t = np.tensordot(t, ey, axes=([1], [0]))        # (nz, kx, ny)
    t = np.tensordot(t, ex, axes=([1], [0]))        # (nz, ny, nx)
    return t.real


def turbulence_block(spec, vrms_norm, ib, nb=(8, 8, 8), n=64, g=3,
                     gamma=5.0 / 3.0, beta=0.1):
    """Block ib = (bx, by, bz) of the decaying-turbulence workload on the
    unit cube: rho = 1, p = 1/gamma (c_s = 1), the solenoidal velocity of
    `spec` scaled by vrms_norm, uniform B_x with beta = 2 p / B_x^2."""
    K, C = spec
    dx = 1.0 / (nb[0] * n)
    m = n + 2 * g
    coord = [(ib[a] * n + np.arange(m) - g + 0.5) * dx for a in range(3)]
    vel = [vrms_norm * _eval_modes(K, C[:, a], coord[2], coord[1], coord[0])
           for a in range(3)]
    rho = np.ones((m, m, m))
    p0 = 1.0 / gamma
    p = np.full((m, m, m), p0)
    bx0 = np.sqrt(2.0 * p0 / beta)
    bfi = (np.full((m, m, m + 1), bx0), np.zeros((m, m + 1, m)),
           np.zeros((m + 1, m, m)))
    blk = _empty_block((n, n, n), g)
    _fill_state(blk, rho, vel, p, bfi, gamma)
    return blk, (dx, dx, dx)


def turbulence_domain_velocity(spec, vrms_norm, N):
    """The velocity of `spec` (scaled by vrms_norm) at the N^3 cell centres
    of the unit cube, evaluated once for the whole domain (three
    matrix products over the dense mode cube instead of one per block).
    Returns an array (3, N, N, N) indexed (component, z, y, x)."""
    K, C = spec
    x = (np.arange(N) + 0.5) / N
    return np.stack([vrms_norm * _eval_modes(K, C[:, a], x, x, x) for a in range(3)])


def turbulence_block_from_domain(vel, ib, n, g=3, gamma=5.0 / 3.0, beta=0.1):
    """Block ib of the decaying-turbulence workload cut from a whole-domain
    velocity (turbulence_domain_velocity), ghost zones included (periodic
    wrap): the state of turbulence_block up to the summation order of the
    modes."""
    N = vel.shape[1]
    m = n + 2 * g
    idx = [(ib[a] * n + np.arange(m) - g) % N for a in range(3)]
    v = [vel[c][np.ix_(idx[2], idx[1], idx[0])] for c in range(3)]
    rho = np.ones((m, m, m))
    p0 = 1.0 / gamma
    p = np.full((m, m, m), p0)
    bx0 = np.sqrt(2.0 * p0 / beta)
    bfi = (np.full((m, m, m + 1), bx0), np.zeros((m, m + 1, m)),
           np.zeros((m + 1, m, m)))
    blk = _empty_block((n, n, n), g)
    _fill_state(blk, rho, v, p, bfi, gamma)
    return blk, (1.0 / N,) * 3


def turbulence_vrms_norm(spec, mach=5.0, samples=32):
    """Factor that brings the rms speed of `spec` to `mach` (c_s = 1),
    measured on a samples^3 grid of the unit cube."""
    K, C = spec
    x = (np.arange(samples) + 0.5) / samples
    v2 = sum(np.mean(_eval_modes(K, C[:, a], x, x, x) ** 2) for a in range(3))
    return mach / np.sqrt(v2)


# -------------------------------------------------- Kelvin-Helmholtz

def kelvin_helmholtz_block(ib, nb=(8, 8, 8), n=128, g=3, gamma=1.4, seed=7,
                           dv=1e-2):
    """Block ib of the hydro Kelvin-Helmholtz slab on [0,1]^3 (y measured
    from the domain centre): v_x = +0.5, rho = 2 for |y| < 0.25, v_x = -0.5,
    rho = 1 outside; p = 2.5; seeded uniform +-dv perturbation of v_y in the
    active cells (ghosts are filled by the refresh)."""
    dx = 1.0 / (nb[0] * n)
    m = n + 2 * g
    yc = (ib[1] * n + np.arange(m) - g + 0.5) * dx - 0.5
    inside = (np.abs(yc) < 0.25)[None, :, None]
    shp = (m, m, m)
    rho = np.where(inside, 2.0, 1.0) * np.ones(shp)
    vx = np.where(inside, 0.5, -0.5) * np.ones(shp)
    rng = np.random.default_rng(seed + 1000003 * (ib[0] + nb[0] * (ib[1] + nb[1] * ib[2])))
    vy = dv * (2.0 * rng.random(shp) - 1.0)
    vz = np.zeros(shp)
    p = np.full(shp, 2.5)
    blk = _empty_block((n, n, n), g, mhd=False)
    _fill_state(blk, rho, [vx, vy, vz], p, None, gamma)
    return blk, (dx, dx, dx)


# ------------------------------- two refinement levels (flux corrections)

def smooth_hydro_block(lower, dx, n=16, g=4, gamma=5.0 / 3.0):
    """Hydro block whose cell centres (ghosts included) sample a smooth
    state: rho = 1 + 0.3 sin(2 pi x) cos(2 pi y) + 0.1 sin(2 pi z),
    v = (0.5 + 0.2 sin 2 pi (y + z), 0.1 cos 2 pi x, 0.1 sin 2 pi y),
    p = 1 + 0.1 cos 2 pi (x + y + z).  Each block is a self-contained
    Method input (its ghost zones hold the same analytic state)."""
    m = n + 2 * g
    c = [lower[a] + (np.arange(m) - g + 0.5) * dx for a in range(3)]
    z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    tp = 2.0 * np.pi
    rho = 1.0 + 0.3 * np.sin(tp * x) * np.cos(tp * y) + 0.1 * np.sin(tp * z)
    vx = 0.5 + 0.2 * np.sin(tp * (y + z))
    vy = 0.1 * np.cos(tp * x)
    vz = 0.1 * np.sin(tp * y)
    p = 1.0 + 0.1 * np.cos(tp * (x + y + z))
    blk = _empty_block((n, n, n), g, mhd=False)
    _fill_state(blk, rho, [vx, vy, vz], p, None, gamma)
    return blk


def two_level_composite(n=16, g=4, gamma=5.0 / 3.0):
    """A coarse block on [0,1]^3 (cell width 1/n) and, across its high-x
    face, the four fine blocks (cell width 1/(2n)) of x in [1, 1.5] that
    cover it: children (0, cy, cz) of the coarse-level block [1,2] x [0,1]^2.
    Returns [(kind, child, lower, dx, arrays)], the coarse block first."""
    dxc = 1.0 / n
    out = [("coarse", None, (0.0, 0.0, 0.0), dxc,
            smooth_hydro_block((0.0, 0.0, 0.0), dxc, n, g, gamma))]
    for cz in range(2):
        for cy in range(2):
            lower = (1.0, 0.5 * cy, 0.5 * cz)
            out.append(("fine", (0, cy, cz), lower, 0.5 * dxc,
                        smooth_hydro_block(lower, 0.5 * dxc, n, g, gamma)))
    return out
