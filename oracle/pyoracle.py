"""ctypes binding of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle_vlct.so`` (built by ``oracle/Makefile``) and exposes
the grid harness used by the parity tests: initial conditions of the
reference's regression decks, the Cello cycle loop, and the reference's L1
error metric (tools/l1_error_norm.py:195-234).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_vlct.so")

O_RHO, O_VX, O_VY, O_VZ, O_ETOT, O_BX, O_BY, O_BZ = range(8)
PRESSURE = 8
FIELD_NAMES = ["density", "velocity_x", "velocity_y", "velocity_z",
               "total_energy", "bfield_x", "bfield_y", "bfield_z", "pressure",
               "internal_energy"]
RIEMANN = {"hlle": 0, "hllc": 1, "hlld": 2}
RECON = {"nn": 0, "plm": 1, "plm_enzo": 1, "plm_athena": 2}
BC = {"periodic": 0, "outflow": 1}


class OConfig(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double),
                ("density_floor", ctypes.c_double),
                ("pressure_floor", ctypes.c_double),
                ("theta_limiter", ctypes.c_double),
                ("riemann", ctypes.c_int),
                ("recon", ctypes.c_int * 2),
                ("nstages", ctypes.c_int),
                ("mhd", ctypes.c_int),
                ("courant", ctypes.c_double),
                ("dual_energy", ctypes.c_int),
                ("eta", ctypes.c_double)]


class OBlock(ctypes.Structure):
    _fields_ = [("mx", ctypes.c_int), ("my", ctypes.c_int),
                ("mz", ctypes.c_int),
                ("cw", ctypes.c_double * 3),
                ("cc", ctypes.POINTER(ctypes.c_double) * 8),
                ("bfi", ctypes.POINTER(ctypes.c_double) * 3),
                ("pressure", ctypes.POINTER(ctypes.c_double)),
                ("eint", ctypes.POINTER(ctypes.c_double))]


class OGrid(ctypes.Structure):
    _fields_ = [("nbx", ctypes.c_int), ("nby", ctypes.c_int),
                ("nbz", ctypes.c_int),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int),
                ("g", ctypes.c_int),
                ("lower", ctypes.c_double * 3),
                ("upper", ctypes.c_double * 3),
                ("bc", ctypes.c_int * 3),
                ("blocks", ctypes.POINTER(OBlock))]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.o_grid_alloc.argtypes = [ctypes.POINTER(OGrid)]
        L.o_grid_free.argtypes = [ctypes.POINTER(OGrid)]
        L.o_grid_refresh.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OGrid)]
        L.o_grid_cycle.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OGrid),
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.o_grid_cycle.restype = ctypes.c_double
        L.o_grid_evolve.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OGrid),
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    dp, ctypes.c_int]
        L.o_grid_evolve.restype = ctypes.c_int
        L.o_init_inclined_wave.argtypes = [
            ctypes.POINTER(OConfig), ctypes.POINTER(OGrid), ctypes.c_char_p,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_int]
        L.o_init_shock_tube_rj2a.argtypes = [ctypes.POINTER(OConfig),
                                             ctypes.POINTER(OGrid), ctypes.c_int]
        L.o_init_shock_tube.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OGrid),
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int]
        L.o_grid_gather.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OGrid),
                                    ctypes.c_int, dp]
        L.o_vlct_timestep.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OBlock)]
        L.o_vlct_timestep.restype = ctypes.c_double
        L.o_vlct_compute.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OBlock),
                                     ctypes.c_double]
        L.o_vlct_compute.restype = ctypes.c_int
        L.o_hlld_face.argtypes = [dp, dp, ctypes.c_double, dp]
        L.o_hlle_mhd_face.argtypes = [dp, dp, ctypes.c_double, dp]
        L.o_hllc_face.argtypes = [dp, dp, ctypes.c_double, dp]
        L.o_hlld_face_de.argtypes = [dp, dp, ctypes.c_double, dp, dp]
        L.o_hlle_mhd_face_de.argtypes = [dp, dp, ctypes.c_double, dp, dp]
        L.o_hllc_face_de.argtypes = [dp, dp, ctypes.c_double, dp, dp]
        L.o_floor_energy_and_sync.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(dp),
                                              dp] + [ctypes.c_int] * 4
        L.o_internal_energy_source.argtypes = [ctypes.POINTER(OConfig), ctypes.c_int,
                                               ctypes.c_double, ctypes.c_double, dp, dp,
                                               dp] + [ctypes.c_int] * 4
        L.o_limiter_enzo.argtypes = [ctypes.c_double] * 4
        L.o_limiter_enzo.restype = ctypes.c_double
        L.o_limiter_athena.argtypes = [ctypes.c_double] * 3
        L.o_limiter_athena.restype = ctypes.c_double
        L.o_reconstruct.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp, dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double]
        L.o_update_bfield.argtypes = [dp, ctypes.c_int,
                                      ctypes.POINTER(dp), dp, dp,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int]
        L.o_compute_center_bfield.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int)
        L.o_flux_data_size.argtypes = [ctypes.c_int] * 4
        L.o_flux_data_size.restype = ctypes.c_long
        L.o_flux_data_offset.argtypes = [ctypes.c_int] * 7
        L.o_flux_data_offset.restype = ctypes.c_long
        L.o_vlct_compute_fc.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OBlock),
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int, ip, dp]
        L.o_vlct_compute_fc.restype = ctypes.c_int
        L.o_vlct_compute_accel.argtypes = [ctypes.POINTER(OConfig), ctypes.POINTER(OBlock),
                                           ctypes.c_double, ctypes.POINTER(dp)]
        L.o_vlct_compute_accel.restype = ctypes.c_int
        L.o_face_fluxes_coarsen.argtypes = [dp, dp] + [ctypes.c_int] * 7
        L.o_flux_send_to_coarse.argtypes = [dp, dp] + [ctypes.c_int] * 6 + [ip]
        L.o_flux_correct.argtypes = [ctypes.POINTER(dp), ctypes.c_int, ctypes.c_int, ip,
                                     dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ip]
        L.o_flux_correct.restype = ctypes.c_int
        L.o_prolong_linear.argtypes = [dp, ip, ip, ip, dp, ip, ip, ip, ctypes.c_int]
        L.o_prolong_linear.restype = ctypes.c_int
        L.o_restrict_linear.argtypes = [dp, ip, ip, ip, dp, ip, ip, ip, ctypes.c_int]
        L.o_restrict_linear.restype = ctypes.c_int
        L.o_enzo_prolong.argtypes = [ctypes.c_int, ctypes.c_int, dp, ip, ip, ip, dp, ip, ip,
                                     ip]
        L.o_enzo_prolong.restype = ctypes.c_int
        _lib = L
    return _lib


def ints(v):
    return (ctypes.c_int * len(v))(*v)


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_config(gamma=1.6666666666666667, riemann="hlld", recon=("nn", "plm"),
                theta=2.0, mhd=True, courant=0.4, density_floor=1.0e-200,
                pressure_floor=1.0e-200, dual_energy_eta=None):
    """Method and fluid parameters of input/vlct/vl.incl and vlct.incl;
    dual_energy_eta selects the "modern" dual-energy formalism with that
    eta (input/vlct/vlct_de.incl)."""
    c = OConfig()
    c.dual_energy = 0 if dual_energy_eta is None else 1
    c.eta = 0.0 if dual_energy_eta is None else dual_energy_eta
    c.gamma = gamma
    c.density_floor = density_floor
    c.pressure_floor = pressure_floor
    c.theta_limiter = theta
    c.riemann = RIEMANN[riemann]
    c.recon[0] = RECON[recon[0]]
    c.recon[1] = RECON[recon[1]] if len(recon) > 1 else 0
    c.nstages = len(recon)
    c.mhd = 1 if mhd else 0
    c.courant = courant
    return c


class Grid:
    """A uniform grid of blocks owned by the oracle library."""

    def __init__(self, cfg, nblocks, ncells, lower, upper, bc=("periodic",) * 3,
                 ghost=3):
        self.cfg = cfg
        g = OGrid()
        g.nbx, g.nby, g.nbz = nblocks
        g.nx, g.ny, g.nz = ncells
        g.g = ghost
        for a in range(3):
            g.lower[a] = lower[a]
            g.upper[a] = upper[a]
            g.bc[a] = BC[bc[a]]
        if lib().o_grid_alloc(ctypes.byref(g)) != 0:
            raise MemoryError("o_grid_alloc failed")
        self.g = g
        self.nblocks = nblocks
        self.ncells = ncells
        self.ghost = ghost

    def __del__(self):
        if getattr(self, "g", None) is not None and _lib is not None:
            _lib.o_grid_free(ctypes.byref(self.g))
            self.g = None

    def block(self, i):
        return self.g.blocks[i]

    def block_arrays(self, i):
        """Numpy views (no copy) of block i's fields."""
        b = self.g.blocks[i]
        mz, my, mx = b.mz, b.my, b.mx
        out = {}
        for q in range(8):
            out[FIELD_NAMES[q]] = np.ctypeslib.as_array(b.cc[q], shape=(mz, my, mx))
        out["pressure"] = np.ctypeslib.as_array(b.pressure, shape=(mz, my, mx))
        out["internal_energy"] = np.ctypeslib.as_array(b.eint, shape=(mz, my, mx))
        out["bfieldi_x"] = np.ctypeslib.as_array(b.bfi[0], shape=(mz, my, mx + 1))
        out["bfieldi_y"] = np.ctypeslib.as_array(b.bfi[1], shape=(mz, my + 1, mx))
        out["bfieldi_z"] = np.ctypeslib.as_array(b.bfi[2], shape=(mz + 1, my, mx))
        return out

    def gather(self, field):
        N = [self.nblocks[a] * self.ncells[a] for a in range(3)]
        out = np.zeros((N[2], N[1], N[0]))
        idx = FIELD_NAMES.index(field)
        lib().o_grid_gather(ctypes.byref(self.cfg), ctypes.byref(self.g), idx,
                            dptr(out))
        return out

    def refresh(self):
        lib().o_grid_refresh(ctypes.byref(self.cfg), ctypes.byref(self.g))

    def cycle(self, t, t_stop, nthreads=1):
        return lib().o_grid_cycle(ctypes.byref(self.cfg), ctypes.byref(self.g),
                                  t, t_stop, nthreads)

    def evolve(self, t0, t_stop, nthreads=1, max_log=100000):
        log = np.zeros(max_log)
        n = lib().o_grid_evolve(ctypes.byref(self.cfg), ctypes.byref(self.g), t0,
                                t_stop, nthreads, dptr(log), max_log)
        return n, log[:min(n, max_log)].copy()


def l1_norm(a_fields, b_fields, res=None):
    """Norm of the L1 error vector (tools/l1_error_norm.py:195-234): per field
    sum|a-b| / normalisation (res**3 when res is given, else the cell count),
    then the Euclidean norm over fields."""
    errs = []
    for a, b in zip(a_fields, b_fields):
        norm = (res ** 3) if res is not None else a.size
        errs.append(np.sum(np.abs(a - b)) / float(norm))
    errs = np.array(errs)
    return float(np.sqrt(np.sum(np.square(errs))))


MHD_L1_FIELDS = ["density", "velocity_x", "velocity_y", "velocity_z",
                 "pressure", "bfield_x", "bfield_y", "bfield_z"]
HD_L1_FIELDS = ["density", "velocity_x", "velocity_y", "velocity_z", "pressure"]

# inclined-wave geometry of the reference decks
# (input/vlct/MHD_linear_wave/initial_*.in)
WAVE_ALPHA = 0.7297276562269663
WAVE_BETA = 1.1071487177940904


def linear_wave_grid(wave, res, mhd=True, nblocks=(1, 1, 1), positive=True,
                     riemann=None):
    """Grid + initial state of the reference's inclined linear-wave decks:
    domain [0,3]x[0,1.5]^2 with (2N, N, N) cells split over nblocks."""
    cfg = make_config(riemann=riemann or ("hlld" if mhd else "hllc"), mhd=mhd)
    N = (2 * res, res, res)
    ncells = tuple(N[a] // nblocks[a] for a in range(3))
    grid = Grid(cfg, nblocks, ncells, (0.0, 0.0, 0.0), (3.0, 1.5, 1.5))
    rc = lib().o_init_inclined_wave(ctypes.byref(cfg), ctypes.byref(grid.g),
                                    wave.encode(), WAVE_ALPHA, WAVE_BETA, 1.e-6,
                                    1.0, 1 if positive else 0)
    if rc != 0:
        raise ValueError("bad wave %r" % wave)
    return grid


def run_linear_wave(wave, res, t_stop, mhd=True, nblocks=(1, 1, 1),
                    nthreads=1, positive=True):
    """Runs the reference linear-wave regression problem and returns
    (L1 error norm, number of cycles)."""
    grid = linear_wave_grid(wave, res, mhd=mhd, nblocks=nblocks,
                            positive=positive)
    fields = MHD_L1_FIELDS if mhd else HD_L1_FIELDS
    init = [grid.gather(f) for f in fields]
    ncyc, _ = grid.evolve(0.0, t_stop, nthreads=nthreads)
    final = [grid.gather(f) for f in fields]
    return l1_norm(init, final, res=res), ncyc


def load_reference_table(fname):
    """Reads a 1D reference table in the reference's CSV format ('#' header
    lines, then named columns) as a dict of arrays
    (tools/l1_error_norm.py:390-440)."""
    nhead = 0
    with open(fname) as f:
        for line in f:
            if line.startswith("#"):
                nhead += 1
            else:
                break
    rec = np.genfromtxt(fname, skip_header=nhead, delimiter=",", names=True,
                        dtype=None, encoding="utf-8")
    return {name: np.asarray(rec[name], dtype=np.float64)
            for name in rec.dtype.names}


def table_l1(grid, table, axis, fields=MHD_L1_FIELDS):
    """L1 error norm of a 3D grid against a 1D table along ``axis``, with the
    vector components permuted as the reference's --permute option does
    (tools/l1_error_norm.py:560-640)."""
    errs = []
    for f in fields:
        src = f
        if "_" in f and f.split("_")[-1] in ("x", "y", "z"):
            pre, comp = f.rsplit("_", 1)
            a = "xyz".index(comp)
            src = pre + "_" + "xyz"[(a - axis) % 3]
        shp = [1, 1, 1]
        shp[2 - axis] = table[src].size
        ref = table[src].reshape(shp)
        sim = grid.gather(f)
        errs.append(np.sum(np.abs(sim - ref)) / float(sim.size))
    return float(np.sqrt(np.sum(np.square(errs))))


def shifted_table_l1(gather, table, axis, lower, cell_width, offset=0.0,
                     bkg_velocity=(0.0, 0.0, 0.0), reverse=False, per_column=False,
                     fields=("density", "velocity_x", "velocity_y", "velocity_z")):
    """The reference's table-mode L1 error norm with --permute, --reverse,
    --offset_soln and --bkg_velocity (compare_to_1D_reference,
    tools/l1_error_norm.py:537-627): the table's positions are shifted by
    `offset`, the simulation is compared on the cells the shifted table
    covers along `axis` (every transverse cell), the table is optionally
    reversed with its vector components negated, vectors are permuted onto
    `axis`, and the background velocity is added (total energy adjusted when
    tabulated).  `gather(name)` returns a simulation field as a (z, y, x)
    array of the whole domain; `lower`/`cell_width` describe that axis.
    Fields missing from the table are skipped (find_common_fields, :274-330).
    per_column: the norms of every 1D slice along `axis` (the --std input,
    L1_error_vector with sum_axis, :195-234) instead of the total."""
    pos_key = [k for k in table if k in ("position", "pos", "x", "y", "z")][0]
    pos = np.asarray(table[pos_key], dtype=np.float64).copy()
    cw = np.diff(pos)[0]
    if offset != 0.0:
        if offset % abs(cw) != 0.0:
            raise ValueError("offset is not a multiple of the table's cell width")
        pos += offset
    left = pos[0] - cw * 0.5
    i0 = int((left - lower) / cell_width + 0.5)
    if not np.isclose(lower + i0 * cell_width, left, atol=1e-13, rtol=0):
        raise ValueError("the table does not start at a cell boundary")
    n = pos.size
    ref = {k: np.asarray(v, dtype=np.float64).copy() for k, v in table.items()
           if k != pos_key}

    def vectors(d):
        cand = {}
        for k in d:
            t = k.split("_")
            if len(t) == 2 and t[1] in "xyz" and len(t[1]) == 1:
                cand.setdefault(t[0], [None, None, None])["xyz".index(t[1])] = k
        return cand

    if reverse:
        for k in ref:
            ref[k] = ref[k][::-1].copy()
        for names in vectors(ref).values():
            for k in names:
                if k is not None:
                    ref[k] = ref[k] * -1.0
    check = ["velocity_x", "velocity_y", "velocity_z"]
    if any(b in ref for b in ("bfield_x", "bfield_y", "bfield_z")):
        check += ["bfield_x", "bfield_y", "bfield_z"]
    for k in check:
        if k not in ref:
            ref[k] = np.zeros_like(pos)
    for prefix, names in vectors(ref).items():
        if None in names:
            continue
        tmp = [ref[names[0]], ref[names[1]], ref[names[2]]]
        ref[names[axis]] = tmp[0]
        ref[names[(axis + 1) % 3]] = tmp[1]
        ref[names[(axis + 2) % 3]] = tmp[2]
    for i, k in enumerate(["velocity_x", "velocity_y", "velocity_z"]):
        old = ref[k].copy()
        ref[k] = ref[k] + bkg_velocity[i]
        if "total_energy" in ref:
            ref["total_energy"] = ref["total_energy"] + 0.5 * (np.square(ref[k]) -
                                                               np.square(old))
    shp = [1, 1, 1]
    shp[2 - axis] = n
    ax_np = 2 - axis
    errs = []
    for f in fields:
        if f not in ref:
            continue
        sim = np.take(gather(f), np.arange(i0, i0 + n), axis=ax_np)
        diff = np.abs(sim - ref[f].reshape(shp))
        if per_column:
            errs.append(np.sum(diff, axis=ax_np) / float(n))
        else:
            errs.append(np.sum(diff) / float(sim.size))
    errs = np.array(errs)
    if per_column:
        return np.sqrt(np.sum(np.square(errs), axis=0))
    return float(np.sqrt(np.sum(np.square(errs))))


# the dual-energy shock tube decks (input/vlct/dual_energy_shock_tube/
# method_vlct_sod_*_de_M10*.in with vlct_de.incl and
# initial_sod_shock_tube.incl): 508 cells along the tube over 4 blocks,
# 4 across, a Mach ~10 background velocity along the tube
DE_TUBE_ETA = 0.00769
DE_TUBE_AXIS_VELOCITY = 11.875
DE_TUBE_OFFSET = 2.96875


def sod_de_grid(axis, reverse=False, n=508, ntrans=4, nblocks=4, eta=DE_TUBE_ETA):
    """Grid + initial state of the dual-energy Sod shock tube deck along
    `axis` (the _reverse deck: flip_initialize on [-2.96875, 1])."""
    cfg = make_config(gamma=1.4, dual_energy_eta=eta)
    nc = [ntrans, ntrans, ntrans]
    nc[axis] = n // nblocks
    nb = [1, 1, 1]
    nb[axis] = nblocks
    lower = [0.0, 0.0, 0.0]
    upper = [0.03125, 0.03125, 0.03125]
    if reverse:
        lower[axis], upper[axis] = -2.96875, 1.0
    else:
        upper[axis] = 3.96875
    bc = ["periodic"] * 3
    bc[axis] = "outflow"
    grid = Grid(cfg, tuple(nb), tuple(nc), tuple(lower), tuple(upper), bc=tuple(bc))
    rc = lib().o_init_shock_tube(ctypes.byref(cfg), ctypes.byref(grid.g), 1, axis,
                                 DE_TUBE_AXIS_VELOCITY, 0.0, 1 if reverse else 0)
    if rc != 0:
        raise ValueError("sod init failed")
    grid.lower, grid.upper = lower, upper
    return grid


def rj2a_grid(axis, n=256, ntrans=4):
    """Grid + initial state of input/vlct/MHD_shock_tube/method_vlct_*_rj2a_N256.in."""
    cfg = make_config()
    nc = [ntrans, ntrans, ntrans]
    nc[axis] = n
    bc = ["periodic"] * 3
    bc[axis] = "outflow"
    grid = Grid(cfg, (1, 1, 1), tuple(nc), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                bc=tuple(bc))
    rc = lib().o_init_shock_tube_rj2a(ctypes.byref(cfg), ctypes.byref(grid.g),
                                      axis)
    if rc != 0:
        raise ValueError("rj2a init failed")
    return grid
