"""Python binding of the gfx950 VL+CT library (``libenzoe_vlct.so``).

Thin ctypes layer over the C ABI declared in ``include/enzoe_vlct.h`` plus
the few HIP runtime calls needed to own device memory.  The product path is
always the HIP library: importing this module on a machine without the built
library or without a HIP device raises instead of falling back to the CPU.

Blocks are stored the way Cello stores them (src/Cello/data_FieldData.cpp:
319-395): one device allocation per block holding every field at its own
offset, each field a dense (z, y, x) array with ghost zones.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libenzoe_vlct.so")

ENZOE_PREC_DOUBLE = 0
ENZOE_PREC_MIXED = 1
ENZOE_PREC_DOUBLE_FMA = 2  # FP64, multiply-adds contracted (not bitwise)
BC = {"periodic": 0, "outflow": 1}

CC_FIELDS = ["density", "velocity_x", "velocity_y", "velocity_z",
             "total_energy", "bfield_x", "bfield_y", "bfield_z"]
FACE_FIELDS = ["bfieldi_x", "bfieldi_y", "bfieldi_z"]


class EnzoeError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("mhd_choice", ctypes.c_char_p),
                ("riemann_solver", ctypes.c_char_p),
                ("reconstruct_method", ctypes.c_char_p),
                ("time_scheme", ctypes.c_char_p),
                ("theta_limiter", ctypes.c_double),
                ("courant", ctypes.c_double),
                ("gamma", ctypes.c_double),
                ("density_floor", ctypes.c_double),
                ("pressure_floor", ctypes.c_double),
                ("precision", ctypes.c_int),
                ("dual_energy", ctypes.c_char_p),
                ("dual_energy_eta", ctypes.c_double)]


_dp = ctypes.POINTER(ctypes.c_double)


class Block(ctypes.Structure):
    _fields_ = [("mx", ctypes.c_int), ("my", ctypes.c_int), ("mz", ctypes.c_int),
                ("ghost", ctypes.c_int),
                ("cell_width", ctypes.c_double * 3),
                ("density", ctypes.c_void_p),
                ("velocity", ctypes.c_void_p * 3),
                ("total_energy", ctypes.c_void_p),
                ("bfield", ctypes.c_void_p * 3),
                ("bfieldi", ctypes.c_void_p * 3),
                ("pressure", ctypes.c_void_p),
                ("internal_energy", ctypes.c_void_p)]


class Diagnostics(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_double), ("momentum", ctypes.c_double * 3),
                ("energy", ctypes.c_double), ("min_density", ctypes.c_double),
                ("min_pressure", ctypes.c_double), ("max_abs_divb", ctypes.c_double),
                ("max_abs_bface", ctypes.c_double), ("nonfinite", ctypes.c_longlong)]


class StateL1(ctypes.Structure):
    _fields_ = [("diff", ctypes.c_double * 8), ("ref", ctypes.c_double * 8),
                ("velocity_magnitude", ctypes.c_double),
                ("bfield_magnitude", ctypes.c_double),
                ("max_rel", ctypes.c_double * 8)]


DEBUG_COUNTS = ("nonfinite", "density_floor", "energy_floor", "reconstruction_floor")

_lib = None
_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise EnzoeError("HIP runtime (libamdhip64.so) not found")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float),
                                             ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return _hip


def hip_check(rc, what):
    if rc != 0:
        raise EnzoeError("%s failed with hipError %d" % (what, rc))


def device_count():
    n = ctypes.c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def lib():
    """The HIP library.  Raises if it was not built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EnzoeError("libenzoe_vlct.so is not built (run "
                             "__graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.enzoe_version.restype = ctypes.c_char_p
        L.enzoe_vlct_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.enzoe_vlct_destroy.argtypes = [vp]
        L.enzoe_vlct_courant.argtypes = [vp]
        L.enzoe_vlct_courant.restype = ctypes.c_double
        L.enzoe_vlct_timestep.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                          vp, _dp]
        L.enzoe_vlct_compute.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                         ctypes.c_double, vp]
        L.enzoe_vlct_compute_timestep.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                                  ctypes.c_double, vp, _dp]
        L.enzoe_vlct_compute_dt.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                            _dp, ctypes.c_int, vp]
        L.enzoe_vlct_autotune.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                          ctypes.c_double, vp, ctypes.POINTER(ctypes.c_int)]
        L.enzoe_vlct_timestep_device.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                                 _dp, ctypes.c_double, _dp, ctypes.c_int,
                                                 vp]
        L.enzoe_vlct_graph_create.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int), ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int,
                                              ctypes.POINTER(vp)]
        L.enzoe_vlct_levels_create.argtypes = [vp, ctypes.POINTER(Block),
                                               ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                               ctypes.POINTER(vp)]
        L.enzoe_vlct_levels_step.argtypes = [vp, vp]
        L.enzoe_vlct_levels_capture.argtypes = [vp, ctypes.c_int]
        L.enzoe_vlct_levels_replay.argtypes = [vp, vp]
        L.enzoe_vlct_levels_state.argtypes = [vp, vp, _dp, _dp]
        L.enzoe_vlct_levels_destroy.argtypes = [vp]
        L.enzoe_vlct_graph_launch.argtypes = [vp, vp]
        L.enzoe_vlct_graph_time.argtypes = [vp, vp, _dp]
        L.enzoe_vlct_graph_destroy.argtypes = [vp]
        L.enzoe_grid_refresh.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), vp]
        L.enzoe_vlct_set_debug_checks.argtypes = [vp, ctypes.c_int]
        L.enzoe_vlct_set_fused_ct.argtypes = [vp, ctypes.c_int]
        L.enzoe_vlct_check_persistent.argtypes = [vp]
        L.enzoe_vlct_set_flux_shape.argtypes = [vp, ctypes.c_int]
        L.enzoe_vlct_run_cycles.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int), vp, ctypes.c_double,
                                            ctypes.c_int, vp]
        L.enzoe_vlct_launch_choice.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int)]
        L.enzoe_vlct_compute_noalloc.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                                 ctypes.c_double, vp]
        L.enzoe_vlct_noalloc_bytes.restype = ctypes.c_size_t
        L.enzoe_vlct_debug_counts.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_ulonglong)]
        L.enzoe_vlct_diagnostics.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int, vp,
                                             ctypes.POINTER(Diagnostics)]
        L.enzoe_vlct_state_l1.argtypes = [vp, ctypes.POINTER(Block), ctypes.POINTER(Block),
                                          ctypes.c_int, vp, ctypes.POINTER(StateL1)]
        L.enzoe_set_profiling.argtypes = [ctypes.c_int]
        L.enzoe_profile_report.argtypes = [ctypes.c_char_p, ctypes.c_int, _dp,
                                           ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        raise EnzoeError("%s returned status %d" % (what, rc))


# ------------------------------------------------------------ device memory

class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        hip_check(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 8)),
                  "hipMalloc(%d)" % self.nbytes)

    def free(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_device(dst_ptr, arr):
    a = np.ascontiguousarray(arr, dtype=np.float64)
    hip_check(hip().hipMemcpy(ctypes.c_void_p(dst_ptr), a.ctypes.data_as(ctypes.c_void_p),
                              a.nbytes, 1), "hipMemcpy H2D")


def from_device(src_ptr, shape):
    a = np.empty(shape, dtype=np.float64)
    hip_check(hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src_ptr),
                              a.nbytes, 2), "hipMemcpy D2H")
    return a


def field_shapes(mx, my, mz, mhd, dual_energy=False):
    shapes = {}
    names = CC_FIELDS if mhd else CC_FIELDS[:5]
    for f in names:
        shapes[f] = (mz, my, mx)
    shapes["pressure"] = (mz, my, mx)
    if dual_energy:
        shapes["internal_energy"] = (mz, my, mx)
    if mhd:
        shapes["bfieldi_x"] = (mz, my, mx + 1)
        shapes["bfieldi_y"] = (mz, my + 1, mx)
        shapes["bfieldi_z"] = (mz + 1, my, mx)
    return shapes


class DeviceBlock:
    """One block's permanent fields in one device allocation."""

    on_device = True

    def __init__(self, mx, my, mz, ghost, cell_width, mhd=True, dual_energy=False):
        self.mx, self.my, self.mz, self.ghost = mx, my, mz, ghost
        self.mhd = mhd
        self.dual_energy = dual_energy
        self.shapes = field_shapes(mx, my, mz, mhd, dual_energy)
        self.offsets = {}
        off = 0
        for name, shp in self.shapes.items():
            self.offsets[name] = off
            n = int(np.prod(shp)) * 8
            off += (n + 255) // 256 * 256  # 256-byte aligned fields
        self.buf = DeviceBuffer(off)
        hip_check(hip().hipMemset(self.buf.ptr, 0, off), "hipMemset")
        base = self.buf.ptr.value
        b = Block()
        b.mx, b.my, b.mz, b.ghost = mx, my, mz, ghost
        for a in range(3):
            b.cell_width[a] = cell_width[a]
        b.density = base + self.offsets["density"]
        for a, c in enumerate("xyz"):
            b.velocity[a] = base + self.offsets["velocity_" + c]
            if mhd:
                b.bfield[a] = base + self.offsets["bfield_" + c]
                b.bfieldi[a] = base + self.offsets["bfieldi_" + c]
        b.total_energy = base + self.offsets["total_energy"]
        b.pressure = base + self.offsets["pressure"]
        if dual_energy:
            b.internal_energy = base + self.offsets["internal_energy"]
        self.desc = b

    def ptr(self, name):
        return self.buf.ptr.value + self.offsets[name]

    def upload(self, arrays):
        for name in self.shapes:
            if name in arrays:
                assert arrays[name].shape == self.shapes[name], name
                to_device(self.ptr(name), arrays[name])

    def download(self, names=None):
        names = names or list(self.shapes)
        return {n: from_device(self.ptr(n), self.shapes[n]) for n in names}


# ------------------------------------------------------------------- method

class VlctMethod:
    """EnzoMethodMHDVlct on the GPU (through the C ABI)."""

    def __init__(self, mhd=True, riemann="hlld", reconstruct="plm",
                 time_scheme="vl", theta=1.5, courant=None, gamma=5.0 / 3.0,
                 density_floor=1e-200, pressure_floor=1e-200,
                 precision=ENZOE_PREC_DOUBLE, dual_energy_eta=None):
        """dual_energy_eta: None disables the dual-energy formalism; a
        number selects the "modern" formulation with that eta."""
        p = Params()
        p.mhd_choice = b"constrained_transport" if mhd else b"no_bfield"
        p.riemann_solver = riemann.encode() if riemann else None
        p.reconstruct_method = reconstruct.encode() if reconstruct else None
        p.time_scheme = time_scheme.encode() if time_scheme else None
        p.theta_limiter = theta
        p.courant = -1.0 if courant is None else courant
        p.gamma = gamma
        p.density_floor = density_floor
        p.pressure_floor = pressure_floor
        p.precision = precision
        p.dual_energy = b"disabled" if dual_energy_eta is None else b"modern"
        p.dual_energy_eta = -1.0 if dual_energy_eta is None else dual_energy_eta
        self.mhd = mhd
        self.dual_energy = dual_energy_eta is not None
        self.handle = ctypes.c_void_p()
        check(lib().enzoe_vlct_create(ctypes.byref(p), ctypes.byref(self.handle)),
              "enzoe_vlct_create")

    def __del__(self):
        if getattr(self, "handle", None) is not None and self.handle.value and _lib:
            _lib.enzoe_vlct_destroy(self.handle)
            self.handle = None

    @property
    def courant(self):
        return lib().enzoe_vlct_courant(self.handle)

    @staticmethod
    def _blocks(blocks):
        arr = (Block * len(blocks))(*[b.desc for b in blocks])
        return arr

    def timestep(self, blocks, stream=None):
        dt = ctypes.c_double(0)
        check(lib().enzoe_vlct_timestep(self.handle, self._blocks(blocks), len(blocks),
                                        stream, ctypes.byref(dt)), "enzoe_vlct_timestep")
        return dt.value

    def compute(self, blocks, dt, stream=None):
        check(lib().enzoe_vlct_compute(self.handle, self._blocks(blocks), len(blocks),
                                       dt, stream), "enzoe_vlct_compute")

    def compute_timestep(self, blocks, dt, stream=None):
        """compute(blocks, dt), then the CFL dt of the result, fused; returns
        the next dt (equal to timestep() after compute())."""
        nxt = ctypes.c_double(0)
        check(lib().enzoe_vlct_compute_timestep(self.handle, self._blocks(blocks), len(blocks),
                                                dt, stream, ctypes.byref(nxt)),
              "enzoe_vlct_compute_timestep")
        return nxt.value

    def set_debug_checks(self, on=True):
        """Counts non-finite updates and applied floors on the device (the
        reference's CHECK_FINITE_ELEMENTS / RAISE_FLOOR_ERROR builds)."""
        check(lib().enzoe_vlct_set_debug_checks(self.handle, 1 if on else 0),
              "enzoe_vlct_set_debug_checks")

    def debug_counts(self, stream=None):
        """The counters since the last call, as a dict; clears them."""
        c = (ctypes.c_ulonglong * 4)()
        check(lib().enzoe_vlct_debug_counts(self.handle, stream, c),
              "enzoe_vlct_debug_counts")
        return dict(zip(DEBUG_COUNTS, [int(x) for x in c]))

    def diagnostics(self, blocks, stream=None):
        """Device summary of the blocks: conserved totals over the active
        cells, minimum density and pressure, max |div B| dx, max |face B|,
        and the number of non-finite values (ghosts included)."""
        d = Diagnostics()
        check(lib().enzoe_vlct_diagnostics(self.handle, self._blocks(blocks), len(blocks),
                                           stream, ctypes.byref(d)), "enzoe_vlct_diagnostics")
        return {"mass": d.mass, "momentum": tuple(d.momentum), "energy": d.energy,
                "min_density": d.min_density, "min_pressure": d.min_pressure,
                "max_abs_divb": d.max_abs_divb, "max_abs_bface": d.max_abs_bface,
                "nonfinite": int(d.nonfinite)}

    def set_fused_ct(self, variant):
        """Launch structure of the stages: 0 (or False) the flux launch and
        (MHD) three launches through memory, 1..3 (True = 1) one fused CT +
        update launch with tile 16x4x4 / 32x4x2 / 8x8x4, 4 the edge EMFs then
        one face + cell launch, 5 each stage as one persistent launch, -1 the
        autotuner's choice."""
        v = int(variant) if not isinstance(variant, bool) else (1 if variant else 0)
        check(lib().enzoe_vlct_set_fused_ct(self.handle, v), "enzoe_vlct_set_fused_ct")

    def set_flux_shape(self, shape):
        """Shape of the flux launch: 0 three faces per thread, 1 one face per
        thread, 2 z pencils for the nearest-neighbour stage, -1 the
        autotuner's choice.  Every shape gives the same bits."""
        check(lib().enzoe_vlct_set_flux_shape(self.handle, int(shape)),
              "enzoe_vlct_set_flux_shape")

    def check_persistent(self):
        """Raises if a barrier wait of a persistent stage launch ran out
        since the last check (synchronises)."""
        check(lib().enzoe_vlct_check_persistent(self.handle), "enzoe_vlct_check_persistent")

    def launch_choice(self, shape):
        """(XCD mask, launch structure, tile rows, flux-launch shape) used for
        blocks of shape (mx, my, mz)."""
        mask, ctu, rows, split = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().enzoe_vlct_launch_choice(self.handle, shape[0], shape[1], shape[2],
                                             ctypes.byref(mask), ctypes.byref(ctu),
                                             ctypes.byref(rows), ctypes.byref(split)),
              "enzoe_vlct_launch_choice")
        return mask.value, ctu.value, rows.value, split.value

    def state_l1(self, blocks, ref_blocks, stream=None, max_rel=False):
        """L1 of the blocks' cell-centred fields against ref_blocks' (same
        shapes), over the active cells, on the device.  Returns (per_field,
        by_magnitude): per field sum|a - b| / sum|b|, and the same with the
        denominator of a velocity or B component the sum of its vector's
        magnitude |b_vec|.  max_rel=True: also, third, per field the largest
        per-cell relative difference |a - b| / |b| (a velocity or B
        component over the cell's |b_vec|)."""
        out = StateL1()
        check(lib().enzoe_vlct_state_l1(self.handle, self._blocks(blocks),
                                        self._blocks(ref_blocks), len(blocks), stream,
                                        ctypes.byref(out)), "enzoe_vlct_state_l1")
        names = CC_FIELDS if self.mhd else CC_FIELDS[:5]
        per, mag = {}, {}
        for q, nm in enumerate(names):
            d, r = out.diff[q], out.ref[q]
            per[nm] = d / r if r > 0 else d
            vec = {"velocity": out.velocity_magnitude,
                   "bfield": out.bfield_magnitude}.get(nm.rsplit("_", 1)[0])
            den = vec if vec else r
            mag[nm] = d / den if den > 0 else d
        if max_rel:
            return per, mag, {nm: out.max_rel[q] for q, nm in enumerate(names)}
        return per, mag

    def autotune(self, blocks, dt, stream=None):
        """Picks the launch order of the stage kernels for this block shape
        by timing the candidates on a copy of the blocks; returns the mask."""
        mask = ctypes.c_int(-1)
        check(lib().enzoe_vlct_autotune(self.handle, self._blocks(blocks), len(blocks), dt,
                                        stream, ctypes.byref(mask)), "enzoe_vlct_autotune")
        return mask.value

    def compute_dt(self, blocks, dts, stream=None, block_array=None):
        """One step with one dt per block (sub-cycled levels advanced
        together).  `block_array` may be a prepared ctypes array of the
        blocks' descriptors (it is rebuilt otherwise)."""
        arr = block_array if block_array is not None else self._blocks(blocks)
        d = (ctypes.c_double * len(dts))(*dts)
        check(lib().enzoe_vlct_compute_dt(self.handle, arr, len(dts), d, 0, stream),
              "enzoe_vlct_compute_dt")

    def refresh(self, blocks, nblocks, bc, stream=None):
        bcs = (ctypes.c_int * 3)(*[BC[b] if isinstance(b, str) else b for b in bc])
        check(lib().enzoe_grid_refresh(self.handle, self._blocks(blocks), nblocks[0],
                                       nblocks[1], nblocks[2], bcs, stream),
              "enzoe_grid_refresh")


def synchronize():
    hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def profiling(on):
    lib().enzoe_set_profiling(1 if on else 0)


def profile_report():
    cap, nl = 64, 64
    names = ctypes.create_string_buffer(cap * nl)
    ms = (ctypes.c_double * cap)()
    cnt = (ctypes.c_int * cap)()
    n = lib().enzoe_profile_report(names, nl, ms, cnt, cap)
    out = {}
    for i in range(n):
        nm = names.raw[i * nl:(i + 1) * nl].split(b"\0", 1)[0].decode()
        out[nm] = (ms[i], cnt[i])
    return out


class DeviceGrid:
    """A uniform grid of equally sized blocks resident on the device."""

    def __init__(self, method, nblocks, ncells, cell_width, ghost=3,
                 bc=("periodic",) * 3):
        self.method = method
        self.nblocks = tuple(nblocks)
        self.ncells = tuple(ncells)
        self.ghost = ghost
        self.bc = tuple(bc)
        m = [ncells[a] + 2 * ghost for a in range(3)]
        nb = nblocks[0] * nblocks[1] * nblocks[2]
        self.blocks = [DeviceBlock(m[0], m[1], m[2], ghost, cell_width[i] if
                                   isinstance(cell_width[0], (list, tuple)) else cell_width,
                                   mhd=method.mhd,
                                   dual_energy=getattr(method, "dual_energy", False))
                       for i in range(nb)]

    def upload_from(self, arrays_per_block):
        for blk, arrs in zip(self.blocks, arrays_per_block):
            blk.upload(arrs)

    def download(self, names=None):
        return [b.download(names) for b in self.blocks]

    def refresh(self):
        self.method.refresh(self.blocks, self.nblocks, self.bc)

    def cycle(self, t, t_stop):
        """One Cello cycle: dt on current data, refresh, compute."""
        dt = self.method.timestep(self.blocks)
        if t < t_stop < t + dt:
            dt = t_stop - t
        rem = t_stop - t
        dt = dt if dt < rem else rem
        self.refresh()
        self.method.compute(self.blocks, dt)
        return dt

    def graph(self, t0, t_stop, ncycles):
        """A captured graph of `ncycles` cycles of this grid (device dt)."""
        return CycleGraph(self, t0, t_stop, ncycles)

    def evolve_graph(self, t0, t_stop, ncycles=16, max_replays=1 << 20):
        """evolve() with device-driven cycles: replays a graph of `ncycles`
        cycles until the device time reaches t_stop.  Returns the number of
        replays and the final time."""
        g = self.graph(t0, t_stop, ncycles)
        t, k = t0, 0
        while t < t_stop and k < max_replays:
            g.launch()
            t = g.time()
            k += 1
        g.destroy()
        return k, t

    def evolve_device(self, t0, t_stop, ncycles=16, max_calls=1 << 20):
        """evolve() with device-driven cycles enqueued without a graph
        (enzoe_vlct_run_cycles): `ncycles` per call, the device time read
        back after each call.  Returns the number of calls and the final
        time."""
        n = len(self.blocks)
        state = DeviceBuffer((n + 1) * 8)
        to_device(state.ptr.value, np.array([t0] + [0.0] * n))
        arr = (Block * n)(*[b.desc for b in self.blocks])
        bcs = (ctypes.c_int * 3)(*[BC[b] if isinstance(b, str) else b for b in self.bc])
        t, k = t0, 0
        while t < t_stop and k < max_calls:
            check(lib().enzoe_vlct_run_cycles(self.method.handle, arr, self.nblocks[0],
                                              self.nblocks[1], self.nblocks[2], bcs,
                                              state.ptr, t_stop, ncycles, None),
                  "enzoe_vlct_run_cycles")
            t = float(from_device(state.ptr.value, (1,))[0])
            k += 1
        state.free()
        return k, t

    def evolve(self, t0, t_stop, max_cycles=1 << 30):
        t, n, log = t0, 0, []
        while t < t_stop and n < max_cycles:
            dt = self.cycle(t, t_stop)
            log.append(dt)
            t = t + dt
            n += 1
        synchronize()
        return n, log


class Levels:
    """enzoe_vlct_levels_*: sub-cycled refinement levels stepped on the
    device (per-level CFL dt, dt_0 = min_l 2^l dt_l, 2^(L-1) batched
    substeps), eagerly or from a captured HIP graph.  levels[l] is the list
    of DeviceBlocks of level l (0 = coarsest)."""

    def __init__(self, method, levels):
        self.method = method
        self.levels = [list(lv) for lv in levels]
        blocks = [b for lv in self.levels for b in lv]
        level_of = [l for l, lv in enumerate(self.levels) for _ in lv]
        arr = (Block * len(blocks))(*[b.desc for b in blocks])
        lev = (ctypes.c_int * len(blocks))(*level_of)
        self.handle = ctypes.c_void_p()
        check(lib().enzoe_vlct_levels_create(method.handle, arr, lev, len(blocks),
                                             ctypes.byref(self.handle)),
              "enzoe_vlct_levels_create")

    def step(self, stream=None):
        check(lib().enzoe_vlct_levels_step(self.handle, stream), "enzoe_vlct_levels_step")

    def capture(self, nsteps):
        check(lib().enzoe_vlct_levels_capture(self.handle, nsteps), "enzoe_vlct_levels_capture")

    def replay(self, stream=None):
        check(lib().enzoe_vlct_levels_replay(self.handle, stream), "enzoe_vlct_levels_replay")

    def state(self, stream=None):
        """(device time, dt_0 of the last step)."""
        t, d = ctypes.c_double(), ctypes.c_double()
        check(lib().enzoe_vlct_levels_state(self.handle, stream, ctypes.byref(t),
                                            ctypes.byref(d)), "enzoe_vlct_levels_state")
        return t.value, d.value

    def destroy(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().enzoe_vlct_levels_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class CycleGraph:
    """enzoe_vlct_graph_*: cycles of a DeviceGrid captured in a HIP graph."""

    def __init__(self, grid, t0, t_stop, ncycles):
        self.grid = grid
        self.arr = (Block * len(grid.blocks))(*[b.desc for b in grid.blocks])
        bcs = (ctypes.c_int * 3)(*[BC[b] if isinstance(b, str) else b for b in grid.bc])
        self.handle = ctypes.c_void_p()
        check(lib().enzoe_vlct_graph_create(grid.method.handle, self.arr, grid.nblocks[0],
                                            grid.nblocks[1], grid.nblocks[2], bcs, t0,
                                            t_stop, ncycles, ctypes.byref(self.handle)),
              "enzoe_vlct_graph_create")

    def launch(self, stream=None):
        check(lib().enzoe_vlct_graph_launch(self.handle, stream), "enzoe_vlct_graph_launch")

    def time(self, stream=None):
        t = ctypes.c_double()
        check(lib().enzoe_vlct_graph_time(self.handle, stream, ctypes.byref(t)),
              "enzoe_vlct_graph_time")
        return t.value

    def destroy(self):
        if self.handle is not None and self.handle.value:
            lib().enzoe_vlct_graph_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ------------------------------------------------ single-operation views

SLOTS = {"density": 0, "velocity_x": 1, "velocity_y": 2, "velocity_z": 3,
         "total_energy": 4, "pressure": 4, "bfield_x": 5, "bfield_y": 6,
         "bfield_z": 7, "internal_energy": 16}
MAP_SLOTS = 17


class FArray(ctypes.Structure):
    """enzoe_farray: one device view (pointer, (z,y,x) shape, strides)."""
    _fields_ = [("data", ctypes.c_void_p), ("shape", ctypes.c_int * 3),
                ("stride", ctypes.c_longlong * 3)]


class FMap(ctypes.Structure):
    """enzoe_fmap: same-shape device views by slot."""
    _fields_ = [("data", ctypes.c_void_p * MAP_SLOTS), ("shape", ctypes.c_int * 3),
                ("stride", ctypes.c_longlong * 3)]


def farray(ptr, shape, stride=None):
    """View of device memory at `ptr` (bytes) with (z,y,x) shape and element
    strides (dense when omitted)."""
    a = FArray()
    a.data = ptr
    if stride is None:
        stride = (shape[1] * shape[2], shape[2], 1)
    for i in range(3):
        a.shape[i] = shape[i]
        a.stride[i] = stride[i]
    return a


def fmap(ptrs, shape, stride=None):
    """Map from {key or slot: device pointer} with a common shape/stride."""
    m = FMap()
    if stride is None:
        stride = (shape[1] * shape[2], shape[2], 1)
    for k, p in ptrs.items():
        m.data[SLOTS[k] if isinstance(k, str) else k] = p
    for i in range(3):
        m.shape[i] = shape[i]
        m.stride[i] = stride[i]
    return m


_ops_bound = False


def ops():
    """The library with the single-operation entry points bound."""
    global _ops_bound
    L = lib()
    if not _ops_bound:
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        PM, PA = ctypes.POINTER(FMap), ctypes.POINTER(FArray)
        L.enzoe_reconstruct_interface.argtypes = [ctypes.c_char_p, d, d, d, PM, PM, PM,
                                                  i, i, vp]
        L.enzoe_reconstructor_staling.argtypes = [ctypes.c_char_p,
                                                  ctypes.POINTER(i), ctypes.POINTER(i)]
        L.enzoe_riemann_solve.argtypes = [ctypes.c_char_p, i, d, PM, PM, PM, i, i, vp]
        L.enzoe_ct_correct_reconstructed_bfield.argtypes = [i, PA, PM, PM, i, vp]
        L.enzoe_ct_identify_upwind.argtypes = [i, PM, PA, i, vp]
        L.enzoe_ct_compute_center_efield.argtypes = [i, PA, PM, i, vp]
        L.enzoe_ct_compute_edge_efield.argtypes = [i, PA, PA, PM, PM, PA, i, vp]
        L.enzoe_ct_update_bfield.argtypes = [_dp, i, PA, PA, PA, d, i, vp]
        L.enzoe_ct_compute_center_bfield.argtypes = [i, PA, PA, i, vp]
        L.enzoe_fill_map.argtypes = [PM, d, vp]
        L.enzoe_accumulate_flux_component.argtypes = [i, d, d, PM, PM, i, vp]
        L.enzoe_update_quantities.argtypes = [PM, PM, PM, i, d, d, d, i, vp]
        L.enzoe_primitive_from_integration.argtypes = [PM, PM, i, d, i, vp]
        L.enzoe_pressure_from_integration.argtypes = [PM, PA, i, d, i, vp]
        L.enzoe_riemann_solve_dual_energy.argtypes = [ctypes.c_char_p, i, d, PM, PM, PM,
                                                      PA, i, i, vp]
        L.enzoe_source_internal_energy.argtypes = [i, d, d, PM, PM, PA, d, i, vp]
        L.enzoe_update_quantities_dual_energy.argtypes = [PM, PM, PM, i, d, d, d, d, i, vp]
        L.enzoe_primitive_from_integration_dual_energy.argtypes = [PM, PM, i, d, i, vp]
        L.enzoe_pressure_from_integration_dual_energy.argtypes = [PM, PA, d, i, vp]
        L.enzoe_apply_floor_to_energy_and_sync.argtypes = [PM, i, d, d, i, d, i, vp]
        L.enzoe_integrator_timestep.argtypes = [PM, PA, i, d, d, d, d, vp, _dp]
        PI = ctypes.POINTER(i)
        L.enzoe_flux_data_size.argtypes = [i, i, i, i]
        L.enzoe_flux_data_size.restype = ctypes.c_longlong
        L.enzoe_vlct_compute_store_fluxes.argtypes = [vp, ctypes.POINTER(Block), i, d, i,
                                                      PI, ctypes.POINTER(vp), vp]
        L.enzoe_save_fluxes_for_corrections.argtypes = [PM, i, d, d, i, i, PI, vp, vp]
        L.enzoe_flux_data_send_to_coarse.argtypes = [vp, vp, i, i, i, i, i, i, PI, vp]
        L.enzoe_flux_correct.argtypes = [ctypes.POINTER(vp), i, i, PI, vp, i, i, i, i,
                                         PI, vp]
        L.enzoe_prolong_linear.argtypes = [vp, PI, PI, PI, vp, PI, PI, PI, i, vp]
        L.enzoe_enzo_prolong.argtypes = [i, i, vp, PI, PI, PI, vp, PI, PI, PI,
                                         ctypes.POINTER(i), vp]
        L.enzoe_vlct_transfer_batch.argtypes = [vp, ctypes.POINTER(Transfer), i, vp]
        L.enzoe_restrict_linear.argtypes = [vp, PI, PI, PI, vp, PI, PI, PI, i, vp]
        L.enzoe_vlct_compute_gravity.argtypes = [vp, ctypes.POINTER(Block), i, d,
                                                 ctypes.POINTER(vp), vp]
        L.enzoe_source_gravity.argtypes = [d, PM, PM, PM, i, vp]
        L.enzoe_vlct_flux_send_to_coarse_batch.argtypes = [vp, ctypes.POINTER(FluxLink), i,
                                                           i, i, i, i, vp]
        L.enzoe_vlct_flux_correct_blocks.argtypes = [vp, ctypes.POINTER(FluxCorrectBlock),
                                                     i, i, i, PI, i, i, i, i, vp]
        _ops_bound = True
    return L


# ----------------------------------------------------- flux corrections

# The "conserved" group of the reference's flux-correction decks
# (input/FluxCorrect/inclined_contact.incl) and its "make_field_conservative"
# members, as slots.
CONSERVED_FIELDS = ("density", "total_energy", "velocity_x", "velocity_y", "velocity_z")
SLOT = {"density": 0, "velocity_x": 1, "velocity_y": 2, "velocity_z": 3,
        "total_energy": 4}
MAKE_CONSERVATIVE = ("velocity_x", "velocity_y", "velocity_z", "total_energy")


def _ints(v):
    return (ctypes.c_int * len(v))(*v)


TRANSFER_PROLONG, TRANSFER_RESTRICT = 0, 1


class Transfer(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("accumulate", ctypes.c_int),
                ("dst", ctypes.c_void_p), ("src", ctypes.c_void_p),
                ("m_dst", ctypes.c_int * 3), ("o_dst", ctypes.c_int * 3),
                ("n_dst", ctypes.c_int * 3), ("m_src", ctypes.c_int * 3),
                ("o_src", ctypes.c_int * 3), ("n_src", ctypes.c_int * 3)]


def transfer_table(items):
    """ctypes table of [(kind, dst, m_dst, o_dst, n_dst, src, m_src, o_src,
    n_src, accumulate)] (build once, launch many times)."""
    arr = (Transfer * len(items))()
    for k, (kind, dst, md, od, nd, src, ms, os_, ns, acc) in enumerate(items):
        t = arr[k]
        t.kind, t.accumulate, t.dst, t.src = kind, int(acc), dst, src
        t.m_dst[:], t.o_dst[:], t.n_dst[:] = list(md), list(od), list(nd)
        t.m_src[:], t.o_src[:], t.n_src[:] = list(ms), list(os_), list(ns)
    return arr


def transfer_batch(method, items, stream=None):
    """The transfers (a list, or a table from transfer_table) in one launch."""
    arr = items if isinstance(items, ctypes.Array) else transfer_table(items)
    check(ops().enzoe_vlct_transfer_batch(method.handle, arr, len(arr), stream),
          "enzoe_vlct_transfer_batch")


class FluxLink(ctypes.Structure):
    _fields_ = [("fine_flux_data", ctypes.c_void_p), ("coarse_flux_data", ctypes.c_void_p),
                ("axis", ctypes.c_int), ("face", ctypes.c_int),
                ("child", ctypes.c_int * 3)]


class FluxCorrectBlock(ctypes.Structure):
    _fields_ = [("fields", ctypes.c_void_p * 16), ("flux_data", ctypes.c_void_p),
                ("correct", ctypes.c_int * 6)]


class FluxLinks:
    """Table of [(fine FluxData, coarse FluxData, axis, face, child)]."""

    def __init__(self, links):
        self.arr = (FluxLink * len(links))()
        for k, (f, c, axis, face, child) in enumerate(links):
            self.arr[k].fine_flux_data, self.arr[k].coarse_flux_data = f.ptr, c.ptr
            self.arr[k].axis, self.arr[k].face = axis, face
            self.arr[k].child[:] = list(child)
        self.nf, self.n = len(links[0][0].names), links[0][0].n


def flux_send_to_coarse_batch(method, links, stream=None):
    """links: a list as FluxLinks takes, or a FluxLinks table."""
    t = links if isinstance(links, FluxLinks) else FluxLinks(links)
    check(ops().enzoe_vlct_flux_send_to_coarse_batch(method.handle, t.arr, len(t.arr),
                                                     t.nf, *t.n, stream),
          "enzoe_vlct_flux_send_to_coarse_batch")


class FluxCorrections:
    """Table of [(DeviceBlock, FluxData, correct[6])], blocks of one size."""

    def __init__(self, entries, conservative=MAKE_CONSERVATIVE):
        names = entries[0][1].names
        self.arr = (FluxCorrectBlock * len(entries))()
        for k, (blk, fd, corr) in enumerate(entries):
            for i, name in enumerate(names):
                self.arr[k].fields[i] = blk.ptr(name)
            self.arr[k].flux_data = fd.ptr
            self.arr[k].correct[:] = list(corr)
        self.nf = len(names)
        self.i_den = names.index("density") if "density" in names else -1
        self.cons = _ints([1 if k in conservative else 0 for k in names])
        self.n, self.ghost = entries[0][1].n, entries[0][0].ghost


def flux_correct_blocks(method, entries, stream=None, conservative=MAKE_CONSERVATIVE):
    """entries: a list as FluxCorrections takes, or a FluxCorrections table."""
    t = entries if isinstance(entries, FluxCorrections) else \
        FluxCorrections(entries, conservative)
    check(ops().enzoe_vlct_flux_correct_blocks(method.handle, t.arr, len(t.arr), t.nf,
                                               t.i_den, t.cons, *t.n, t.ghost, stream),
          "enzoe_vlct_flux_correct_blocks")


class FluxData:
    """One block's FluxData (single device array, data_FluxData.cpp:14-110)
    for the fields `names` and nx*ny*nz active cells."""

    def __init__(self, n, names=CONSERVED_FIELDS):
        self.n = tuple(n)
        self.names = tuple(names)
        self.size = ops().enzoe_flux_data_size(len(names), *self.n)
        if self.size < 0:
            raise EnzoeError("invalid FluxData size")
        self.buf = DeviceBuffer(self.size * 8)
        hip_check(hip().hipMemset(self.buf.ptr, 0, self.size * 8), "hipMemset")

    @property
    def ptr(self):
        return self.buf.ptr.value

    def download(self):
        return from_device(self.ptr, (self.size,))

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size == self.size
        to_device(self.ptr, a)

    def slots(self):
        return _ints([SLOT[k] for k in self.names])


def compute_store_fluxes(method, blocks, dt, flux_data, stream=None):
    """enzoe_vlct_compute_store_fluxes over `blocks` with one FluxData each."""
    L = ops()
    names = flux_data[0].names
    ptrs = (ctypes.c_void_p * len(blocks))(*[f.ptr for f in flux_data])
    check(L.enzoe_vlct_compute_store_fluxes(method.handle, method._blocks(blocks),
                                            len(blocks), dt, len(names),
                                            flux_data[0].slots(), ptrs, stream),
          "enzoe_vlct_compute_store_fluxes")


def flux_send_to_coarse(fine, coarse, axis, face, child, stream=None):
    """Fine block's FluxData on (axis, face) into the coarse neighbour's."""
    check(ops().enzoe_flux_data_send_to_coarse(fine.ptr, coarse.ptr, len(fine.names),
                                               *fine.n, axis, face, _ints(child), stream),
          "enzoe_flux_data_send_to_coarse")


def flux_correct(block, flux_data, correct, stream=None,
                 conservative=MAKE_CONSERVATIVE):
    """MethodFluxCorrect::flux_correct_ on a DeviceBlock; correct[2*axis+face]."""
    names = flux_data.names
    fields = (ctypes.c_void_p * len(names))(*[block.ptr(k) for k in names])
    i_den = names.index("density") if "density" in names else -1
    cons = _ints([1 if k in conservative else 0 for k in names])
    check(ops().enzoe_flux_correct(fields, len(names), i_den, cons, flux_data.ptr,
                                   *flux_data.n, block.ghost, _ints(correct), stream),
          "enzoe_flux_correct")


# ------------------------------------------------------ C++ drop-in layer

HOST_LIB_PATH = os.path.join(HERE, "libenzoe_vlct_host.so")
_host = None


def host_lib():
    """libenzoe_vlct_host.so: the C++ EnzoMethodMHDVlct and its plugins."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise EnzoeError("libenzoe_vlct_host.so is not built (run "
                             "__graft_entry__.build())")
        lib()
        H = ctypes.CDLL(HOST_LIB_PATH)
        vp = ctypes.c_void_p
        H.enzoe_host_set_fluid_props.argtypes = [ctypes.c_double] * 3
        H.enzoe_host_set_dual_energy.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int]
        H.enzoe_host_method_create.argtypes = [ctypes.c_char_p]
        H.enzoe_host_method_create.restype = vp
        H.enzoe_host_method_destroy.argtypes = [vp]
        H.enzoe_host_method_name.argtypes = [vp]
        H.enzoe_host_method_name.restype = ctypes.c_char_p
        H.enzoe_host_method_courant.argtypes = [vp]
        H.enzoe_host_method_courant.restype = ctypes.c_double
        H.enzoe_host_method_compute.argtypes = [vp, ctypes.POINTER(Block), ctypes.c_int,
                                                ctypes.c_double, vp]
        H.enzoe_host_method_compute.restype = ctypes.c_int
        H.enzoe_host_method_compute_passive.argtypes = [
            vp, ctypes.POINTER(Block), ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
            ctypes.POINTER(vp), ctypes.c_double, vp]
        H.enzoe_host_method_compute_passive.restype = ctypes.c_int
        H.enzoe_host_method_compute_accel.argtypes = [vp, ctypes.POINTER(Block),
                                                      ctypes.POINTER(vp), ctypes.c_double, vp]
        H.enzoe_host_method_compute_accel.restype = ctypes.c_int
        H.enzoe_host_method_timestep.argtypes = [vp, ctypes.POINTER(Block), vp]
        H.enzoe_host_method_timestep.restype = ctypes.c_double
        PI = ctypes.POINTER(ctypes.c_int)
        H.enzoe_host_amr_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                            ctypes.POINTER(Block), PI, ctypes.c_int]
        H.enzoe_host_amr_create.restype = vp
        H.enzoe_host_amr_set_face.argtypes = [vp] + [ctypes.c_int] * 5 + [PI]
        H.enzoe_host_amr_cycle.argtypes = [vp, ctypes.c_double, vp, _dp, ctypes.c_int]
        H.enzoe_host_amr_destroy.argtypes = [vp]
        H.enzoe_host_initial_inclined_wave.argtypes = [ctypes.c_char_p, ctypes.POINTER(Block),
                                                       _dp]
        H.enzoe_host_initial_shock_tube.argtypes = [ctypes.c_char_p, ctypes.POINTER(Block),
                                                    _dp, _dp]
        H.enzoe_host_enzo_prolong_apply.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                                    ctypes.c_int, vp, PI, PI, PI, vp, PI,
                                                    PI, PI, ctypes.c_int, vp]
        H.enzoe_host_prolong_apply.argtypes = [ctypes.c_char_p, ctypes.c_int, vp, PI, PI, PI,
                                               vp, PI, PI, PI, ctypes.c_int, vp]
        H.enzoe_host_restrict_apply.argtypes = [ctypes.c_char_p, vp, PI, PI, PI, vp, PI, PI,
                                                PI, ctypes.c_int, vp]
        _host = H
    return _host


class HostAmr:
    """The C++ Method list ["mhd_vlct", "flux_correct"] over a static set
    of hydro DeviceBlocks of given refinement levels."""

    def __init__(self, vlct_parameters, flux_correct_parameters, blocks, levels):
        self.blocks = blocks  # keeps the device memory alive
        arr = (Block * len(blocks))(*[b.desc for b in blocks])
        self.handle = host_lib().enzoe_host_amr_create(
            vlct_parameters.encode(), flux_correct_parameters.encode(), arr,
            _ints(levels), len(blocks))
        if not self.handle:
            raise EnzoeError("enzoe_host_amr_create failed")

    def set_face(self, block, axis, face, neighbor_level, coarse_neighbor=-1,
                 child=(0, 0, 0)):
        check(host_lib().enzoe_host_amr_set_face(self.handle, block, axis, face,
                                                 neighbor_level, coarse_neighbor,
                                                 _ints(child)), "enzoe_host_amr_set_face")

    def cycle(self, dt, stream=None):
        """One cycle; returns (conserved digits per field, min_digits met)."""
        d = (ctypes.c_double * 8)()
        rc = host_lib().enzoe_host_amr_cycle(self.handle, dt, stream, d, 8)
        return list(d)[:len(CONSERVED_FIELDS)], rc == 0

    def __del__(self):
        if getattr(self, "handle", None) and _host is not None:
            _host.enzoe_host_amr_destroy(self.handle)
            self.handle = None


def set_dual_energy(kind, eta=()):
    """Physics:fluid_props:dual_energy of the C++ layer's fluid properties
    (set after enzoe_host_set_fluid_props)."""
    e = (ctypes.c_double * max(1, len(eta)))(*eta)
    host_lib().enzoe_host_set_dual_energy(kind.encode(), e, len(eta))


class HostMethod:
    """EnzoMethodMHDVlct of the C++ drop-in layer, built from the text of a
    "Method:mhd_vlct" parameter group."""

    def __init__(self, parameters, gamma=5.0 / 3.0, density_floor=1e-200,
                 pressure_floor=1e-200, dual_energy=None):
        """dual_energy: None, or (type, [eta...]) of the fluid properties'
        "dual_energy" group."""
        H = host_lib()
        H.enzoe_host_set_fluid_props(gamma, density_floor, pressure_floor)
        if dual_energy is not None:
            set_dual_energy(*dual_energy)
        self.handle = H.enzoe_host_method_create(parameters.encode())

    def __del__(self):
        if getattr(self, "handle", None) and _host:
            _host.enzoe_host_method_destroy(self.handle)
            self.handle = None

    @property
    def name(self):
        return host_lib().enzoe_host_method_name(self.handle).decode()

    @property
    def courant(self):
        return host_lib().enzoe_host_method_courant(self.handle)

    def compute(self, blocks, dt, stream=None):
        arr = (Block * len(blocks))(*[b.desc for b in blocks])
        done = host_lib().enzoe_host_method_compute(self.handle, arr, len(blocks), dt,
                                                    stream)
        if done != len(blocks):
            raise EnzoeError("compute_done called %d times for %d blocks"
                             % (done, len(blocks)))

    def compute_passive(self, block, passives, dt, stream=None):
        """compute() of one block with passive scalars {name: device ptr}."""
        names = (ctypes.c_char_p * len(passives))(*[k.encode() for k in passives])
        ptrs = (ctypes.c_void_p * len(passives))(*passives.values())
        done = host_lib().enzoe_host_method_compute_passive(
            self.handle, ctypes.byref(block.desc), len(passives), names, ptrs, dt, stream)
        if done != 1:
            raise EnzoeError("compute_done called %d times" % done)

    def compute_accel(self, block, accel, dt, stream=None):
        """compute() of one block carrying acceleration_x/y/z (device ptrs)."""
        ptrs = (ctypes.c_void_p * 3)(*accel)
        done = host_lib().enzoe_host_method_compute_accel(
            self.handle, ctypes.byref(block.desc), ptrs, dt, stream)
        if done != 1:
            raise EnzoeError("compute_done called %d times" % done)

    def timestep(self, block, stream=None):
        return host_lib().enzoe_host_method_timestep(self.handle, ctypes.byref(block.desc),
                                                     stream)


# ------------------------------------------------ coarse/fine transfers

def prolong_linear(values_f, mf3, of3, nf3, values_c, mc3, oc3, nc3, accumulate=False,
                   stream=None):
    """ProlongLinear::apply on device pointers (dims x, y, z)."""
    check(ops().enzoe_prolong_linear(values_f, _ints(mf3), _ints(of3), _ints(nf3), values_c,
                                     _ints(mc3), _ints(oc3), _ints(nc3), int(accumulate),
                                     stream), "enzoe_prolong_linear")


ENZO_PROLONG_METHODS = {"3A": 0, "2A": 1, "2B": 2}


def enzo_prolong(method, values_f, mf3, of3, nf3, values_c, mc3, oc3, nc3, positive=False,
                 stream=None):
    """EnzoProlong's interpolate() ("3A", "2A", "2B") on device pointers
    (dims x, y, z; nc3 = nf3/2 + 2).  Returns interp3d's error flag (2A;
    0 otherwise)."""
    ierr = ctypes.c_int(0)
    check(ops().enzoe_enzo_prolong(ENZO_PROLONG_METHODS[method], int(bool(positive)), values_f,
                                   _ints(mf3), _ints(of3), _ints(nf3), values_c, _ints(mc3),
                                   _ints(oc3), _ints(nc3), ctypes.byref(ierr), stream),
          "enzoe_enzo_prolong")
    return ierr.value


def restrict_linear(values_c, nd3_c, im3_c, n3_c, values_f, nd3_f, im3_f, n3_f,
                    accumulate=False, stream=None):
    """RestrictLinear::apply on device pointers (dims x, y, z)."""
    check(ops().enzoe_restrict_linear(values_c, _ints(nd3_c), _ints(im3_c), _ints(n3_c),
                                      values_f, _ints(nd3_f), _ints(im3_f), _ints(n3_f),
                                      int(accumulate), stream), "enzoe_restrict_linear")


def initial_inclined_wave(parameters, block, lower):
    """EnzoInitialInclinedWave of the C++ layer on a DeviceBlock whose
    active region starts at `lower`."""
    lo = (ctypes.c_double * 3)(*lower)
    check(host_lib().enzoe_host_initial_inclined_wave(parameters.encode(),
                                                      ctypes.byref(block.desc), lo),
          "enzoe_host_initial_inclined_wave")


def initial_shock_tube(parameters, block, lower, upper):
    """EnzoInitialShockTube of the C++ layer on a DeviceBlock whose active
    region is [lower, upper)."""
    lo = (ctypes.c_double * 3)(*lower)
    hi = (ctypes.c_double * 3)(*upper)
    check(host_lib().enzoe_host_initial_shock_tube(parameters.encode(),
                                                   ctypes.byref(block.desc), lo, hi),
          "enzoe_host_initial_shock_tube")


def compute_gravity(method, blocks, dt, accel, stream=None):
    """enzoe_vlct_compute_gravity: accel[i] = (ax, ay, az) device pointers."""
    ptrs = (ctypes.c_void_p * (3 * len(blocks)))(*[p for a in accel for p in a])
    check(ops().enzoe_vlct_compute_gravity(method.handle, method._blocks(blocks),
                                           len(blocks), dt, ptrs, stream),
          "enzoe_vlct_compute_gravity")
