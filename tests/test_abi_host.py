"""CPU checks of the shared libraries: every function declared in
include/*.h is exported, configurations are validated as the reference
validates them (error codes at the C ABI; message + exit(1) in the C++
drop-in layer, as the reference's ERROR/ASSERT), and the single-operation
entry points reject malformed views before launching anything.  Nothing here
touches a GPU."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "enzo-e_amd"))
import enzoe_vlct as ev  # noqa: E402

ERR_CONFIG, ERR_ARG = 1, 2


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(enzoe_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,path", [
    ("enzoe_vlct.h", ev.LIB_PATH), ("enzoe_vlct_host.h", ev.HOST_LIB_PATH),
    ("enzoe_output.h", os.path.join(ROOT, "enzo-e_amd", "libenzoe_io.so"))])
def test_library_exports_every_declared_symbol(header, path):
    names = declared(header)
    assert len(names) > 5
    L = ctypes.CDLL(path)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def _create(**kw):
    p = ev.Params()
    p.mhd_choice = kw.get("mhd_choice", b"constrained_transport")
    p.riemann_solver = kw.get("riemann", b"hlld")
    p.reconstruct_method = kw.get("recon", b"plm")
    p.time_scheme = kw.get("time_scheme", b"vl")
    p.theta_limiter = kw.get("theta", 2.0)
    p.courant = kw.get("courant", -1.0)
    p.gamma = kw.get("gamma", 5.0 / 3.0)
    p.density_floor = kw.get("dfloor", 1e-200)
    p.pressure_floor = kw.get("pfloor", 1e-200)
    p.precision = kw.get("precision", 0)
    p.dual_energy = kw.get("dual_energy", None)
    p.dual_energy_eta = kw.get("eta", -1.0)
    h = ctypes.c_void_p()
    rc = ev.lib().enzoe_vlct_create(ctypes.byref(p), ctypes.byref(h))
    return rc, h


@pytest.mark.parametrize("kw", [
    dict(riemann=b"hll"), dict(riemann=b"roe"), dict(riemann=b"hllc"),
    dict(mhd_choice=b"no_bfield", riemann=b"hlld"),
    dict(mhd_choice=b"no_bfield", riemann=b"hlle"),
    dict(mhd_choice=b"unsafe_constant_uniform"), dict(mhd_choice=None),
    dict(recon=b"ppm"), dict(time_scheme=b"rk3"), dict(theta=2.5), dict(theta=0.5),
    dict(gamma=1.0), dict(dfloor=0.0), dict(precision=7),
    dict(time_scheme=b"euler"),  # CT needs two partial timesteps
    # the VL+CT integrator takes the "modern" dual energy only
    # (EnzoMHDIntegratorStageCommands.cpp:31-34), in FP64
    dict(dual_energy=b"bryan95"), dict(dual_energy=b"on"),
    dict(dual_energy=b"modern", precision=1),
])
def test_abi_rejects_configurations_the_reference_rejects(kw):
    rc, h = _create(**kw)
    assert rc == ERR_CONFIG and not h.value


@pytest.mark.parametrize("kw", [dict(precision=2), dict(precision=2, dual_energy=b"modern"),
                                dict(precision=2, mhd_choice=b"no_bfield", riemann=b"hllc")])
def test_abi_accepts_contracted_fp64(kw):
    rc, h = _create(**kw)
    assert rc == 0 and h.value
    ev.lib().enzoe_vlct_destroy(h)


def test_abi_accepts_and_defaults_courant():
    rc, h = _create()
    assert rc == 0
    assert ev.lib().enzoe_vlct_courant(h) == 0.3  # EnzoMethodMHDVlct.cpp:96
    ev.lib().enzoe_vlct_destroy(h)
    rc, h = _create(mhd_choice=b"no_bfield", riemann=b"HLLC", recon=b"PLM_Athena",
                    time_scheme=b"euler", courant=0.7)
    assert rc == 0 and ev.lib().enzoe_vlct_courant(h) == 0.7
    ev.lib().enzoe_vlct_destroy(h)


def test_staling_rates():
    L = ev.ops()
    t, i = ctypes.c_int(), ctypes.c_int()
    for name, rates in [(b"nn", (1, 0)), (b"plm", (2, 1)), (b"PLM_ENZO", (2, 1)),
                        (b"plm_athena", (2, 1))]:
        assert L.enzoe_reconstructor_staling(name, ctypes.byref(t), ctypes.byref(i)) == 0
        assert (t.value, i.value) == rates
    assert L.enzoe_reconstructor_staling(b"weno5", ctypes.byref(t), ctypes.byref(i)) \
        == ERR_CONFIG


def test_single_operations_validate_views_without_launching():
    """Wrong shapes / missing keys return ERR_ARG (no device call is made, so
    the fake addresses are never touched)."""
    L = ev.ops()
    fake = 1 << 40
    keys = {k: fake + 8 * i for i, k in enumerate(
        ["density", "velocity_x", "velocity_y", "velocity_z", "pressure",
         "bfield_x", "bfield_y", "bfield_z"])}
    m = (12, 11, 10)
    prim = ev.fmap(keys, m)
    face_x = ev.fmap(keys, (12, 11, 9), stride=(110, 10, 1))
    wrong = ev.fmap(keys, (12, 10, 10))
    assert L.enzoe_reconstruct_interface(b"plm", 2.0, 1e-200, 1e-200, prim, wrong,
                                         wrong, 0, 0, None) == ERR_ARG
    assert L.enzoe_reconstruct_interface(b"weno", 2.0, 1e-200, 1e-200, prim, face_x,
                                         face_x, 0, 0, None) == ERR_CONFIG
    assert L.enzoe_riemann_solve(b"hlld", 0, 5 / 3, face_x, face_x, face_x, 0, 0,
                                 None) == ERR_CONFIG
    partial = ev.fmap({"density": fake}, (12, 11, 9))
    assert L.enzoe_riemann_solve(b"hlld", 1, 5 / 3, face_x, face_x, partial, 0, 0,
                                 None) == ERR_ARG
    bfi_bad = ev.farray(fake, (12, 11, 10))
    assert L.enzoe_ct_correct_reconstructed_bfield(0, bfi_bad, face_x, face_x, 0,
                                                   None) == ERR_ARG
    assert L.enzoe_ct_identify_upwind(3, face_x, ev.farray(fake, (12, 11, 9)), 0,
                                      None) == ERR_ARG
    assert L.enzoe_accumulate_flux_component(0, 0.1, 0.1, wrong, prim, 1,
                                             None) == ERR_ARG
    assert L.enzoe_update_quantities(prim, prim, partial, 1, 5 / 3, 1e-200, 1e-200,
                                     2, None) == ERR_ARG


def _run_host(params, dual_energy=None):
    code = ("import sys; sys.path.insert(0, %r); import enzoe_vlct as ev; "
            "m = ev.HostMethod(%r, dual_energy=%r); print('constructed', m.name, m.courant)"
            % (os.path.join(ROOT, "enzo-e_amd"), params, dual_energy))
    return subprocess.run([sys.executable, "-c", code], capture_output=True,
                          text=True, timeout=120)


def test_host_method_defaults():
    r = _run_host("mhd_choice = constrained_transport")
    assert r.returncode == 0, r.stderr
    assert "constructed mhd_vlct 0.3" in r.stdout
    r = _run_host('mhd_choice = "no_bfield"; riemann_solver = "HLLC"; '
                  'time_scheme = "euler"; reconstruct_method = "plm_athena"')
    assert r.returncode == 0, r.stderr
    assert "constructed mhd_vlct 1.0" in r.stdout
    assert "USE WITH CAUTION" in r.stderr


@pytest.mark.parametrize("params,message", [
    ("riemann_solver = hlld", "mhd_choice wasn't specified"),
    ("mhd_choice = constrained_transport; riemann_solver = roe",
     "The only known solvers are HLL, HLLE, HLLC, & HLLD"),
    ("mhd_choice = constrained_transport; riemann_solver = hllc",
     "can't support mhd"),
    ("mhd_choice = no_bfield; riemann_solver = hlld", "requires magnetic fields"),
    ("mhd_choice = no_bfield; riemann_solver = hlle", "untested"),
    ("mhd_choice = constrained_transport; riemann_solver = hll", "HLL"),
    ("mhd_choice = constrained_transport; theta_limiter = 2.5",
     "theta_limiter must satisfy 1<=theta_limiter<=2"),
    ("mhd_choice = constrained_transport; reconstruct_method = ppm",
     "The only allowed solvers are NN, PLM, PLM_ENZO, & PLM_ATHENA"),
    ("mhd_choice = constrained_transport; time_scheme = rk3",
     "must be \"vl\" or \"euler\""),
    ("mhd_choice = constrained_transport; half_dt_reconstruct_method = nn",
     "have been removed"),
    ("mhd_choice = magnetic", "Unrecognized choice"),
    ("mhd_choice = constrained_transport; time_scheme = euler",
     "num_partial_timesteps!=2"),
    ("mhd_choice = constrained_transport; device_fused = false; precision = mixed",
     "mixed precision is only provided by the fused device path"),
    ("mhd_choice = constrained_transport; device_fused = false; precision = double_fma",
     "contracted FP64 is only provided by the fused device path"),
    ("mhd_choice = constrained_transport; precision = single",
     "must be \"double\", \"double_fma\" or \"mixed\""),
])
def test_host_method_errors_exit_like_the_reference(params, message):
    r = _run_host(params)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "ERROR in" in r.stderr and message in r.stderr, r.stderr
    assert "constructed" not in r.stdout


def test_fused_entry_points_reject_bad_arguments_without_a_device():
    """Argument checks of the fused entry points run before any device call,
    so they can be exercised here."""
    L = ev.lib()
    rc, h = _create()
    assert rc == 0
    empty = (ev.Block * 1)()
    assert L.enzoe_vlct_compute(h, empty, 0, 0.1, None) == ERR_ARG
    assert L.enzoe_vlct_compute(None, empty, 1, 0.1, None) == ERR_ARG
    dts = (ctypes.c_double * 1)(-1.0)
    assert L.enzoe_vlct_compute_dt(h, empty, 1, dts, 0, None) == ERR_ARG
    dtp = ctypes.c_double()
    assert L.enzoe_vlct_timestep(h, empty, 0, None, ctypes.byref(dtp)) == ERR_ARG
    assert L.enzoe_vlct_compute_timestep(h, empty, 0, 0.1, None, ctypes.byref(dtp)) == ERR_ARG
    assert L.enzoe_vlct_compute_timestep(h, empty, 1, 0.1, None, None) == ERR_ARG
    mask = ctypes.c_int()
    assert L.enzoe_vlct_autotune(h, empty, 1, 0.0, None, ctypes.byref(mask)) == ERR_ARG
    g = ctypes.c_void_p()
    bcs = (ctypes.c_int * 3)(0, 0, 0)
    assert L.enzoe_vlct_graph_create(h, empty, 0, 1, 1, bcs, 0.0, 1.0, 2,
                                     ctypes.byref(g)) == ERR_ARG
    assert L.enzoe_vlct_graph_create(h, empty, 1, 1, 1, bcs, 0.0, 1.0, 0,
                                     ctypes.byref(g)) == ERR_ARG
    assert L.enzoe_vlct_graph_launch(None, None) == ERR_ARG
    # a block too shallow for the staling of NN + PLM (ghost 2 < 3) or
    # without its face fields is rejected before anything is launched
    fake = 1 << 40
    b = ev.Block()
    b.mx = b.my = b.mz = 12
    b.ghost = 2
    b.density = fake
    for a in range(3):
        b.velocity[a] = b.bfield[a] = b.bfieldi[a] = fake
    b.total_energy = fake
    arr = (ev.Block * 1)(b)
    assert L.enzoe_vlct_compute(h, arr, 1, 0.1, None) == ERR_ARG
    assert L.enzoe_vlct_compute_timestep(h, arr, 1, 0.1, None, ctypes.byref(dtp)) == ERR_ARG
    # cell indices are 32-bit: a block of 2^31 cells or more is refused
    b.ghost = 3
    b.mx = b.my = b.mz = 1300
    arr = (ev.Block * 1)(b)
    assert L.enzoe_vlct_compute(h, arr, 1, 0.1, None) == ERR_ARG
    b.mx = b.my = b.mz = 12
    b.bfieldi[1] = None
    arr = (ev.Block * 1)(b)
    assert L.enzoe_vlct_compute(h, arr, 1, 0.1, None) == ERR_ARG
    assert L.enzoe_vlct_compute_timestep(h, arr, 1, 0.1, None, ctypes.byref(dtp)) == ERR_ARG
    L.enzoe_vlct_destroy(h)


def test_launch_settings_and_device_cycles_reject_bad_arguments_without_a_device():
    """The launch-structure, flux-shape and device-cycle entry points check
    their arguments before any device call; a method that never launched a
    persistent stage has nothing to report."""
    L = ev.lib()
    rc, h = _create()
    assert rc == 0
    for v in (-1, 0, 1, 2, 3, 4, 5):
        assert L.enzoe_vlct_set_fused_ct(h, v) == 0
    for v in (-2, 6):
        assert L.enzoe_vlct_set_fused_ct(h, v) == ERR_ARG
    assert L.enzoe_vlct_set_fused_ct(None, 0) == ERR_ARG
    for v in (-1, 0, 1, 2):
        assert L.enzoe_vlct_set_flux_shape(h, v) == 0
    for v in (-2, 3):
        assert L.enzoe_vlct_set_flux_shape(h, v) == ERR_ARG
    assert L.enzoe_vlct_set_flux_shape(None, 0) == ERR_ARG
    assert L.enzoe_vlct_check_persistent(h) == 0
    assert L.enzoe_vlct_check_persistent(None) == ERR_ARG
    # the launch choice of a shape the autotuner has not seen: the settings
    L.enzoe_vlct_set_fused_ct(h, 5)
    L.enzoe_vlct_set_flux_shape(h, 2)
    out = [ctypes.c_int(-9) for _ in range(4)]
    assert L.enzoe_vlct_launch_choice(h, 20, 20, 20, *[ctypes.byref(o) for o in out]) == 0
    assert out[1].value == 5 and out[3].value == 2
    assert L.enzoe_vlct_launch_choice(None, 20, 20, 20, None, None, None, None) == ERR_ARG
    empty = (ev.Block * 1)()
    bcs = (ctypes.c_int * 3)(0, 0, 0)
    state = ctypes.c_void_p(1 << 40)
    assert L.enzoe_vlct_run_cycles(h, empty, 0, 1, 1, bcs, state, 1.0, 2, None) == ERR_ARG
    assert L.enzoe_vlct_run_cycles(h, empty, 1, 1, 1, bcs, None, 1.0, 2, None) == ERR_ARG
    assert L.enzoe_vlct_run_cycles(h, empty, 1, 1, 1, bcs, state, 1.0, 0, None) == ERR_ARG
    assert L.enzoe_vlct_run_cycles(None, empty, 1, 1, 1, bcs, state, 1.0, 2, None) == ERR_ARG
    L.enzoe_vlct_destroy(h)
    # a dual-energy method has no persistent stage: its choice stays the
    # separate launches
    rc, h = _create(dual_energy=b"modern")
    assert rc == 0
    L.enzoe_vlct_set_fused_ct(h, 5)
    assert L.enzoe_vlct_launch_choice(h, 20, 20, 20, *[ctypes.byref(o) for o in out]) == 0
    assert out[1].value == 0
    L.enzoe_vlct_destroy(h)


def test_flux_correction_entry_points_reject_bad_arguments_without_a_device():
    """The flux-correction entry points validate before any device call:
    the MHD method refuses to store fluxes (EnzoMethodMHDVlct.cpp:137-141),
    FaceFluxes needs even block dimensions (data_FaceFluxes.cpp:38-45), a
    conservative field needs density (problem_MethodFluxCorrect.cpp:520)."""
    L = ev.ops()
    assert L.enzoe_flux_data_size(5, 16, 16, 16) == 5 * 12 * 256
    assert L.enzoe_flux_data_size(3, 4, 6, 8) == 3 * 4 * (48 + 32 + 24)
    assert L.enzoe_flux_data_size(1, 0, 4, 4) == -1
    fake = 1 << 40
    b = ev.Block()
    b.mx = b.my = b.mz = 24
    b.ghost = 4
    b.density = b.total_energy = fake
    for a in range(3):
        b.velocity[a] = fake
    arr = (ev.Block * 1)(b)
    slots = (ctypes.c_int * 5)(0, 4, 1, 2, 3)
    fd = (ctypes.c_void_p * 1)(fake)
    rc, h = _create()                      # constrained_transport
    assert L.enzoe_vlct_compute_store_fluxes(h, arr, 1, 0.1, 5, slots, fd, None) == ERR_CONFIG
    L.enzoe_vlct_destroy(h)
    rc, h = _create(mhd_choice=b"no_bfield", riemann=b"hllc")
    assert rc == 0
    bad = (ctypes.c_int * 1)(5)            # bfield_x has no hydro flux
    assert L.enzoe_vlct_compute_store_fluxes(h, arr, 1, 0.1, 1, bad, fd, None) == ERR_ARG
    b.mx = 25                              # 17 active cells along x
    arr = (ev.Block * 1)(b)
    assert L.enzoe_vlct_compute_store_fluxes(h, arr, 1, 0.1, 5, slots, fd, None) == ERR_ARG
    L.enzoe_vlct_destroy(h)
    child = (ctypes.c_int * 3)(0, 1, 0)
    assert L.enzoe_flux_data_send_to_coarse(fake, fake, 5, 16, 15, 16, 0, 0, child,
                                            None) == ERR_ARG
    assert L.enzoe_flux_data_send_to_coarse(fake, fake, 5, 16, 16, 16, 0, 2, child,
                                            None) == ERR_ARG
    child2 = (ctypes.c_int * 3)(0, 2, 0)
    assert L.enzoe_flux_data_send_to_coarse(fake, fake, 5, 16, 16, 16, 0, 0, child2,
                                            None) == ERR_ARG
    fields = (ctypes.c_void_p * 5)(*([fake] * 5))
    cons = (ctypes.c_int * 5)(0, 1, 1, 1, 1)
    corr = (ctypes.c_int * 6)(0, 1, 0, 0, 0, 0)
    assert L.enzoe_flux_correct(fields, 5, -1, cons, fake, 16, 16, 16, 4, corr,
                                None) == ERR_ARG
    assert L.enzoe_flux_correct(fields, 5, 0, cons, fake, 16, 16, 15, 4, corr,
                                None) == ERR_ARG
    rc, h = _create(mhd_choice=b"no_bfield", riemann=b"hllc")
    links = (ev.FluxLink * 1)()
    links[0].fine_flux_data = links[0].coarse_flux_data = fake
    links[0].axis, links[0].face = 0, 0
    links[0].child[:] = [0, 2, 0]                       # not a child index
    assert L.enzoe_vlct_flux_send_to_coarse_batch(h, links, 1, 5, 16, 16, 16,
                                                  None) == ERR_ARG
    assert L.enzoe_vlct_flux_send_to_coarse_batch(h, links, 70000, 5, 16, 16, 16,
                                                  None) == ERR_ARG
    blocks = (ev.FluxCorrectBlock * 1)()
    blocks[0].flux_data = fake
    for i in range(5):
        blocks[0].fields[i] = fake
    assert L.enzoe_vlct_flux_correct_blocks(h, blocks, 1, 5, -1, cons, 16, 16, 16, 4,
                                            None) == ERR_ARG
    blocks[0].fields[3] = None
    assert L.enzoe_vlct_flux_correct_blocks(h, blocks, 1, 5, 0, cons, 16, 16, 16, 4,
                                            None) == ERR_ARG
    L.enzoe_vlct_destroy(h)


def _run_host_amr(vlct_params, fc_params):
    """Constructs the C++ Method list ["mhd_vlct", "flux_correct"] (no
    device call is made before the constructors' checks)."""
    code = ("import sys; sys.path.insert(0, %r); import enzoe_vlct as ev\n"
            "ev.host_lib().enzoe_host_set_fluid_props(5.0 / 3.0, 1e-30, 1e-30)\n"
            "class B: desc = ev.Block()\n"
            "a = ev.HostAmr(%r, %r, [B()], [0]); print('constructed')"
            % (os.path.join(ROOT, "enzo-e_amd"), vlct_params, fc_params))
    return subprocess.run([sys.executable, "-c", code], capture_output=True,
                          text=True, timeout=120)


@pytest.mark.parametrize("vlct,fc,message", [
    ("mhd_choice = constrained_transport", "",
     "Flux corrections are currently only supported in hydro-mode"),
    ("mhd_choice = no_bfield; riemann_solver = hllc", "min_digits = [density, 1, pressure]",
     "must have a non-negative, even length"),
    ("mhd_choice = no_bfield; riemann_solver = hllc", "min_digits = [pressure, 12.0]",
     "since it's not in the field group"),
    ("mhd_choice = no_bfield; riemann_solver = hllc", "group = colour",
     "Must add field(s) to the field group"),
])
def test_host_flux_correct_errors_exit_like_the_reference(vlct, fc, message):
    r = _run_host_amr(vlct, fc)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "ERROR in" in r.stderr and message in r.stderr, r.stderr


def test_host_method_list_with_flux_correct_constructs():
    r = _run_host_amr("mhd_choice = no_bfield; riemann_solver = hllc",
                      "min_digits = [density, 16.5, total_energy, 16.2]")
    assert r.returncode == 0, r.stderr
    assert "constructed" in r.stdout


def test_host_amr_entry_points_reject_bad_arguments():
    H = ev.host_lib()
    child = (ctypes.c_int * 3)(0, 1, 1)
    assert H.enzoe_host_amr_set_face(None, 0, 0, 1, 1, -1, child) == ERR_ARG
    d = (ctypes.c_double * 5)()
    assert H.enzoe_host_amr_cycle(None, 0.1, None, d, 5) == ERR_ARG
    assert H.enzoe_host_amr_create(b"", b"", None, None, 0) is None


def test_transfer_entry_points_reject_bad_arguments_without_a_device():
    L = ev.ops()
    fake = 1 << 40
    I = lambda *v: (ctypes.c_int * 3)(*v)  # noqa: E731
    # nf = 2*nc or 2*(nc-2) only (problem_ProlongLinear.cpp:97-103)
    assert L.enzoe_prolong_linear(fake, I(12, 12, 12), I(0, 0, 0), I(6, 8, 8), fake,
                                  I(8, 8, 8), I(1, 1, 1), I(4, 4, 4), 0, None) == ERR_ARG
    # regions outside their arrays
    assert L.enzoe_prolong_linear(fake, I(8, 8, 8), I(1, 0, 0), I(8, 8, 8), fake,
                                  I(8, 8, 8), I(0, 0, 0), I(4, 4, 4), 0, None) == ERR_ARG
    assert L.enzoe_restrict_linear(fake, I(8, 8, 8), I(0, 0, 0), I(4, 4, 4), fake,
                                   I(8, 8, 8), I(1, 0, 0), I(8, 8, 8), 0, None) == ERR_ARG
    rc, h = _create(mhd_choice=b"no_bfield", riemann=b"hllc")
    items = (ev.Transfer * 1)()
    items[0].kind = 2
    items[0].dst = items[0].src = fake
    assert L.enzoe_vlct_transfer_batch(h, items, 1, None) == ERR_ARG
    items[0].kind = ev.TRANSFER_RESTRICT
    items[0].m_dst[:] = [8, 8, 8]
    items[0].n_dst[:] = [4, 4, 4]
    items[0].m_src[:] = [8, 8, 8]
    items[0].o_src[:] = [1, 0, 0]              # 2*4 fine cells from x = 1: outside
    assert L.enzoe_vlct_transfer_batch(h, items, 1, None) == ERR_ARG
    L.enzoe_vlct_destroy(h)


def _run_host_initial(params):
    code = ("import sys, ctypes; sys.path.insert(0, %r); import enzoe_vlct as ev\n"
            "ev.host_lib().enzoe_host_set_fluid_props(5.0 / 3.0, 1e-30, 1e-30)\n"
            "b = ev.Block(); lo = (ctypes.c_double * 3)(0, 0, 0)\n"
            "print(ev.host_lib().enzoe_host_initial_inclined_wave(%r, ctypes.byref(b), lo))"
            % (os.path.join(ROOT, "enzo-e_amd"), params.encode()))
    return subprocess.run([sys.executable, "-c", code], capture_output=True,
                          text=True, timeout=120)


@pytest.mark.parametrize("params,message", [
    ("alpha = 0.5", "wave_type must be specified"),
    ("wave_type = shear", "has invalid value, \"shear\""),
    ("wave_type = jeans", "is not provided"),
    ("wave_type = circ_alfven", "is not provided"),
    ("wave_type = sound; parallel_vel = 0.5", "parallel_vel is not provided"),
    ("wave_type = fast", "A MHD wave requires fields to store magnetic field values."),
])
def test_host_initial_inclined_wave_errors_exit_like_the_reference(params, message):
    """Parameter checks and the MHD-fields assert come before any device
    copy (the block here has no fields at all)."""
    r = _run_host_initial(params)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "ERROR in" in r.stderr and message in r.stderr, r.stderr


def _run_host_tube(params):
    code = ("import sys, ctypes; sys.path.insert(0, %r); import enzoe_vlct as ev\n"
            "ev.host_lib().enzoe_host_set_fluid_props(5.0 / 3.0, 1e-30, 1e-30)\n"
            "b = ev.Block(); lo = (ctypes.c_double * 3)(0, 0, 0)\n"
            "print(ev.host_lib().enzoe_host_initial_shock_tube(%r, ctypes.byref(b), lo, lo))"
            % (os.path.join(ROOT, "enzo-e_amd"), params.encode()))
    return subprocess.run([sys.executable, "-c", code], capture_output=True,
                          text=True, timeout=120)


@pytest.mark.parametrize("params,message", [
    ("setup_name = brio_wu", "specifies an invalid name: (must be rj2a, sod), not brio_wu."),
    ("setup_name = rj2a; aligned_ax = w", "must specify \"x\", \"y\", or \"z\", not \"w\""),
    ("setup_name = rj2a", "the \"rj2a\" tube needs the magnetic fields"),
])
def test_host_shock_tube_errors_exit_like_the_reference(params, message):
    r = _run_host_tube(params)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "ERROR in" in r.stderr and message in r.stderr, r.stderr


def test_gravity_entry_points_reject_bad_arguments_without_a_device():
    L = ev.ops()
    rc, h = _create()
    b = (ev.Block * 1)()
    acc = (ctypes.c_void_p * 3)(1 << 40, None, 1 << 40)
    assert L.enzoe_vlct_compute_gravity(h, b, 1, 0.1, acc, None) == ERR_ARG
    assert L.enzoe_vlct_compute_gravity(h, b, 1, 0.1, None, None) == ERR_ARG
    L.enzoe_vlct_destroy(h)
    m = ev.FMap()
    assert L.enzoe_source_gravity(0.1, ctypes.byref(m), ctypes.byref(m), ctypes.byref(m), 0,
                                  None) == ERR_ARG


def test_dual_energy_configuration():
    """The "modern" formulation is accepted (eta default 0.001,
    EnzoConfig.cpp:988-995); a dual-energy method rejects blocks without
    internal_energy before anything is launched."""
    L = ev.lib()
    rc, h = _create(dual_energy=b"modern")
    assert rc == 0
    rc2, h2 = _create(dual_energy=b"MODERN", eta=0.0)
    assert rc2 == 0
    fake = 1 << 40
    b = ev.Block()
    b.mx = b.my = b.mz = 12
    b.ghost = 3
    b.density = b.total_energy = fake
    for a in range(3):
        b.velocity[a] = b.bfield[a] = b.bfieldi[a] = fake
    arr = (ev.Block * 1)(b)
    assert L.enzoe_vlct_compute(h, arr, 1, 0.1, None) == ERR_ARG
    bcs = (ctypes.c_int * 3)(0, 0, 0)
    assert L.enzoe_grid_refresh(h, arr, 1, 1, 1, bcs, None) == ERR_ARG
    L.enzoe_vlct_destroy(h)
    L.enzoe_vlct_destroy(h2)


@pytest.mark.parametrize("params,dual,message", [
    ("mhd_choice = constrained_transport", ("bryan95", []),
     "selected formulation of dual energy formalism is incompatible"),
    ("mhd_choice = constrained_transport; precision = mixed", ("modern", []),
     "the dual-energy formalism needs"),
    ("mhd_choice = constrained_transport", ("disabled", [0.01]), "can't be specified"),
    ("mhd_choice = constrained_transport", ("modern", [0.1, 0.2]), "must provide 1 value"),
    ("mhd_choice = constrained_transport", ("modern", [-1.0]), "eta must be non-negative"),
    ("mhd_choice = constrained_transport", ("dual", []), "is invalid"),
])
def test_host_dual_energy_errors_exit_like_the_reference(params, dual, message):
    r = _run_host(params, dual)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "ERROR in" in r.stderr and message in r.stderr, r.stderr


def test_host_dual_energy_modern_constructs():
    r = _run_host("mhd_choice = constrained_transport", ("modern", [0.00769]))
    assert r.returncode == 0, r.stderr
    assert "constructed mhd_vlct 0.3" in r.stdout
    # the stage-by-stage plugin path carries the formalism too
    r = _run_host("mhd_choice = no_bfield; riemann_solver = hllc; device_fused = false",
                  ("modern", []))
    assert r.returncode == 0, r.stderr


def test_dual_energy_single_operations_validate_without_a_device():
    """The dual-energy entry points check their views before launching: the
    flux map needs the internal_energy slot and the interface velocity view
    (EnzoRiemannUtils.hpp:381-384), the source needs the pressure and the
    internal-energy change and a face-trimmed velocity view, the update and
    the synchronisation need the internal-energy slot and eta >= 0, and the
    plain update refuses an internal-energy change."""
    L = ev.ops()
    fake = 1 << 40
    keys = {k: fake for k in ("density", "velocity_x", "velocity_y", "velocity_z",
                              "total_energy", "bfield_x", "bfield_y", "bfield_z")}
    face = ev.fmap(keys, (8, 8, 7))
    with_e = ev.fmap(dict(keys, internal_energy=fake), (8, 8, 7))
    vel = ev.farray(fake, (8, 8, 7))
    assert L.enzoe_riemann_solve_dual_energy(b"hlld", 1, 5 / 3, face, face, face, vel, 0, 0,
                                             None) == ERR_ARG
    assert L.enzoe_riemann_solve_dual_energy(b"hlld", 1, 5 / 3, face, face, with_e, None, 0,
                                             0, None) == ERR_ARG
    assert L.enzoe_riemann_solve_dual_energy(b"hllc", 1, 5 / 3, face, face, with_e, vel, 0,
                                             0, None) == ERR_CONFIG
    cell = ev.fmap(dict(keys, internal_energy=fake), (8, 8, 8))
    prim = ev.fmap({"pressure": fake}, (8, 8, 8))
    assert L.enzoe_source_internal_energy(0, 0.1, 0.1, prim, cell, ev.farray(fake, (8, 8, 8)),
                                          1e-30, 0, None) == ERR_ARG
    assert L.enzoe_source_internal_energy(0, 0.1, 0.1, prim, ev.fmap(keys, (8, 8, 8)),
                                          vel, 1e-30, 0, None) == ERR_ARG
    plain = ev.fmap(keys, (8, 8, 8))
    assert L.enzoe_update_quantities_dual_energy(plain, plain, plain, 1, 5 / 3, 1e-30,
                                                 1e-30, 0.01, 0, None) == ERR_ARG
    assert L.enzoe_update_quantities_dual_energy(cell, cell, cell, 1, 5 / 3, 1e-30, 1e-30,
                                                 -1.0, 0, None) == ERR_ARG
    assert L.enzoe_update_quantities(cell, cell, cell, 1, 5 / 3, 1e-30, 1e-30, 0,
                                     None) == ERR_ARG
    assert L.enzoe_apply_floor_to_energy_and_sync(plain, 1, 5 / 3, 1e-30, 1, 0.01, 0,
                                                  None) == ERR_ARG
    assert L.enzoe_pressure_from_integration_dual_energy(plain, ev.farray(fake, (8, 8, 8)),
                                                         5 / 3, 0, None) == ERR_ARG
    assert L.enzoe_primitive_from_integration_dual_energy(plain, plain, 1, 5 / 3, 0,
                                                          None) == ERR_ARG


ERR_STATE = 5


def test_debug_counts_need_the_checks_on():
    L = ev.lib()
    L.enzoe_vlct_debug_counts.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_ulonglong)]
    rc, h = _create()
    assert rc == 0
    try:
        c = (ctypes.c_ulonglong * 4)()
        assert L.enzoe_vlct_debug_counts(h, None, c) == ERR_STATE
    finally:
        L.enzoe_vlct_destroy(h)
