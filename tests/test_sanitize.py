"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only): the oracle driven through every entry point on small grids
(oracle/sanitize_main.c), and the C++ drop-in layer's construction,
factories, parameter parsing and error paths, each built with
-fsanitize=address,undefined (oracle/Makefile.sanitize,
enzo-e_amd/host/Makefile.sanitize).  A sanitizer report fails the run."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(makefile):
    # one build at a time (tests of this module may run in parallel workers)
    import fcntl
    import tempfile
    with open(os.path.join(tempfile.gettempdir(), "enzoe_sanitize_build.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        r = subprocess.run(["make", "-s", "-f", os.path.join(ROOT, makefile)],
                           capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


def _asan_runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                       text=True)
    path = r.stdout.strip()
    assert os.path.isabs(path) and os.path.exists(path), path
    return path


def test_oracle_under_sanitizers():
    _make("oracle/Makefile.sanitize")
    exe = os.path.join(ROOT, "oracle", "_san", "oracle_sanitize")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert "oracle sanitizer run clean" in r.stdout


HOST_SCRIPT = r"""
import sys
sys.path.insert(0, %(pkg)r)
import enzoe_vlct as ev
ev.HOST_LIB_PATH = %(lib)r
H = ev.host_lib()
for params in [
    "mhd_choice = constrained_transport",
    "mhd_choice = constrained_transport; riemann_solver = hlle; reconstruct_method = plm_athena; theta_limiter = 1.2; courant = 0.4",
    "mhd_choice = no_bfield; riemann_solver = hllc; reconstruct_method = nn; device_fused = false",
    "mhd_choice = constrained_transport; precision = mixed; debug_checks = true",
    'mhd_choice = "no_bfield"; riemann_solver = "HLLC"; time_scheme = "euler"',
]:
    m = ev.HostMethod(params)
    print("constructed", m.name, m.courant)
    del m
ev.set_dual_energy("modern", (0.01,))
m = ev.HostMethod("mhd_choice = constrained_transport")
print("constructed", m.name, m.courant)
del m
print("host layer clean")
"""


def _run_host_san(script):
    _make("enzo-e_amd/host/Makefile.sanitize")
    lib = os.path.join(ROOT, "enzo-e_amd", "_san", "libenzoe_vlct_host_san.so")
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(),
               # the interpreter itself is not instrumented and keeps memory
               # until exit: leaks are not checked here
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    code = script % {"pkg": os.path.join(ROOT, "enzo-e_amd"), "lib": lib}
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          timeout=300, env=env)


def test_host_layer_under_sanitizers():
    r = _run_host_san(HOST_SCRIPT)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert "host layer clean" in r.stdout
    assert r.stdout.count("constructed mhd_vlct") == 6


def test_host_layer_error_path_under_sanitizers():
    """An ERROR path (exit(1) with the reference's message) is clean too."""
    script = HOST_SCRIPT.replace(
        '"mhd_choice = constrained_transport",\n',
        '"mhd_choice = constrained_transport; theta_limiter = 2.5",\n', 1)
    r = _run_host_san(script)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "theta_limiter must satisfy" in r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
