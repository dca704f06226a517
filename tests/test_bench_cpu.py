"""bench.py's host-side parts on the CPU: the one-GPU guard and the oracle
baselines it reports beside the GPU figures (small samples here)."""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test",
                                                  os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


@pytest.mark.parametrize("workload", ["turbulence", "kelvin_helmholtz"])
def test_oracle_baseline_on_a_sample_of_the_workload_blocks(workload, monkeypatch):
    import problems
    monkeypatch.setenv("OMP_NUM_THREADS", "2")
    b = _bench()
    n = 8
    if workload == "turbulence":
        spec = problems.turbulence_spectrum()
        norm = problems.turbulence_vrms_norm(spec, mach=5.0)
        out = b.cpu_baseline_sample(
            lambda ib: problems.turbulence_block(spec, norm, ib, nb=(8, 8, 8), n=n),
            n, True, 5.0 / 3.0, "hlld", 2)
    else:
        out = b.cpu_baseline_sample(
            lambda ib: problems.kelvin_helmholtz_block(ib, nb=(8, 8, 8), n=n),
            n, False, 1.4, "hllc", 2)
    assert out["value"] > 0 and out["cores"] == 2 and out["kind"] == "port"
    assert "2x2x2" in out["sample"] and "%d^3" % n in out["sample"]
