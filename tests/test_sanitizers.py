"""The oracle (oracle/tdp_oracle.c, tdp_ewald.c), the host tables (tdp_tables.cpp) and the C++ host
mirror (tholedipoleplugin_amd/host/TholeDipoleForce.cpp) under AddressSanitizer and
UndefinedBehaviorSanitizer, on the CPU: tests/cpp/sanitize_driver (built by build(), see
tholedipoleplugin_amd/host/Makefile) runs them on the reference's ammonia pair and small water
boxes (all-pairs passes, Ewald with smooth PME and exact sums, FFT, B-splines, cutoff pairs,
exclusion tiles, validation messages, the kernel mirror's initialize).  Any sanitizer report
aborts the driver.  No GPU is used: the device list is hidden from the process.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "sanitize_driver")


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["HIP_VISIBLE_DEVICES"] = ""
    env["OMP_NUM_THREADS"] = "2"
    return env


def test_driver_is_instrumented():
    assert os.path.exists(EXE), "build() has not built tests/cpp/sanitize_driver"
    env = _env()
    env["ASAN_OPTIONS"] = "help=1"
    r = subprocess.run([EXE], capture_output=True, text=True, env=env, timeout=120)
    assert "AddressSanitizer" in (r.stdout + r.stderr)


def test_oracle_tables_and_host_mirror_clean_under_sanitizers():
    r = subprocess.run([EXE], capture_output=True, text=True, env=_env(), timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "sanitize driver: all passed" in out
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out
