"""Allocation-free entry point (tdp_execute_device_noalloc) against tdp_execute_device (GPU).

The check runs in a fresh process (tests/noalloc_check.py) in which torch initialises the HIP
runtime and owns the device buffers: two noalloc calls in a row, the second with a smaller system,
each against the existing entry point at the parity bar, the same-order results bit for bit, and
the refusal of a context too large for the scratch.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_matches_device_entry_point():
    print("case noalloc entry point", flush=True)
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "noalloc_check.py")
    r = subprocess.run([sys.executable, script], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and "OK" in r.stdout
