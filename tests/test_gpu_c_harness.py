"""A C program calls the ABI with per-box, column-major, ghosted fabs as an FBoxLib host lays
them out (tests/abi/fab_harness.c): 8 boxes agree bit for bit with one box for apply_matrix and
gmres, no ghost cell of an output is written, and faces two boxes hold are written to both."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HARNESS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi", "fab_harness")


def test_c_caller_multibox():
    if not os.path.exists(HARNESS):
        raise RuntimeError(f"{HARNESS} is missing: run __graft_entry__.build()")
    r = subprocess.run([HARNESS], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
