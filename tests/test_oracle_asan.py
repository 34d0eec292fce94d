"""The C oracle under AddressSanitizer + UndefinedBehaviorSanitizer (CPU):
oracle/asan.sh builds sailor_oracle.c with a driver that calls every oracle
function on a small synthetic scene and runs it; any out-of-bounds access,
use after free or undefined behaviour aborts with a non-zero status."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_runs_clean_under_sanitizers():
    r = subprocess.run(["bash", os.path.join(ROOT, "oracle", "asan.sh")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "asan driver:" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
