"""The oracle's C restatement under AddressSanitizer and
UndefinedBehaviorSanitizer: oracle/asan_check.c runs small forward and
gradient cases of every physics in both precisions, with points on the grid
edges and in the C-PML, free surface on and off and a partial last replay
segment (oracle/Makefile target `asan`); any out-of-bounds access or
undefined behaviour fails the run."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_runs_clean_under_sanitizers():
    r = subprocess.run(["make", "-B", "-C", os.path.join(REPO, "oracle"), "asan"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "all cases ran clean" in r.stdout
