"""The edit path's host thread pool (vgi_pool.h), built alone with g++ under
ThreadSanitizer: several host threads submitting jobs at once each get every
index of their own job written, with no data race reported."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pool_serves_concurrent_callers_without_races(tmp_path):
    exe = str(tmp_path / "pool_check")
    src = os.path.join(ROOT, "tests", "host", "pool_check.cpp")
    inc = os.path.join(ROOT, "global-illumination_amd", "csrc")
    subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-fsanitize=thread", "-I", inc, "-o", exe, src], check=True)
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "pool ok" in run.stdout
    assert "WARNING: ThreadSanitizer" not in run.stderr
