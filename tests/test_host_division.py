"""The traversal divides by a ray's direction components through a binary64
reciprocal taken once per ray: RN32(RN64(n * RN64(1/d))) (vgi_trace.h
div_by), and the light tree's walk divides a node's offset and radius by
its distance the same way. Built alone with g++ (IEEE SSE arithmetic, no
contraction) and checked bit for bit against binary32 division, with the
reciprocal exact and up to 3 binary64 ulps off it (the device refines a
reciprocal estimate by Newton steps), on random operands over the traversal's and the walk's ranges, quotients next to
binary32 midpoints, neighbouring operands, powers of two, signed zeros and
the still axis's divisor 1."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_reciprocal_division_equals_ieee_division(tmp_path):
    exe = str(tmp_path / "div_check")
    src = os.path.join(ROOT, "tests", "host", "div_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-msse2", "-mfpmath=sse", "-o", exe, src],
                   check=True)
    run = subprocess.run([exe, "20000000"], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "div ok" in run.stdout
