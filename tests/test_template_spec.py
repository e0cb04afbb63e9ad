"""Template word specs (tspec, hvm-64_amd/csrc/hvm64_rules.h) built for the
host and checked case by case against the spec format the host compiler
writes (hvm64_engine.h): literals, slot ports with their labels, var-cell
halves and registers read as var endpoints, for both register areas and both
slot strides (tests/tspec_check.cpp).  CPU only; the same function runs in
the kernels, where the GPU snapshot tests check the nets templates build."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "hvm-64_amd", "csrc")


def test_template_specs_resolve_as_the_format_says(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "tspec_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "tspec_check.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["ok", "100"]
