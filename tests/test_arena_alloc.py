"""The per-net arena allocator of planned batches (arena_take,
hvm-64_amd/csrc/hvm64_rules.h) built for the host and driven by threads
(tests/arena_stress.cpp): chunks never overlap, each lies in its own net's
granules (or in the overflow region), and every take is served.  CPU only;
the same function runs in the kernels."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "hvm-64_amd", "csrc")


@pytest.fixture(scope="module")
def stress(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("arena") / "arena_stress")
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread", "-I", CSRC, os.path.join(HERE, "arena_stress.cpp"),
                    "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("threads,iters", [(1, 200000), (8, 100000), (16, 50000)])
def test_arena_takes_never_overlap(stress, threads, iters):
    r = subprocess.run([stress, str(threads), str(iters)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    fields = dict(zip(r.stdout.split()[0::2], map(int, r.stdout.split()[1::2])))
    assert fields["takes"] == threads * iters
    assert fields["pool_granules"] > 0  # extents were installed
