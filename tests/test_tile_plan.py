"""The owner-computes tile plan (host/tiles.cpp) against the contract the operator kernels'
fixed LDS layout relies on (owned-node capacity, node budget, ownership, halo completeness,
vertex ids, ids and class within the 16-byte device record), checked on the CPU by
tests/tile_plan_check.cpp for the budgets and capacities of the 256-, 512- and 1024-thread
kernels (8192-element tiles included) and of the batched kernels."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = ["stencils.cpp", "octree.cpp", "da.cpp", "tiles.cpp", "host_api.cpp"]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("tiles") / "tile_plan_check")
    objs = [os.path.join(ROOT, "dendro_amd", "_build", s + ".o") for s in HOST]
    if not all(os.path.exists(o) for o in objs):   # no library build yet: compile the sources
        objs = [os.path.join(ROOT, "dendro_amd", "csrc", "host", s) for s in HOST]
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "dendro_amd", "csrc"), os.path.join(ROOT, "tests", "tile_plan_check.cpp"),
           *objs, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("npts,depth", [(20000, 9), (150000, 11)])
def test_tile_plan_contract(checker, npts, depth):
    r = subprocess.run([checker, str(npts), str(depth)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK (0 failures)")
