"""The C++ facade (include/dendro_hip.hpp) with the reference's DAMG names, driven by the
compiled example dendro_amd/bin/tst_mg (the flow of the reference's tstMg / tstMatVec
drivers): the host hierarchy on the CPU, the whole solve on the GPU."""
import os
import re
import subprocess

import pytest

import dendro_amd as da

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "dendro_amd", "bin", "tst_mg")


def _run(*args, timeout=300):
    assert os.path.exists(EXE), f"{EXE} not built: run python -m dendro_amd.build"
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_facade_host_hierarchy_matches_python_bindings():
    r = _run(3000, 6, 3, 0, "--host-only")
    assert r.returncode == 0, r.stderr
    sizes = [tuple(map(int, m)) for m in re.findall(r"level \d+: (\d+) elements, (\d+) nodes", r.stdout)]
    pts = da.points(da.PTS_GAUSS, 3000, 1, 31)
    damg = da.DAMG(da.Octree.from_points(pts, 6), 3, 500)
    assert sizes == [(damg.info(l).n_elem, damg.info(l).n_node) for l in range(damg.nlevels)]


@pytest.mark.gpu
def test_facade_solve_on_gpu():
    r = _run(100000, 8, 3, 0)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"its=(\d+) converged=1 rel_true=([0-9.e+-]+)", r.stdout)
    assert m and float(m.group(2)) <= 1e-8
