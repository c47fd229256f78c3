"""The C ABI library: it loads, exports every symbol include/dendro_hip.h declares, and its
host-side error paths return status codes (no GPU calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import dendro_amd as da
from dendro_amd import core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "dendro_hip.h")).read()
    return sorted(set(re.findall(r"^int (dgh_\w+)\(", txt, re.M)))


def test_exports_every_declared_symbol():
    lib = da.lib()
    names = declared()
    assert len(names) >= 30
    assert sorted(core.EXPORTS) == names
    for n in names:
        assert hasattr(lib, n), n


def test_error_codes_and_messages():
    lib = da.lib()
    assert lib.dgh_etype(0, 1) == -5            # vertex 0 of child 0 cannot hang
    assert "invalid" in core.last_error()
    assert lib.dgh_etype(9, 0) == -5
    # an incomplete octree is rejected
    a = np.zeros((1, 3), np.uint32)
    l = np.array([1], np.uint8)
    h = C.c_void_p()
    assert lib.dgh_octree_from_leaves(core.ptr(a), core.ptr(l), C.c_uint64(1), 3, C.byref(h)) == -1
    # device entry points before dgh_init fail cleanly without touching a GPU
    if not core._initialised:
        m = C.c_void_p()
        nl = np.zeros((1, 8), np.uint32)
        z = np.zeros(1, np.uint8)
        assert lib.dgh_mesh_create(1, 1, 4, core.ptr(nl), core.ptr(z), core.ptr(z), core.ptr(z), C.byref(m)) == -3
    with pytest.raises(da.DendroError):
        da.Octree.from_points(np.zeros((4, 3)), 30)


def test_default_params_are_reference_options():
    p = da.default_params()
    assert p.rtol == 1e-8 and p.omega == 0.667 and p.max_it == 10
    assert p.nu_pre == 4 and p.nu_post == 4 and p.norm_type == da.NORM_PRECONDITIONED


def test_point_generators_deterministic_and_in_range():
    for kind in range(5):
        a = da.points(kind, 1000, 7, 3)
        b = da.points(kind, 1000, 7, 3)
        assert np.array_equal(a, b)
        assert a.shape == (1000, 3) and np.isfinite(a).all()
    g = da.points(da.PTS_GAUSS, 200000, 1)
    assert abs(g.mean() - 0.5) < 2e-3 and abs(g.std() - 0.1) < 2e-3
    ln = da.points(da.PTS_LOGNORMAL, 10000, 2)
    assert ln.min() == 0.0 and ln.max() == 1.0


def test_noalloc_limit_matches_header():
    txt = open(os.path.join(ROOT, "include", "dendro_hip.h")).read()
    m = re.search(r"#define DGH_NOALLOC_MAX_TILES (\d+)", txt)
    assert m and int(m.group(1)) == core.NOALLOC_MAX_TILES


def test_solver_default_params():
    """dgh_solver_default_params (no GPU): the reference's bin/options values, the dense coarse
    limit and the iterative coarse solve's defaults, through the Python mirror of the struct
    (a layout mismatch between the header and core.SolverParams shows up here)."""
    p = da.default_params()
    assert p.rtol == 1e-8 and p.omega == 0.667 and p.max_it == 10
    assert p.nu_pre == 4 and p.nu_post == 4
    assert p.norm_type == da.NORM_PRECONDITIONED and p.null_space == 0
    assert p.vcycle_prec == da.F64 and p.use_graph == 0 and p.coarse_op == da.COARSE_JAC2
    assert p.coarse_dense_max == 40000 and p.coarse_rtol == 1e-12 and p.coarse_max_it == 5000
    txt = open(os.path.join(ROOT, "include", "dendro_hip.h")).read()
    body = txt[txt.index("typedef struct {\n  double rtol;"):txt.index("} dgh_solver_params;")]
    fields = re.findall(r"^\s+(?:double|int) (\w+);", body, re.M)
    assert fields == [f[0] for f in core.SolverParams._fields_]


STRUCTS = {"dgh_level_info": "LevelInfo", "dgh_function": "Function", "dgh_solver_params": "SolverParams",
           "dgh_solve_stats": "SolveStats", "dgh_learn_params": "LearnParams", "dgh_learn_stats": "LearnStats"}


def test_struct_layouts_match_header(tmp_path):
    """Every struct of include/dendro_hip.h that core.py mirrors has the same size and field
    offsets in C (gcc on the header) as in ctypes."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dendro_hip.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        cls = getattr(core, pyname)
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for l in out:
        if l:
            a, b, c = l.split()
            got[(a, b)] = int(c)
    for cname, pyname in STRUCTS.items():
        cls = getattr(core, pyname)
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for f in cls._fields_:
            assert got[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])
