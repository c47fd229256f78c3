"""The C and C++ snippets of INTEGRATION.md that use only this library's headers compile
against include/dendro_hip.h and include/dendro_hip.hpp (g++ -fsyntax-only), so the guide
keeps the entry points' current signatures.  The snippets written against the reference's
own types (ot::TreeNode, ot::DA, PETSc's Mat and Vec) are not compiled."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PREAMBLE = r"""
#include <cstdint>
#include <cstddef>
#include <vector>
#include "dendro_hip.hpp"
#define iC(expr) do { if ((expr) != 0) return 1; } while (0)
"""


def _blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, flags=re.S)


def _compile(src, tmp_path):
    f = tmp_path / "snippet.cpp"
    f.write_text(src)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wno-unused-variable",
                        "-Wno-unused-but-set-variable", "-I" + os.path.join(ROOT, "include"), str(f)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + "\n" + src


def _own(blocks, marker):
    hits = [b for b in blocks if marker in b]
    assert len(hits) == 1, marker
    return hits[0]


def test_snippets_found():
    assert len(_blocks()) >= 6


def test_facade_snippet_compiles(tmp_path):
    b = _own(_blocks(), "DAMGCreateAndSetDA(")
    b = re.sub(r"=\s*/\*.*?\*/\s*;", "{};", b)         # the octree the caller builds
    b = b.replace('#include "dendro_hip.hpp"', "")
    head, _, body = b.partition("using namespace dgh::omg;")
    src = PREAMBLE + "using namespace dgh::omg;\nint ComputeRHS(DAMG, double*);\n" + head + \
        "int driver(unsigned maxDepth) {\n" + body + "  return 0;\n}\n"
    _compile(src, tmp_path)


def test_solve_and_batched_snippets_compile(tmp_path):
    blocks = _blocks()
    solve = _own(blocks, "dgh_solver_create(")
    multi = _own(blocks, "dgh_solver_solve_multi(")
    src = PREAMBLE + ("int driver(const double* b_dev, double* x_dev, const double* B_dev, double* X_dev) {\n"
                      + solve + "{\n" + multi + "}\n  return 0;\n}\n")   # multi reuses s, meshes
    _compile(src, tmp_path)
