"""Build-level guards of the operator kernels (CPU: hipcc cross-compiles gfx950 assembly).

tile_op_kernel and elas_tile_kernel address their LDS at fixed byte offsets from address 0
(sums, max |x| words, staged x, stencil copy), which holds only while the kernel declares no
static LDS: the dynamic LDS of a launch then starts at 0.  Every instantiation must also run
without scratch (no register spills)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


def _compile(tmp_path_factory, src):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("asm") / (src + ".s"))
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
           "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "dendro_amd", "csrc"), "--cuda-device-only", "-S",
           os.path.join(ROOT, "dendro_amd", "csrc", "device", src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-2000:]
    return open(out).read()


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    return _compile(tmp_path_factory, "device_api.hip")


@pytest.fixture(scope="module")
def asm_elas(tmp_path_factory):
    return _compile(tmp_path_factory, "elas_api.hip")


def _kernels(asm, pattern):
    ks = re.findall(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", asm, re.S)
    return [(n, b) for n, b in ks if pattern in n]


def test_operator_kernels_have_no_static_lds(asm):
    ks = _kernels(asm, "tile_op_kernel")
    assert len(ks) > 100
    bad = [n for n, b in ks if int(re.search(r"\.amdhsa_group_segment_fixed_size (\d+)", b).group(1)) != 0]
    assert not bad, bad[:3]


def test_operator_kernels_do_not_spill(asm):
    ks = _kernels(asm, "tile_op_kernel")
    bad = [n for n, b in ks if int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", b).group(1)) != 0]
    assert not bad, bad[:3]


def test_elasticity_kernels_have_no_static_lds_and_do_not_spill(asm_elas):
    ks = _kernels(asm_elas, "elas_tile_kernel")
    assert len(ks) == 1
    for n, b in ks:
        assert int(re.search(r"\.amdhsa_group_segment_fixed_size (\d+)", b).group(1)) == 0, n
        assert int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", b).group(1)) == 0, n


def test_1024_thread_scalar_row_kernels_fit_two_workgroups_per_cu(asm):
    """The large-mesh tiling runs 1024-thread workgroups, two per CU (32 wavefronts, 8 per
    SIMD), which holds only while a kernel needs at most 64 VGPRs; above that a CU runs one
    workgroup and the operator loses about half its latency hiding (a held-value variant of
    the FP64 sweeps measured 17% slower PCG).  Guarded for the product, CG, residual,
    smoothing and diagonal modes with stencil rows through the scalar cache (per type or per
    class): tile_op_kernel<T, TC, MASS, MODE, 1024, UNI, LANE = false, COMB>."""
    ks = [(n, b) for n, b in _kernels(asm, "tile_op_kernel")
          if re.search(r"ELi[0-4]ELi1024ELb[01]ELb0ELb[01]E", n)]
    assert len(ks) >= 30
    over = [(n, int(re.search(r"\.amdhsa_next_free_vgpr (\d+)", b).group(1))) for n, b in ks]
    over = [(n, v) for n, v in over if v > 64]
    assert not over, over[:3]
