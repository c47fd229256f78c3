"""dendro_amd: MI355X-native matrix-free octree FEM solve path of Dendro.

The compute path is libdendrohip.so (HIP kernels for gfx950 behind the C ABI of
include/dendro_hip.h); this package holds its build recipe and Python bindings.
"""
from .core import (  # noqa: F401
    COARSE_JAC2, COARSE_JAC3, DAMG, ROWS_CLASS, ROWS_LANE, ROWS_TYPE, F32, F64, FN_COS_SUM, FN_GAUSSIANS, FN_POLY, LEARN_NPARAM, Mesh, NORM_PRECONDITIONED, NORM_UNPRECONDITIONED, Octree,
    PTS_CLUSTERS, PTS_GAUSS, PTS_LOGNORMAL, PTS_MIXED, PTS_UNIFORM, SMOOTHER_JACOBI,
    SMOOTHER_LEARNED, Solver, DendroError, Transfer, default_params, etype, function, init, lib, points,
    stencils,
)
