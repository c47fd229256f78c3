"""The allocation-free fused MatVec + dot (dgh_matvec_dot_noalloc) against the allocating
dgh_matvec_dot, through the C ABI: two calls in a row, the second on a smaller mesh (fewer
tiles than the first call left in the library's fixed partials array), both bit-identical to
the allocating entry point; a mesh with more tiles than the array holds is rejected."""
import ctypes as C

import numpy as np
import pytest

import dendro_amd as da
from dendro_amd import core

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    da.init(0)
    return torch.device("cuda:0")


def _mesh(level, seed):
    rng = np.random.default_rng(seed)
    m = da.Mesh(level)
    m.set_coefficients(10.0 ** rng.uniform(-1, 1, level.n_elem), np.ones(level.n_elem))
    return m


def test_noalloc_matches_allocating_entry_point(dev, hier_small, hier_tiny):
    big, small = _mesh(hier_small.fine, 1), _mesh(hier_tiny.fine, 2)
    assert small.tile_info()["n_tiles"] < big.tile_info()["n_tiles"]
    for m, L, seed in ((big, hier_small.fine, 3), (small, hier_tiny.fine, 4)):
        x = torch.tensor(np.random.default_rng(seed).uniform(-1, 1, L.n_node), dtype=torch.float64, device=dev)
        y_ref, y = torch.empty_like(x), torch.full_like(x, float("nan"))
        d_ref = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
        d = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
        m.matvec_dot(x, y_ref, d_ref)
        m.matvec_dot_noalloc(x, y, d)
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref)
        assert torch.equal(d, d_ref)
        assert np.isfinite(d.item())


def test_noalloc_rejects_too_many_tiles(dev):
    pts = da.points(da.PTS_LOGNORMAL, 1000000, 2)
    damg = da.DAMG(da.Octree.from_points(pts, 12), 1, 500)
    L = damg.level(damg.nlevels - 1)
    m = _mesh(L, 5)
    m.set_tile(64, 256)
    assert m.tile_info()["n_tiles"] > core.NOALLOC_MAX_TILES
    x = torch.zeros(L.n_node, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    d = torch.empty(1, dtype=torch.float64, device=dev)
    rc = core.lib().dgh_matvec_dot_noalloc(m.h, core.ptr(x), core.ptr(y), core.ptr(d), C.c_void_p(None))
    assert rc == -1 and "at most" in core.last_error()
