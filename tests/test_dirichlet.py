"""Dirichlet-masked operators (examples/src/backend/omgJac.C:859-1160): boundary flags of
the DA nodes, the oracle's masked element loop, and the GPU entry points against it.

The reference's Dirichlet MatVec drops the columns of boundary nodes from every element
product and sets the rows of boundary nodes to the identity; its diagonal is 1 there.
DirichletLaplacianMatMult is the uniform operator with lap_fac 1 and mass_fac 0,
DirichletJacobianMatMult the one with mass_fac 1.

CPU: the flags against the node anchors' coordinates, and the oracle's masked loop against
the unmasked oracle product of the masked vector (a different loop).  GPU (marked): the
C-ABI entry points against the oracle, FP64 to 1e-12 and FP32 with its error stated.
"""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import rel_err, var_coefficients

MODES = {
    "laplacian": dict(lap_fac=1.0, mass_fac=0.0),   # DirichletLaplacianMatMult
    "jacobian": dict(lap_fac=1.0, mass_fac=1.0),    # DirichletJacobianMatMult
}


def _coef(h, mode):
    if mode in MODES:
        return None, MODES[mode]
    c = var_coefficients(h, mass_base=1.0)[-1]
    return c, dict(kappa=c[0], mass=c[1])


def test_boundary_flags_from_anchors(hier_small):
    """A node is flagged iff its anchor lies on a face of the cube [0, 2^D]^3; nodes are
    numbered in the order of the regular-node octants."""
    h = hier_small
    for l in range(h.n):
        info = h.damg.info(l)
        anc, _, fl = h.damg.octants(l)
        lim = 1 << (info.max_depth - 1)
        reg = anc[(fl & 128) != 0]
        assert len(reg) == info.n_node
        expect = (np.any(reg == 0, axis=1) | np.any(reg == lim, axis=1)).astype(np.uint8)
        b = h.damg.boundary(l)
        np.testing.assert_array_equal(b, expect)
        assert 0 < int(b.sum()) < info.n_node


@pytest.mark.parametrize("mode", ["laplacian", "jacobian", "var"])
def test_oracle_dirichlet_matches_masked_product(hier_small, mode):
    h = hier_small
    F = h.fine
    b = h.damg.boundary(h.n - 1)
    _, kw = _coef(h, mode)
    x = np.random.default_rng(7).uniform(-1, 1, F.n_node)
    y = oracle.matvec_dirichlet(F, b, x, **kw)
    y_ref = oracle.matvec(F, np.where(b != 0, 0.0, x), **kw)
    y_ref[b != 0] = x[b != 0]
    assert rel_err(y, y_ref) <= 1e-15
    d = oracle.diagonal_dirichlet(F, b, **kw)
    d_ref = oracle.diagonal(F, **kw)
    d_ref[b != 0] = 1.0
    assert rel_err(d, d_ref) <= 1e-15


def test_oracle_dirichlet_interior_block_symmetric(hier_small):
    """u.A v = v.A u for vectors that vanish on the boundary (the interior block of A)."""
    h = hier_small
    F = h.fine
    b = h.damg.boundary(h.n - 1)
    rng = np.random.default_rng(3)
    u, v = (np.where(b != 0, 0.0, rng.uniform(-1, 1, F.n_node)) for _ in range(2))
    _, kw = _coef(h, "var")
    a = u @ oracle.matvec_dirichlet(F, b, v, **kw)
    c = v @ oracle.matvec_dirichlet(F, b, u, **kw)
    assert abs(a - c) <= 1e-12 * abs(a)


def test_oracle_dirichlet_identity_rows(hier_small):
    """Boundary rows return x itself whatever the element products."""
    h = hier_small
    F = h.fine
    b = h.damg.boundary(h.n - 1)
    x = np.random.default_rng(5).uniform(-1, 1, F.n_node)
    y = oracle.matvec_dirichlet(F, b, x, lap_fac=1.0, mass_fac=1.0)
    np.testing.assert_array_equal(y[b != 0], x[b != 0])


# ----------------------------------------------------------------------------- GPU
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    da.init(0)
    return torch.device("cuda:0")


def _mesh(h, mode):
    F = h.fine
    c, kw = _coef(h, mode)
    m = da.Mesh(F)
    if c is None:
        m.set_coefficients(None, None, kw["lap_fac"], kw["mass_fac"])
    else:
        m.set_coefficients(c[0], c[1])
    return m, kw


@pytest.mark.gpu
@pytest.mark.parametrize("mode,lane,block", [("laplacian", 0, 0), ("jacobian", 0, 0), ("var", 0, 0),
                                             ("var", 1, 0), ("jacobian", 1, 256), ("var", 0, 1024)])
def test_matvec_dirichlet_gpu(dev, hier_small, mode, lane, block):
    """lane: stencil rows per lane from LDS; block: 256 / 1024 threads per tile (0: default)"""
    h = hier_small
    F = h.fine
    b = h.damg.boundary(h.n - 1)
    m, kw = _mesh(h, mode)
    if block:
        m.set_tile(2048, block)
    m.set_lane_stencils(lane)
    assert m.tile_info()["lane_stencils"] == lane
    m.set_boundary(b)
    x = np.random.default_rng(11).uniform(-1, 1, F.n_node)
    y_ref = oracle.matvec_dirichlet(F, b, x, **kw)
    xt = torch.as_tensor(x, dtype=torch.float64, device=dev)
    yt = torch.empty_like(xt)
    m.matvec_dirichlet(xt, yt)
    torch.cuda.synchronize()
    y = yt.cpu().numpy()
    assert rel_err(y, y_ref) <= 1e-12
    np.testing.assert_array_equal(y[b != 0], x[b != 0])
    x32 = xt.float()
    y32 = torch.empty_like(x32)
    m.matvec_dirichlet(x32, y32, prec=da.F32)
    torch.cuda.synchronize()
    err = rel_err(y32.cpu().numpy().astype(np.float64), y_ref)
    print(f"FP32 Dirichlet MatVec ({mode}) relative error vs FP64 oracle: {err:.3e}")
    assert err <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["laplacian", "var"])
def test_diagonal_dirichlet_gpu(dev, hier_small, mode):
    h = hier_small
    F = h.fine
    b = h.damg.boundary(h.n - 1)
    m, kw = _mesh(h, mode)
    m.set_boundary(b)
    d_ref = oracle.diagonal_dirichlet(F, b, **kw)
    dt = torch.empty(F.n_node, dtype=torch.float64, device=dev)
    m.diagonal_dirichlet(dt)
    torch.cuda.synchronize()
    d = dt.cpu().numpy()
    assert rel_err(d, d_ref) <= 1e-12
    assert np.all(d[b != 0] == 1.0)


@pytest.mark.gpu
def test_dirichlet_needs_boundary(dev, hier_small):
    """Host-side check: no flags, no product (nothing is launched)."""
    m, _ = _mesh(hier_small, "laplacian")
    xt = torch.zeros(hier_small.fine.n_node, dtype=torch.float64, device=dev)
    yt = torch.empty_like(xt)
    with pytest.raises(da.DendroError):
        m.matvec_dirichlet(xt, yt)
    m.set_boundary(hier_small.damg.boundary(hier_small.n - 1))
    m.set_boundary(None)
    with pytest.raises(da.DendroError):
        m.diagonal_dirichlet(yt)
