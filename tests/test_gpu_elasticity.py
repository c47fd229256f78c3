"""Linear elasticity operator (3 dofs per node, ElasticityMatMult of
examples/src/backend/elasticityJac.C:129-211) on the GPU against the oracle: natural and
Dirichlet boundaries, FP64 <= 1e-12; translations are in the kernel of the natural operator;
repeated calls are bit-identical (no atomics)."""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    da.init(0)
    return torch.device("cuda:0")


@pytest.mark.parametrize("dirichlet", [False, True])
def test_elasticity_matches_oracle(dev, hier_small, dirichlet):
    L = hier_small.fine
    mu, lam = 0.8, 2.5
    m = da.Mesh(L)
    m.set_coefficients(None, None, 1.0, 0.0)
    bdy = hier_small.damg.boundary(hier_small.n - 1) if dirichlet else None
    if dirichlet:
        m.set_boundary(bdy)
    m.set_elasticity(mu, lam)
    x = np.random.default_rng(70).uniform(-1, 1, (L.n_node, 3))
    ref = oracle.matvec_elasticity(L, x, mu, lam, bdy=bdy)
    xt = torch.tensor(x, dtype=torch.float64, device=dev)
    y1, y2 = torch.empty_like(xt), torch.empty_like(xt)
    m.matvec_elasticity(xt, y1)
    m.matvec_elasticity(xt, y2)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    assert rel_err(y1.cpu().numpy().ravel(), ref) <= 1e-12
    if dirichlet:
        b = bdy != 0
        assert np.array_equal(y1.cpu().numpy()[b], x[b])
    else:
        t = torch.zeros_like(xt)
        t[:, 1] = 1.0
        m.matvec_elasticity(t, y1)
        torch.cuda.synchronize()
        assert float(y1.abs().max()) <= 1e-13 * float(torch.tensor(ref).abs().max())


def test_elasticity_throughput_lognormal(dev):
    """The 1M log-normal octree (2.0M elements): parity against the oracle, with and without
    Dirichlet flags, and the time per product (printed)."""
    pts = da.points(da.PTS_LOGNORMAL, 1000000, 2)
    damg = da.DAMG(da.Octree.from_points(pts, 12), 1, 500)
    L = damg.level(damg.nlevels - 1)
    m = da.Mesh(L)
    m.set_elasticity(1.0, 1.0)
    bdy = damg.boundary(damg.nlevels - 1)
    xb = np.random.default_rng(72).uniform(-1, 1, (L.n_node, 3))
    mb = da.Mesh(L)
    mb.set_boundary(bdy)
    mb.set_elasticity(0.7, 1.9)
    xbt = torch.tensor(xb, dtype=torch.float64, device=dev)
    ybt = torch.empty_like(xbt)
    mb.matvec_elasticity(xbt, ybt)
    torch.cuda.synchronize()
    assert rel_err(ybt.cpu().numpy().ravel(), oracle.matvec_elasticity(L, xb, 0.7, 1.9, bdy=bdy)) <= 1e-12
    del mb, xbt, ybt
    x = np.random.default_rng(71).uniform(-1, 1, (L.n_node, 3))
    xt = torch.tensor(x, dtype=torch.float64, device=dev)
    yt = torch.empty_like(xt)
    for _ in range(3):
        m.matvec_elasticity(xt, yt)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        m.matvec_elasticity(xt, yt)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / 20 * 1e3
    alg = L.n_elem * 34 + L.n_node * 3 * 8 * 2
    print(f"elasticity E={L.n_elem} N={L.n_node}: {us:.1f} us per product, {L.n_node * 3 / us / 1e3:.2f} GDoF/s, "
          f"{alg / us / 1e3:.0f} GB/s algorithmic ({100 * alg / us / 1e3 / 8000:.1f}% of 8 TB/s)")
    assert rel_err(yt.cpu().numpy().ravel(), oracle.matvec_elasticity(L, x, 1.0, 1.0)) <= 1e-12
