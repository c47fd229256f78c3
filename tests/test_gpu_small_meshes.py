"""Smallest meshes through every GPU path: a single element (8 nodes, one partial wavefront,
one tile), 8 + 15 elements with hanging nodes, and 92 + 211 and 155-element octrees
(records that fill no wavefront and no tile evenly).  Each against the oracle: the operator
(both precisions, per-type and lane stencil rows), its diagonal, the fused product + dot and
its allocation-free variant, the batched operator, the elasticity operator, and the
multigrid PCG (the first iterate against the oracle's, the solve to 1e-10, and a one-level
solve whose preconditioner is the coarse solve alone)."""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import hierarchy, rel_err, var_coefficients

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MESHES = {
    "one_element": ("uniform", 1, 1, 1),
    "hanging_15": ("uniform", 5, 3, 2),
    "gauss_211": ("gauss", 30, 4, 2),
    "uniform_155": ("uniform", 60, 3, 1),
}


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    da.init(0)
    return torch.device("cuda:0")


@pytest.fixture(scope="module", params=sorted(MESHES))
def small(request):
    kind, npts, depth, nlev = MESHES[request.param]
    h = hierarchy(kind, npts, depth, nlev, seed=4, min_coarse=1)
    return request.param, h


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def test_small_operator(dev, small):
    name, h = small
    F = h.fine
    kappa, mass = var_coefficients(h, mass_base=1.0)[-1]
    m = da.Mesh(F)
    m.set_coefficients(kappa, mass)
    x = np.random.default_rng(80).uniform(-1, 1, F.n_node)
    ref = oracle.matvec(F, x, kappa=kappa, mass=mass)
    xt = torch.tensor(x, dtype=torch.float64, device=dev)
    for rows in (da.ROWS_TYPE, da.ROWS_LANE):
        m.set_lane_stencils(rows)
        y = torch.empty_like(xt)
        m.matvec(xt, y)
        assert rel_err(_np(y), ref) <= 1e-12, (name, rows)
        y32 = torch.empty(F.n_node, dtype=torch.float32, device=dev)
        m.matvec(xt.float(), y32, prec=da.F32)
        assert rel_err(_np(y32), ref) <= 1e-5, (name, rows)
    d = torch.empty_like(xt)
    m.diagonal(d)
    assert rel_err(_np(d), oracle.diagonal(F, kappa=kappa, mass=mass)) <= 1e-12
    y = torch.empty_like(xt)
    xy = torch.zeros(1, dtype=torch.float64, device=dev)
    m.matvec_dot(xt, y, xy)
    assert rel_err(_np(y), ref) <= 1e-12
    assert abs(float(_np(xy)[0]) - float(x @ ref)) <= 1e-12 * abs(x) @ np.abs(ref)
    y2 = torch.empty_like(xt)
    xy2 = torch.zeros(1, dtype=torch.float64, device=dev)
    m.matvec_dot_noalloc(xt, y2, xy2)
    assert torch.equal(y, y2) and torch.equal(xy, xy2)


def test_small_batched(dev, small):
    name, h = small
    F = h.fine
    kappa, mass = var_coefficients(h, mass_base=1.0)[-1]
    m = da.Mesh(F)
    m.set_coefficients(kappa, mass)
    k = 8
    X = np.random.default_rng(81).uniform(-1, 1, (F.n_node, k))
    refs = [oracle.matvec(F, X[:, j], kappa=kappa, mass=mass) for j in range(k)]
    Xt = torch.tensor(X, dtype=torch.float64, device=dev)
    Y = torch.empty_like(Xt)
    m.matvec_multi(Xt, Y, k, prec=da.F64)
    Yn = _np(Y)
    for j in range(k):
        assert rel_err(Yn[:, j], refs[j]) <= 1e-12, (name, j)
    for variant in (0, 1, 4, 5):
        m.set_tile_multi(256, variant)
        Y32 = torch.empty((F.n_node, k), dtype=torch.float32, device=dev)
        m.matvec_multi(Xt.float(), Y32, k, prec=da.F32)
        Yn = _np(Y32)
        for j in (0, k - 1):
            assert rel_err(Yn[:, j], refs[j]) <= 1e-5, (name, variant, j)


def test_small_elasticity(dev, small):
    name, h = small
    F = h.fine
    m = da.Mesh(F)
    m.set_coefficients(None, None, 1.0, 0.0)
    m.set_elasticity(0.6, 1.7)
    x = np.random.default_rng(82).uniform(-1, 1, (F.n_node, 3))
    xt = torch.tensor(x, dtype=torch.float64, device=dev)
    y = torch.empty_like(xt)
    m.matvec_elasticity(xt, y)
    assert rel_err(_np(y).ravel(), oracle.matvec_elasticity(F, x, 0.6, 1.7)) <= 1e-12, name


def test_small_pcg(dev, small):
    name, h = small
    coefs = var_coefficients(h, mass_base=1.0)
    ms = []
    for l in range(h.n):
        mm = da.Mesh(h.levels[l])
        mm.set_coefficients(coefs[l][0], coefs[l][1])
        ms.append(mm)
    ts = [da.Transfer(ms[l - 1], ms[l], *h.transfers[l - 1]) for l in range(1, h.n)]
    lv = [(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)]
    mg = oracle.MG(lv, omega=2.0 / 3.0, nu_pre=2, nu_post=2)
    b = np.random.default_rng(83).uniform(-1, 1, h.fine.n_node)
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    # the first iterate (later ones of the tiniest meshes sit at rounding level)
    xo, its, _, _ = mg.pcg(b, rtol=1e-30, max_it=1, norm_type=da.NORM_UNPRECONDITIONED)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=1, rtol=1e-30,
                          norm_type=da.NORM_UNPRECONDITIONED)
    x = torch.zeros_like(bt)
    st, _ = da.Solver(ms, ts, p).solve(bt, x)
    assert st.iterations == its == 1
    assert rel_err(_np(x), xo) <= 1e-11, name
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=50, rtol=1e-10,
                          norm_type=da.NORM_UNPRECONDITIONED, use_graph=1)
    x = torch.zeros_like(bt)
    st, _ = da.Solver(ms, ts, p).solve(bt, x)
    assert st.converged and st.rel_true <= 1e-10, name
    # one level: the preconditioner is the coarse solve itself (dense inverse)
    x1 = torch.zeros_like(bt)
    st1, _ = da.Solver(ms[-1:], [], p).solve(bt, x1)
    assert st1.converged and st1.iterations <= 2 and st1.rel_true <= 1e-10, (name, st1.iterations)
