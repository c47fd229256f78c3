"""Quadrature right-hand sides, nodal values and L2 errors on the GPU (the reference's
ComputeRHS2 / ComputeSol / ComputeError5, examples/src/backend/omgRhs.C) against the oracle,
and a known-answer test of the whole solve path: the Helmholtz problem -div grad u + u = f on
the unit cube with natural boundary conditions and u = sum_k cos(k pi x) cos(k pi y)
cos(k pi z), whose L2 error must fall as h^2 under uniform refinement."""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

POLY = [0.3, 1.0, -2.0, 0.5, 1.5, -0.7, 0.25, 2.0]
FUNCS = {"gaussians": dict(kind=da.FN_GAUSSIANS), "cos9": dict(kind=da.FN_COS_SUM, nfreq=9, a0=1.0, a1=1.0),
         "poly": dict(kind=da.FN_POLY, poly=POLY)}


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    da.init(0)
    return torch.device("cuda:0")


def _mesh(L):
    m = da.Mesh(L)
    m.set_coefficients(None, None, 1.0, 1.0)
    m.set_geometry(L.anchor)
    return m


@pytest.mark.parametrize("name", sorted(FUNCS))
def test_rhs_nodal_error_match_oracle(dev, hier_small, name):
    L = hier_small.fine
    kw = FUNCS[name]
    m = _mesh(L)
    f = da.function(**kw)
    b = torch.empty(L.n_node, dtype=torch.float64, device=dev)
    m.rhs_quadrature(f, b)
    torch.cuda.synchronize()
    b_ref = oracle.rhs_quadrature(L, **kw)
    assert rel_err(b.cpu().numpy(), b_ref) <= 1e-12
    u = torch.empty_like(b)
    m.nodal_values(f, u)
    torch.cuda.synchronize()
    u_ref = oracle.nodal_values(L, **kw)
    assert rel_err(u.cpu().numpy(), u_ref) <= 1e-13
    # error of a perturbed field, so that neither side is zero
    w = torch.tensor(u_ref + 0.01 * np.sin(np.arange(L.n_node)), device=dev)
    e = m.l2_error(f, w)
    e_ref = oracle.l2_error(L, w.cpu().numpy(), **kw)
    assert abs(e - e_ref) <= 1e-12 * e_ref


def test_geometry_required(dev, hier_small):
    m = da.Mesh(hier_small.fine)
    b = torch.empty(hier_small.fine.n_node, dtype=torch.float64, device=dev)
    with pytest.raises(da.DendroError):
        m.rhs_quadrature(da.function(da.FN_GAUSSIANS), b)


def _uniform_hierarchy(d, nlevels=4):
    n = 1 << d
    c = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    damg = da.DAMG(da.Octree.from_points(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).ravel(), d), nlevels, 20)
    return damg, [damg.level(l) for l in range(damg.nlevels)]


def _solve_helmholtz(damg, levels, dev, nfreq, prec=da.F64):
    ms = []
    for L in levels:
        ms.append(_mesh(L))
    ts = [da.Transfer(ms[l - 1], ms[l], *damg.transfer(l)) for l in range(1, len(levels))]
    F, mf = levels[-1], ms[-1]
    b = torch.empty(F.n_node, dtype=torch.float64, device=dev)
    mf.rhs_quadrature(da.function(da.FN_COS_SUM, nfreq=nfreq, a0=1.0, a1=1.0), b)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=200, rtol=1e-12 if prec == da.F64 else 1e-10,
                          norm_type=da.NORM_UNPRECONDITIONED, vcycle_prec=prec, use_graph=1)
    x = torch.zeros_like(b)
    st, _ = da.Solver(ms, ts, p).solve(b, x)
    assert st.converged
    return mf.l2_error(da.function(da.FN_COS_SUM, nfreq=nfreq, a0=1.0, a1=0.0), x), st


@pytest.mark.parametrize("nfreq,depths", [(2, (4, 5, 6)), (9, (4, 5, 6))])
def test_known_answer_second_order(dev, nfreq, depths):
    """Uniform octrees of increasing depth: the L2 error of the GPU solve against the exact
    solution (ComputeSol's cosine sum, nfreq = 9 as in the reference, and a smoother nfreq = 2)
    falls by about 4 per refinement (the oracle's solves give 4.12 and 4.03 for nfreq = 9)."""
    errs = []
    for d in depths:
        damg, levels = _uniform_hierarchy(d)
        e, st = _solve_helmholtz(damg, levels, dev, nfreq)
        errs.append(e)
        print(f"nfreq={nfreq} depth {d}: E={levels[-1].n_elem} L2 error {e:.4e} ({st.iterations} PCG its)")
    rates = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
    print("error ratios per refinement:", [f"{r:.3f}" for r in rates])
    assert all(3.8 <= r <= 4.3 for r in rates), rates


def test_known_answer_adaptive(dev):
    """An adaptive octree (Gaussian points, hanging vertices): the mixed-precision solve of the
    same problem reaches the FP64 solve's discretisation error, and that error lies between
    those of the uniform meshes at the octree's finest and coarsest element sizes."""
    pts = da.points(da.PTS_GAUSS, 20000, 7)
    damg = da.DAMG(da.Octree.from_points(pts, 7), 6, 500)
    levels = [damg.level(l) for l in range(damg.nlevels)]
    assert levels[-1].n_hanging > 0
    e64, st64 = _solve_helmholtz(damg, levels, dev, 2)
    e32, st32 = _solve_helmholtz(damg, levels, dev, 2, prec=da.F32)
    lv = levels[-1].lev
    damg_f, lf = _uniform_hierarchy(int(lv.max()) - 1)
    damg_c, lc = _uniform_hierarchy(int(lv.min()) - 1)
    ef, _ = _solve_helmholtz(damg_f, lf, dev, 2)
    ec, _ = _solve_helmholtz(damg_c, lc, dev, 2)
    print(f"adaptive E={levels[-1].n_elem}: L2 error FP64 {e64:.4e} ({st64.iterations} its), mixed {e32:.4e} "
          f"({st32.iterations} its); uniform finest {ef:.4e}, coarsest {ec:.4e}")
    assert abs(e32 - e64) <= 1e-6 * e64
    assert ef <= e64 <= ec
