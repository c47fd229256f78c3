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
                                             ("var", 1, 0), ("jacobian", 1, 256), ("var", 0, 1024),
                                             ("laplacian", 2, 0), ("jacobian", 2, 1024)])
def test_matvec_dirichlet_gpu(dev, hier_small, mode, lane, block):
    """lane: stencil-row source (1 per lane from LDS, 2 per (type, level) class matrix);
    block: 256 / 1024 threads per tile (0: default)"""
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


# ------------------------------------------------------------------ the Dirichlet solve
def _dirichlet_hierarchy(h, mode):
    ms, coefs = [], var_coefficients(h, mass_base=1.0) if mode == "var" else None
    for l in range(h.n):
        m = da.Mesh(h.levels[l])
        if coefs is None:
            m.set_coefficients(None, None, **MODES[mode])
        else:
            m.set_coefficients(coefs[l][0], coefs[l][1])
        m.set_boundary(h.damg.boundary(l))
        ms.append(m)
    ts = [da.Transfer(ms[l - 1], ms[l], *h.transfers[l - 1]) for l in range(1, h.n)]
    lv = [(h.levels[l], None if coefs is None else coefs[l][0], None if coefs is None else coefs[l][1],
           MODES.get(mode, {}).get("lap_fac", 1.0), MODES.get(mode, {}).get("mass_fac", 1.0)) for l in range(h.n)]
    return ms, ts, lv


@pytest.mark.gpu
def test_dirichlet_transfers_gpu(dev, hier_small):
    """With boundary flags on both meshes, restriction and prolongation drop every entry that
    touches a boundary node (the suppressed DOFs of RestrictMatVec.C:75-80 and
    ProlongMatVec.C:76-80): against the oracle's masked transfer pairs, through the V-cycle's
    own kernels."""
    h = hier_small
    ms, ts, lv = _dirichlet_hierarchy(h, "laplacian")
    C, F = h.levels[-2], h.levels[-1]
    bc, bf = h.damg.boundary(h.n - 2) != 0, h.damg.boundary(h.n - 1) != 0
    f = np.random.default_rng(5).uniform(-1, 1, F.n_node)
    c = np.random.default_rng(6).uniform(-1, 1, C.n_node)
    ct = torch.empty(C.n_node, dtype=torch.float64, device=dev)
    ts[-1].restrict(torch.tensor(f, device=dev), ct)
    ref = oracle.restrict(C, F, np.where(bf, 0.0, f))
    ref[bc] = 0.0
    assert rel_err(ct.cpu().numpy(), ref) <= 1e-12
    fo = torch.empty(F.n_node, dtype=torch.float64, device=dev)
    ts[-1].prolong_add(torch.tensor(c, device=dev), torch.tensor(f, device=dev), fo)
    p = oracle.prolong(C, F, np.where(bc, 0.0, c))
    p[bf] = 0.0
    assert rel_err(fo.cpu().numpy(), f + p) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["laplacian", "var"])
def test_dirichlet_solve_gpu(dev, hier_small, mode):
    """The Dirichlet problem through the multigrid-preconditioned CG (SetDirichletJacContexts,
    examples/src/backend/omgJac.C:508-586): V-cycle and PCG iterates k = 1, 3, 5 against the
    oracle's Dirichlet hierarchy (<= 1e-11 / 1e-12), the solution returns b on boundary nodes,
    and the FP32 V-cycle solve meets the same true residual."""
    h = hier_small
    ms, ts, lv = _dirichlet_hierarchy(h, mode)
    bd = [h.damg.boundary(l) for l in range(h.n)]
    mg = oracle.MG(lv, omega=2.0 / 3.0, nu_pre=2, nu_post=2, boundary=bd)
    b = np.random.default_rng(51).uniform(-1, 1, h.fine.n_node)
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    s = da.Solver(ms, ts, da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2))
    z = torch.empty_like(bt)
    s.vcycle(bt, z)
    torch.cuda.synchronize()
    assert rel_err(z.cpu().numpy(), mg.vcycle(b)) <= 1e-11
    for k in (1, 3, 5):
        xo, its, _, ho = mg.pcg(b, rtol=1e-30, max_it=k, norm_type=da.NORM_UNPRECONDITIONED)
        p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=k, rtol=1e-30,
                              norm_type=da.NORM_UNPRECONDITIONED)
        x = torch.zeros_like(bt)
        st, hist = da.Solver(ms, ts, p).solve(bt, x, history_len=k + 1)
        torch.cuda.synchronize()
        assert st.iterations == its == k
        assert rel_err(x.cpu().numpy(), xo) <= 1e-12, k
        assert np.allclose(hist, ho, rtol=1e-10, atol=0)
    out = {}
    for prec in (da.F64, da.F32):
        p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=200, rtol=1e-10,
                              norm_type=da.NORM_UNPRECONDITIONED, vcycle_prec=prec, use_graph=1)
        x = torch.zeros_like(bt)
        st, _ = da.Solver(ms, ts, p).solve(bt, x)
        assert st.converged and st.rel_true <= 1e-10, (prec, st.rel_true)
        xs = x.cpu().numpy()
        assert np.abs(xs[bd[-1] != 0] - b[bd[-1] != 0]).max() <= 1e-9
        out[prec] = (xs, st.iterations)
    print(f"Dirichlet {mode}: FP64 {out[da.F64][1]} its, FP32 V-cycle {out[da.F32][1]} its, "
          f"solution difference {rel_err(out[da.F32][0], out[da.F64][0]):.2e}")


@pytest.mark.gpu
def test_dirichlet_solver_rules(dev, hier_small):
    """Flags on some levels only, a null space with flags, or flags changed after the solver
    was created are errors; new coefficients after creation are taken up (diagonals and the
    coarsest inverse rebuilt)."""
    h = hier_small
    ms, ts, lv = _dirichlet_hierarchy(h, "laplacian")
    ms[0].set_boundary(None)
    with pytest.raises(da.DendroError):
        da.Solver(ms, ts, da.default_params())
    ms[0].set_boundary(h.damg.boundary(0))
    with pytest.raises(da.DendroError):
        da.Solver(ms, ts, da.default_params(null_space=1))
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=100, rtol=1e-10,
                          norm_type=da.NORM_UNPRECONDITIONED)
    s = da.Solver(ms, ts, p)
    b = torch.tensor(np.random.default_rng(2).uniform(-1, 1, h.fine.n_node), device=dev)
    x = torch.zeros_like(b)
    s.solve(b, x)
    # Laplacian -> Jacobian (mass added) on every level after creation: same solver, new operator
    for m in ms:
        m.set_coefficients(None, None, 1.0, 1.0)
    x2 = torch.zeros_like(b)
    st2, _ = s.solve(b, x2)
    _, _, lvj = _dirichlet_hierarchy(h, "jacobian")
    xo, its, conv, _ = oracle.MG(lvj, omega=2.0 / 3.0, nu_pre=2, nu_post=2,
                                 boundary=[h.damg.boundary(l) for l in range(h.n)]).pcg(
        b.cpu().numpy(), rtol=1e-10, max_it=100, norm_type=0)
    assert st2.converged and st2.iterations == its
    assert rel_err(x2.cpu().numpy(), xo) <= 1e-9
    ms[-1].set_boundary(h.damg.boundary(h.n - 1))
    with pytest.raises(da.DendroError):
        s.solve(b, x)
