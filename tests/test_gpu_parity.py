"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Bar (FP64): relative 2-norm error <= 1e-12 for the MatVec, diagonal, restriction,
prolongation, one V-cycle and the first PCG iterates; node and element indexing are shared
bit-exactly (both sides consume the same DA arrays).  FP32 paths state their error
against the FP64 oracle.
"""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import hierarchy, matvec_longdouble, node_coords, rel_err, var_coefficients

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL64 = 1e-12


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    da.init(0)
    return torch.device("cuda:0")


def _t(a, dev, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def _meshes(h, coefs=None, lap_fac=1.0, mass_fac=1.0, lane=0):
    ms = []
    for l in range(h.n):
        m = da.Mesh(h.levels[l])
        if coefs is None:
            m.set_coefficients(None, None, lap_fac, mass_fac)
        else:
            m.set_coefficients(coefs[l][0], coefs[l][1])
        m.set_lane_stencils(lane)
        ms.append(m)
    ts = [da.Transfer(ms[l - 1], ms[l], *h.transfers[l - 1]) for l in range(1, h.n)]
    return ms, ts


@pytest.mark.parametrize("lane", [da.ROWS_TYPE, da.ROWS_LANE, da.ROWS_CLASS])
@pytest.mark.parametrize("mode", ["jac1", "jac2", "jac2_nomass"])
def test_matvec_fp64(dev, hier_small, mode, lane):
    """lane: the stencil-row source (per wavefront type, per lane from the LDS copy of the
    table, per (type, level) class matrix; per-element coefficients use the per-type rows
    for the last)."""
    h = hier_small
    F = h.fine
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, F.n_node)
    if mode == "jac1":
        coef, kw = None, dict(lap_fac=1.0, mass_fac=1.0)
    else:
        c = var_coefficients(h, mass_base=None if mode == "jac2_nomass" else 1.0)
        coef, kw = c[-1], dict(kappa=c[-1][0], mass=c[-1][1])
    y_ref = oracle.matvec(F, x, **kw)
    m = da.Mesh(F)
    if coef is None:
        m.set_coefficients(None, None, 1.0, 1.0)
    else:
        m.set_coefficients(coef[0], coef[1])
    m.set_lane_stencils(lane)
    assert m.tile_info()["lane_stencils"] == lane
    xt = _t(x, dev)
    yt = torch.empty_like(xt)
    m.matvec(xt, yt)
    assert rel_err(_np(yt), y_ref) <= TOL64


@pytest.mark.parametrize("case", ["cos_mass", "near_constant"])
def test_matvec_smooth_x(dev, hier_small, case):
    """Smooth inputs, where A x is much smaller than the terms each row sums.  The fixed-point
    row sums round every element-row term to 2^-50 of its tile's bound (the largest element
    factor times stencil row sum in the tile, times max |x| over the tile's nodes, each rounded
    up to a power of two), not to 2^-53 of the result: with kappa spread over 10^(-1..1) in a
    tile that is about 2^-45 of a row's |A||x|, where FP64 summation keeps about 2^-54.
    Against an extended-precision product (64-bit significand) the GPU error is bounded by
    2^-43 of |A||x| (a 4x margin on the derivation), and by 1e-12 of |A x| when A x is not
    dominated by cancellation.  cos_mass: x = cos(pi X) cos(pi Y) cos(pi Z) with mass;
    near_constant: x = 1 + 1e-3 of the same with zero mass (constants are in the Laplacian's
    null space, |A x| ~ 1e-5 |A||x|: no FP64 evaluation, the oracle's included, keeps 1e-12 of
    |A x| there)."""
    h = hier_small
    F = h.fine
    xyz = node_coords(F)
    w = np.cos(np.pi * xyz[:, 0]) * np.cos(np.pi * xyz[:, 1]) * np.cos(np.pi * xyz[:, 2])
    c = var_coefficients(h, mass_base=1.0 if case == "cos_mass" else None)[-1]
    x = w if case == "cos_mass" else 1.0 + 1e-3 * w
    ex = matvec_longdouble(F, x, kappa=c[0], mass=c[1])
    scale = np.linalg.norm(matvec_longdouble(F, x, kappa=c[0], mass=c[1], absolute=True).astype(np.float64))
    y_or = oracle.matvec(F, x, kappa=c[0], mass=c[1])
    m = da.Mesh(F)
    m.set_coefficients(c[0], c[1])
    xt = _t(x, dev)
    yt = torch.empty_like(xt)
    m.matvec(xt, yt)
    y = _np(yt)
    e_gpu = float(np.linalg.norm((y - ex).astype(np.float64)))
    e_or = float(np.linalg.norm((y_or - ex).astype(np.float64)))
    nrm = float(np.linalg.norm(ex.astype(np.float64)))
    print(f"{case}: |Ax|/(|A||x|) = {nrm / scale:.2e}; error vs extended precision relative to |A||x|: "
          f"GPU {e_gpu / scale:.2e} (2^{np.log2(e_gpu / scale):.1f}), oracle {e_or / scale:.2e}; "
          f"relative to |Ax|: GPU {e_gpu / nrm:.2e}, oracle {e_or / nrm:.2e}")
    assert e_gpu <= 2.0 ** -43 * scale
    if case == "cos_mass":
        assert e_gpu <= 1e-12 * nrm


@pytest.mark.parametrize("prec", [da.F64, da.F32])
def test_matvec_reproducible_and_scale_exact(dev, hier_small, prec):
    """The operator sums owned-node rows in 64-bit fixed point with a per-tile scale taken
    from max |x|: repeated launches are bit-identical, scaling x by a power of two scales
    y by exactly the same power (the integers are unchanged), and a non-finite x turns the
    rows it touches into NaN instead of wrapping around."""
    h = hier_small
    F = h.fine
    c = var_coefficients(h, mass_base=1.0)
    m = da.Mesh(F)
    m.set_coefficients(c[-1][0], c[-1][1])
    dt = torch.float64 if prec == da.F64 else torch.float32
    s = 300 if prec == da.F64 else 60
    x = _t(np.random.default_rng(7).uniform(-1, 1, F.n_node), dev, dt)
    ys = []
    for xv in (x, x, x * 2.0 ** s, x * 2.0 ** -s):
        y = torch.empty_like(xv)
        m.matvec(xv, y, prec=prec)
        torch.cuda.synchronize()
        ys.append(y)
    assert torch.equal(ys[0], ys[1])
    assert torch.equal(ys[2], ys[0] * 2.0 ** s)
    assert torch.equal(ys[3], ys[0] * 2.0 ** -s)
    xn = x.clone()
    xn[F.n_node // 2] = float("nan")
    yn = torch.empty_like(xn)
    m.matvec(xn, yn, prec=prec)
    torch.cuda.synchronize()
    assert torch.isnan(yn[F.n_node // 2])


@pytest.mark.parametrize("lane", [da.ROWS_TYPE, da.ROWS_LANE, da.ROWS_CLASS])
def test_matvec_dot(dev, hier_small, lane):
    """y = A x with x.y from the operator's tile partials: y bit-identical to dgh_matvec,
    both against the oracle (<= 1e-12), and x.y against the FP64 dot of the oracle's y."""
    h = hier_small
    F = h.fine
    # the class matrices serve uniform coefficients (lambda = 1.5, mu = 0.5)
    c = var_coefficients(h, mass_base=1.0)[-1] if lane != da.ROWS_CLASS else (None, None)
    kw = dict(kappa=c[0], mass=c[1]) if c[0] is not None else dict(lap_fac=1.5, mass_fac=0.5)
    x = np.random.default_rng(17).uniform(-1, 1, F.n_node)
    y_ref = oracle.matvec(F, x, **kw)
    m = da.Mesh(F)
    m.set_coefficients(c[0], c[1], kw.get("lap_fac", 1.0), kw.get("mass_fac", 1.0))
    m.set_lane_stencils(lane)
    xt = _t(x, dev)
    y1, y2 = torch.empty_like(xt), torch.empty_like(xt)
    xy = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
    m.matvec(xt, y1)
    m.matvec_dot(xt, y2, xy)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    assert rel_err(_np(y2), y_ref) <= TOL64
    d_ref = float(np.dot(x, y_ref))
    assert abs(float(xy.item()) - d_ref) <= 1e-12 * abs(d_ref)


@pytest.mark.parametrize("lane", [da.ROWS_TYPE, da.ROWS_LANE, da.ROWS_CLASS])
def test_diagonal_fp64(dev, hier_small, lane):
    h = hier_small
    c = var_coefficients(h, mass_base=1.0)[-1] if lane != da.ROWS_CLASS else (None, None)
    kw = dict(kappa=c[0], mass=c[1]) if c[0] is not None else dict(lap_fac=1.5, mass_fac=0.5)
    m = da.Mesh(h.fine)
    m.set_coefficients(c[0], c[1], kw.get("lap_fac", 1.0), kw.get("mass_fac", 1.0))
    m.set_lane_stencils(lane)
    d = torch.empty(h.fine.n_node, dtype=torch.float64, device=dev)
    m.diagonal(d)
    assert rel_err(_np(d), oracle.diagonal(h.fine, **kw)) <= TOL64


def test_transfer_fp64(dev, hier_small):
    h = hier_small
    ms, ts = _meshes(h)
    rng = np.random.default_rng(3)
    for l in range(1, h.n):
        C, F = h.levels[l - 1], h.levels[l]
        f = rng.uniform(-1, 1, F.n_node)
        c = rng.uniform(-1, 1, C.n_node)
        ct = torch.empty(C.n_node, dtype=torch.float64, device=dev)
        ts[l - 1].restrict(_t(f, dev), ct)
        assert rel_err(_np(ct), oracle.restrict(C, F, f)) <= TOL64
        fin = _t(f, dev)
        fout = torch.empty_like(fin)
        ts[l - 1].prolong_add(_t(c, dev), fin, fout)
        assert rel_err(_np(fout), f + oracle.prolong(C, F, c)) <= TOL64


@pytest.mark.parametrize("lane", [0, 1])
def test_fp32_operators_error_stated(dev, hier_small, lane):
    h = hier_small
    F = h.fine
    x = np.random.default_rng(4).uniform(-1, 1, F.n_node)
    m = da.Mesh(F)
    m.set_coefficients(None, None, 1.0, 1.0)
    m.set_lane_stencils(lane)
    xt = _t(x, dev, torch.float32)
    yt = torch.empty_like(xt)
    m.matvec(xt, yt, prec=da.F32)
    err = rel_err(_np(yt), oracle.matvec(F, x))
    print(f"FP32 MatVec relative error vs FP64 oracle: {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("lane", [0, 1])
@pytest.mark.parametrize("null_space", [False, True])
def test_vcycle_fp64(dev, hier_tiny, null_space, lane):
    h = hier_tiny
    coefs = var_coefficients(h, mass_base=None if null_space else 1.0)
    ms, ts = _meshes(h, coefs, lane=lane)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=int(null_space))
    s = da.Solver(ms, ts, p)
    mg = oracle.MG([(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)],
                   omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=null_space)
    r = np.random.default_rng(7).uniform(-1, 1, h.fine.n_node)
    if null_space:
        r -= r.mean()
    z = torch.empty(h.fine.n_node, dtype=torch.float64, device=dev)
    s.vcycle(_t(r, dev), z)
    assert rel_err(_np(z), mg.vcycle(r)) <= 1e-11


@pytest.mark.parametrize("lane", [0, 1])
@pytest.mark.parametrize("norm_type", [da.NORM_PRECONDITIONED, da.NORM_UNPRECONDITIONED])
def test_pcg_iterates_fp64(dev, hier_tiny, norm_type, lane):
    h = hier_tiny
    ms, ts = _meshes(h, None, 1.0, 1.0, lane=lane)
    b = np.random.default_rng(12).uniform(-1, 1, h.fine.n_node)
    mg = oracle.MG([(h.levels[l], None, None, 1.0, 1.0) for l in range(h.n)],
                   omega=2.0 / 3.0, nu_pre=2, nu_post=2)
    for k in (1, 3, 5):
        p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=k, rtol=1e-30,
                              norm_type=norm_type)
        s = da.Solver(ms, ts, p)
        x = torch.zeros(h.fine.n_node, dtype=torch.float64, device=dev)
        st, hist = s.solve(_t(b, dev), x, history_len=k + 1)
        xo, its, _, ho = mg.pcg(b, rtol=1e-30, max_it=k, norm_type=norm_type)
        assert st.iterations == its == k
        assert rel_err(_np(x), xo) <= TOL64, k
        assert np.allclose(hist, ho, rtol=1e-10, atol=0)


@pytest.mark.parametrize("norm_type", [da.NORM_PRECONDITIONED, da.NORM_UNPRECONDITIONED])
def test_pcg_iterates_jac2_null_space(dev, hier_small, norm_type):
    """The operator the benchmark solves: Jacobian2 with kappa = 10^U(-1,1), zero mass and the
    constant null space projected, FP64.  V-cycle and PCG iterates k = 1, 3, 5 against the
    oracle (<= 1e-12), stopping-test norms to 1e-10; the FP32 V-cycle's iterates and its
    converged solution are stated against the same oracle run."""
    h = hier_small
    coefs = var_coefficients(h, mass_base=None)
    ms, ts = _meshes(h, coefs)
    mg = oracle.MG([(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)],
                   omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=True)
    b = np.random.default_rng(22).uniform(-1, 1, h.fine.n_node)
    bt = _t(b, dev)
    r = b - b.mean()
    s = da.Solver(ms, ts, da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=1))
    z = torch.empty_like(bt)
    s.vcycle(_t(r, dev), z)
    assert rel_err(_np(z), mg.vcycle(r)) <= 1e-11
    for k in (1, 3, 5):
        xo, its, _, ho = mg.pcg(b, rtol=1e-30, max_it=k, norm_type=norm_type)
        for prec in (da.F64, da.F32):
            p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=k, rtol=1e-30,
                                  norm_type=norm_type, null_space=1, vcycle_prec=prec)
            s = da.Solver(ms, ts, p)
            x = torch.zeros_like(bt)
            st, hist = s.solve(bt, x, history_len=k + 1)
            err = rel_err(_np(x), xo)
            assert st.iterations == its == k
            if prec == da.F64:
                assert err <= TOL64, (k, err)
                assert np.allclose(hist, ho, rtol=1e-10, atol=0)
            else:
                print(f"k={k} norm_type={norm_type}: FP32 V-cycle iterate vs FP64 oracle {err:.2e}")
                assert err <= 1e-5
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=200, rtol=1e-10,
                          norm_type=norm_type, null_space=1, vcycle_prec=da.F32)
    x = torch.zeros_like(bt)
    st, _ = da.Solver(ms, ts, p).solve(bt, x)
    xo, its, conv, _ = mg.pcg(b, rtol=1e-10, max_it=200, norm_type=norm_type)
    err = rel_err(_np(x), xo)
    print(f"converged to 1e-10: mixed {st.iterations} its, oracle FP64 {its} its, solution error {err:.2e}")
    assert st.converged and conv and err <= 1e-6


@pytest.mark.parametrize("lane", [0, 1])
def test_pcg_converges_fp64_and_mixed(dev, hier_small, lane):
    h = hier_small
    coefs = var_coefficients(h, mass_base=None)
    ms, ts = _meshes(h, coefs, lane=lane)
    b = np.random.default_rng(22).uniform(-1, 1, h.fine.n_node)
    out = {}
    for prec in (da.F64, da.F32):
        p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=200, rtol=1e-8,
                              norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=prec)
        s = da.Solver(ms, ts, p)
        x = torch.zeros(h.fine.n_node, dtype=torch.float64, device=dev)
        st, _ = s.solve(_t(b, dev), x)
        assert st.converged and st.rel_true <= 1.5e-8, (prec, st.iterations, st.rel_true)
        out[prec] = _np(x)
        print(f"prec={prec} its={st.iterations} rel_true={st.rel_true:.2e} t={st.seconds*1e3:.2f} ms")
    err = rel_err(out[da.F32], out[da.F64])
    print(f"mixed-precision solution error vs FP64 solve: {err:.2e}")
    assert err <= 1e-6


def test_pcg_graph_replays_match_eager(dev, hier_small):
    """The hipGraph-captured CG iteration gives the eager iterates on every replay."""
    h = hier_small
    coefs = var_coefficients(h, mass_base=1.0)
    ms, ts = _meshes(h, coefs)
    b = np.random.default_rng(31).uniform(-1, 1, h.fine.n_node)
    bt = _t(b, dev)
    res = {}
    for g in (0, 1):
        p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=50, rtol=1e-10,
                              norm_type=da.NORM_UNPRECONDITIONED, use_graph=g)
        s = da.Solver(ms, ts, p)
        outs = []
        for rep in range(3 if g else 1):
            x = torch.zeros(h.fine.n_node, dtype=torch.float64, device=dev)
            st, hist = s.solve(bt, x, history_len=51)
            _ = (bt * 2.0).sum().item()          # eager work between replays
            outs.append((_np(x), hist.copy(), st.iterations))
        res[g] = outs
    x0, h0, it0 = res[0][0]
    for x1, h1, it1 in res[1]:
        assert it1 == it0
        assert rel_err(x1, x0) <= 1e-12
        assert np.allclose(h1, h0, rtol=1e-9, atol=0)


def test_pcg_graph_recaptured_after_retiling(dev, hier_small):
    """A captured CG iteration holds the meshes' device buffers: re-tiling a mesh or setting
    new coefficients after a solve must not replay a stale graph (the solver re-captures)."""
    h = hier_small
    coefs = var_coefficients(h, mass_base=1.0)
    ms, ts = _meshes(h, coefs)
    b = _t(np.random.default_rng(32).uniform(-1, 1, h.fine.n_node), dev)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=50, rtol=1e-10,
                          norm_type=da.NORM_UNPRECONDITIONED, use_graph=1, vcycle_prec=da.F32)
    s = da.Solver(ms, ts, p)
    x0 = torch.zeros_like(b)
    st0, _ = s.solve(b, x0)
    ms[-1].set_tile(128, 256)            # new tiling of the finest operator
    ms[-2].set_tile(256, 512)
    x1 = torch.zeros_like(b)
    st1, _ = s.solve(b, x1)
    assert st1.converged and st1.iterations == st0.iterations
    assert rel_err(_np(x1), _np(x0)) <= 1e-6
    ms[-1].set_coefficients(coefs[-1][0], coefs[-1][1])   # same values, new device buffers
    x2 = torch.zeros_like(b)
    st2, _ = s.solve(b, x2)
    # the operator sums are order-independent: the same operator gives the same iterates
    assert st2.converged and st2.iterations == st1.iterations
    assert torch.equal(x2, x1)


@pytest.mark.parametrize("prec", [da.F64, da.F32])
def test_class_rows_agree(dev, hier_small, prec):
    """Uniform coefficients: the (type, level) class matrices and the per-type rows agree to
    rounding, the default source of a small tiling is the per-type rows, and a PCG solve takes
    the same iterations with every source."""
    h = hier_small
    dt = torch.float64 if prec == da.F64 else torch.float32
    x = _t(np.random.default_rng(43).uniform(-1, 1, h.fine.n_node), dev, dt)
    m0 = da.Mesh(h.fine)   # no source set: a tiling of few tiles takes the per-type rows
    m0.set_coefficients(None, None, 1.0, 1.0)
    assert m0.tile_info()["lane_stencils"] == da.ROWS_TYPE
    with pytest.raises(da.DendroError):
        m0.set_lane_stencils(3)
    del m0
    ms, ts = _meshes(h, lap_fac=1.0, mass_fac=1.0)
    ys = []
    for rows in (da.ROWS_TYPE, da.ROWS_CLASS, da.ROWS_LANE):
        ms[-1].set_lane_stencils(rows)
        y = torch.empty_like(x)
        ms[-1].matvec(x, y, prec=prec)
        ys.append(_np(y))
    for y in ys[1:]:
        err = rel_err(y, ys[0])
        print(f"class / lane vs type stencil rows: {err:.2e}")
        assert err <= (1e-14 if prec == da.F64 else 1e-6)
    b = _t(np.random.default_rng(44).uniform(-1, 1, h.fine.n_node), dev)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=50, rtol=1e-10,
                          norm_type=da.NORM_UNPRECONDITIONED, use_graph=1, vcycle_prec=prec)
    s = da.Solver(ms, ts, p)
    its = []
    for rows in (da.ROWS_TYPE, da.ROWS_CLASS):
        for m in ms:
            m.set_lane_stencils(rows)
        xo = torch.zeros_like(b)
        st, _ = s.solve(b, xo)
        assert st.converged and st.rel_true <= 1e-10
        its.append(st.iterations)
    assert its[0] == its[1]


@pytest.mark.parametrize("prec", [da.F64, da.F32])
def test_lane_stencils_agree(dev, hier_small, prec):
    """The two stencil-row sources agree to rounding (the row sums may be scheduled
    differently), and a solver captured with one source re-captures after the switch and
    reproduces the solve."""
    h = hier_small
    coefs = var_coefficients(h, mass_base=1.0)
    ms, ts = _meshes(h, coefs)
    dt = torch.float64 if prec == da.F64 else torch.float32
    x = _t(np.random.default_rng(41).uniform(-1, 1, h.fine.n_node), dev, dt)
    ys = []
    for lane in (0, 1):
        ms[-1].set_lane_stencils(lane)
        y = torch.empty_like(x)
        ms[-1].matvec(x, y, prec=prec)
        ys.append(_np(y))
    err = rel_err(ys[1], ys[0])
    print(f"lane vs wavefront-type stencil rows: {err:.2e}")
    assert err <= (1e-14 if prec == da.F64 else 1e-6)
    b = _t(np.random.default_rng(42).uniform(-1, 1, h.fine.n_node), dev)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=50, rtol=1e-10,
                          norm_type=da.NORM_UNPRECONDITIONED, use_graph=1, vcycle_prec=prec)
    s = da.Solver(ms, ts, p)
    out = []
    for lane in (0, 1):
        for m in ms:
            m.set_lane_stencils(lane)
        xo = torch.zeros_like(b)
        st, _ = s.solve(b, xo)
        assert st.converged
        out.append((st.iterations, _np(xo)))
    assert out[0][0] == out[1][0]
    assert rel_err(out[1][1], out[0][1]) <= 1e-8
