"""The benchmark configurations at their full sizes on the GPU.

Full-size parity: the FP64 MatVec against the oracle on the whole finest mesh (<= 1e-12)
in every configuration; on the 1M log-normal octree also the V-cycle and the PCG iterates
k = 1, 3, 5 against the oracle's (<= 1e-12) and the 32 batched MatVec columns; the FP64
true residual of every solve; agreement between the FP64, mixed-precision and
learned-damping solves; convergence of every right-hand side of a batch.  The 8M-point and
4M-point octrees take about a minute of host time each to build.
"""
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


def _build(kind, npts, depth, nlevels, seed, cluster_seed=0, min_coarse=500):
    pts = da.points(kind, npts, seed, cluster_seed)
    damg = da.DAMG(da.Octree.from_points(pts, depth), nlevels, min_coarse)
    return damg, [damg.level(l) for l in range(damg.nlevels)]


def _device_hierarchy(damg, levels, kappa_fine=None, lap_fac=1.0, mass_fac=1.0):
    ms = []
    for l, L in enumerate(levels):
        m = da.Mesh(L)
        if kappa_fine is None:
            m.set_coefficients(None, None, lap_fac, mass_fac)
        else:
            m.set_coefficients(damg.coarsen_coefficients(kappa_fine, None, l)[0], None)
        ms.append(m)
    ts = [da.Transfer(ms[l - 1], ms[l], *damg.transfer(l)) for l in range(1, len(levels))]
    return ms, ts


def _solve(ms, ts, b, dev, **kw):
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=500, rtol=1e-8,
                          norm_type=da.NORM_UNPRECONDITIONED, use_graph=1)
    for k, v in kw.items():
        setattr(p, k, v)
    s = da.Solver(ms, ts, p)
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    x = torch.zeros_like(bt)
    st, _ = s.solve(bt, x)
    torch.cuda.synchronize()
    return s, st, x.cpu().numpy()


def test_config1_gaussian_helmholtz(dev):
    """100k Gaussian points, depth 8, Jacobian1 (lambda = mu = 1): MatVec parity on the full
    mesh, 3-level PCG to 1e-8 (FP64 and mixed)."""
    damg, levels = _build(da.PTS_GAUSS, 100000, 8, 3, 1)
    F = levels[-1]
    ms, ts = _device_hierarchy(damg, levels)
    x = np.random.default_rng(11).uniform(-1, 1, F.n_node)
    xt = torch.tensor(x, device=dev)
    yt = torch.empty_like(xt)
    ms[-1].matvec(xt, yt)
    assert rel_err(yt.cpu().numpy(), oracle.matvec(F, x)) <= 1e-12
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(100):                      # the configuration's 100 MatVecs
        ms[-1].matvec(xt, yt)
    e1.record()
    torch.cuda.synchronize()
    print(f"config 1: 100 FP64 MatVecs in {e0.elapsed_time(e1):.3f} ms (default tiling)")
    b = np.random.default_rng(12).uniform(-1, 1, F.n_node)
    _, st64, x64 = _solve(ms, ts, b, dev)
    _, st32, x32 = _solve(ms, ts, b, dev, vcycle_prec=da.F32)
    print(f"config 1: E={F.n_elem} N={F.n_node}; PCG FP64 {st64.iterations} its {st64.seconds*1e3:.2f} ms, "
          f"mixed {st32.iterations} its {st32.seconds*1e3:.2f} ms, |x32-x64|/|x64| {rel_err(x32, x64):.2e}")
    assert st64.converged and st64.rel_true <= 1e-8
    assert st32.converged and st32.rel_true <= 1e-8
    assert rel_err(x32, x64) <= 1e-6


def test_config2_lognormal_poisson(dev):
    """1M log-normal points, depth 12, kappa = 10^U(-1,1), 5 levels, PCG to 1e-10."""
    damg, levels = _build(da.PTS_LOGNORMAL, 1000000, 12, 5, 2)
    F = levels[-1]
    kf = 10.0 ** np.random.default_rng(21).uniform(-1, 1, F.n_elem)
    ms, ts = _device_hierarchy(damg, levels, kf)
    x = np.random.default_rng(11).uniform(-1, 1, F.n_node)
    xt = torch.tensor(x, device=dev)
    yt = torch.empty_like(xt)
    ms[-1].matvec(xt, yt)
    assert rel_err(yt.cpu().numpy(), oracle.matvec(F, x, kappa=kf)) <= 1e-12
    b = np.random.default_rng(22).uniform(-1, 1, F.n_node)
    _, st64, x64 = _solve(ms, ts, b, dev, rtol=1e-10, null_space=1)
    _, st32, x32 = _solve(ms, ts, b, dev, rtol=1e-10, null_space=1, vcycle_prec=da.F32)
    print(f"config 2: E={F.n_elem} N={F.n_node}; PCG(1e-10) FP64 {st64.iterations} its "
          f"{st64.seconds*1e3:.2f} ms, mixed {st32.iterations} its {st32.seconds*1e3:.2f} ms, "
          f"|x32-x64|/|x64| {rel_err(x32, x64):.2e}")
    assert st64.converged and st64.rel_true <= 1e-10
    assert st32.converged and st32.rel_true <= 1e-10
    assert rel_err(x32, x64) <= 1e-8


@pytest.fixture(scope="module")
def config2():
    damg, levels = _build(da.PTS_LOGNORMAL, 1000000, 12, 5, 2)
    kf = 10.0 ** np.random.default_rng(21).uniform(-1, 1, levels[-1].n_elem)
    return damg, levels, kf


def test_config2_iterates_vs_oracle(dev, config2):
    """The benchmark's solve on the whole 1M log-normal octree (Jacobian2, zero mass,
    constant null space, 5 levels): one V-cycle and the PCG iterates k = 1, 3, 5 of the
    plain FP64 GPU path against the oracle (<= 1e-12), with the stopping-test norms; the
    FP32 V-cycle's iterates stated against the same oracle iterates."""
    damg, levels, kf = config2
    F = levels[-1]
    ms, ts = _device_hierarchy(damg, levels, kf)
    co = [damg.coarsen_coefficients(kf, None, l)[0] for l in range(len(levels))]
    mg = oracle.MG([(levels[l], co[l], None, 1.0, 1.0) for l in range(len(levels))],
                   omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=True)
    b = np.random.default_rng(22).uniform(-1, 1, F.n_node)
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    r = b - b.mean()
    s = da.Solver(ms, ts, da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=1))
    z = torch.empty_like(bt)
    s.vcycle(torch.tensor(r, device=dev), z)
    torch.cuda.synchronize()
    ev = rel_err(z.cpu().numpy(), mg.vcycle(r))
    print(f"config 2 V-cycle vs oracle: {ev:.2e}")
    assert ev <= 1e-11
    for k in (1, 3, 5):
        xo, its, _, ho = mg.pcg(b, rtol=1e-30, max_it=k, norm_type=da.NORM_UNPRECONDITIONED)
        for prec in (da.F64, da.F32):
            p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=k, rtol=1e-30,
                                  norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=prec)
            s = da.Solver(ms, ts, p)
            x = torch.zeros_like(bt)
            st, hist = s.solve(bt, x, history_len=k + 1)
            torch.cuda.synchronize()
            err = rel_err(x.cpu().numpy(), xo)
            print(f"config 2 PCG iterate k={k} prec={prec}: rel err vs oracle {err:.2e}")
            assert st.iterations == its == k
            if prec == da.F64:
                assert err <= 1e-12
                assert np.allclose(hist, ho, rtol=1e-10, atol=0)
            else:
                assert err <= 1e-5


def test_config4_batched_32rhs(dev, config2):
    """Config 4: the 1M log-normal octree with 32 right-hand sides (seeds 400..431) in one
    batched solve (FP32 batched V-cycle captured in a hipGraph, FP64 outer CG per column):
    every column meets its own 1e-8 true relative residual; the batched FP64 MatVec's 32
    columns match the oracle (<= 1e-12) and the FP32 ones are stated."""
    damg, levels, kf = config2
    F = levels[-1]
    N = F.n_node
    k = 32
    ms, ts = _device_hierarchy(damg, levels, kf)
    X = np.random.default_rng(11).uniform(-1, 1, (N, k))
    Xt = torch.tensor(X, dtype=torch.float64, device=dev)
    Yt = torch.empty_like(Xt)
    ms[-1].matvec_multi(Xt, Yt, k, prec=da.F64)
    X32 = Xt.float()
    Y32 = torch.empty_like(X32)
    ms[-1].matvec_multi(X32, Y32, k, prec=da.F32)
    torch.cuda.synchronize()
    Y, Y32 = Yt.cpu().numpy(), Y32.double().cpu().numpy()
    worst64 = worst32 = 0.0
    for j in range(k):
        ref = oracle.matvec(F, X[:, j], kappa=kf)
        worst64 = max(worst64, rel_err(Y[:, j], ref))
        worst32 = max(worst32, rel_err(Y32[:, j], ref))
    print(f"config 4 batched MatVec, 32 columns vs oracle: FP64 worst {worst64:.2e}, FP32 worst {worst32:.2e}")
    assert worst64 <= 1e-12 and worst32 <= 1e-5
    del Xt, Yt, X32
    B = np.stack([np.random.default_rng(400 + j).uniform(-1, 1, N) for j in range(k)], 1)
    Bt = torch.tensor(B, dtype=torch.float64, device=dev)
    Xs = torch.zeros_like(Bt)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=500, rtol=1e-8,
                          norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=da.F32, use_graph=1)
    s = da.Solver(ms, ts, p)
    st = s.solve_multi(Bt, Xs, k)
    its = [x.iterations for x in st]
    print(f"config 4: 32 RHS in {st[0].seconds*1e3:.1f} ms, iterations {min(its)}-{max(its)}, "
          f"max rel_true {max(x.rel_true for x in st):.2e}")
    assert all(x.converged and x.rel_true <= 1e-8 for x in st)
    # a column of the batch against the single-RHS FP64 solve of the same column
    _, st1, x1 = _solve(ms, ts, B[:, 7], dev, null_space=1)
    torch.cuda.synchronize()
    e = rel_err(Xs[:, 7].cpu().numpy(), x1)
    print(f"config 4 column 7 vs single FP64 solve: {e:.2e}")
    assert e <= 1e-6


def test_config3_mixed_8m(dev):
    """8M points (4M uniform + 4M in 8 Gaussian clusters), depth 14: mixed-precision PCG
    against the FP64 solve on the same octree."""
    damg, levels = _build(da.PTS_MIXED, 8000000, 14, 8, 3, 31)
    F = levels[-1]
    ms, ts = _device_hierarchy(damg, levels, None, 1.0, 1.0)
    # uniform coefficients on a tiling of >= 1024 tiles: the class matrices by default
    assert ms[-1].tile_info()["lane_stencils"] == da.ROWS_CLASS
    x = np.random.default_rng(11).uniform(-1, 1, F.n_node)
    xt = torch.tensor(x, device=dev)
    yt = torch.empty_like(xt)
    ms[-1].matvec(xt, yt)
    y_ref = oracle.matvec(F, x)
    assert rel_err(yt.cpu().numpy(), y_ref) <= 1e-12
    ms[-1].set_lane_stencils(da.ROWS_TYPE)    # and the per-type rows at full size
    ms[-1].matvec(xt, yt)
    assert rel_err(yt.cpu().numpy(), y_ref) <= 1e-12
    ms[-1].set_lane_stencils(da.ROWS_CLASS)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(100):                      # the configuration's 100 MatVecs
        ms[-1].matvec(xt, yt)
    e1.record()
    torch.cuda.synchronize()
    print(f"config 3: 100 FP64 MatVecs in {e0.elapsed_time(e1):.3f} ms (default tiling)")
    b = np.random.default_rng(12).uniform(-1, 1, F.n_node)
    _, st64, x64 = _solve(ms, ts, b, dev)
    _, st32, x32 = _solve(ms, ts, b, dev, vcycle_prec=da.F32)
    print(f"config 3: E={F.n_elem} N={F.n_node} levels {len(levels)} coarsest N={levels[0].n_node}; "
          f"PCG FP64 {st64.iterations} its {st64.seconds*1e3:.1f} ms, mixed {st32.iterations} its "
          f"{st32.seconds*1e3:.1f} ms (x{st64.seconds/st32.seconds:.2f}), |x32-x64|/|x64| {rel_err(x32, x64):.2e}")
    assert st64.converged and st64.rel_true <= 1e-8
    assert st32.converged and st32.rel_true <= 1e-8
    assert rel_err(x32, x64) <= 1e-6


def test_config5_learned_4m(dev):
    """4M points in 8 Gaussian clusters, depth 13, log-normal kappa with contrast 1e3:
    learned damping (trained on 200 solver-generated sample batches, seeds 5000..5199)
    against plain Jacobi."""
    damg, levels = _build(da.PTS_CLUSTERS, 4000000, 13, 8, 4, 51)
    F = levels[-1]
    z = np.random.default_rng(52).standard_normal(F.n_elem)
    kf = np.clip(np.exp(np.log(10 ** 1.5) / 2.0 * z), 10 ** -1.5, 10 ** 1.5)
    ms, ts = _device_hierarchy(damg, levels, kf)
    x = np.random.default_rng(11).uniform(-1, 1, F.n_node)
    xt = torch.tensor(x, device=dev)
    yt = torch.empty_like(xt)
    ms[-1].matvec(xt, yt)
    y_ref = oracle.matvec(F, x, kappa=kf)
    y = yt.cpu().numpy()
    e = rel_err(y, y_ref)
    # per node, against the scale of the terms its row sums: for these zero-row-sum rows
    # sum_j |A_nj| = 2 A_nn, and |x| <= 1.  Every term is rounded to 2^-50 of its tile's bound,
    # so a low-kappa row in a tile with kappa 1e3 times larger keeps about 2^-50 * 1e3 ~ 1e-12
    # of its own scale: bounded here at 1e-11.
    row_scale = 2.0 * np.abs(oracle.diagonal(F, kappa=kf))
    worst = float(np.max(np.abs(y - y_ref) / row_scale))
    print(f"config 5 MatVec vs oracle: {e:.2e}; worst node error / (sum_j |A_nj|) {worst:.2e}")
    assert e <= 1e-12
    assert worst <= 1e-11
    del xt, yt
    b = np.random.default_rng(22).uniform(-1, 1, F.n_node)
    s, stj, xj = _solve(ms, ts, b, dev, null_space=1, vcycle_prec=da.F32)
    s.learned_init(1)
    tr = s.learned_train(steps=200, samples=8, nu=2, refresh=1, lr=3e-3, seed=5000)
    mask = s.learned_apply(True, samples=8, seed=6000)
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    x = torch.zeros_like(bt)
    s.solve(bt, x)
    stl, _ = s.solve(bt, x)
    print(f"config 5: E={F.n_elem} N={F.n_node}; Jacobi {stj.iterations} its {stj.seconds*1e3:.1f} ms; learned "
          f"(levels {mask:#x}, training {tr.seconds:.1f}s, objective {tr.loss_first:.4f} -> {tr.loss_last:.4f}) "
          f"{stl.iterations} its {stl.seconds*1e3:.1f} ms")
    assert stj.converged and stl.converged and stl.rel_true <= 1e-8
    assert tr.loss_last < tr.loss_first
    del s
    # Jacobian3 below the finest level (the next finer level's coefficients; Jacobi smoothing)
    _, st3, _ = _solve(ms, ts, b, dev, null_space=1, vcycle_prec=da.F32, coarse_op=da.COARSE_JAC3)
    print(f"config 5: Jacobian3 levels, Jacobi {st3.iterations} its {st3.seconds*1e3:.1f} ms")
    assert st3.converged and st3.rel_true <= 1e-8
