"""Batched right-hand sides on the GPU: the k-column operator against the FP64 oracle column
by column, and the k-column CG (batched FP32 V-cycle, FP64 outer CG) against single solves.

Bars: FP64 batched MatVec <= 1e-12 relative per column; FP32 batched MatVec (matrix-core
element products) error stated and <= 1e-5; every column of the batched solve meets its own
1e-8 true relative residual, and its solution agrees with the single right-hand-side solve
of that column to the accuracy both solves guarantee.
"""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import rel_err, var_coefficients

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    da.init(0)
    return torch.device("cuda:0")


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def _meshes(h, coefs):
    ms = []
    for l in range(h.n):
        m = da.Mesh(h.levels[l])
        m.set_coefficients(coefs[l][0], coefs[l][1])
        ms.append(m)
    ts = [da.Transfer(ms[l - 1], ms[l], *h.transfers[l - 1]) for l in range(1, h.n)]
    return ms, ts


@pytest.mark.parametrize("k,tile,mass,variant", [(8, 256, True, 0), (16, 512, False, 0), (32, 1024, True, 0),
                                                  (8, 512, True, 1), (32, 1024, False, 1), (16, 512, True, 2),
                                                  (32, 512, False, 3), (32, 512, True, 4), (16, 256, False, 4),
                                                  (32, 512, True, 5), (8, 1024, False, 5), (32, 512, False, 6),
                                                  (8, 256, True, 6)])
def test_matvec_multi_matches_oracle(dev, hier_small, k, tile, mass, variant):
    h = hier_small
    F = h.fine
    c = var_coefficients(h, mass_base=1.0 if mass else None)
    m = da.Mesh(F)
    m.set_coefficients(c[-1][0], c[-1][1])
    m.set_tile_multi(tile, variant)
    X = np.random.default_rng(100 + k).uniform(-1, 1, (F.n_node, k))
    Xt = torch.tensor(X, dtype=torch.float64, device=dev)
    Yt = torch.empty_like(Xt)
    m.matvec_multi(Xt, Yt, k, prec=da.F64)
    Y = _np(Yt)
    X32 = Xt.float()
    Y32 = torch.empty_like(X32)
    m.matvec_multi(X32, Y32, k, prec=da.F32)
    Y32 = _np(Y32)
    worst32 = 0.0
    for j in (0, k // 2, k - 1):
        ref = oracle.matvec(F, X[:, j], kappa=c[-1][0], mass=c[-1][1])
        assert rel_err(Y[:, j], ref) <= 1e-12, j
        worst32 = max(worst32, rel_err(Y32[:, j], ref))
    print(f"k={k} tile={tile} variant={variant}: FP32 batched MatVec relative error vs FP64 oracle {worst32:.3e}")
    assert worst32 <= 1e-5


@pytest.mark.parametrize("prec,variant", [(da.F64, 0), (da.F32, 0), (da.F32, 4), (da.F32, 5), (da.F32, 6)])
def test_matvec_multi_column_scales(dev, hier_small, prec, variant):
    """The vector-ALU batched operator scales its fixed-point sums per column: columns of
    very different magnitude in one slice keep their own precision (each column equals the
    single-column product bit for bit up to its power-of-two scale), and a NaN in one column
    leaves the other columns finite."""
    h = hier_small
    F = h.fine
    c = var_coefficients(h, mass_base=1.0)
    m = da.Mesh(F)
    m.set_coefficients(c[-1][0], c[-1][1])
    m.set_tile_multi(512, variant)
    k = 8
    dt = torch.float64 if prec == da.F64 else torch.float32
    base = np.random.default_rng(5).uniform(-1, 1, F.n_node)
    s = 200 if prec == da.F64 else 40
    scales = [2.0 ** (s * ((j % 3) - 1)) for j in range(k)]   # 2^-s, 1, 2^s, ...
    X = torch.tensor(np.stack([base * sc for sc in scales], axis=1), dtype=dt, device=dev).contiguous()
    Y = torch.empty_like(X)
    m.matvec_multi(X, Y, k, prec=prec)
    torch.cuda.synchronize()
    for j in range(1, k):   # same column up to the power of two: bit-identical after rescaling
        assert torch.equal(Y[:, j] / scales[j], Y[:, 0] / scales[0]), j
    ref = oracle.matvec(F, base, kappa=c[-1][0], mass=c[-1][1])
    assert rel_err(_np(Y[:, 1]), ref) <= (1e-12 if prec == da.F64 else 1e-5)
    X[F.n_node // 3, 2] = float("nan")
    m.matvec_multi(X, Y, k, prec=prec)
    torch.cuda.synchronize()
    assert torch.isnan(Y[F.n_node // 3, 2])
    for j in (0, 1, 3, 7):
        assert torch.isfinite(Y[:, j]).all(), j


def test_matvec_multi_rejects_bad_k(dev, hier_tiny):
    m = da.Mesh(hier_tiny.fine)
    x = torch.zeros(hier_tiny.fine.n_node * 12, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    with pytest.raises(da.DendroError):
        m.matvec_multi(x, y, 12)


def test_matvec_multi_rejects_misaligned(dev, hier_tiny):
    """The batched kernels move 16-byte vectors of a node's row: X, Y (and B) must be 16-byte
    aligned, checked on the host before any launch."""
    m = da.Mesh(hier_tiny.fine)
    n, k = hier_tiny.fine.n_node, 8
    base = torch.zeros(n * k + 4, dtype=torch.float32, device=dev)
    y = torch.empty(n * k, dtype=torch.float32, device=dev)
    with pytest.raises(da.DendroError):
        m.matvec_multi(base[1:1 + n * k], y, k, prec=da.F32)   # 4 bytes past an aligned start
    m.matvec_multi(base[4:4 + n * k], y, k, prec=da.F32)       # 16 bytes past: accepted
    torch.cuda.synchronize()


@pytest.mark.parametrize("graph", [0, 1])
def test_solve_multi_matches_single(dev, hier_small, graph):
    h = hier_small
    coefs = var_coefficients(h, mass_base=None)
    ms, ts = _meshes(h, coefs)
    for i, m in enumerate(ms[1:]):
        m.set_tile_multi(512, (0, 1, 5, 6)[i % 4])
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=200, rtol=1e-8,
                          norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=da.F32,
                          use_graph=graph)
    s = da.Solver(ms, ts, p)
    N, k = h.fine.n_node, 8
    B = np.random.default_rng(400).uniform(-1, 1, (N, k))
    Bt = torch.tensor(B, dtype=torch.float64, device=dev)
    Xt = torch.zeros_like(Bt)
    stats = s.solve_multi(Bt, Xt, k)
    X = _np(Xt)
    for j in range(k):
        assert stats[j].converged and stats[j].rel_true <= 1.05e-8, (j, stats[j].rel_true)
    # the same columns one at a time through the single right-hand-side solver
    for j in (0, k - 1):
        bj = torch.tensor(B[:, j], dtype=torch.float64, device=dev)
        xj = torch.zeros_like(bj)
        st, _ = s.solve(bj, xj)
        assert st.converged
        xs = _np(xj)
        assert abs(stats[j].iterations - st.iterations) <= 2
        # both are 1e-8 solutions of the same (mean-free) system
        assert rel_err(X[:, j] - X[:, j].mean(), xs - xs.mean()) <= 1e-5
    # a second call reuses the buffers (and the captured graph); LDS additions may run in a
    # different order, so the iterates agree to rounding
    Xt2 = torch.zeros_like(Bt)
    stats2 = s.solve_multi(Bt, Xt2, k)
    for a, b in zip(stats, stats2):
        assert abs(a.iterations - b.iterations) <= 1 and b.rel_true <= 1.05e-8
    assert rel_err(_np(Xt2), X) <= 1e-6
