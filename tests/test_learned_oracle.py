"""Pins the CPU oracle of the learned Jacobi damping (oracle/learned.c): network gradients,
the nodal averaging and the smoothing objective against central finite differences, and
closed forms (zero network, zero damping, the damping cap on a uniform grid)."""
import numpy as np
import pytest

import oracle
from tests.meshes import hierarchy, var_coefficients


@pytest.fixture(scope="module")
def small():
    h = hierarchy("gauss", 400, 5, 2, seed=4, min_coarse=1)
    c = var_coefficients(h)
    return h.fine, c[-1][0]


def _theta(seed, scale=0.3):
    return np.random.default_rng(seed).normal(0, scale, oracle.NPARAM).astype(np.float32)


def test_mlp_backward_matches_finite_differences(small):
    L, kap = small
    X = oracle.learn_features(L, kap)[:64]
    th = _theta(1)
    dO = np.random.default_rng(2).normal(0, 1, (X.shape[0], oracle.NO))
    g = oracle.mlp_backward(th, X, dO, round_bf16=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(oracle.NPARAM, 80, replace=False), [oracle.NPARAM - 1, 0]])

    def fd(p, h):
        tp, tm = th.copy(), th.copy()
        tp[p] += np.float32(h)
        tm[p] -= np.float32(h)
        step = float(tp[p]) - float(tm[p])
        return (np.sum(dO * oracle.mlp_forward(tp, X, False)) - np.sum(dO * oracle.mlp_forward(tm, X, False))) / step

    checked = 0
    for p in idx:
        a, b = fd(p, 2e-4), fd(p, 1e-4)
        if abs(a - b) > 1e-7 * max(1.0, abs(a)):
            continue        # a ReLU kink inside the stencil: the difference quotient is not a derivative
        checked += 1
        assert abs(b - g[p]) <= 1e-6 * max(1.0, abs(g[p])), (p, b, g[p])
    assert checked >= 0.9 * len(idx)


def test_zero_network_gives_constant_logits(small):
    L, kap = small
    X = oracle.learn_features(L, kap)
    th = np.zeros(oracle.NPARAM, np.float32)
    b3 = np.arange(oracle.NO, dtype=np.float32) * 0.25 - 1.0
    th[-oracle.NO:] = b3
    O = oracle.mlp_forward(th, X)
    assert np.array_equal(O, np.tile(b3.astype(np.float64), (L.n_elem, 1)))
    cap = np.full(L.n_node, 0.9)
    theta_n, om = oracle.learn_nodal(L, O, cap)
    # theta_n = mean over (e, k) at node n of b3[k]
    s = np.zeros(L.n_node)
    c = np.zeros(L.n_node)
    np.add.at(s, L.nlist.ravel(), np.tile(b3, L.n_elem))
    np.add.at(c, L.nlist.ravel(), 1)
    assert np.allclose(theta_n, s / c, rtol=0, atol=1e-15)
    assert np.allclose(om, 0.9 / (1 + np.exp(-s / c)), rtol=1e-15)


def test_nodal_grad_matches_finite_differences(small):
    L, kap = small
    rng = np.random.default_rng(5)
    O = rng.normal(0, 1, (L.n_elem, oracle.NO))
    cap = rng.uniform(0.5, 1.0, L.n_node)
    gw = rng.normal(0, 1, L.n_node)
    th, _ = oracle.learn_nodal(L, O, cap)
    dO = oracle.learn_nodal_grad(L, th, cap, gw)
    for (e, k) in [(0, 0), (7, 3), (L.n_elem - 1, 7), (100, 5)]:
        h = 1e-6
        Op, Om = O.copy(), O.copy()
        Op[e, k] += h
        Om[e, k] -= h
        fd = (gw @ oracle.learn_nodal(L, Op, cap)[1] - gw @ oracle.learn_nodal(L, Om, cap)[1]) / (2 * h)
        assert abs(fd - dO[e, k]) <= 1e-7 * max(1.0, abs(fd))


def test_loss_gradient_matches_finite_differences(small):
    L, kap = small
    rng = np.random.default_rng(6)
    E0 = rng.uniform(-1, 1, (L.n_node, 2))
    om = rng.uniform(0.4, 0.9, L.n_node)
    loss, g = oracle.learn_loss(L, om, E0, nu=2, kappa=kap)
    assert 0 < loss < 1
    for n in [0, 5, L.n_node // 2, L.n_node - 1]:
        h = 1e-6
        op, omm = om.copy(), om.copy()
        op[n] += h
        omm[n] -= h
        fd = (oracle.learn_loss(L, op, E0, 2, kap, grad=False) - oracle.learn_loss(L, omm, E0, 2, kap, grad=False)) / (2 * h)
        assert abs(fd - g[n]) <= 1e-6 * max(abs(fd), 1e-8), (n, fd, g[n])
    # no damping: the error is untouched
    assert oracle.learn_loss(L, np.zeros(L.n_node), E0, 2, kap, grad=False) == pytest.approx(1.0, abs=1e-14)


def test_cap_is_one_on_uniform_laplacian():
    h = hierarchy("uniform", 60, 3, 1, seed=4, min_coarse=1)
    L = h.fine
    if L.n_hanging == 0:
        cap = oracle.learn_cap(L, None, None, 1.0, 0.0)
        assert np.allclose(cap, 1.0, atol=1e-13)
    # in general the cap makes 2 diag(A)/cap - A diagonally dominant (checked on the
    # assembled matrix): every damped sweep with weights in (0, cap] is an A-norm contraction
    h2 = hierarchy("gauss", 400, 5, 2, seed=4, min_coarse=1)
    F = h2.fine
    for mf in (0.0, 1.0):
        cap2 = oracle.learn_cap(F, None, None, 1.0, mf)
        A = np.stack([oracle.matvec(F, np.eye(F.n_node)[j], None, None, 1.0, mf) for j in range(F.n_node)], 1)
        lhs = 2 * np.diag(A) / cap2 - np.diag(A)
        off = np.abs(A).sum(1) - np.abs(np.diag(A))
        assert np.all(lhs >= off * (1 - 1e-12)) and cap2.min() > 0.3
