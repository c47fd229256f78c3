"""Learned Jacobi damping on the GPU against the CPU oracle (oracle/learned.c).

Bars (FP32 device arithmetic against the FP64 oracle with the same BF16-rounded weights):
network outputs <= 1e-5 relative, parameter gradients <= 1e-4 relative (2-norm), smoothing
objective <= 1e-5 and its gradient <= 1e-4.  Training lowers the objective, and on fresh
samples the installed weights beat plain damped Jacobi wherever they are kept (elsewhere
the level falls back to Jacobi); the learned-smoother PCG still meets 1e-8.
"""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import hierarchy, rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    da.init(0)
    return torch.device("cuda:0")


def _contrast_coefficients(h, seed=52):
    """log-normal kappa with contrast ~1e3 (clipped to [10^-1.5, 10^1.5]), per level"""
    z = np.random.default_rng(seed).standard_normal(h.fine.n_elem)
    kf = np.clip(np.exp(np.log(10 ** 1.5) / 2.0 * z), 10 ** -1.5, 10 ** 1.5)
    return [h.damg.coarsen_coefficients(kf, None, l)[0] for l in range(h.n)]


@pytest.fixture(scope="module")
def setup(dev):
    h = hierarchy("clusters", 20000, 8, 4, seed=4, min_coarse=100, cluster_seed=51)
    kap = _contrast_coefficients(h)
    ms = []
    for l in range(h.n):
        m = da.Mesh(h.levels[l])
        m.set_coefficients(kap[l], None)
        ms.append(m)
    ts = [da.Transfer(ms[l - 1], ms[l], *h.transfers[l - 1]) for l in range(1, h.n)]
    return h, kap, ms, ts


def _solver(setup, **kw):
    h, kap, ms, ts = setup
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=300, rtol=1e-8,
                          norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=da.F32, use_graph=1)
    for k, v in kw.items():
        setattr(p, k, v)
    return da.Solver(ms, ts, p)


def test_mlp_forward_backward_match_oracle(setup):
    h, kap, ms, ts = setup
    s = _solver(setup)
    l = h.n - 1
    L = h.levels[l]
    theta = np.random.default_rng(7).normal(0, 0.2, da.LEARN_NPARAM).astype(np.float32)
    s.set_learned_params(theta)
    assert np.array_equal(s.learned_params(), theta)
    X = oracle.learn_features(L, kap[l])
    dO = np.random.default_rng(8).normal(0, 1, (L.n_elem, 8)).astype(np.float32)
    O, g = s.learned_eval(l, L.n_elem, dO)
    O_ref = oracle.mlp_forward(theta, X, round_bf16=True)
    g_ref = oracle.mlp_backward(theta, X, dO.astype(np.float64), round_bf16=True)
    eo, eg = rel_err(O, O_ref), rel_err(g, g_ref)
    print(f"MLP forward rel err {eo:.2e}, gradient rel err {eg:.2e}")
    assert eo <= 1e-5 and eg <= 1e-4
    # the forward alone (no dO) gives the same outputs
    assert np.array_equal(s.learned_eval(l, L.n_elem), O)


def test_objective_and_gradient_match_oracle(setup):
    h, kap, ms, ts = setup
    s = _solver(setup)
    for l in (1, h.n - 1):
        L = h.levels[l]
        rng = np.random.default_rng(10 + l)
        om = rng.uniform(0.3, 0.9, L.n_node).astype(np.float32)
        E0 = rng.uniform(-1, 1, (L.n_node, 8)).astype(np.float32)
        loss, g = s.learned_loss(l, L.n_node, omega=om, samples=E0, nu=2)
        loss_ref, g_ref = oracle.learn_loss(L, om.astype(np.float64), E0.astype(np.float64), nu=2, kappa=kap[l])
        print(f"level {l}: objective {loss:.6f} (oracle {loss_ref:.6f}), gradient rel err {rel_err(g, g_ref):.2e}")
        assert abs(loss - loss_ref) <= 1e-5 * loss_ref
        assert rel_err(g, g_ref) <= 1e-4


def test_network_weights_match_oracle_nodal_map(setup):
    """objective with the network's own weights = oracle objective of the oracle's nodal map
    (averaging, cap, sigmoid) applied to the device outputs"""
    h, kap, ms, ts = setup
    s = _solver(setup)
    s.learned_init(123)
    l = h.n - 1
    L = h.levels[l]
    O = s.learned_eval(l, L.n_elem).astype(np.float64)
    cap = oracle.learn_cap(L, kap[l])
    _, om = oracle.learn_nodal(L, O, cap)
    E0 = np.random.default_rng(3).uniform(-1, 1, (L.n_node, 8)).astype(np.float32)
    loss = s.learned_loss(l, L.n_node, omega=None, samples=E0, nu=2, grad=False)
    loss_ref = oracle.learn_loss(L, om, E0.astype(np.float64), nu=2, kappa=kap[l], grad=False)
    assert abs(loss - loss_ref) <= 2e-5 * loss_ref


def test_training_gain_and_fallback(setup, dev):
    h, kap, ms, ts = setup
    s = _solver(setup)
    N = h.fine.n_node
    b = torch.tensor(np.random.default_rng(22).uniform(-1, 1, N), dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    st_j, _ = s.solve(b, x)
    assert st_j.converged
    # objective of plain Jacobi and of the untrained network on the same fresh samples
    top = h.n - 1
    jac = s.learned_loss(top, N, omega=np.full(N, 2.0 / 3.0, np.float32), S=8, nu=2, seed=777, grad=False)
    s.learned_init(5)
    untrained = s.learned_loss(top, N, omega=None, S=8, nu=2, seed=777, grad=False)
    st = s.learned_train(steps=90, samples=8, nu=2, refresh=3, lr=3e-3, seed=5000)
    trained = s.learned_loss(top, N, omega=None, S=8, nu=2, seed=777, grad=False)
    print(f"objective (2 sweeps, A-norm error reduction) Jacobi {jac:.4f}, untrained {untrained:.4f}, "
          f"trained {trained:.4f}; training {st.loss_first:.4f} -> {st.loss_last:.4f} in {st.seconds:.2f}s")
    assert st.loss_last < st.loss_first
    assert trained < jac
    mask = s.learned_apply(True, samples=8, seed=6000)
    assert mask & (1 << top)
    x.zero_()
    st_l, _ = s.solve(b, x)
    print(f"PCG to 1e-8: Jacobi {st_j.iterations} its {st_j.seconds*1e3:.2f} ms, learned (levels {mask:#x}) "
          f"{st_l.iterations} its {st_l.seconds*1e3:.2f} ms, rel_true {st_l.rel_true:.2e}")
    assert st_l.converged and st_l.rel_true <= 1.05e-8
    # the gain: fewer PCG iterations than plain damped Jacobi on this high-contrast mesh
    # (training and sampling are seeded and every reduction runs in a fixed order)
    assert st_l.iterations < st_j.iterations
    # fallback: disabling restores the Jacobi solve exactly (same iterations)
    s.learned_apply(False)
    x.zero_()
    st_b, _ = s.solve(b, x)
    assert st_b.iterations == st_j.iterations
