"""Jacobian3 (the coarse operator formed from the next finer level's element coefficients) on
the GPU against the oracle's element-by-element restatement of Jacobian3MatMult
(examples/src/backend/omgJac.C:2139-2339): the product itself (<= 1e-12), and the V-cycle and
PCG iterates of a solver whose levels below the finest use it (-jacType 3,
examples/src/drivers/checkError.C:658-667), FP64 <= 1e-12 / 1e-11, FP32 stated."""
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


@pytest.mark.parametrize("mass", [True, False])
def test_matvec_jac3(dev, hier_small, mass):
    h = hier_small
    coefs = var_coefficients(h, mass_base=1.0 if mass else None)
    ms, ts = _meshes(h, coefs)
    for l in range(1, h.n):
        C, F = h.levels[l - 1], h.levels[l]
        x = np.random.default_rng(30 + l).uniform(-1, 1, C.n_node)
        ref = oracle.matvec_jac3(C, F, x, coefs[l][0], coefs[l][1])
        xt = torch.tensor(x, dtype=torch.float64, device=dev)
        yt = torch.empty_like(xt)
        ts[l - 1].matvec_jac3(xt, yt)
        assert rel_err(_np(yt), ref) <= 1e-12, l
        x32 = xt.float()
        y32 = torch.empty_like(x32)
        ts[l - 1].matvec_jac3(x32, y32, prec=da.F32)
        e32 = rel_err(_np(y32), ref)
        print(f"level {l - 1}: FP32 Jacobian3 vs FP64 oracle {e32:.2e}")
        assert e32 <= 1e-5
        # Jacobian3 differs from the level's own Jacobian2 (coarsened coefficients)
        assert rel_err(ref, oracle.matvec(C, x, kappa=coefs[l - 1][0], mass=coefs[l - 1][1])) > 1e-3


@pytest.mark.parametrize("ns", [False, True])
def test_vcycle_and_pcg_jac3(dev, hier_small, ns):
    h = hier_small
    coefs = var_coefficients(h, mass_base=None if ns else 1.0)
    ms, ts = _meshes(h, coefs)
    lv = [(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)]
    mg = oracle.MG(lv, omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=ns, jac3=True)
    b = np.random.default_rng(44).uniform(-1, 1, h.fine.n_node)
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    r = b - b.mean() if ns else b
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=int(ns), coarse_op=da.COARSE_JAC3)
    s = da.Solver(ms, ts, p)
    z = torch.empty_like(bt)
    s.vcycle(torch.tensor(r, device=dev), z)
    assert rel_err(_np(z), mg.vcycle(r)) <= 1e-11
    for k in (1, 3, 5):
        xo, its, _, ho = mg.pcg(b, rtol=1e-30, max_it=k, norm_type=da.NORM_UNPRECONDITIONED)
        for prec in (da.F64, da.F32):
            p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=k, rtol=1e-30,
                                  norm_type=da.NORM_UNPRECONDITIONED, null_space=int(ns), vcycle_prec=prec,
                                  coarse_op=da.COARSE_JAC3)
            x = torch.zeros_like(bt)
            st, hist = da.Solver(ms, ts, p).solve(bt, x, history_len=k + 1)
            err = rel_err(_np(x), xo)
            assert st.iterations == its == k
            if prec == da.F64:
                assert err <= 1e-12, (k, err)
                assert np.allclose(hist, ho, rtol=1e-10, atol=0)
            else:
                print(f"k={k}: FP32 V-cycle (Jacobian3) iterate vs FP64 oracle {err:.2e}")
                assert err <= 1e-5
    # both coarse operators converge; iteration counts stated
    out = {}
    for op in (da.COARSE_JAC2, da.COARSE_JAC3):
        p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=200, rtol=1e-8,
                              norm_type=da.NORM_UNPRECONDITIONED, null_space=int(ns), use_graph=1, coarse_op=op)
        x = torch.zeros_like(bt)
        st, _ = da.Solver(ms, ts, p).solve(bt, x)
        assert st.converged and st.rel_true <= 1e-8
        out[op] = st.iterations
    print(f"PCG to 1e-8: Jacobian2 levels {out[da.COARSE_JAC2]} its, Jacobian3 levels {out[da.COARSE_JAC3]} its")


def test_jac3_rejected_where_unsupported(dev, hier_small):
    h = hier_small
    ms, ts = _meshes(h, var_coefficients(h, mass_base=1.0))
    with pytest.raises(da.DendroError):
        da.Solver(ms, ts, da.default_params(coarse_op=da.COARSE_JAC3, smoother=da.SMOOTHER_LEARNED))
    with pytest.raises(da.DendroError):
        da.Solver(ms, ts, da.default_params(coarse_op=7))
    s = da.Solver(ms, ts, da.default_params(coarse_op=da.COARSE_JAC3))
    B = torch.zeros(h.fine.n_node, 8, dtype=torch.float64, device=dev)
    X = torch.zeros_like(B)
    with pytest.raises(da.DendroError):
        s.solve_multi(B, X, 8)
