"""Iterative coarsest-level solve: Jacobi-preconditioned CG on the device for coarsest levels
above the dense limit (the reference factors the assembled coarsest matrix with PETSc's
sparse LU, bin/options:204-206; the dense Cholesky inverse here stops at 40,000 nodes).

Forced on a small hierarchy (coarse_dense_max = 0) it must give the V-cycle and the PCG of
the dense path, and the oracle's (dense LU), to the coarse tolerance: plain, constant null
space, Jacobian3 levels, Dirichlet, FP32 V-cycle.  On a hierarchy whose coarsest level is
above 40,000 nodes (which the dense path refuses) the solve converges.  The batched V-cycle
keeps the dense inverse and refuses such a solver."""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import hierarchy, rel_err, var_coefficients

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


def _meshes(h, coefs, bdy=False):
    ms = []
    for l in range(h.n):
        m = da.Mesh(h.levels[l])
        m.set_coefficients(coefs[l][0], coefs[l][1])
        if bdy:
            m.set_boundary(h.damg.boundary(l))
        ms.append(m)
    ts = [da.Transfer(ms[l - 1], ms[l], *h.transfers[l - 1]) for l in range(1, h.n)]
    return ms, ts


CASES = {
    "jac2_mass": dict(ns=False, mass=True, op="jac2", bdy=False),
    "null_space": dict(ns=True, mass=False, op="jac2", bdy=False),
    "jacobian3": dict(ns=True, mass=False, op="jac3", bdy=False),
    "dirichlet": dict(ns=False, mass=True, op="jac2", bdy=True),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_iterative_coarse_matches_dense(dev, hier_small, case):
    c = CASES[case]
    h = hier_small
    coefs = var_coefficients(h, mass_base=1.0 if c["mass"] else None)
    ms, ts = _meshes(h, coefs, bdy=c["bdy"])
    op = da.COARSE_JAC3 if c["op"] == "jac3" else da.COARSE_JAC2
    base = dict(omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=int(c["ns"]), coarse_op=op)
    dense = da.Solver(ms, ts, da.default_params(**base))
    itv = da.Solver(ms, ts, da.default_params(coarse_dense_max=0, coarse_rtol=1e-13, **base))
    assert dense.coarse_info()[0] == 0 and itv.coarse_info()[0] == 1
    b = np.random.default_rng(61).uniform(-1, 1, h.fine.n_node)
    r = b - b.mean() if c["ns"] else b
    rt = torch.tensor(r, dtype=torch.float64, device=dev)
    zd, zi = torch.empty_like(rt), torch.empty_like(rt)
    dense.vcycle(rt, zd)
    itv.vcycle(rt, zi)
    its = itv.coarse_info()[1]
    e_v = rel_err(_np(zi), _np(zd))
    print(f"{case}: coarsest {h.levels[0].n_node} nodes, {its} CG iterations; V-cycle vs dense {e_v:.2e}")
    assert its > 0
    assert e_v <= 1e-10
    if not c["bdy"] and c["op"] == "jac2":   # the oracle's V-cycle (dense LU) too
        lv = [(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)]
        mg = oracle.MG(lv, omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=c["ns"])
        assert rel_err(_np(zi), mg.vcycle(r)) <= 1e-10
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    for prec in (da.F64, da.F32):
        p = dict(base, max_it=200, rtol=1e-8, norm_type=da.NORM_UNPRECONDITIONED, vcycle_prec=prec, use_graph=1)
        xd, xi = torch.zeros_like(bt), torch.zeros_like(bt)
        sd, _ = da.Solver(ms, ts, da.default_params(**p)).solve(bt, xd)
        si, _ = da.Solver(ms, ts, da.default_params(coarse_dense_max=0, coarse_rtol=1e-13, **p)).solve(bt, xi)
        assert sd.converged and si.converged and si.rel_true <= 1e-8
        assert abs(si.iterations - sd.iterations) <= (0 if prec == da.F64 else 1)
        e_x = rel_err(_np(xi), _np(xd))
        print(f"{case} {'FP64' if prec == da.F64 else 'FP32'} V-cycle: PCG {si.iterations} its "
              f"(dense {sd.iterations}), solutions within {e_x:.2e}")
        assert e_x <= (1e-9 if prec == da.F64 else 1e-6)


def test_coarsest_above_dense_limit(dev):
    """Two levels from 60k Gaussian points, depth 8: the coarsest level is above the dense
    limit, so the default solver takes the iterative coarse solve and PCG converges."""
    h = hierarchy("gauss", 60000, 8, 2, seed=3, min_coarse=50)
    nc = h.levels[0].n_node
    assert nc > 40000, nc
    coefs = var_coefficients(h, mass_base=None)
    ms, ts = _meshes(h, coefs)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=1, max_it=100, rtol=1e-8,
                          norm_type=da.NORM_UNPRECONDITIONED, coarse_rtol=1e-10, use_graph=1)
    s = da.Solver(ms, ts, p)
    assert s.coarse_info()[0] == 1
    b = np.random.default_rng(62).uniform(-1, 1, h.fine.n_node)
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    x = torch.zeros_like(bt)
    st, _ = s.solve(bt, x)
    print(f"coarsest {nc} nodes (fine {h.fine.n_node}): PCG {st.iterations} its, "
          f"{s.coarse_info()[1]} coarse CG iterations in the last coarse solve, {st.seconds * 1e3:.1f} ms")
    assert st.converged and st.rel_true <= 1e-8
    # the batched V-cycle needs the dense inverse
    k = 8
    B = torch.zeros((h.fine.n_node, k), dtype=torch.float64, device=dev)
    with pytest.raises(da.DendroError):
        s.solve_multi(B, torch.zeros_like(B), k)


def test_coarse_parameters_checked(dev, hier_tiny):
    h = hier_tiny
    coefs = var_coefficients(h, mass_base=1.0)
    ms, ts = _meshes(h, coefs)
    for bad in (dict(coarse_rtol=0.0), dict(coarse_max_it=0), dict(coarse_dense_max=-1)):
        with pytest.raises(da.DendroError):
            da.Solver(ms, ts, da.default_params(**bad))
