"""Pins the CPU oracle against independent mathematics on tiny meshes.

- The hanging-node operator equals the constrained trilinear FEM: assemble the textbook Q1
  element matrices (Kronecker products of 1D matrices) over all element vertices, then
  eliminate every hanging vertex by interpolation from the vertices of the larger element
  whose edge or face it lies on (resolved recursively).  This fixes the stencil tables,
  the node lists, the hanging masks and the element factors h/2, h^3/8 together.
- The prolongation equals trilinear interpolation of the constrained coarse FE function at
  the fine nodes; restriction is its exact transpose.
- The V-cycle with equal pre/post sweeps is a symmetric operator; PCG converges to the
  dense solution.
PETSc's own arithmetic is not in the reference tree: beyond these checks the solver
iterates are "parity unpinned".
"""
import numpy as np
import pytest

import dendro_amd as da
import oracle
from tests.meshes import hierarchy, rel_err, var_coefficients

K1 = np.array([[0.5, -0.5], [-0.5, 0.5]])
M1 = np.array([[2.0, 1.0], [1.0, 2.0]]) / 3.0
Q1_LAP = np.kron(np.kron(M1, M1), K1) + np.kron(np.kron(M1, K1), M1) + np.kron(np.kron(K1, M1), M1)
Q1_MASS = np.kron(np.kron(M1, M1), M1)


class Constrained:
    """Vertex value = weights over regular nodes, by recursive edge/face interpolation."""

    def __init__(self, h, level):
        L = h.levels[level]
        anc, lev, fl = h.damg.octants(level)
        self.D = L.max_depth
        self.node_id = {}
        for a, f in zip(anc, fl):
            if f & 128:
                self.node_id[tuple(int(v) for v in a)] = len(self.node_id)
        self.elems = [(tuple(int(v) for v in a), 1 << (self.D - int(l))) for a, l in zip(L.anchor, L.lev)]
        self.cache = {}

    def weights(self, q):
        if q in self.node_id:
            return {self.node_id[q]: 1.0}
        if q in self.cache:
            return self.cache[q]
        for (a, s) in self.elems:
            t = [q[d] - a[d] for d in range(3)]
            if not all(0 <= t[d] <= s for d in range(3)):
                continue
            if all(t[d] in (0, s) for d in range(3)):
                continue     # q is a vertex of this element
            mids = [d for d in range(3) if t[d] == s // 2]
            if any(t[d] not in (0, s // 2, s) for d in range(3)) or not mids:
                continue
            corners = [[]]
            for d in range(3):
                opts = [a[d], a[d] + s] if d in mids else [q[d]]
                corners = [c + [o] for c in corners for o in opts]
            w = {}
            for c in corners:
                for n, v in self.weights(tuple(c)).items():
                    w[n] = w.get(n, 0.0) + v / len(corners)
            self.cache[q] = w
            return w
        raise AssertionError(f"cannot resolve vertex {q}")

    def element_rows(self, a, s):
        rows = []
        for k in range(8):
            q = (a[0] + (k & 1) * s, a[1] + ((k >> 1) & 1) * s, a[2] + ((k >> 2) & 1) * s)
            rows.append(self.weights(q))
        return rows


def dense_from_oracle(level, n, **kw):
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = oracle.matvec(level, e, **kw)
    return A


@pytest.fixture(scope="module")
def tiny():
    return hierarchy("gauss", 400, 5, 2, seed=4, min_coarse=1)


def test_operator_is_constrained_q1(tiny):
    h = tiny
    L = h.fine
    cons = Constrained(h, h.n - 1)
    kappa = np.random.default_rng(1).uniform(0.5, 2.0, L.n_elem)
    mass = np.random.default_rng(2).uniform(0.5, 2.0, L.n_elem)
    N = L.n_node
    A = np.zeros((N, N))
    unit = 1 << (L.max_depth - 1)
    for e, (a, s) in enumerate(cons.elems):
        hphys = s / unit
        Ke = kappa[e] * hphys / 2.0 * Q1_LAP + mass[e] * hphys ** 3 / 8.0 * Q1_MASS
        P = np.zeros((8, N))
        for k, w in enumerate(cons.element_rows(a, s)):
            for n, v in w.items():
                P[k, n] += v
        A += P.T @ Ke @ P
    Ao = dense_from_oracle(L, N, kappa=kappa, mass=mass)
    assert np.abs(Ao - A).max() <= 1e-13 * np.abs(A).max()
    assert np.abs(Ao - Ao.T).max() <= 1e-15 * np.abs(Ao).max()


def test_prolongation_is_constrained_interpolation(tiny):
    h = tiny
    C, F = h.levels[0], h.levels[1]
    cc = Constrained(h, 0)
    anc, lev, fl = h.damg.octants(1)
    fnodes = [tuple(int(v) for v in a) for a, f in zip(anc, fl) if f & 128]
    Pb = np.zeros((F.n_node, C.n_node))
    for j, x in enumerate(fnodes):
        for (a, s) in cc.elems:
            if all(a[d] <= x[d] <= a[d] + s for d in range(3)):
                xi = [(x[d] - a[d]) / s for d in range(3)]
                for k, w in enumerate(cc.element_rows(a, s)):
                    wk = 1.0
                    for d in range(3):
                        wk *= xi[d] if (k >> d) & 1 else 1.0 - xi[d]
                    for n, v in w.items():
                        Pb[j, n] += wk * v
                break
    Po = np.zeros_like(Pb)
    for i in range(C.n_node):
        e = np.zeros(C.n_node)
        e[i] = 1.0
        Po[:, i] = oracle.prolong(C, F, e)
    assert np.abs(Po - Pb).max() <= 1e-14
    Ro = np.zeros((C.n_node, F.n_node))
    for j in range(F.n_node):
        e = np.zeros(F.n_node)
        e[j] = 1.0
        Ro[:, j] = oracle.restrict(C, F, e)
    assert np.array_equal(Ro, Po.T)
    # the product's transfer map (fine node -> coarse element, lattice position) gives the
    # same matrix through the closed-form weights
    st = da.stencils()
    ce, pos = h.transfers[0]
    Pp = np.zeros_like(Pb)
    for j in range(F.n_node):
        c = int(ce[j])
        t = da.etype(int(C.child[c]), int(C.hnmask[c]))
        p = int(pos[j])
        tx, ty, tz = p % 3, (p // 3) % 3, p // 9
        cf = (tx >= 1) | ((ty >= 1) << 1) | ((tz >= 1) << 2)
        f = (tx == 2) | ((ty == 2) << 1) | ((tz == 2) << 2)
        for k in range(8):
            Pp[j, C.nlist[c, k]] += st["rmat1"][C.child[c], cf, t, k, f]
    assert np.abs(Pp - Pb).max() <= 1e-14


def test_vcycle_symmetric_and_pcg_matches_dense():
    h = hierarchy("gauss", 3000, 6, 3, seed=5, min_coarse=20)
    coefs = var_coefficients(h, mass_base=1.0)
    mg = oracle.MG([(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)],
                   omega=2.0 / 3.0, nu_pre=2, nu_post=2)
    rng = np.random.default_rng(0)
    u, v = rng.uniform(-1, 1, (2, h.fine.n_node))
    lhs, rhs = np.dot(mg.vcycle(u), v), np.dot(u, mg.vcycle(v))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1e-300)
    # PCG solution vs a dense solve of the assembled operator
    F = h.fine
    b = rng.uniform(-1, 1, F.n_node)
    x, its, conv, hist = mg.pcg(b, rtol=1e-12, max_it=200, norm_type=0)
    assert conv and np.all(np.diff(hist) < 0)
    A = dense_from_oracle(F, F.n_node, kappa=coefs[-1][0], mass=coefs[-1][1])
    assert rel_err(x, np.linalg.solve(A, b)) <= 1e-10


def test_vcycle_null_space_mean_free():
    h = hierarchy("gauss", 3000, 6, 3, seed=5, min_coarse=20)
    coefs = var_coefficients(h, mass_base=None)
    mg = oracle.MG([(h.levels[l], coefs[l][0], None, 1.0, 1.0) for l in range(h.n)],
                   omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=True)
    b = np.random.default_rng(3).uniform(-1, 1, h.fine.n_node)
    x, its, conv, _ = mg.pcg(b, rtol=1e-10, max_it=300, norm_type=0)
    assert conv and abs(x.mean()) <= 1e-12
    r = (b - b.mean()) - mg.apply(h.n - 1, x)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(b - b.mean())
