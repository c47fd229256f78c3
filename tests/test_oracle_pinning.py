"""Pins the CPU oracle against independent mathematics on tiny meshes.

- The hanging-node operator equals the constrained trilinear FEM: assemble the textbook Q1
  element matrices (Kronecker products of 1D matrices) over all element vertices, then
  eliminate every hanging vertex by interpolation from the vertices of the larger element
  whose edge or face it lies on (resolved recursively).  This fixes the stencil tables,
  the node lists, the hanging masks and the element factors h/2, h^3/8 together.
- The prolongation equals trilinear interpolation of the constrained coarse FE function at
  the fine nodes; restriction is its exact transpose.
- The diagonal equals diag(A) of that assembled matrix; the Dirichlet-masked product and
  diagonal equal the masked matrix (boundary rows identity, boundary columns dropped).
- One V-cycle equals a dense textbook multiplicative V-cycle (Briggs, Henson and McCormick,
  "A Multigrid Tutorial", ch. 3; PETSc PCMG with Richardson + Jacobi smoothing, zero initial
  guesses on every level, exact coarsest solve) built from the assembled level matrices
  and the interpolation matrices pinned above, for several omega / sweep counts and with a
  constant null space; a wrong omega or sweep count fails (negative controls).
- The PCG iterate x_k equals the A-norm minimiser of the error over the Krylov space
  K_k(BA, B r_0) of the dense V-cycle B (Saad, "Iterative Methods for Sparse Linear
  Systems", Alg. 9.1 and Prop. 6.11), for k = 1..5, and its stopping-test norms are
  |r_k| or |B r_k| of those iterates; the minimiser of another preconditioner fails.
PETSc's own floating-point order (ddot, LU pivoting) is not in the reference tree; the
iterates are pinned to these definitions up to rounding.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

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
    # the Jacobi diagonal is diag(A) of the constrained assembly
    d = oracle.diagonal(L, kappa=kappa, mass=mass)
    assert np.abs(d - np.diag(A)).max() <= 1e-13 * np.abs(A).max()
    # Dirichlet masking: boundary rows are the identity, boundary columns are dropped
    bdy = h.damg.boundary(h.n - 1).astype(bool)
    assert 0 < bdy.sum() < N
    Ad = np.where(bdy[:, None] | bdy[None, :], 0.0, A)
    Ad[bdy, bdy] = 1.0
    dd = oracle.diagonal_dirichlet(L, bdy, kappa=kappa, mass=mass)
    assert np.abs(dd - np.diag(Ad)).max() <= 1e-13 * np.abs(A).max()
    x = np.random.default_rng(3).uniform(-1, 1, N)
    yd = oracle.matvec_dirichlet(L, bdy, x, kappa=kappa, mass=mass)
    assert np.abs(yd - Ad @ x).max() <= 1e-13 * np.abs(A).max() * np.abs(x).sum() / N * 8
    assert np.array_equal(yd[bdy], x[bdy])


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


# ------------------------------------------------------------------ dense V-cycle and PCG
def _sparse_from_columns(apply, n, m=None):
    rows, cols, vals = [], [], []
    e = np.zeros(n)
    for j in range(n):
        e[:] = 0.0
        e[j] = 1.0
        c = apply(e)
        nz = np.nonzero(c)[0]
        rows.append(nz)
        cols.append(np.full(len(nz), j))
        vals.append(c[nz])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(m or n, n))


class DenseMG:
    """Textbook multiplicative V-cycle on assembled matrices: A_l from the oracle operator
    (pinned to the constrained Q1 assembly), P_l from the oracle prolongation (pinned to
    constrained interpolation), R_l = P_l^T, Jacobi-preconditioned Richardson smoothing
    x <- x + omega D^-1 (b - A x) from a zero guess, an exact coarsest solve (the mean-free
    one with a constant null space, whose right-hand side is projected first)."""

    def __init__(self, h, coefs, null_space, P=None):
        self.h, self.ns = h, null_space
        self.A = []
        for l in range(h.n):
            L = h.levels[l]
            k, m = coefs[l]
            self.A.append(_sparse_from_columns(lambda e, L=L, k=k, m=m: oracle.matvec(L, e, kappa=k, mass=m),
                                               L.n_node))
        self.P = P or []
        for l in range(len(self.P) + 1, h.n):
            C, F = h.levels[l - 1], h.levels[l]
            self.P.append(_sparse_from_columns(lambda e, C=C, F=F: oracle.prolong(C, F, e), C.n_node, F.n_node))
        A0 = self.A[0].toarray()
        self.A0inv = np.linalg.pinv(A0) if null_space else np.linalg.inv(A0)

    def vcycle(self, B, omega, nu_pre, nu_post, l=None):
        top = l is None
        l = self.h.n - 1 if top else l
        if l == 0:
            X = self.A0inv @ ((B - B.mean(0)) if self.ns else B)
        else:
            A = self.A[l]
            dinv = (1.0 / A.diagonal())[:, None]
            X = np.zeros_like(B)
            for _ in range(nu_pre):
                X = X + omega * dinv * (B - A @ X)
            Xc = self.vcycle(self.P[l - 1].T @ (B - A @ X), omega, nu_pre, nu_post, l - 1)
            X = X + self.P[l - 1] @ Xc
            for _ in range(nu_post):
                X = X + omega * dinv * (B - A @ X)
        if top and self.ns:
            X = X - X.mean(0)
        return X

    def krylov_minimisers(self, b, kmax, omega, nu):
        """x_k = argmin |x - x*|_A over K_k(BA, B b) for k = 1..kmax (B: the V-cycle with
        nu + nu sweeps, symmetric), from an orthonormal basis grown one vector at a time."""
        A = self.A[-1]
        if self.ns:
            b = b - b.mean()
        Bv = lambda v: self.vcycle(v[:, None], omega, nu, nu)[:, 0]
        q = Bv(b)
        Q = [q / np.linalg.norm(q)]
        xs = []
        for k in range(1, kmax + 1):
            V = np.stack(Q, 1)
            xs.append(V @ np.linalg.solve(V.T @ (A @ V), V.T @ b))
            w = Bv(A @ Q[-1])
            for _ in range(2):   # Gram-Schmidt, twice
                for qq in Q:
                    w = w - (qq @ w) * qq
            Q.append(w / np.linalg.norm(w))
        return xs


@pytest.fixture(scope="module")
def dense_tiny(hier_tiny):
    h = hier_tiny
    coefs = var_coefficients(h, mass_base=1.0)
    return h, coefs, DenseMG(h, coefs, False)


@pytest.fixture(scope="module")
def dense_tiny_ns(hier_tiny, dense_tiny):
    h = hier_tiny
    coefs = var_coefficients(h, mass_base=None)
    return h, coefs, DenseMG(h, coefs, True, P=dense_tiny[2].P)


def _probe(n, seed):
    """8 basis vectors and 8 random vectors as the columns of one block"""
    rng = np.random.default_rng(seed)
    E = np.zeros((n, 8))
    E[rng.choice(n, 8, replace=False), np.arange(8)] = 1.0
    return np.concatenate([E, rng.uniform(-1, 1, (n, 8))], 1)


def _oracle_vcycle(mg, B):
    return np.stack([mg.vcycle(B[:, j]) for j in range(B.shape[1])], 1)


@pytest.mark.parametrize("ns", [False, True])
@pytest.mark.parametrize("omega,nu_pre,nu_post", [(2.0 / 3.0, 2, 2), (0.6, 1, 3), (0.8, 3, 0)])
def test_vcycle_equals_dense_textbook(dense_tiny, dense_tiny_ns, ns, omega, nu_pre, nu_post):
    h, coefs, D = dense_tiny_ns if ns else dense_tiny
    mg = oracle.MG([(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)],
                   omega=omega, nu_pre=nu_pre, nu_post=nu_post, null_space=ns)
    B = _probe(h.fine.n_node, 1)
    if ns:
        B = B - B.mean(0)
    Zo = _oracle_vcycle(mg, B)
    Zd = D.vcycle(B, omega, nu_pre, nu_post)
    err = np.abs(Zo - Zd).max() / np.abs(Zd).max()
    assert err <= 1e-12, err


@pytest.mark.parametrize("wrong", [dict(omega=0.5), dict(nu_pre=1), dict(nu_post=3)])
def test_vcycle_negative_controls(dense_tiny, wrong):
    """An oracle V-cycle with another omega or sweep count differs from the textbook one
    with omega 2/3 and 2 + 2 sweeps far beyond the tolerance above."""
    h, coefs, D = dense_tiny
    kw = dict(omega=2.0 / 3.0, nu_pre=2, nu_post=2)
    kw.update(wrong)
    mg = oracle.MG([(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)], **kw)
    B = _probe(h.fine.n_node, 2)
    Zd = D.vcycle(B, 2.0 / 3.0, 2, 2)
    assert np.abs(_oracle_vcycle(mg, B) - Zd).max() / np.abs(Zd).max() > 1e-4


@pytest.mark.parametrize("ns", [False, True])
@pytest.mark.parametrize("norm_type", [0, 1])
def test_pcg_iterates_are_krylov_minimisers(dense_tiny, dense_tiny_ns, ns, norm_type):
    h, coefs, D = dense_tiny_ns if ns else dense_tiny
    mg = oracle.MG([(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)],
                   omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=ns)
    b = np.random.default_rng(12).uniform(-1, 1, h.fine.n_node)
    xs = D.krylov_minimisers(b, 5, 2.0 / 3.0, 2)
    bp = b - b.mean() if ns else b
    A = D.A[-1]
    for k in range(1, 6):
        xo, its, _, hist = mg.pcg(b, rtol=1e-30, max_it=k, norm_type=norm_type)
        assert its == k
        assert rel_err(xo, xs[k - 1]) <= 1e-10, (k, rel_err(xo, xs[k - 1]))
        r = bp - A @ xs[k - 1]
        nrm = np.linalg.norm(D.vcycle(r[:, None], 2.0 / 3.0, 2, 2)[:, 0]) if norm_type else np.linalg.norm(r)
        # recurrence residual against the residual of the minimiser: rounding of order
        # eps |A| |x|, stated relative to the initial norm
        assert abs(hist[k] - nrm) <= 1e-10 * hist[0], (k, hist[k], nrm)
    r0 = np.linalg.norm(D.vcycle(bp[:, None], 2.0 / 3.0, 2, 2)[:, 0]) if norm_type else np.linalg.norm(bp)
    assert abs(hist[0] - r0) <= 1e-12 * r0


def test_pcg_negative_control(dense_tiny):
    """The Krylov minimisers of a different preconditioner (omega 0.5) are not the
    iterates of the oracle's PCG with omega 2/3."""
    h, coefs, D = dense_tiny
    mg = oracle.MG([(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)],
                   omega=2.0 / 3.0, nu_pre=2, nu_post=2)
    b = np.random.default_rng(12).uniform(-1, 1, h.fine.n_node)
    xs = D.krylov_minimisers(b, 3, 0.5, 2)
    for k in (2, 3):
        xo, _, _, _ = mg.pcg(b, rtol=1e-30, max_it=k, norm_type=0)
        assert rel_err(xo, xs[k - 1]) > 1e-6


def test_threaded_matvec_equals_serial(hier_small):
    """or_matvec_mt (the all-core CPU baseline) is the serial element loop split over element
    ranges: the same products, rows summed in another order."""
    F = hier_small.fine
    c = var_coefficients(hier_small, mass_base=1.0)[-1]
    x = np.random.default_rng(3).uniform(-1, 1, F.n_node)
    ref = oracle.matvec(F, x, kappa=c[0], mass=c[1])
    for t in (1, 3, 8, 64):
        assert rel_err(oracle.matvec_mt(F, x, kappa=c[0], mass=c[1], threads=t), ref) <= 1e-14, t


@pytest.mark.parametrize("ns", [False, True])
def test_lapack_coarse_factor_equals_oracle_lu(hier_tiny, ns, monkeypatch):
    """A large coarsest level is factored by LAPACK getrf (scipy) instead of the oracle's
    unblocked loop; both factor the same assembled matrix with partial pivoting, so the
    V-cycles agree to rounding."""
    h = hier_tiny
    coefs = var_coefficients(h, mass_base=None if ns else 1.0)
    lv = [(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)]
    r = np.random.default_rng(9).uniform(-1, 1, h.fine.n_node)
    if ns:
        r -= r.mean()
    z_c = oracle.MG(lv, null_space=ns).vcycle(r)
    monkeypatch.setattr(oracle.MG, "LAPACK_ABOVE", 0)
    z_l = oracle.MG(lv, null_space=ns).vcycle(r)
    assert rel_err(z_l, z_c) <= 1e-12


# ------------------------------------------------------------------ Jacobian3
def test_type1_and_shape_function_stencils_match_reference_tables():
    """or_stencils_type1 / or_shfn against the reference's LapType1Stencils.inp,
    MassType1Stencils.inp and ShFnStencils.inp (fixture), and what fixes them: the Type1
    stencils of a coarse element sum over its 8 fine octants to its Type2 stencils, and the
    shape functions are a partition of unity at every Gauss node."""
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "stencils_ref2.npz"))
    s1 = oracle.stencils_type1()
    s2 = oracle.stencils()
    assert np.abs(s1["lap1"] - ref["lap1"]).max() <= 1e-15
    assert np.abs(s1["mass1"] - ref["mass1"]).max() <= 1e-15
    assert np.abs(s1["shfn"] - ref["shfn"]).max() <= 1e-15
    assert np.abs(s1["lap1"].sum(2) - s2["lap2"]).max() <= 1e-14
    assert np.abs(s1["mass1"].sum(2) - s2["mass2"]).max() <= 1e-14
    assert np.abs(s1["shfn"].sum(2) - 1.0).max() <= 1e-14


def test_jacobian3_is_one_level_galerkin(hier_tiny):
    """Jacobian3 on a coarse level with the fine level's element coefficients equals the
    one-level Galerkin product P^T A_f P (A_f: the fine level's Jacobian2 with its own
    coefficients, P: the prolongation), since the coarse space is nested in the fine one.
    With a huge tolerance every refined element takes the mean of its children: the result is
    then Jacobian2 on the coarse level with the coarsened coefficients.  Negative control:
    Jacobian2 with the coarsened coefficients differs from the Galerkin product."""
    h = hier_tiny
    l = h.n - 1
    C, F = h.levels[l - 1], h.levels[l]
    kf = 10.0 ** np.random.default_rng(8).uniform(-1, 1, F.n_elem)
    mf = np.random.default_rng(9).uniform(0.5, 2.0, F.n_elem)
    Af = _sparse_from_columns(lambda e: oracle.matvec(F, e, kappa=kf, mass=mf), F.n_node)
    P = _sparse_from_columns(lambda e: oracle.prolong(C, F, e), C.n_node, F.n_node)
    G = (P.T @ Af @ P).toarray()
    A3 = _sparse_from_columns(lambda e: oracle.matvec_jac3(C, F, e, kf, mf), C.n_node).toarray()
    scale = np.abs(G).max()
    assert np.abs(A3 - G).max() <= 1e-13 * scale
    assert np.abs(A3 - A3.T).max() <= 1e-14 * scale
    kc, mc = h.damg.coarsen_coefficients(kf, mf, l - 1)
    A2c = _sparse_from_columns(lambda e: oracle.matvec(C, e, kappa=kc, mass=mc), C.n_node).toarray()
    assert np.abs(A2c - G).max() > 1e-3 * scale
    x = np.random.default_rng(10).uniform(-1, 1, C.n_node)
    assert rel_err(oracle.matvec_jac3(C, F, x, kf, mf, tol=1e300), A2c @ x) <= 1e-14
