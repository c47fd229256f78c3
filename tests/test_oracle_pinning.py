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


class DenseMGJac3(DenseMG):
    """DenseMG with the Jacobian3 level operators of -jacType 3: below the finest level the
    residuals and sweeps use J_l = P_l^T A_{l+1} P_l (the next finer level's Jacobian2 with its
    own coefficients), the Jacobi diagonals stay diag(A_l), and the coarsest solve uses J_0."""

    def __init__(self, h, coefs, null_space):
        super().__init__(h, coefs, null_space)
        self.J = [(self.P[l].T @ self.A[l + 1] @ self.P[l]).tocsr() for l in range(h.n - 1)] + [self.A[-1]]
        J0 = self.J[0].toarray()
        self.A0inv = np.linalg.pinv(J0) if null_space else np.linalg.inv(J0)

    def vcycle(self, B, omega, nu_pre, nu_post, l=None):
        top = l is None
        l = self.h.n - 1 if top else l
        if l == 0:
            X = self.A0inv @ ((B - B.mean(0)) if self.ns else B)
        else:
            J = self.J[l]
            dinv = (1.0 / self.A[l].diagonal())[:, None]
            X = np.zeros_like(B)
            for _ in range(nu_pre):
                X = X + omega * dinv * (B - J @ X)
            Xc = self.vcycle(self.P[l - 1].T @ (B - J @ X), omega, nu_pre, nu_post, l - 1)
            X = X + self.P[l - 1] @ Xc
            for _ in range(nu_post):
                X = X + omega * dinv * (B - J @ X)
        if top and self.ns:
            X = X - X.mean(0)
        return X


@pytest.mark.parametrize("ns", [False, True])
def test_vcycle_jacobian3_equals_dense_textbook(hier_tiny, ns):
    """The oracle's V-cycle with Jacobian3 level operators (including the element-by-element
    coarsest Type3 assembly) against the dense V-cycle on Galerkin products; negative control:
    the Jacobian2 V-cycle differs."""
    h = hier_tiny
    coefs = var_coefficients(h, mass_base=None if ns else 1.0)
    lv = [(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)]
    d3 = DenseMGJac3(h, coefs, ns)
    B = _probe(h.fine.n_node, 5)
    if ns:
        B = B - B.mean(0)
    mg3 = oracle.MG(lv, omega=0.6, nu_pre=1, nu_post=3, null_space=ns, jac3=True)
    Z3 = _oracle_vcycle(mg3, B)
    Zd = d3.vcycle(B, 0.6, 1, 3)
    assert np.abs(Z3 - Zd).max() <= 1e-11 * np.abs(Zd).max()
    Z2 = _oracle_vcycle(oracle.MG(lv, omega=0.6, nu_pre=1, nu_post=3, null_space=ns), B)
    assert np.abs(Z2 - Zd).max() > 1e-4 * np.abs(Zd).max()


# ------------------------------------------------------------------ right-hand sides and errors
POLY = [0.3, 1.0, -2.0, 0.5, 1.5, -0.7, 0.25, 2.0]


def test_function_special_values():
    """The Gaussians of ComputeRHS2 are 1 at their centres, cut at their boxes (the first box
    open at 0.2, the others closed), and the cosine sum is the closed form at chosen points."""
    G = 0
    assert oracle.eval_fn(G, 0.1, 0.1, 0.1) == 1.0
    assert oracle.eval_fn(G, 0.3, 0.3, 0.3) == 1.0
    assert oracle.eval_fn(G, 0.6, 0.6, 0.6) == 1.0
    assert oracle.eval_fn(G, 0.2, 0.2, 0.2) == pytest.approx(np.exp(-0.03 / 0.005), rel=1e-15)
    assert oracle.eval_fn(G, 0.25, 0.1, 0.1) == 0.0
    assert oracle.eval_fn(G, 0.45, 0.45, 0.45) == 0.0
    assert oracle.eval_fn(1, 0.0, 0.0, 0.0, nfreq=9, a0=1.0, a1=0.0) == pytest.approx(9.0, rel=1e-15)
    v = sum((2 + 3 * 3 * k * k * np.pi ** 2) * np.cos(k * np.pi * 0.3) * np.cos(k * np.pi * 0.7) *
            np.cos(k * np.pi * 0.1) for k in range(1, 5))
    assert oracle.eval_fn(1, 0.3, 0.7, 0.1, nfreq=4, a0=2.0, a1=3.0) == pytest.approx(v, rel=1e-13)


def test_quadrature_rhs_and_l2_error_exact_on_trilinear(tiny):
    """On an adaptive mesh with hanging vertices: the quadrature right-hand side of a trilinear
    g is M g (g lies in the finite element space and 3-point Gauss is exact for the products),
    so f = 1 gives the mass matrix's row sums; nodal values of g are g at the node
    coordinates; the L2 error of g's own nodal values is zero and that of u = 0 is
    sqrt(g^T M g)."""
    from tests.meshes import node_coords
    L = tiny.fine
    assert L.n_hanging > 0
    xyz = node_coords(L)
    g = oracle.nodal_values(L, 2, poly=POLY)
    gx = (POLY[0] + POLY[1] * xyz[:, 0] + POLY[2] * xyz[:, 1] + POLY[3] * xyz[:, 2] + POLY[4] * xyz[:, 0] * xyz[:, 1]
          + POLY[5] * xyz[:, 1] * xyz[:, 2] + POLY[6] * xyz[:, 2] * xyz[:, 0] + POLY[7] * np.prod(xyz, 1))
    assert np.abs(g - gx).max() <= 1e-14
    Mg = oracle.matvec(L, g, lap_fac=0.0, mass_fac=1.0)
    b = oracle.rhs_quadrature(L, 2, poly=POLY)
    assert np.abs(b - Mg).max() <= 1e-14 * np.abs(Mg).max()
    one = [1.0] + [0.0] * 7
    b1 = oracle.rhs_quadrature(L, 2, poly=one)
    M1 = oracle.matvec(L, np.ones(L.n_node), lap_fac=0.0, mass_fac=1.0)
    assert np.abs(b1 - M1).max() <= 1e-15 and abs(b1.sum() - 1.0) <= 1e-14
    assert oracle.l2_error(L, g, 2, poly=POLY) <= 1e-14
    assert oracle.l2_error(L, np.zeros(L.n_node), 2, poly=POLY) == pytest.approx(np.sqrt(g @ Mg), rel=1e-13)


def test_cosine_sum_norm_and_convergence():
    """||sum_{k<=3} cos(k pi x) cos(k pi y) cos(k pi z)||_L2 = sqrt(3/8) (orthogonal terms), to
    quadrature accuracy on a uniform mesh; the Helmholtz solve with the quadrature right-hand
    side of (1 + 3 k^2 pi^2) cos cos cos converges to it at second order in h."""
    errs = []
    for d in (3, 4, 5):
        n = 1 << d
        c = (np.arange(n) + 0.5) / n
        X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
        damg = da.DAMG(da.Octree.from_points(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).ravel(), d), 3, 20)
        Ls = [damg.level(l) for l in range(damg.nlevels)]
        F = Ls[-1]
        assert F.n_elem == n ** 3
        if d == 5:
            assert oracle.l2_error(F, np.zeros(F.n_node), 1, nfreq=3) == pytest.approx(np.sqrt(3 / 8), rel=1e-6)
        b = oracle.rhs_quadrature(F, 1, nfreq=2, a0=1.0, a1=1.0)
        x, its, conv, _ = oracle.MG([(L, None, None, 1.0, 1.0) for L in Ls]).pcg(b, rtol=1e-12, max_it=100)
        assert conv
        errs.append(oracle.l2_error(F, x, 1, nfreq=2))
    rates = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(3.8 <= r <= 4.3 for r in rates), rates


class DenseMGDirichlet(DenseMG):
    """DenseMG for the Dirichlet problem: A_l the Dirichlet-masked operators (identity rows and
    columns on boundary nodes), P_l with every entry that touches a boundary node removed."""

    def __init__(self, h, coefs):
        self.h, self.ns = h, False
        self.bdy = [h.damg.boundary(l).astype(bool) for l in range(h.n)]
        self.A = []
        for l in range(h.n):
            L, (k, m), bd = h.levels[l], coefs[l], self.bdy[l]
            self.A.append(_sparse_from_columns(
                lambda e, L=L, k=k, m=m, bd=bd: oracle.matvec_dirichlet(L, bd, e, kappa=k, mass=m), L.n_node))
        self.P = []
        for l in range(1, h.n):
            C, F = h.levels[l - 1], h.levels[l]
            P = _sparse_from_columns(lambda e, C=C, F=F: oracle.prolong(C, F, e), C.n_node, F.n_node).tolil()
            P[np.nonzero(self.bdy[l])[0], :] = 0.0
            P[:, np.nonzero(self.bdy[l - 1])[0]] = 0.0
            self.P.append(P.tocsr())
        self.A0inv = np.linalg.inv(self.A[0].toarray())


def test_vcycle_dirichlet_equals_dense_textbook(hier_tiny):
    """The oracle's Dirichlet V-cycle (masked operators and diagonals, suppressed-DOF transfers,
    identity rows in the coarsest matrix) against the dense textbook V-cycle on the masked
    matrices; negative control: unmasked transfers differ."""
    h = hier_tiny
    coefs = var_coefficients(h, mass_base=1.0)
    lv = [(h.levels[l], coefs[l][0], coefs[l][1], 1.0, 1.0) for l in range(h.n)]
    dd = DenseMGDirichlet(h, coefs)
    B = _probe(h.fine.n_node, 6)
    mg = oracle.MG(lv, omega=0.6, nu_pre=2, nu_post=1, boundary=dd.bdy)
    Z = _oracle_vcycle(mg, B)
    Zd = dd.vcycle(B, 0.6, 2, 1)
    assert np.abs(Z - Zd).max() <= 1e-11 * np.abs(Zd).max()
    P_full = [_sparse_from_columns(lambda e, C=h.levels[l - 1], F=h.levels[l]: oracle.prolong(C, F, e),
                                   h.levels[l - 1].n_node, h.levels[l].n_node) for l in range(1, h.n)]
    dd.P = P_full
    assert np.abs(dd.vcycle(B, 0.6, 2, 1) - Zd).max() > 1e-6 * np.abs(Zd).max()
    # a Dirichlet PCG solve returns x = b on the boundary nodes and satisfies the masked system
    b = np.random.default_rng(3).uniform(-1, 1, h.fine.n_node)
    x, its, conv, _ = mg.pcg(b, rtol=1e-12, max_it=100)
    assert conv
    bd = dd.bdy[-1]
    assert np.abs(x[bd] - b[bd]).max() <= 1e-10
    assert np.linalg.norm(dd.A[-1] @ x - b) <= 1e-10 * np.linalg.norm(b)


# ------------------------------------------------------------------ elasticity
def _q1_graddiv():
    """Regular-element grad-div block [3k+a][3j+b] = int_{[-1,1]^3} d_a phi_k d_b phi_j in closed
    form: the 1-D factors are K1 (derivative-derivative), M1 (value-value) and
    D1 = int phi_p dphi_q, one derivative per direction a / b."""
    D1 = np.array([[-0.5, 0.5], [-0.5, 0.5]])   # int phi_p phi_q' on [-1, 1]
    G = np.zeros((24, 24))
    for a in range(3):
        for b in range(3):
            f = []
            for d in range(3):   # Kronecker order z, y, x (vertex bit 0 = x)
                if d == a and d == b:
                    f.append(K1)
                elif d == a:
                    f.append(D1.T)
                elif d == b:
                    f.append(D1)
                else:
                    f.append(M1)
            blk = np.kron(np.kron(f[2], f[1]), f[0])
            G[a::3, b::3] = blk
    return G


def test_elasticity_stencils_and_operator(tiny):
    """or_stencils_graddiv against the reference's GDType2Stencils.inp (fixture) and the
    closed-form regular element; the elasticity product (mu Laplacian on every component plus
    (mu + lambda) grad-div) against the constrained vector-Q1 assembly on an adaptive mesh;
    translations in its kernel; with Dirichlet flags, identity rows and dropped columns."""
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "stencils_ref2.npz"))
    gd = oracle.stencils_graddiv()
    assert np.abs(gd - ref["gd2"]).max() <= 1e-15
    assert np.abs(gd[0, 0] - _q1_graddiv()).max() <= 1e-15
    L = tiny.fine
    cons = Constrained(tiny, tiny.n - 1)
    N = L.n_node
    mu, lam = 0.7, 1.9
    A = np.zeros((3 * N, 3 * N))
    unit = 1 << (L.max_depth - 1)
    Ke0 = np.kron(Q1_LAP, np.eye(3)) * mu + _q1_graddiv() * (mu + lam)
    for e, (a, s) in enumerate(cons.elems):
        rows = cons.element_rows(a, s)
        nodes = sorted({n for w in rows for n in w})
        P = np.zeros((8, len(nodes)))
        for k, w in enumerate(rows):
            for n, v in w.items():
                P[k, nodes.index(n)] += v
        P3 = np.kron(P, np.eye(3))
        ix = (3 * np.array(nodes)[:, None] + np.arange(3)).ravel()
        A[np.ix_(ix, ix)] += (s / unit) / 2.0 * (P3.T @ Ke0 @ P3)
    Ao = _sparse_from_columns(lambda v: oracle.matvec_elasticity(L, v, mu, lam), 3 * N).toarray()
    assert np.abs(Ao - A).max() <= 1e-13 * np.abs(A).max()
    for d in range(3):
        t = np.zeros((N, 3))
        t[:, d] = 1.0
        assert np.abs(Ao @ t.ravel()).max() <= 1e-14 * np.abs(A).max()
    bdy = tiny.damg.boundary(tiny.n - 1).astype(bool)
    b3 = np.repeat(bdy, 3)
    Ad = np.where(b3[:, None] | b3[None, :], 0.0, A)
    Ad[b3, b3] = 1.0
    x = np.random.default_rng(4).uniform(-1, 1, 3 * N)
    yd = oracle.matvec_elasticity(L, x, mu, lam, bdy=bdy)
    assert np.abs(yd - Ad @ x).max() <= 1e-12 * np.abs(A).max()
    assert np.array_equal(yd[b3], x[b3])
