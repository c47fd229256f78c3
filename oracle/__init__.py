"""CPU oracle of the reference's solve path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings of oracle/liboracle.so (built from oracle/oracle.c by oracle/Makefile).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
checker; the product (dendro_amd/) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-B", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built: run make -C oracle")
        _lib = C.CDLL(LIB_PATH)
        _lib.or_mg_create.restype = C.c_void_p
        _lib.or_mg_create.argtypes = [C.c_int]
        _lib.or_mg_free.argtypes = [C.c_void_p]
        for f in ["or_mg_set_level", "or_mg_setup", "or_mg_setup_ex", "or_mg_vcycle", "or_pcg",
                  "or_mg_apply", "or_mg_coarse_matrix", "or_mg_coarse_set_lu", "or_mg_coarse_size",
                  "or_mg_set_threads", "or_matvec_mt", "or_mg_set_jac3", "or_matvec_jac3", "or_mg_set_boundary"]:
            getattr(_lib, f).argtypes = None
        _lib.or_learn_loss.restype = C.c_double
        _lib.or_l2_error.restype = C.c_double
        _lib.or_eval_fn.restype = C.c_double
    return _lib


def _p(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)


def etype(child, mask):
    return lib().or_etype(int(child), int(mask))


def stencils():
    L, M, R2, R1 = np.zeros(9216), np.zeros(9216), np.zeros(9216), np.zeros(73728)
    lib().or_stencils(_p(L), _p(M), _p(R2), _p(R1))
    return {"lap2": L.reshape(8, 18, 8, 8), "mass2": M.reshape(8, 18, 8, 8),
            "rmat2": R2.reshape(8, 18, 8, 8), "rmat1": R1.reshape(8, 8, 18, 8, 8)}


_ST = None


def _st():
    global _ST
    if _ST is None:
        s = stencils()
        _ST = {k: np.ascontiguousarray(v.reshape(-1)) for k, v in s.items()}
    return _ST


def _coef(level, kappa, mass):
    k = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
    m = None if mass is None else np.ascontiguousarray(mass, np.float64)
    return k, m


def matvec(level, x, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0):
    st = _st()
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(level.n_node)
    k, m = _coef(level, kappa, mass)
    lib().or_matvec(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                    _p(level.nlist), _p(level.child), _p(level.hnmask), _p(level.lev), _p(k), _p(m),
                    C.c_double(lap_fac), C.c_double(mass_fac), _p(st["lap2"]), _p(st["mass2"]), _p(x), _p(y))
    return y


def matvec_mt(level, x, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0, threads=None):
    """or_matvec on `threads` host threads (default: all CPUs of this process)."""
    st = _st()
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(level.n_node)
    k, m = _coef(level, kappa, mass)
    nt = threads or len(os.sched_getaffinity(0))
    rc = lib().or_matvec_mt(C.c_int(nt), C.c_uint32(level.n_elem), C.c_uint32(level.n_node),
                            C.c_uint32(level.max_depth), _p(level.nlist), _p(level.child), _p(level.hnmask),
                            _p(level.lev), _p(k), _p(m), C.c_double(lap_fac), C.c_double(mass_fac),
                            _p(st["lap2"]), _p(st["mass2"]), _p(x), _p(y))
    if rc:
        raise RuntimeError(f"or_matvec_mt failed ({rc})")
    return y


def diagonal(level, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0):
    st = _st()
    d = np.zeros(level.n_node)
    k, m = _coef(level, kappa, mass)
    lib().or_diagonal(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                      _p(level.nlist), _p(level.child), _p(level.hnmask), _p(level.lev), _p(k), _p(m),
                      C.c_double(lap_fac), C.c_double(mass_fac), _p(st["lap2"]), _p(st["mass2"]), _p(d))
    return d


def matvec_dirichlet(level, bdy, x, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0):
    """Dirichlet-masked product (or_matvec_dirichlet); bdy[n_node] nonzero on boundary nodes."""
    st = _st()
    x = np.ascontiguousarray(x, np.float64)
    b = np.ascontiguousarray(bdy, np.uint8)
    y = np.zeros(level.n_node)
    k, m = _coef(level, kappa, mass)
    lib().or_matvec_dirichlet(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                              _p(level.nlist), _p(level.child), _p(level.hnmask), _p(level.lev), _p(b), _p(k),
                              _p(m), C.c_double(lap_fac), C.c_double(mass_fac), _p(st["lap2"]), _p(st["mass2"]),
                              _p(x), _p(y))
    return y


def diagonal_dirichlet(level, bdy, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0):
    st = _st()
    b = np.ascontiguousarray(bdy, np.uint8)
    d = np.zeros(level.n_node)
    k, m = _coef(level, kappa, mass)
    lib().or_diagonal_dirichlet(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                                _p(level.nlist), _p(level.child), _p(level.hnmask), _p(level.lev), _p(b), _p(k),
                                _p(m), C.c_double(lap_fac), C.c_double(mass_fac), _p(st["lap2"]),
                                _p(st["mass2"]), _p(d))
    return d


class _OrLevel(C.Structure):
    _fields_ = [("E", C.c_uint32), ("N", C.c_uint32), ("max_depth", C.c_uint32),
                ("nlist", C.c_void_p), ("child", C.c_void_p), ("hnmask", C.c_void_p),
                ("lev", C.c_void_p), ("anchor", C.c_void_p)]


def _orlevel(L):
    return _OrLevel(L.n_elem, L.n_node, L.max_depth, L.nlist.ctypes.data, L.child.ctypes.data,
                    L.hnmask.ctypes.data, L.lev.ctypes.data, L.anchor.ctypes.data)


_ST1 = None


def stencils_type1():
    """Type1 element stencils lap1/mass1 [8][18][8][8][8] and the shape functions at the
    3-point Gauss nodes shfn [8][18][8][3][3][3]."""
    global _ST1
    if _ST1 is None:
        L, M, S = np.zeros(73728), np.zeros(73728), np.zeros(31104)
        lib().or_stencils_type1(_p(L), _p(M))
        lib().or_shfn(_p(S))
        _ST1 = {"lap1": L, "mass1": M, "shfn": S}
    return {"lap1": _ST1["lap1"].reshape(8, 18, 8, 8, 8), "mass1": _ST1["mass1"].reshape(8, 18, 8, 8, 8),
            "shfn": _ST1["shfn"].reshape(8, 18, 8, 3, 3, 3)}


def matvec_jac3(coarse, fine, x, kappa_fine, mass_fine=None, tol=1e-12):
    """Jacobian3 product on the coarse level with the fine level's element coefficients."""
    st = _st()
    s1 = stencils_type1()
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(coarse.n_node)
    k, m = _coef(fine, kappa_fine, mass_fine)
    oc, of = _orlevel(coarse), _orlevel(fine)
    rc = lib().or_matvec_jac3(C.byref(oc), C.byref(of), _p(k), _p(m), C.c_double(tol), _p(st["lap2"]),
                              _p(st["mass2"]), _p(np.ascontiguousarray(s1["lap1"].reshape(-1))),
                              _p(np.ascontiguousarray(s1["mass1"].reshape(-1))), _p(x), _p(y))
    if rc:
        raise RuntimeError(f"or_matvec_jac3 failed ({rc})")
    return y


def _fn_args(kind, nfreq, a0, a1, poly):
    pv = np.zeros(8) if poly is None else np.ascontiguousarray(poly, np.float64)
    return [C.c_int(kind), C.c_int(nfreq), C.c_double(a0), C.c_double(a1), _p(pv)], pv


_GD = None


def stencils_graddiv():
    global _GD
    if _GD is None:
        _GD = np.zeros(82944)
        lib().or_stencils_graddiv(_p(_GD))
    return _GD.reshape(8, 18, 24, 24)


def matvec_elasticity(level, x, mu=1.0, lam=1.0, bdy=None):
    """ElasticityMatMult: x, y [N][3] interleaved; bdy flags or None."""
    st = _st()
    gd = np.ascontiguousarray(stencils_graddiv().reshape(-1))
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    y = np.zeros(3 * level.n_node)
    b = None if bdy is None else np.ascontiguousarray(bdy, np.uint8)
    lib().or_matvec_elasticity(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                               _p(level.nlist), _p(level.child), _p(level.hnmask), _p(level.lev), _p(b),
                               C.c_double(mu), C.c_double(lam), _p(st["lap2"]), _p(gd), _p(x), _p(y))
    return y


def eval_fn(kind, x, y, z, nfreq=9, a0=1.0, a1=0.0, poly=None):
    args, keep = _fn_args(kind, nfreq, a0, a1, poly)
    return lib().or_eval_fn(*args, C.c_double(x), C.c_double(y), C.c_double(z))


def rhs_quadrature(level, kind, nfreq=9, a0=1.0, a1=0.0, poly=None):
    """ComputeRHS2's quadrature right-hand side of a function (kinds as dgh_function)."""
    shfn = np.ascontiguousarray(stencils_type1()["shfn"].reshape(-1))
    args, keep = _fn_args(kind, nfreq, a0, a1, poly)
    b = np.zeros(level.n_node)
    lib().or_rhs_quadrature(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                            _p(level.nlist), _p(level.child), _p(level.hnmask), _p(level.lev), _p(level.anchor),
                            *args, _p(shfn), _p(b))
    return b


def nodal_values(level, kind, nfreq=9, a0=1.0, a1=0.0, poly=None):
    args, keep = _fn_args(kind, nfreq, a0, a1, poly)
    u = np.zeros(level.n_node)
    lib().or_nodal_values(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                          _p(level.nlist), _p(level.hnmask), _p(level.lev), _p(level.anchor), *args, _p(u))
    return u


def l2_error(level, u, kind, nfreq=9, a0=1.0, a1=0.0, poly=None):
    shfn = np.ascontiguousarray(stencils_type1()["shfn"].reshape(-1))
    args, keep = _fn_args(kind, nfreq, a0, a1, poly)
    u = np.ascontiguousarray(u, np.float64)
    return lib().or_l2_error(C.c_uint32(level.n_elem), C.c_uint32(level.max_depth), _p(level.nlist),
                             _p(level.child), _p(level.hnmask), _p(level.lev), _p(level.anchor), *args,
                             _p(shfn), _p(u))


def restrict(coarse, fine, f):
    st = _st()
    f = np.ascontiguousarray(f, np.float64)
    c = np.zeros(coarse.n_node)
    oc, of = _orlevel(coarse), _orlevel(fine)
    rc = lib().or_restrict(C.byref(oc), C.byref(of), _p(st["rmat2"]), _p(st["rmat1"]), _p(f), _p(c))
    if rc:
        raise RuntimeError(f"or_restrict failed ({rc})")
    return c


def prolong(coarse, fine, c):
    st = _st()
    c = np.ascontiguousarray(c, np.float64)
    f = np.zeros(fine.n_node)
    oc, of = _orlevel(coarse), _orlevel(fine)
    rc = lib().or_prolong(C.byref(oc), C.byref(of), _p(st["rmat2"]), _p(st["rmat1"]), _p(c), _p(f))
    if rc:
        raise RuntimeError(f"or_prolong failed ({rc})")
    return f


class MG:
    """Oracle multigrid hierarchy; levels: list (coarsest first) of (Level, kappa, mass,
    lap_fac, mass_fac)."""

    # coarsest levels above this size are factored by LAPACK (scipy) instead of the oracle's
    # unblocked loop; both are partial-pivoting LU of the same assembled matrix
    LAPACK_ABOVE = 2500

    def __init__(self, levels, omega=2.0 / 3.0, nu_pre=2, nu_post=2, null_space=False, threads=1,
                 jac3=False, jac3_tol=1e-12, boundary=None):
        self._keep = []
        self.h = lib().or_mg_create(len(levels))
        for i, (L, kappa, mass, lf, mf) in enumerate(levels):
            k, m = _coef(L, kappa, mass)
            self._keep += [L, k, m]
            rc = lib().or_mg_set_level(C.c_void_p(self.h), C.c_int(i), C.c_uint32(L.n_elem),
                                       C.c_uint32(L.n_node), C.c_uint32(L.max_depth), _p(L.nlist),
                                       _p(L.child), _p(L.hnmask), _p(L.lev), _p(L.anchor), _p(k), _p(m),
                                       C.c_double(lf), C.c_double(mf))
            assert rc == 0
        if threads > 1:
            lib().or_mg_set_threads(C.c_void_p(self.h), C.c_int(threads))
        if jac3:
            lib().or_mg_set_jac3(C.c_void_p(self.h), C.c_int(1), C.c_double(jac3_tol))
        if boundary is not None:   # one flag array per level (coarsest first)
            for i, bd in enumerate(boundary):
                b8 = np.ascontiguousarray(bd, np.uint8)
                self._keep.append(b8)
                assert lib().or_mg_set_boundary(C.c_void_p(self.h), C.c_int(i), _p(b8)) == 0
        big = levels[0][0].n_node > self.LAPACK_ABOVE
        rc = lib().or_mg_setup_ex(C.c_void_p(self.h), C.c_double(omega), C.c_int(nu_pre), C.c_int(nu_post),
                                  C.c_int(int(null_space)), C.c_int(0 if big else 1))
        if rc:
            raise RuntimeError(f"or_mg_setup failed ({rc})")
        if big:
            import scipy.linalg
            n = lib().or_mg_coarse_size(C.c_void_p(self.h))
            A = np.zeros((n, n))
            assert lib().or_mg_coarse_matrix(C.c_void_p(self.h), _p(A)) == 0
            lu, piv = scipy.linalg.lu_factor(A, overwrite_a=True, check_finite=False)
            lu = np.ascontiguousarray(lu)
            piv = np.ascontiguousarray(piv, np.int32)
            assert lib().or_mg_coarse_set_lu(C.c_void_p(self.h), _p(lu), _p(piv)) == 0
        self.n = levels[-1][0].n_node

    def vcycle(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros(self.n)
        rc = lib().or_mg_vcycle(C.c_void_p(self.h), _p(r), _p(z))
        if rc:
            raise RuntimeError(f"or_mg_vcycle failed ({rc})")
        return z

    def apply(self, level, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(len(x))
        lib().or_mg_apply(C.c_void_p(self.h), C.c_int(level), _p(x), _p(y))
        return y

    def pcg(self, b, rtol=1e-8, max_it=100, norm_type=0):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n)
        hist = np.zeros(max_it + 1)
        conv = C.c_int()
        its = lib().or_pcg(C.c_void_p(self.h), _p(b), _p(x), C.c_double(rtol), C.c_int(max_it),
                           C.c_int(norm_type), _p(hist), C.c_int(max_it + 1), C.byref(conv))
        return x, its, bool(conv.value), hist[: its + 1]

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mg_free(C.c_void_p(self.h))
            self.h = None


# ------------------------------------------------------------------ learned Jacobi damping
NF, NH, NO = 35, 64, 8
NPARAM = NH * NF + NH + NH * NH + NH + NO * NH + NO


def learn_node_logk(level, kappa=None, lap_fac=1.0):
    k = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
    lkbar = np.zeros(level.n_node, np.float32)
    lke = np.zeros(level.n_elem, np.float32)
    lib().or_learn_node_logk(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), _p(level.nlist), _p(k),
                             C.c_double(lap_fac), _p(lkbar), _p(lke))
    return lkbar, lke


def learn_features(level, kappa=None, lap_fac=1.0):
    lkbar, lke = learn_node_logk(level, kappa, lap_fac)
    X = np.zeros((level.n_elem, NF), np.float32)
    lib().or_learn_features(C.c_uint32(level.n_elem), C.c_uint32(level.max_depth), _p(level.nlist),
                            _p(level.child), _p(level.hnmask), _p(level.lev), _p(lkbar), _p(lke), _p(X))
    return X


def mlp_forward(theta, X, round_bf16=True):
    theta = np.ascontiguousarray(theta, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    O = np.zeros((X.shape[0], NO))
    lib().or_mlp_forward(_p(theta), C.c_int(int(round_bf16)), C.c_uint32(X.shape[0]), _p(X), _p(O))
    return O


def mlp_backward(theta, X, dO, round_bf16=True):
    theta = np.ascontiguousarray(theta, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    dO = np.ascontiguousarray(dO, np.float64)
    g = np.zeros(NPARAM)
    lib().or_mlp_backward(_p(theta), C.c_int(int(round_bf16)), C.c_uint32(X.shape[0]), _p(X), _p(dO), _p(g))
    return g


def learn_cap(level, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0):
    st = _st()
    k, m = _coef(level, kappa, mass)
    cap = np.zeros(level.n_node)
    lib().or_learn_cap(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                       _p(level.nlist), _p(level.child), _p(level.hnmask), _p(level.lev), _p(k), _p(m),
                       C.c_double(lap_fac), C.c_double(mass_fac), _p(st["lap2"]), _p(st["mass2"]), _p(cap))
    return cap


def learn_nodal(level, O, cap):
    O = np.ascontiguousarray(O, np.float64)
    cap = np.ascontiguousarray(cap, np.float64)
    th = np.zeros(level.n_node)
    om = np.zeros(level.n_node)
    lib().or_learn_nodal(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), _p(level.nlist), _p(O), _p(cap),
                         _p(th), _p(om))
    return th, om


def learn_nodal_grad(level, theta_n, cap, g_omega):
    dO = np.zeros((level.n_elem, NO))
    lib().or_learn_nodal_grad(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), _p(level.nlist),
                              _p(np.ascontiguousarray(theta_n, np.float64)), _p(np.ascontiguousarray(cap, np.float64)),
                              _p(np.ascontiguousarray(g_omega, np.float64)), _p(dO))
    return dO


def learn_loss(level, omega, E0, nu=2, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0, grad=True):
    """L = mean_s |S_w^nu e_s|_A^2 / |e_s|_A^2 and dL/domega (E0: [N][S])."""
    st = _st()
    k, m = _coef(level, kappa, mass)
    invd = 1.0 / diagonal(level, kappa, mass, lap_fac, mass_fac)
    E0 = np.ascontiguousarray(E0, np.float64)
    S = E0.shape[1]
    g = np.zeros(level.n_node) if grad else None
    L = lib().or_learn_loss(C.c_uint32(level.n_elem), C.c_uint32(level.n_node), C.c_uint32(level.max_depth),
                            _p(level.nlist), _p(level.child), _p(level.hnmask), _p(level.lev), _p(k), _p(m),
                            C.c_double(lap_fac), C.c_double(mass_fac), _p(st["lap2"]), _p(st["mass2"]),
                            _p(np.ascontiguousarray(omega, np.float64)), _p(invd), C.c_int(nu), C.c_int(S),
                            _p(E0), _p(g))
    return (L, g) if grad else L
