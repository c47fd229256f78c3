"""Seeded test meshes and coefficients (shared by the CPU and GPU tests)."""
import numpy as np

import dendro_amd as da

KINDS = {"gauss": da.PTS_GAUSS, "lognormal": da.PTS_LOGNORMAL, "mixed": da.PTS_MIXED,
         "clusters": da.PTS_CLUSTERS, "uniform": da.PTS_UNIFORM}


class Hier:
    def __init__(self, damg):
        self.damg = damg
        self.n = damg.nlevels
        self.levels = [damg.level(l) for l in range(self.n)]
        self.transfers = [damg.transfer(l) for l in range(1, self.n)]

    @property
    def fine(self):
        return self.levels[-1]


def hierarchy(kind, npts, depth, nlevels, seed=1, min_coarse=500, cluster_seed=31):
    pts = da.points(KINDS[kind], npts, seed, cluster_seed)
    oct_ = da.Octree.from_points(pts, depth)
    return Hier(da.DAMG(oct_, nlevels, min_coarse))


def var_coefficients(hier, seed=21, mass_base=None):
    """kappa_e = 10^U(-1,1) on the finest level, averaged to coarser levels."""
    rng = np.random.default_rng(seed)
    kf = 10.0 ** rng.uniform(-1, 1, hier.fine.n_elem)
    mf = None if mass_base is None else np.full(hier.fine.n_elem, float(mass_base))
    out = []
    for l in range(hier.n):
        k, m = hier.damg.coarsen_coefficients(kf, mf, l)
        out.append((k, m))
    return out


def node_coords(level):
    """Physical coordinates [N][3] of the nodes, read off the elements' non-hanging vertices
    (vertex k at anchor + h (k&1, k>>1&1, k>>2&1); DA coordinate c is c / 2^(maxD-1) in the
    unit cube).  Raises if two elements place one node at different points."""
    nl = level.nlist.reshape(-1, 8)
    an = level.anchor.reshape(-1, 3).astype(np.int64)
    hs = np.left_shift(1, level.max_depth - level.lev.astype(np.int64))
    xyz = np.full((level.n_node, 3), -1, np.int64)
    for k in range(8):
        off = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
        m = ((level.hnmask >> k) & 1) == 0
        c = an[m] + hs[m, None] * off
        ids = nl[m, k]
        prev = xyz[ids]
        if np.any((prev[:, 0] >= 0) & np.any(prev != c, axis=1)):
            raise ValueError("inconsistent node coordinates")
        xyz[ids] = c
    if np.any(xyz < 0):
        raise ValueError("node without a regular vertex")
    return xyz / float(1 << (level.max_depth - 1))


def matvec_longdouble(level, x, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0, absolute=False):
    """y = A x with every product and sum in x87 extended precision (64-bit significand):
    y_n = sum over elements and vertex rows k at n of fac1 (L x_e)_k + fac2 (M x_e)_k.  Its
    error is about 2^-11 of an FP64 evaluation's, so it serves as the exact product when
    FP64 evaluations are compared on inputs where cancellation is large.  absolute=True
    gives |A| |x| (the scale of the terms each row sums)."""
    import oracle
    st = oracle.stencils()
    ld = np.longdouble
    nl = level.nlist.reshape(-1, 8).astype(np.int64)
    E = level.n_elem
    t = np.zeros(E, np.int64)
    for e in np.nonzero(level.hnmask)[0]:
        t[e] = oracle.etype(level.child[e], level.hnmask[e])
    c = level.child.astype(np.int64)
    h = np.ldexp(1.0, (level.max_depth - level.lev.astype(np.int64)) - (level.max_depth - 1))
    k1 = np.full(E, lap_fac) if kappa is None else np.asarray(kappa, np.float64)
    k2 = np.full(E, mass_fac) if kappa is None else (np.zeros(E) if mass is None else np.asarray(mass, np.float64))
    fac1 = (k1 * h / 2.0).astype(ld)
    fac2 = (k2 * h * h * h / 8.0).astype(ld)
    xe = np.asarray(x, np.float64)[nl].astype(ld)                     # [E][8]
    Le = st["lap2"][c, t].astype(ld)                                   # [E][8][8]
    Me = st["mass2"][c, t].astype(ld)
    if absolute:
        xe, Le, Me, fac1, fac2 = np.abs(xe), np.abs(Le), np.abs(Me), np.abs(fac1), np.abs(fac2)
    ye = fac1[:, None] * np.einsum("ekj,ej->ek", Le, xe) + fac2[:, None] * np.einsum("ekj,ej->ek", Me, xe)
    y = np.zeros(level.n_node, ld)
    np.add.at(y, nl.reshape(-1), ye.reshape(-1))
    return y


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
