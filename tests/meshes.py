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


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
