"""ctypes bindings of libdendrohip.so (include/dendro_hip.h).

Host-side objects (octrees, the DA hierarchy) are plain numpy; device vectors are any
objects exposing a device pointer (torch tensors on cuda, via .data_ptr()).  Every call
checks the C status and raises DendroError with the library's message.  There is no
fallback: if the shared library is missing, importing the bindings fails.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdendrohip.so")

F64, F32 = 0, 1
NORM_UNPRECONDITIONED, NORM_PRECONDITIONED = 0, 1
SMOOTHER_JACOBI, SMOOTHER_LEARNED = 0, 1
COARSE_JAC2, COARSE_JAC3 = 0, 1
ROWS_TYPE, ROWS_LANE, ROWS_CLASS = 0, 1, 2   # DGH_ROWS_*: stencil-row source of the operator
FN_GAUSSIANS, FN_COS_SUM, FN_POLY = 0, 1, 2
PTS_GAUSS, PTS_LOGNORMAL, PTS_MIXED, PTS_CLUSTERS, PTS_UNIFORM = range(5)
NOALLOC_MAX_TILES = 16384     # DGH_NOALLOC_MAX_TILES

# every symbol include/dendro_hip.h declares
EXPORTS = [
    "dgh_last_error", "dgh_etype", "dgh_stencils_generate", "dgh_points_generate",
    "dgh_octree_from_points", "dgh_octree_from_leaves", "dgh_octree_balance", "dgh_octree_coarsen",
    "dgh_octree_size", "dgh_octree_get", "dgh_octree_free",
    "dgh_damg_build", "dgh_damg_num_levels", "dgh_damg_level_info", "dgh_damg_level_arrays",
    "dgh_damg_level_octants", "dgh_damg_level_boundary", "dgh_damg_transfer_arrays",
    "dgh_damg_coarsen_coefficients", "dgh_damg_save", "dgh_damg_load", "dgh_damg_free",
    "dgh_init", "dgh_finalize", "dgh_alloc", "dgh_free", "dgh_copy_to_device", "dgh_copy_to_host",
    "dgh_synchronize",
    "dgh_mesh_create", "dgh_mesh_set_coefficients", "dgh_mesh_set_boundary", "dgh_matvec_dirichlet",
    "dgh_diagonal_dirichlet", "dgh_mesh_info", "dgh_mesh_destroy",
    "dgh_mesh_set_tile", "dgh_mesh_tile_info", "dgh_mesh_autotune", "dgh_mesh_set_lane_stencils",
    "dgh_mesh_lane_stencils",
    "dgh_matvec", "dgh_matvec_dot", "dgh_matvec_dot_noalloc", "dgh_diagonal", "dgh_matvec_multi", "dgh_mesh_set_tile_multi",
    "dgh_mesh_autotune_multi",
    "dgh_mesh_set_geometry", "dgh_rhs_quadrature", "dgh_nodal_values", "dgh_l2_error",
    "dgh_mesh_set_elasticity", "dgh_matvec_elasticity",
    "dgh_transfer_create", "dgh_transfer_destroy", "dgh_restrict", "dgh_prolong_add", "dgh_matvec_jac3",
    "dgh_solver_default_params", "dgh_solver_create", "dgh_solver_destroy", "dgh_solver_solve",
    "dgh_solver_solve_multi", "dgh_solver_vcycle", "dgh_solver_coarse_info",
    "dgh_learn_default_params", "dgh_learned_init", "dgh_learned_get_params", "dgh_learned_set_params",
    "dgh_learned_train", "dgh_learned_apply", "dgh_learned_eval", "dgh_learned_loss",
]


class DendroError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class LevelInfo(C.Structure):
    _fields_ = [("n_elem", C.c_uint32), ("n_node", C.c_uint32), ("n_oct", C.c_uint32),
                ("max_depth", C.c_uint32), ("n_hanging", C.c_uint32)]


class SolverParams(C.Structure):
    _fields_ = [("rtol", C.c_double), ("omega", C.c_double), ("max_it", C.c_int),
                ("nu_pre", C.c_int), ("nu_post", C.c_int), ("norm_type", C.c_int),
                ("null_space", C.c_int), ("vcycle_prec", C.c_int), ("smoother", C.c_int),
                ("use_graph", C.c_int), ("coarse_op", C.c_int), ("coarse_dense_max", C.c_int),
                ("coarse_rtol", C.c_double), ("coarse_max_it", C.c_int)]


class Function(C.Structure):
    """dgh_function: FN_GAUSSIANS (ComputeRHS2's source), FN_COS_SUM
    sum_{k<=nfreq} (a0 + 3 a1 k^2 pi^2) cos(k pi x) cos(k pi y) cos(k pi z), FN_POLY."""
    _fields_ = [("kind", C.c_int), ("nfreq", C.c_int), ("a0", C.c_double), ("a1", C.c_double),
                ("poly", C.c_double * 8)]


def function(kind, nfreq=9, a0=1.0, a1=0.0, poly=None):
    f = Function(kind, nfreq, a0, a1)
    for i, v in enumerate(poly or []):
        f.poly[i] = v
    return f


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("rnorm0", C.c_double),
                ("rnorm", C.c_double), ("rel_true", C.c_double), ("seconds", C.c_double)]


class LearnParams(C.Structure):
    _fields_ = [("steps", C.c_int), ("samples", C.c_int), ("nu", C.c_int), ("refresh", C.c_int),
                ("lr", C.c_float), ("seed", C.c_uint64)]


class LearnStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("loss_first", C.c_double), ("loss_last", C.c_double),
                ("seconds", C.c_double)]


LEARN_NPARAM = 6984
_lib = None


def lib():
    """The loaded libdendrohip.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built: run python -m dendro_amd.build")
        _lib = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(_lib, name).restype = C.c_int
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    lib().dgh_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc < 0:
        raise DendroError(rc, last_error())
    return rc


def ptr(a):
    """Host numpy array or device tensor -> void*."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags.c_contiguous
        return C.c_void_p(a.ctypes.data)
    if hasattr(a, "data_ptr"):
        assert a.is_contiguous()
        return C.c_void_p(a.data_ptr())
    if isinstance(a, int):
        return C.c_void_p(a)
    raise TypeError(type(a))


def _stream(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


# ------------------------------------------------------------------ host: tables, points
def etype(child, mask):
    return check(lib().dgh_etype(int(child), int(mask)))


def stencils():
    L = np.zeros(9216)
    M = np.zeros(9216)
    R2 = np.zeros(9216)
    R1 = np.zeros(73728)
    check(lib().dgh_stencils_generate(ptr(L), ptr(M), ptr(R2), ptr(R1)))
    return {"lap2": L.reshape(8, 18, 8, 8), "mass2": M.reshape(8, 18, 8, 8),
            "rmat2": R2.reshape(8, 18, 8, 8), "rmat1": R1.reshape(8, 8, 18, 8, 8)}


def points(kind, n, seed, cluster_seed=0):
    out = np.zeros((n, 3))
    check(lib().dgh_points_generate(int(kind), C.c_uint64(n), C.c_uint64(seed),
                                    C.c_uint64(cluster_seed), ptr(out)))
    return out


# ------------------------------------------------------------------ host: octrees
class Octree:
    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_points(cls, pts, max_depth, max_pts=1, inc_corner=True):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        h = C.c_void_p()
        check(lib().dgh_octree_from_points(ptr(pts), C.c_uint64(pts.size // 3), int(max_depth),
                                           int(max_pts), int(bool(inc_corner)), C.byref(h)))
        return cls(h)

    @classmethod
    def from_leaves(cls, anchors, levels, max_depth):
        a = np.ascontiguousarray(anchors, dtype=np.uint32)
        l = np.ascontiguousarray(levels, dtype=np.uint8)
        h = C.c_void_p()
        check(lib().dgh_octree_from_leaves(ptr(a), ptr(l), C.c_uint64(len(l)), int(max_depth),
                                           C.byref(h)))
        return cls(h)

    def balance(self, inc_corner=True):
        h = C.c_void_p()
        check(lib().dgh_octree_balance(self.h, int(bool(inc_corner)), C.byref(h)))
        return Octree(h)

    def coarsen(self):
        h = C.c_void_p()
        check(lib().dgh_octree_coarsen(self.h, C.byref(h)))
        return Octree(h)

    def size(self):
        n = C.c_uint64()
        d = C.c_uint32()
        check(lib().dgh_octree_size(self.h, C.byref(n), C.byref(d)))
        return n.value, d.value

    def leaves(self):
        n, d = self.size()
        a = np.zeros((n, 3), np.uint32)
        l = np.zeros(n, np.uint8)
        check(lib().dgh_octree_get(self.h, ptr(a), ptr(l)))
        return a, l, d

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dgh_octree_free(self.h)
            self.h = None


class Level:
    """Arrays of one DA level (host numpy)."""

    def __init__(self, info, nlist, child, hnmask, lev, anchor):
        self.n_elem = info.n_elem
        self.n_node = info.n_node
        self.n_oct = info.n_oct
        self.max_depth = info.max_depth
        self.n_hanging = info.n_hanging
        self.nlist, self.child, self.hnmask, self.lev, self.anchor = nlist, child, hnmask, lev, anchor


class DAMG:
    """Host multigrid hierarchy of DAs (level 0 = coarsest)."""

    def __init__(self, octree, nlevels, min_coarse=500):
        h = C.c_void_p()
        nl = C.c_int(int(nlevels))
        check(lib().dgh_damg_build(octree.h, C.byref(nl), int(min_coarse), C.byref(h)))
        self.h = h
        self.nlevels = nl.value
        self._levels = {}

    @classmethod
    def load(cls, path):
        """A hierarchy saved by save() (versioned binary, checked on load)."""
        h = C.c_void_p()
        check(lib().dgh_damg_load(os.fsencode(path), C.byref(h)))
        self = cls.__new__(cls)
        self.h = h
        n = C.c_int()
        check(lib().dgh_damg_num_levels(h, C.byref(n)))
        self.nlevels = n.value
        self._levels = {}
        return self

    def save(self, path):
        check(lib().dgh_damg_save(self.h, os.fsencode(path)))

    def info(self, l):
        li = LevelInfo()
        check(lib().dgh_damg_level_info(self.h, int(l), C.byref(li)))
        return li

    def level(self, l):
        if l not in self._levels:
            li = self.info(l)
            E = li.n_elem
            nlist = np.zeros((E, 8), np.uint32)
            child = np.zeros(E, np.uint8)
            hn = np.zeros(E, np.uint8)
            lev = np.zeros(E, np.uint8)
            anc = np.zeros((E, 3), np.uint32)
            check(lib().dgh_damg_level_arrays(self.h, int(l), ptr(nlist), ptr(child), ptr(hn),
                                              ptr(lev), ptr(anc)))
            self._levels[l] = Level(li, nlist, child, hn, lev, anc)
        return self._levels[l]

    def octants(self, l):
        li = self.info(l)
        anc = np.zeros((li.n_oct, 3), np.uint32)
        lev = np.zeros(li.n_oct, np.uint8)
        fl = np.zeros(li.n_oct, np.uint8)
        check(lib().dgh_damg_level_octants(self.h, int(l), ptr(anc), ptr(lev), ptr(fl)))
        return anc, lev, fl

    def boundary(self, l):
        """uint8 flags per node of level l: 1 on the boundary of the unit cube"""
        b = np.zeros(self.info(l).n_node, np.uint8)
        check(lib().dgh_damg_level_boundary(self.h, int(l), ptr(b)))
        return b

    def transfer(self, fine_level):
        li = self.info(fine_level)
        ce = np.zeros(li.n_node, np.uint32)
        pos = np.zeros(li.n_node, np.uint8)
        check(lib().dgh_damg_transfer_arrays(self.h, int(fine_level), ptr(ce), ptr(pos)))
        return ce, pos

    def coarsen_coefficients(self, kappa, mass, level):
        li = self.info(level)
        k = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        m = None if mass is None else np.ascontiguousarray(mass, np.float64)
        ok = np.zeros(li.n_elem) if k is not None else None
        om = np.zeros(li.n_elem) if m is not None else None
        check(lib().dgh_damg_coarsen_coefficients(self.h, ptr(k), ptr(m), int(level), ptr(ok), ptr(om)))
        return ok, om

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dgh_damg_free(self.h)
            self.h = None


# ------------------------------------------------------------------ device
_initialised = False


def init(device=0):
    global _initialised
    check(lib().dgh_init(int(device)))
    _initialised = True


class Mesh:
    """One DA level uploaded to the device."""

    def __init__(self, level):
        self.level = level
        h = C.c_void_p()
        check(lib().dgh_mesh_create(level.n_elem, level.n_node, level.max_depth, ptr(level.nlist),
                                    ptr(level.child), ptr(level.hnmask), ptr(level.lev), C.byref(h)))
        self.h = h
        self.n_elem, self.n_node = level.n_elem, level.n_node

    def set_coefficients(self, kappa=None, mass=None, lap_fac=1.0, mass_fac=1.0):
        k = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        m = None if mass is None else np.ascontiguousarray(mass, np.float64)
        check(lib().dgh_mesh_set_coefficients(self.h, ptr(k), ptr(m), C.c_double(lap_fac),
                                              C.c_double(mass_fac)))

    def set_tile(self, tile_elems, block=0):
        check(lib().dgh_mesh_set_tile(self.h, int(tile_elems), int(block)))

    def tile_info(self):
        te, nt, lds = C.c_uint32(), C.c_uint32(), C.c_uint32()
        nr, nn = C.c_uint64(), C.c_uint64()
        check(lib().dgh_mesh_tile_info(self.h, C.byref(te), C.byref(nt), C.byref(nr), C.byref(nn),
                                       C.byref(lds)))
        on = C.c_int()
        check(lib().dgh_mesh_lane_stencils(self.h, C.byref(on)))
        return {"tile_elems": te.value, "n_tiles": nt.value, "n_records": nr.value,
                "n_node_entries": nn.value, "lds_bytes_f64": lds.value, "lane_stencils": on.value}

    def set_lane_stencils(self, on):
        """Stencil-row source: ROWS_LANE (1) per lane from an LDS copy of the table;
        ROWS_TYPE (0) per stencil type of a wavefront through the scalar cache; ROWS_CLASS (2)
        the element matrix of each (type, level) class (uniform coefficients only)."""
        check(lib().dgh_mesh_set_lane_stencils(self.h, int(on)))

    def autotune(self, prec=F64):
        te, blk, us = C.c_uint32(), C.c_uint32(), C.c_float()
        check(lib().dgh_mesh_autotune(self.h, int(prec), C.byref(te), C.byref(blk), C.byref(us)))
        on = C.c_int()
        check(lib().dgh_mesh_lane_stencils(self.h, C.byref(on)))
        return {"tile_elems": te.value, "block": blk.value, "us": us.value, "lane_stencils": on.value}

    def matvec(self, x, y, prec=F64, stream=None):
        check(lib().dgh_matvec(self.h, int(prec), ptr(x), ptr(y), _stream(stream)))

    def set_elasticity(self, mu, lam):
        check(lib().dgh_mesh_set_elasticity(self.h, C.c_double(mu), C.c_double(lam)))

    def matvec_elasticity(self, x, y, stream=None):
        """y = A x for the elasticity operator, x and y [N][3] FP64"""
        check(lib().dgh_matvec_elasticity(self.h, ptr(x), ptr(y), _stream(stream)))

    def set_geometry(self, anchor):
        """element anchors [E][3] (DA coordinates) for the function-evaluating calls"""
        a = np.ascontiguousarray(anchor, np.uint32)
        check(lib().dgh_mesh_set_geometry(self.h, ptr(a)))

    def rhs_quadrature(self, f, b, stream=None):
        check(lib().dgh_rhs_quadrature(self.h, C.byref(f), ptr(b), _stream(stream)))

    def nodal_values(self, f, u, stream=None):
        check(lib().dgh_nodal_values(self.h, C.byref(f), ptr(u), _stream(stream)))

    def l2_error(self, f, u):
        e = C.c_double()
        check(lib().dgh_l2_error(self.h, C.byref(f), ptr(u), C.byref(e)))
        return e.value

    def matvec_dot(self, x, y, xy, stream=None):
        """y = A x and xy[0] = x.y (FP64; xy a one-element device tensor)."""
        check(lib().dgh_matvec_dot(self.h, ptr(x), ptr(y), ptr(xy), _stream(stream)))

    def matvec_dot_noalloc(self, x, y, xy, stream=None):
        """matvec_dot with its tile partials in the library's fixed device array (no allocation)."""
        check(lib().dgh_matvec_dot_noalloc(self.h, ptr(x), ptr(y), ptr(xy), _stream(stream)))

    def matvec_multi(self, x, y, k, prec=F64, stream=None):
        """Y = A X for k columns stored node-major ([N][k])."""
        check(lib().dgh_matvec_multi(self.h, int(prec), int(k), ptr(x), ptr(y), _stream(stream)))

    def set_tile_multi(self, tile_elems, variant=0):
        """FP32 product variant: 0 vector ALUs (8 columns per slice, 512 threads, one column
        per lane), 1 matrix cores (MFMA), 2 as 0 with 1024 threads, 3 as 0 with 4 columns per
        slice, 4 and 5 as 0 with 2 and 4 columns per lane, 6 as 3 with 2 columns per lane"""
        check(lib().dgh_mesh_set_tile_multi(self.h, int(tile_elems), int(variant)))

    def autotune_multi(self, k):
        te, var, us = C.c_uint32(), C.c_int(), C.c_float()
        check(lib().dgh_mesh_autotune_multi(self.h, int(k), C.byref(te), C.byref(var), C.byref(us)))
        return {"tile_elems": te.value, "variant": var.value, "us": us.value}

    def diagonal(self, d, prec=F64, stream=None):
        check(lib().dgh_diagonal(self.h, int(prec), ptr(d), _stream(stream)))

    def set_boundary(self, bdy):
        """boundary flags [N] (nonzero = boundary) for the Dirichlet operators; None removes them"""
        b = None if bdy is None else np.ascontiguousarray(bdy, np.uint8)
        check(lib().dgh_mesh_set_boundary(self.h, ptr(b)))

    def matvec_dirichlet(self, x, y, prec=F64, stream=None):
        check(lib().dgh_matvec_dirichlet(self.h, int(prec), ptr(x), ptr(y), _stream(stream)))

    def diagonal_dirichlet(self, d, prec=F64, stream=None):
        check(lib().dgh_diagonal_dirichlet(self.h, int(prec), ptr(d), _stream(stream)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dgh_mesh_destroy(self.h)
            self.h = None


class Transfer:
    def __init__(self, coarse, fine, celem, pos):
        self.coarse, self.fine = coarse, fine   # keep the meshes alive
        h = C.c_void_p()
        check(lib().dgh_transfer_create(coarse.h, fine.h, ptr(np.ascontiguousarray(celem, np.uint32)),
                                        ptr(np.ascontiguousarray(pos, np.uint8)), C.byref(h)))
        self.h = h

    def restrict(self, f, c, prec=F64, stream=None):
        check(lib().dgh_restrict(self.h, int(prec), ptr(f), ptr(c), _stream(stream)))

    def matvec_jac3(self, x, y, prec=F64, stream=None):
        """Jacobian3 on the coarse level: y = P^T A_fine P x (fine mesh coefficients)."""
        check(lib().dgh_matvec_jac3(self.h, int(prec), ptr(x), ptr(y), _stream(stream)))

    def prolong_add(self, c, f_in, f_out, prec=F64, stream=None):
        check(lib().dgh_prolong_add(self.h, int(prec), ptr(c), ptr(f_in), ptr(f_out), _stream(stream)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dgh_transfer_destroy(self.h)
            self.h = None


def default_params(**kw):
    p = SolverParams()
    check(lib().dgh_solver_default_params(C.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Solver:
    def __init__(self, meshes, transfers, params):
        self.meshes, self.transfers = list(meshes), list(transfers)
        self.params = params
        mh = (C.c_void_p * len(meshes))(*[m.h.value for m in meshes])
        th = (C.c_void_p * max(1, len(transfers)))(*([t.h.value for t in transfers] or [0]))
        h = C.c_void_p()
        check(lib().dgh_solver_create(len(meshes), mh, th if transfers else None, C.byref(params),
                                      C.byref(h)))
        self.h = h

    def coarse_info(self):
        """(iterative, last_iterations): 1 when the coarsest level is solved by CG."""
        it, n = C.c_int(), C.c_int()
        check(lib().dgh_solver_coarse_info(self.h, C.byref(it), C.byref(n)))
        return it.value, n.value

    def solve(self, b, x, history_len=0, stream=None):
        st = SolveStats()
        hist = np.zeros(max(1, history_len))
        check(lib().dgh_solver_solve(self.h, ptr(b), ptr(x), C.byref(st),
                                     ptr(hist) if history_len else None, int(history_len),
                                     _stream(stream)))
        return st, hist[: min(history_len, st.iterations + 1)]

    def solve_multi(self, B, X, k, stream=None):
        """k independent solves; B, X FP64 [N][k] node-major.  Returns k SolveStats."""
        st = (SolveStats * int(k))()
        check(lib().dgh_solver_solve_multi(self.h, int(k), ptr(B), ptr(X), st, _stream(stream)))
        return list(st)

    def vcycle(self, r, z, stream=None):
        check(lib().dgh_solver_vcycle(self.h, ptr(r), ptr(z), _stream(stream)))

    # ---- learned Jacobi damping
    def learned_init(self, seed):
        check(lib().dgh_learned_init(self.h, C.c_uint64(seed)))

    def learned_params(self):
        th = np.zeros(LEARN_NPARAM, np.float32)
        check(lib().dgh_learned_get_params(self.h, ptr(th), LEARN_NPARAM))
        return th

    def set_learned_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        check(lib().dgh_learned_set_params(self.h, ptr(th), int(th.size)))

    def learned_train(self, **kw):
        p = LearnParams()
        check(lib().dgh_learn_default_params(C.byref(p)))
        for k, v in kw.items():
            setattr(p, k, v)
        st = LearnStats()
        check(lib().dgh_learned_train(self.h, C.byref(p), C.byref(st)))
        return st

    def learned_apply(self, enable=True, samples=8, seed=6000):
        mask = C.c_int()
        check(lib().dgh_learned_apply(self.h, int(enable), int(samples), C.c_uint64(seed), C.byref(mask)))
        return mask.value

    def learned_eval(self, level, n_elem, dO=None):
        O = np.zeros((n_elem, 8), np.float32)
        g = np.zeros(LEARN_NPARAM, np.float32)
        d = None if dO is None else np.ascontiguousarray(dO, np.float32)
        check(lib().dgh_learned_eval(self.h, int(level), ptr(d), ptr(O), ptr(g) if d is not None else None))
        return (O, g) if d is not None else O

    def learned_loss(self, level, n_node, omega=None, samples=None, S=8, nu=2, seed=0, grad=True):
        om = None if omega is None else np.ascontiguousarray(omega, np.float32)
        sm = None if samples is None else np.ascontiguousarray(samples, np.float32)
        if sm is not None:
            S = sm.shape[1]
        loss = C.c_double()
        g = np.zeros(n_node, np.float32) if grad else None
        check(lib().dgh_learned_loss(self.h, int(level), ptr(om), ptr(sm), int(S), int(nu), C.c_uint64(seed),
                                     C.byref(loss), ptr(g)))
        return (loss.value, g) if grad else loss.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dgh_solver_destroy(self.h)
            self.h = None
