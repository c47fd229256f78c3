"""Benchmark of the octree FEM solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2] [--no-cpu] ...

Headline workload (default, "c3"): octree from 8M seeded points (4M U(0,1)^3 with seed 3 and
4M in 8 Gaussian clusters of sd 0.02 with centres from seed 31), max depth 14, 2:1 balanced;
Helmholtz operator (Jacobian1, lambda = mu = 1); 8-level GMG-PCG, damped Jacobi omega = 2/3
with 2+2 sweeps, b ~ U[-1,1] (seed 12).  A step is one FP64 operator application over the
whole finest mesh (44.3M elements, 31.5M nodes); the timed region is K steps with x already
in HBM.  The line also carries the PCG time to a 1e-8 true relative residual on the same
octree (plain FP64, FP32 V-cycle, FP32 V-cycle with learned damping), the roofline of the
MatVec kernel and the CPU oracle timed on a bounded sample of the same MatVec (one thread
and all host threads).

Secondary workload ("c2", also run by default under the key "c2"): octree from 1M log-normal
points, max depth 12, kappa_e = 10^U(-1,1), zero mass with the constant null space projected,
5 levels; its MatVec, PCG, Dirichlet-masked MatVec and the 32 right-hand-side batch.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PCG s to 1e-8 rel resid; MatVec GDoF/s + % HBM peak; vs plain FP64 GPU"
HBM_PEAK_GBS = 8000.0
AUTOTUNE_MAX_ELEMS = 8_000_000
LARGE_TILING = (8192, 1024)   # elements per tile (cap), threads per tile, above AUTOTUNE_MAX_ELEMS

WORKLOADS = {
    "c3": dict(name="mixed8M_d14", kind="MIXED", npts=8000000, seed=3, cluster_seed=31, depth=14, nlevels=8,
               coef="helmholtz", null_space=0, rhs_seed=12,
               desc="Helmholtz (Jacobian1, lambda=mu=1) MatVec + 8-level GMG-PCG, octree from 8M points "
                    "(4M uniform + 4M in 8 Gaussian clusters), max depth 14, FP64"),
    "c5": dict(name="clusters4M_d13", kind="CLUSTERS", npts=4000000, seed=4, cluster_seed=51, depth=13, nlevels=8,
               coef="contrast1e3", null_space=1, rhs_seed=22,
               desc="Jacobian2 PCG with learned damping, octree from 4M points in 8 Gaussian clusters, max depth "
                    "13, log-normal kappa with contrast 1e3, zero mass, constant null space, FP32 V-cycle"),
    "c2": dict(name="lognormal1M_d12", kind="LOGNORMAL", npts=1000000, seed=2, cluster_seed=0, depth=12,
               nlevels=5, coef="lognormal", null_space=1, rhs_seed=22,
               desc="Jacobian2 MatVec + 5-level GMG-PCG, octree from 1M log-normal points, max depth 12, "
                    "kappa=10^U(-1,1), zero mass, constant null space, FP64"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Problem:
    """A DA hierarchy with its coefficients, device meshes and transfers."""

    def __init__(self, w):
        import dendro_amd as da
        self.w = w
        t0 = time.time()
        pts = da.points(getattr(da, "PTS_" + w["kind"]), w["npts"], w["seed"], w["cluster_seed"])
        self.damg = da.DAMG(da.Octree.from_points(pts, w["depth"]), w["nlevels"], 500)
        del pts
        self.levels = [self.damg.level(l) for l in range(self.damg.nlevels)]
        F = self.levels[-1]
        if w["coef"] in ("lognormal", "contrast1e3"):
            if w["coef"] == "lognormal":
                kf = 10.0 ** np.random.default_rng(21).uniform(-1, 1, F.n_elem)
            else:   # kappa = exp(sigma z), z ~ N(0,1) seed 52, clipped to [10^-1.5, 10^1.5]
                z = np.random.default_rng(52).standard_normal(F.n_elem)
                kf = np.clip(np.exp(np.log(10 ** 1.5) / 2.0 * z), 10 ** -1.5, 10 ** 1.5)
            self.coefs = [(self.damg.coarsen_coefficients(kf, None, l)[0], None) for l in range(self.damg.nlevels)]
        else:
            self.coefs = [(None, None)] * self.damg.nlevels
        self.host_seconds = time.time() - t0
        self.meshes = []
        for l, L in enumerate(self.levels):
            m = da.Mesh(L)
            if self.coefs[l][0] is None:
                m.set_coefficients(None, None, 1.0, 1.0)
            else:
                m.set_coefficients(self.coefs[l][0], None)
            self.meshes.append(m)
        self.trans = [da.Transfer(self.meshes[l - 1], self.meshes[l], *self.damg.transfer(l))
                      for l in range(1, len(self.levels))]
        log(f"{w['name']}: {self.damg.nlevels} levels, finest E={F.n_elem} N={F.n_node} hanging={F.n_hanging}, "
            f"coarsest N={self.levels[0].n_node}, host build {self.host_seconds:.1f}s")

    @property
    def fine(self):
        return self.levels[-1]

    @property
    def uniform(self):
        return self.coefs[-1][0] is None

    def alg_bytes(self, b_v=8):
        """Algorithmic bytes of one MatVec: u32 nlist[8] + packed stencil id/level (2 B) per
        element, x read once and y written once, plus kappa per element for Jacobian2."""
        F = self.fine
        return F.n_elem * (32 + 2) + F.n_node * 2 * b_v + (0 if self.uniform else F.n_elem * b_v)

    def kernel_name(self, block, rows):
        """The FP64 MatVec kernel instantiation for a stencil-row source (da.ROWS_*)."""
        mass = "true" if self.uniform else "false"
        uni = "true" if self.uniform else "false"
        lane = "true" if rows == 1 else "false"
        comb = "true" if rows == 2 and self.uniform else "false"
        return f"dgh::tile_op_kernel<double, double, {mass}, 0, {block}, {uni}, {lane}, {comb}>"

    def params(self, **kw):
        import dendro_amd as da
        p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=500, rtol=1e-8,
                              norm_type=da.NORM_UNPRECONDITIONED, null_space=self.w["null_space"], use_graph=1)
        for k, v in kw.items():
            setattr(p, k, v)
        return p


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(P, reps=5):
    """The oracle's MatVec (the reference's element loop restated in C, -O3) on the full
    finest mesh of the workload: one warm-up, then the median of `reps` timed calls, on one
    thread and on every host thread this process may use (at most 16, the host share of one
    GPU)."""
    import oracle
    F = P.fine
    kappa = P.coefs[-1][0]
    x = np.random.default_rng(11).uniform(-1, 1, F.n_node)
    threads = min(16, len(os.sched_getaffinity(0)))
    out = {}
    for name, fn in (("1", lambda: oracle.matvec(F, x, kappa=kappa)),
                     ("all", lambda: oracle.matvec_mt(F, x, kappa=kappa, threads=threads))):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        out[name] = sorted(ts)[len(ts) // 2]
    g1, ga = F.n_node / out["1"] / 1e9, F.n_node / out["all"] / 1e9
    log(f"CPU oracle MatVec: 1 thread {out['1']*1e3:.0f} ms ({g1:.4f} GDoF/s), {threads} threads "
        f"{out['all']*1e3:.0f} ms ({ga:.4f} GDoF/s)")
    return {"value": round(ga, 5), "unit": "GDoF/s", "cores": threads, "kind": "port",
            "single_thread_value": round(g1, 5), "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"oracle MatVec (C restatement of the reference element loop, -O3) over the full finest "
                      f"mesh (E={F.n_elem}, N={F.n_node}), median of {reps} calls after a warm-up, on 1 thread "
                      f"({out['1']:.2f} s per call) and on {threads} threads ({out['all']:.3f} s per call)"}


def cpu_table(reps=5):
    """The CPU baseline table of the CPU-runnable workloads, timed by the oracle on this host:
    the 100k Gaussian octree (max depth 8, Helmholtz): 100 MatVecs, and 3-level PCG to 1e-8
    (omega 2/3, 2+2 sweeps, b ~ U[-1,1] seed 12); the 1M log-normal octree (max depth 12,
    kappa = 10^U(-1,1), zero mass, null space): one MatVec, and 5-level PCG to 1e-10
    (b seed 22).  One thread and all host threads (at most 16); median of `reps` runs after one
    warm-up (PCG: median of 3).  Prints one JSON line."""
    import dendro_amd as da
    import oracle
    threads = min(16, len(os.sched_getaffinity(0)))

    def med(fn, n):
        fn()
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2], r

    rows = {"cpu_model": cpu_model(), "threads_all": threads}
    for name, kind, npts, seed, depth, nlev, coef, rtol, rhs_seed in (
            ("c1", da.PTS_GAUSS, 100000, 1, 8, 3, None, 1e-8, 12),
            ("c2", da.PTS_LOGNORMAL, 1000000, 2, 12, 5, "lognormal", 1e-10, 22)):
        damg = da.DAMG(da.Octree.from_points(da.points(kind, npts, seed), depth), nlev, 500)
        levels = [damg.level(l) for l in range(damg.nlevels)]
        F = levels[-1]
        if coef:
            kf = 10.0 ** np.random.default_rng(21).uniform(-1, 1, F.n_elem)
            co = [(damg.coarsen_coefficients(kf, None, l)[0], None, 1.0, 1.0) for l in range(len(levels))]
        else:
            co = [(None, None, 1.0, 1.0)] * len(levels)
        ns = coef is not None
        x = np.random.default_rng(11).uniform(-1, 1, F.n_node)
        nmv = 100 if name == "c1" else 1
        r = {"n_elem": F.n_elem, "n_node": F.n_node}
        for tag, fn in (("1", lambda: [oracle.matvec(F, x, kappa=co[-1][0]) for _ in range(nmv)]),
                        ("all", lambda: [oracle.matvec_mt(F, x, kappa=co[-1][0], threads=threads) for _ in range(nmv)])):
            t, _ = med(fn, reps)
            r[f"matvec_x{nmv}_s_{tag}"] = round(t, 4)
            log(f"{name} {nmv} MatVec(s), {tag} thread(s): {t:.3f} s")
        b = np.random.default_rng(rhs_seed).uniform(-1, 1, F.n_node)
        for tag, nt in (("1", 1), ("all", threads)):
            mg = oracle.MG([(L, *co[l]) for l, L in enumerate(levels)], null_space=ns, threads=nt)
            t, res = med(lambda: mg.pcg(b, rtol=rtol, max_it=500, norm_type=0), 3)
            r[f"pcg_s_{tag}"] = round(t, 3)
            r["pcg_iterations"] = int(res[1])
            log(f"{name} PCG to {rtol:g}, {tag} thread(s) for the MatVecs: {t:.2f} s, {res[1]} its")
        rows[name] = r
    print(json.dumps({"cpu_table": rows}), flush=True)


def time_matvec(mesh, x, y, steps, warmup, stream, prec=None):
    import torch
    import dendro_amd as da
    prec = da.F64 if prec is None else prec
    for _ in range(warmup):
        mesh.matvec(x, y, prec=prec, stream=stream)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        mesh.matvec(x, y, prec=prec, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps, time.perf_counter() - t0


def cold_matvec_ms(mesh, x, y, stream):
    """Each launch alone after reading a 1 GiB buffer, which evicts the 256 MB Infinity
    Cache (a read, not a fill: a fill's dirty lines would be written back inside the timed
    launch)."""
    import torch
    flush = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device=x.device)
    for _ in range(2):
        flush.sum()
    cold = []
    for it in range(12):
        flush.sum()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        mesh.matvec(x, y, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        if it >= 2:
            cold.append(e0.elapsed_time(e1))
    del flush
    return sorted(cold)[len(cold) // 2]


def bench_pcg(P, dev, learned=True):
    """PCG to a 1e-8 true relative residual: plain FP64, FP32 V-cycle, and FP32 V-cycle with
    the learned damping (network trained on the device on 200 batches of solver-generated
    samples, seeds 5000..5199; each level keeps it only where it beats plain Jacobi on fresh
    samples).  The second of two identical solves is reported (the first captures the
    hipGraph)."""
    import torch
    import dendro_amd as da
    N = P.fine.n_node
    b = torch.tensor(np.random.default_rng(P.w["rhs_seed"]).uniform(-1, 1, N), dtype=torch.float64, device=dev)
    out = {}
    xs = {}
    for name, prec in (("fp64", da.F64), ("mixed", da.F32)):
        s = da.Solver(P.meshes, P.trans, P.params(vcycle_prec=prec))
        x = torch.zeros(N, dtype=torch.float64, device=dev)
        s.solve(b, x)
        st, _ = s.solve(b, x)
        out[f"pcg_{name}"] = {"seconds": round(st.seconds, 5), "iterations": st.iterations,
                              "rel_true": st.rel_true, "converged": bool(st.converged)}
        xs[name] = x
        log(f"PCG {name}: {st.iterations} its, {st.seconds*1e3:.2f} ms, rel_true {st.rel_true:.2e}")
        del s
    out["pcg_mixed"]["solution_rel_diff_vs_fp64"] = float(torch.linalg.norm(xs["mixed"] - xs["fp64"]) /
                                                          torch.linalg.norm(xs["fp64"]))
    out["pcg_speedup_mixed_vs_fp64"] = round(out["pcg_fp64"]["seconds"] / out["pcg_mixed"]["seconds"], 3)
    if learned:
        s = da.Solver(P.meshes, P.trans, P.params(vcycle_prec=da.F32, smoother=da.SMOOTHER_LEARNED))
        s.learned_init(1)
        tr = s.learned_train(steps=200, samples=8, nu=2, refresh=1, lr=3e-3, seed=5000)
        mask = s.learned_apply(True, samples=8, seed=6000)
        x = torch.zeros(N, dtype=torch.float64, device=dev)
        s.solve(b, x)
        st, _ = s.solve(b, x)
        out["pcg_mixed_learned"] = {
            "seconds": round(st.seconds, 5), "iterations": st.iterations, "rel_true": st.rel_true,
            "converged": bool(st.converged), "levels_learned_mask": mask, "train_seconds": round(tr.seconds, 3),
            "objective_first": round(tr.loss_first, 5), "objective_last": round(tr.loss_last, 5),
            "speedup_vs_fp64": round(out["pcg_fp64"]["seconds"] / st.seconds, 3)}
        log(f"PCG mixed+learned: {st.iterations} its, {st.seconds*1e3:.2f} ms (levels {mask:#x}); training "
            f"{tr.seconds:.2f}s objective {tr.loss_first:.4f} -> {tr.loss_last:.4f}")
        del s
    return out


def bench_c5(dev, args):
    """Config 5: learned damping against plain Jacobi on the 4M cluster octree (kappa contrast
    1e3), FP32 V-cycle, PCG to 1e-8: iterations and time of both, training time."""
    import torch
    import dendro_amd as da
    P = Problem(WORKLOADS["c5"])
    choose_tilings(P, argparse.Namespace(tile=0, block=0, autotune_all=False))
    N = P.fine.n_node
    b = torch.tensor(np.random.default_rng(22).uniform(-1, 1, N), dtype=torch.float64, device=dev)
    out = {"workload": P.w["desc"], "n_elem": P.fine.n_elem, "n_node": N}
    for name, kw in (("jacobi", {}), ("learned", {"smoother": da.SMOOTHER_LEARNED})):
        s = da.Solver(P.meshes, P.trans, P.params(vcycle_prec=da.F32, **kw))
        extra = {}
        if name == "learned":
            s.learned_init(1)
            tr = s.learned_train(steps=200, samples=8, nu=2, refresh=1, lr=3e-3, seed=5000)
            extra = {"levels_learned_mask": s.learned_apply(True, samples=8, seed=6000),
                     "train_seconds": round(tr.seconds, 2), "objective_first": round(tr.loss_first, 5),
                     "objective_last": round(tr.loss_last, 5)}
        x = torch.zeros_like(b)
        s.solve(b, x)
        st, _ = s.solve(b, x)
        out[name] = {"seconds": round(st.seconds, 5), "iterations": st.iterations, "rel_true": st.rel_true,
                     "converged": bool(st.converged), **extra}
        log(f"c5 {name}: {st.iterations} its, {st.seconds*1e3:.2f} ms {extra}")
        del s
    return out


def calib_copy(dev):
    """Three 256 MiB device-to-device copies: the reference kernel of known bytes for scaling
    the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev)
    c = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        c.copy_(a)
    torch.cuda.synchronize()
    del a, c


def bench_multi(P, dev, stream, steps, solve=True, tile=0, variant=0):
    """32 right-hand sides (seeds 400..431): the batched FP32 operator and the batched PCG
    (one FP32 V-cycle per iteration for all columns, FP64 outer CG, one hipGraph).  tile > 0
    fixes the finest level's batched tiling and product variant instead of autotuning
    (profiler runs); solve=False times the operator only."""
    import torch
    import dendro_amd as da
    k = 32
    N = P.fine.n_node
    if tile:
        P.meshes[-1].set_tile_multi(tile, variant)
        tuned = [{"tile_elems": tile, "variant": variant}]
    else:
        tuned = [m.autotune_multi(k) for m in (P.meshes[1:] if solve else P.meshes[-1:])]
    mf = P.meshes[-1]
    X = torch.tensor(np.random.default_rng(11).uniform(-1, 1, (N, k)), dtype=torch.float32, device=dev)
    Y = torch.empty_like(X)
    for _ in range(5):
        mf.matvec_multi(X, Y, k, prec=da.F32, stream=stream)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(10, steps // 10)
    ev0.record(stream)
    for _ in range(reps):
        mf.matvec_multi(X, Y, k, prec=da.F32, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    del X, Y
    F = P.fine
    alg = F.n_elem * 34 + k * N * 2 * 4 + (0 if P.uniform else F.n_elem * 4)
    out = {"k": k, "matvec_f32_us": round(ms * 1e3, 2), "matvec_f32_gdofs": round(N * k / (ms * 1e-3) / 1e9, 3),
           "matvec_f32_hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "tile_elems": [t["tile_elems"] for t in tuned], "variants": [t["variant"] for t in tuned]}
    log(f"batched k={k}: MatVec FP32 {ms*1e3:.1f} us ({out['matvec_f32_gdofs']} GDoF/s, "
        f"{100*out['matvec_f32_hbm_frac']:.1f}% HBM), tiling {out['tile_elems']} variants {out['variants']}")
    if not solve:
        return out
    B = torch.tensor(np.stack([np.random.default_rng(400 + j).uniform(-1, 1, N) for j in range(k)], 1),
                     dtype=torch.float64, device=dev)
    Xs = torch.zeros_like(B)
    s = da.Solver(P.meshes, P.trans, P.params(vcycle_prec=da.F32))
    s.solve_multi(B, Xs, k)
    st = s.solve_multi(B, Xs, k)
    secs = st[0].seconds
    out.update({"solve_seconds": round(secs, 5), "seconds_per_rhs": round(secs / k, 6),
                "iterations": [x.iterations for x in st], "max_rel_true": max(x.rel_true for x in st),
                "all_converged": all(x.converged for x in st)})
    log(f"batched k={k}: PCG {secs*1e3:.1f} ms for {k} RHS, its "
        f"{min(out['iterations'])}-{max(out['iterations'])}, max rel_true {out['max_rel_true']:.2e}")
    del s
    return out


def measured_traffic(workload, kernel, tile_elems, block):
    """HBM bytes per FP64 MatVec launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes over this workload (already scaled by the copy-kernel calibration), when they were
    taken with the same kernel and tiling; None otherwise."""
    path = os.path.join(ROOT, "profiles", f"tile_op_matvec_f64_{workload}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel or t.get("tile_elems") != tile_elems or t.get("block") != block:
        return None
    return t.get("bytes_per_launch")


def choose_tilings(P, args):
    """Operator tilings of every level; returns the finest level's choice.  The autotuner
    rebuilds the host tiling once per candidate: seconds per candidate on meshes of tens of
    millions of elements.  Above AUTOTUNE_MAX_ELEMS (unless --autotune-all) a level takes the
    fastest tiling of a sweep on the 44.3M-element finest level of the 8M mixed octree
    (at most 8192 elements x 1024 threads per tile, 1.276 records per element: 570 us per
    FP64 MatVec there, against 624 us for 4096 x 512)."""
    import dendro_amd as da
    mf = P.meshes[-1]
    if args.tile:
        for l, m in enumerate(P.meshes[:-1]):
            if getattr(args, "matvec_only", False):
                break   # profiler runs: the finest level only
            m.autotune(da.F64) if P.levels[l].n_elem <= AUTOTUNE_MAX_ELEMS else m.set_tile(*LARGE_TILING)
        mf.set_tile(args.tile, args.block)
        return {"tile_elems": args.tile, "block": args.block or 512}
    t0 = time.time()
    for l, m in enumerate(P.meshes):
        if P.levels[l].n_elem > AUTOTUNE_MAX_ELEMS and not args.autotune_all:
            m.set_tile(*LARGE_TILING)
            tuned = {"tile_elems": LARGE_TILING[0], "block": LARGE_TILING[1], "fixed": True}
        else:
            tuned = m.autotune(da.F64)
        log(f"  level {l} (E={P.levels[l].n_elem}): {tuned}  [{time.time() - t0:.1f}s]")
    log(f"tilings chosen in {time.time() - t0:.1f}s: finest {tuned}")
    return tuned


def matvec_section(P, args, dev, stream):
    """Autotuned FP64 MatVec: K timed steps, cold-cache launches, FP32 rate, roofline."""
    import torch
    import dendro_amd as da
    F = P.fine
    mf = P.meshes[-1]
    tuned = choose_tilings(P, args)
    ti = mf.tile_info()
    log(f"tiling: {ti}")
    x = torch.tensor(np.random.default_rng(11).uniform(-1, 1, F.n_node), dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    ms, wall = time_matvec(mf, x, y, args.steps, args.warmup, stream)
    alg = P.alg_bytes()
    achieved = alg / (ms * 1e-3) / 1e9
    log(f"MatVec FP64: {ms*1e3:.1f} us/launch, {F.n_node / (ms * 1e-3) / 1e9:.2f} GDoF/s, {achieved:.0f} GB/s "
        f"algorithmic ({100*achieved/HBM_PEAK_GBS:.1f}% of HBM peak); wall {wall:.3f}s")
    cold = cold_matvec_ms(mf, x, y, stream)
    log(f"MatVec FP64 after an Infinity Cache flush: {cold*1e3:.1f} us "
        f"({100 * alg / (cold * 1e-3) / 1e9 / HBM_PEAK_GBS:.1f}% of HBM peak)")
    if args.calib_copy:
        calib_copy(dev)
    x32, y32 = x.float(), torch.empty(F.n_node, dtype=torch.float32, device=dev)
    ms32, _ = time_matvec(mf, x32, y32, args.steps, args.warmup, stream, prec=da.F32)
    del x32, y32
    kernel = P.kernel_name(tuned["block"], ti["lane_stencils"])
    res = {
        "ms": ms, "gdofs": F.n_node / (ms * 1e-3) / 1e9, "achieved": achieved, "alg": alg,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(P.w["name"], kernel, ti["tile_elems"], tuned["block"]),
                     "algorithmic_bytes_per_launch": alg, "kernel": kernel,
                     "tile_elems": ti["tile_elems"], "block": tuned["block"], "n_tiles": ti["n_tiles"],
                     "n_records": ti["n_records"], "cold_cache_us": round(cold * 1e3, 2),
                     "cold_cache_frac": round(alg / (cold * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "matvec_fp32_us": round(ms32 * 1e3, 2),
        "matvec_fp32_gdofs": round(F.n_node / (ms32 * 1e-3) / 1e9, 4),
        "x": x, "y": y,
    }
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pcg", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-damping PCG")
    ap.add_argument("--no-c2", action="store_true", help="skip the secondary 1M log-normal workload")
    ap.add_argument("--no-c5", action="store_true", help="skip the 4M cluster learned-damping comparison")
    ap.add_argument("--no-multi", action="store_true", help="skip the 32 right-hand-side batch (secondary)")
    ap.add_argument("--no-dirichlet", action="store_true", help="skip the Dirichlet-masked MatVec (secondary)")
    ap.add_argument("--matvec-only", action="store_true",
                    help="only the headline MatVec (for profiler runs): implies --no-cpu --no-pcg --no-c2")
    ap.add_argument("--multi-only", action="store_true",
                    help="only the batched FP32 MatVec (32 columns) on the 1M log-normal octree (profiler runs)")
    ap.add_argument("--multi-tile", type=int, default=0,
                    help="with --multi-only: batched tile size (0: autotune), with --multi-variant")
    ap.add_argument("--multi-variant", type=int, default=4, help="batched FP32 product variant with --multi-tile")
    ap.add_argument("--multi-trace", action="store_true",
                    help="only three eager (no hipGraph) 32 right-hand-side PCG solves on the 1M log-normal "
                         "octree, for a kernel trace")
    ap.add_argument("--tile", type=int, default=0, help="operator tile size (0: autotune)")
    ap.add_argument("--block", type=int, default=0, help="workgroup size with --tile (0: library default)")
    ap.add_argument("--autotune-all", action="store_true",
                    help=f"autotune levels above {AUTOTUNE_MAX_ELEMS} elements too (minutes of host time)")
    ap.add_argument("--pcg-trace", action="store_true",
                    help="only three eager (no hipGraph) mixed-precision PCG solves, for a kernel trace")
    ap.add_argument("--cpu-table", action="store_true",
                    help="only the CPU baseline table of the 100k Gaussian and 1M log-normal workloads (no GPU)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 256 MiB device copy (reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the solve path is "
              "not sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    if args.matvec_only:
        args.no_cpu = args.no_pcg = args.no_c2 = args.no_c5 = True
    if args.cpu_table:
        cpu_table()
        return

    import torch
    import dendro_amd as da

    dev = torch.device("cuda:0")
    da.init(0)
    stream = torch.cuda.current_stream()
    if args.multi_trace:   # batched solves on the 1M log-normal octree, graph capture off
        Q = Problem(WORKLOADS["c2"])
        tuned = [m.autotune_multi(32) for m in Q.meshes[1:]]
        log("batched tilings: " + ", ".join(f"{t['tile_elems']}/v{t['variant']}" for t in tuned))
        k, N = 32, Q.fine.n_node
        B = torch.tensor(np.stack([np.random.default_rng(400 + j).uniform(-1, 1, N) for j in range(k)], 1),
                         dtype=torch.float64, device=dev)
        X = torch.zeros_like(B)
        s = da.Solver(Q.meshes, Q.trans, Q.params(vcycle_prec=da.F32, use_graph=0))
        for _ in range(3):
            st = s.solve_multi(B, X, k)
            log(f"eager batched PCG: {st[0].iterations} its, {st[0].seconds * 1e3:.2f} ms")
        return
    if args.multi_only:   # the batched operator runs on the 1M log-normal octree whatever --workload
        Q = Problem(WORKLOADS["c2"])
        r = bench_multi(Q, dev, stream, args.steps, solve=False, tile=args.multi_tile, variant=args.multi_variant)
        if args.calib_copy:
            calib_copy(dev)
        print(json.dumps({"workload": Q.w["desc"], "multi_rhs": r}), flush=True)
        return
    P = Problem(WORKLOADS[args.workload])
    F = P.fine

    if args.pcg_trace:
        choose_tilings(P, args)
        log("levels: " + ", ".join(f"{l}: E={L.n_elem} {P.meshes[l].tile_info()}" for l, L in enumerate(P.levels)))
        b = torch.tensor(np.random.default_rng(P.w["rhs_seed"]).uniform(-1, 1, F.n_node), dtype=torch.float64,
                         device=dev)
        s = da.Solver(P.meshes, P.trans, P.params(vcycle_prec=da.F32, use_graph=0))
        xs = torch.zeros(F.n_node, dtype=torch.float64, device=dev)
        for _ in range(3):
            st, _ = s.solve(b, xs)
            log(f"eager mixed PCG: {st.iterations} its, {st.seconds * 1e3:.2f} ms")
        return

    mv = matvec_section(P, args, dev, stream)
    del mv["x"], mv["y"]
    out = {
        "metric": METRIC,
        "value": round(mv["gdofs"], 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mv["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": P.w["desc"], "name": P.w["name"], "n_elem": F.n_elem, "n_node": F.n_node,
                   "n_levels": len(P.levels), "n_hanging_elem": F.n_hanging,
                   "coarsest_nodes": P.levels[0].n_node, "host_build_seconds": round(P.host_seconds, 1)},
        "roofline": mv["roofline"],
        "matvec_fp32_us": mv["matvec_fp32_us"],
        "matvec_fp32_gdofs": mv["matvec_fp32_gdofs"],
    }
    if not args.no_pcg:
        log("PCG on the headline octree ...")
        out.update(bench_pcg(P, dev, learned=not args.no_learned))

    if not args.no_c2 and args.workload != "c2":
        Q = Problem(WORKLOADS["c2"])
        q = matvec_section(Q, argparse.Namespace(**{**vars(args), "tile": 0, "calib_copy": False}), dev, stream)
        c2 = {"workload": Q.w["desc"], "n_elem": Q.fine.n_elem, "n_node": Q.fine.n_node,
              "matvec_f64_us": round(q["ms"] * 1e3, 2), "matvec_f64_gdofs": round(q["gdofs"], 4),
              "hbm_frac": q["roofline"]["frac"], "cold_cache_us": q["roofline"]["cold_cache_us"],
              "cold_cache_frac": q["roofline"]["cold_cache_frac"], "matvec_fp32_us": q["matvec_fp32_us"]}
        if not args.no_dirichlet:
            bdy = Q.damg.boundary(Q.damg.nlevels - 1)
            mq = Q.meshes[-1]
            mq.set_boundary(bdy)
            for _ in range(args.warmup):
                mq.matvec_dirichlet(q["x"], q["y"], stream=stream)
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(args.steps):
                mq.matvec_dirichlet(q["x"], q["y"], stream=stream)
            ev1.record(stream)
            torch.cuda.synchronize()
            msd = ev0.elapsed_time(ev1) / args.steps
            mq.set_boundary(None)
            log(f"c2 Dirichlet MatVec FP64: {msd*1e3:.1f} us ({int(bdy.sum())} boundary nodes)")
            c2["matvec_dirichlet_f64_us"] = round(msd * 1e3, 2)
            c2["n_boundary_nodes"] = int(bdy.sum())
        del q["x"], q["y"]
        if not args.no_pcg:
            c2.update(bench_pcg(Q, dev, learned=False))
        if not args.no_multi:
            c2["multi_rhs"] = bench_multi(Q, dev, stream, args.steps)
        out["c2"] = c2
        del Q

    if not args.no_c5 and not args.no_pcg:
        out["c5"] = bench_c5(dev, args)

    if not args.no_cpu:
        log("CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(P)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
