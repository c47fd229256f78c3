"""Benchmark of the octree FEM solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (one GPU): octree from 1M seeded log-normal points (per-axis lognormal sigma=1,
rescaled to the unit cube), max depth 12, 2:1 balanced; variable-coefficient Poisson with
element coefficient kappa_e = 10^U(-1,1) (seed 21), zero mass (constant null space
projected); 5-level GMG-PCG, damped Jacobi omega=2/3 with 2+2 sweeps, b ~ U[-1,1] (seed 22).

A step is one FP64 operator application (Jacobian2 MatVec) over the whole finest mesh; the
timed region is K steps with the input already in HBM.  The line also carries the PCG
time to a 1e-8 true relative residual (plain FP64 and FP32 V-cycle), the roofline of the
MatVec kernel and the CPU oracle timed on a bounded sample of the same MatVec.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PCG s to 1e-8 rel resid; MatVec GDoF/s + % HBM peak; vs plain FP64 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(npts, depth, nlevels):
    import dendro_amd as da
    t0 = time.time()
    pts = da.points(da.PTS_LOGNORMAL, npts, 2)
    damg = da.DAMG(da.Octree.from_points(pts, depth), nlevels, 500)
    levels = [damg.level(l) for l in range(damg.nlevels)]
    kf = 10.0 ** np.random.default_rng(21).uniform(-1, 1, levels[-1].n_elem)
    coefs = [damg.coarsen_coefficients(kf, None, l) for l in range(damg.nlevels)]
    trans = [damg.transfer(l) for l in range(1, damg.nlevels)]
    log(f"mesh: {damg.nlevels} levels, finest E={levels[-1].n_elem} N={levels[-1].n_node} "
        f"hanging={levels[-1].n_hanging}, coarsest N={levels[0].n_node}, setup {time.time() - t0:.1f}s")
    return damg, levels, coefs, trans


def cpu_baseline(F, kappa, budget_s=15.0):
    import oracle
    x = np.random.default_rng(11).uniform(-1, 1, F.n_node)
    oracle.matvec(F, x, kappa=kappa, mass=None)          # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        oracle.matvec(F, x, kappa=kappa, mass=None)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or n >= 200:
            break
    return {"value": F.n_node * n / dt / 1e9, "unit": "GDoF/s", "cores": 1, "kind": "port",
            "sample": f"{n} single-thread oracle MatVecs (Jacobian2) of the full finest mesh "
                      f"(E={F.n_elem}, N={F.n_node}), {dt:.1f}s"}


def bench_learned(meshes, trans, N, dev, extra):
    """FP32 V-cycle PCG with the learned damping: the network is trained on the device on
    200 batches of solver-generated samples (seeds 5000..5199), then every level keeps the
    learned weights only where they beat plain Jacobi on fresh samples."""
    import torch
    import dendro_amd as da
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=500, rtol=1e-8,
                          norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=da.F32,
                          use_graph=1, smoother=da.SMOOTHER_LEARNED)
    s = da.Solver(meshes, trans, p)
    s.learned_init(1)
    tr = s.learned_train(steps=200, samples=8, nu=2, refresh=1, lr=3e-3, seed=5000)
    mask = s.learned_apply(True, samples=8, seed=6000)
    b = torch.tensor(np.random.default_rng(22).uniform(-1, 1, N), dtype=torch.float64, device=dev)
    xs = torch.zeros(N, dtype=torch.float64, device=dev)
    s.solve(b, xs)
    st, _ = s.solve(b, xs)
    out = {"seconds": round(st.seconds, 5), "iterations": st.iterations, "rel_true": st.rel_true,
           "converged": bool(st.converged), "levels_learned_mask": mask,
           "train_seconds": round(tr.seconds, 3), "objective_first": round(tr.loss_first, 5),
           "objective_last": round(tr.loss_last, 5)}
    if "pcg_fp64" in extra:
        out["speedup_vs_fp64"] = round(extra["pcg_fp64"]["seconds"] / st.seconds, 3)
    log(f"PCG mixed+learned: {st.iterations} its, {st.seconds*1e3:.1f} ms (levels {mask:#x}); training "
        f"{tr.seconds:.2f}s objective {tr.loss_first:.4f} -> {tr.loss_last:.4f}")
    del s
    return out


def bench_multi(args, meshes, trans, N, dev, stream, extra):
    """32 right-hand sides (seeds 400..431): the batched FP32 operator and the batched
    PCG (one FP32 V-cycle per iteration for all columns, FP64 outer CG, one hipGraph)."""
    import torch
    import dendro_amd as da
    k = 32
    tuned = [m.autotune_multi(k) for m in meshes[1:]]
    mf = meshes[-1]
    log(f"batched tiles (autotuned, k={k}): {[t['tile_elems'] for t in tuned]}")
    X = torch.tensor(np.random.default_rng(11).uniform(-1, 1, (N, k)), dtype=torch.float32, device=dev)
    Y = torch.empty_like(X)
    for _ in range(5):
        mf.matvec_multi(X, Y, k, prec=da.F32, stream=stream)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(10, args.steps // 10)
    ev0.record(stream)
    for _ in range(reps):
        mf.matvec_multi(X, Y, k, prec=da.F32, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    del X, Y
    B = torch.tensor(np.stack([np.random.default_rng(400 + j).uniform(-1, 1, N) for j in range(k)], 1),
                     dtype=torch.float64, device=dev)
    Xs = torch.zeros_like(B)
    p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=500, rtol=1e-8,
                          norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=da.F32, use_graph=1)
    s = da.Solver(meshes, trans, p)
    s.solve_multi(B, Xs, k)                    # warm-up (buffers, graph)
    st = s.solve_multi(B, Xs, k)
    secs = st[0].seconds
    single = extra.get("pcg_mixed", {}).get("seconds")
    out = {"k": k, "matvec_f32_us": round(ms * 1e3, 2), "matvec_f32_gdofs": round(N * k / (ms * 1e-3) / 1e9, 3),
           "tile_elems": [t["tile_elems"] for t in tuned], "solve_seconds": round(secs, 5),
           "seconds_per_rhs": round(secs / k, 6), "iterations": [x.iterations for x in st],
           "max_rel_true": max(x.rel_true for x in st), "all_converged": all(x.converged for x in st)}
    if single:
        out["speedup_per_rhs_vs_single_mixed"] = round(single / (secs / k), 2)
    log(f"batched k={k}: MatVec FP32 {ms*1e3:.1f} us ({out['matvec_f32_gdofs']} GDoF/s), "
        f"PCG {secs*1e3:.1f} ms for {k} RHS, its {min(out['iterations'])}-{max(out['iterations'])}, "
        f"max rel_true {out['max_rel_true']:.2e}")
    del s
    return out


def measured_traffic(kernel, tile_elems, block):
    """HBM bytes per FP64 MatVec launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes over this workload (already scaled by the copy-kernel calibration), when they were
    taken with the same kernel and tiling; None otherwise."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "tile_op_matvec_f64_lognormal1M_d12_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel or t.get("tile_elems") != tile_elems or t.get("block") != block:
        return None
    return t.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--npts", type=int, default=1000000)
    ap.add_argument("--depth", type=int, default=12)
    ap.add_argument("--nlevels", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pcg", action="store_true")
    ap.add_argument("--tile", type=int, default=0, help="operator tile size (0: autotune)")
    ap.add_argument("--block", type=int, default=0, help="workgroup size with --tile (0: library default)")
    ap.add_argument("--sweep-tiles", action="store_true", help="time the MatVec for several tile sizes")
    ap.add_argument("--no-multi", action="store_true", help="skip the 32 right-hand-side batch")
    ap.add_argument("--no-dirichlet", action="store_true", help="skip the Dirichlet-masked MatVec")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-damping PCG")
    ap.add_argument("--pcg-trace", action="store_true",
                    help="only three eager (no hipGraph) mixed-precision PCG solves, for a kernel trace")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 256 MiB device copy (reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the solve path is "
              "not sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)

    import torch
    import dendro_amd as da

    dev = torch.device("cuda:0")
    da.init(0)
    damg, levels, coefs, trans_arr = build_problem(args.npts, args.depth, args.nlevels)
    F = levels[-1]
    E, N = F.n_elem, F.n_node
    meshes = []
    for l, L in enumerate(levels):
        m = da.Mesh(L)
        m.set_coefficients(coefs[l][0], None)
        meshes.append(m)
    trans = [da.Transfer(meshes[l - 1], meshes[l], *trans_arr[l - 1]) for l in range(1, len(levels))]
    mf = meshes[-1]
    tuned = {"tile_elems": args.tile, "block": args.block or 512}
    if args.tile:
        mf.set_tile(args.tile, args.block)
    else:
        t0 = time.time()
        for m in meshes:
            tuned = m.autotune(da.F64)
        log(f"autotuned tilings in {time.time() - t0:.1f}s: finest {tuned} {mf.tile_info()}")
    log(f"tiling: {mf.tile_info()}")
    if args.pcg_trace:
        log("levels: " + ", ".join(f"{l}: E={L.n_elem} {meshes[l].tile_info()}" for l, L in enumerate(levels)))
        b = torch.tensor(np.random.default_rng(22).uniform(-1, 1, N), dtype=torch.float64, device=dev)
        p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=500, rtol=1e-8,
                              norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=da.F32,
                              use_graph=0)
        s = da.Solver(meshes, trans, p)
        xs = torch.zeros(N, dtype=torch.float64, device=dev)
        for _ in range(3):
            st, _ = s.solve(b, xs)
            log(f"eager mixed PCG: {st.iterations} its, {st.seconds * 1e3:.2f} ms")
        return

    # ---- MatVec: K steps, inputs resident in HBM, events on the launch stream
    stream = torch.cuda.current_stream()
    x = torch.tensor(np.random.default_rng(11).uniform(-1, 1, N), dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        mf.matvec(x, y, stream=stream)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        mf.matvec(x, y, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    ms_per = ms_total / args.steps
    gdofs = N / (ms_per * 1e-3) / 1e9
    alg_bytes = E * (32 + 2) + N * 2 * 8 + E * 8     # nlist + stencil/level, x read + y written, kappa
    achieved = alg_bytes / (ms_per * 1e-3) / 1e9
    log(f"MatVec FP64: {ms_per*1e3:.1f} us/launch, {gdofs:.2f} GDoF/s, {achieved:.0f} GB/s algorithmic "
        f"({100*achieved/HBM_PEAK_GBS:.1f}% of HBM peak); wall {wall:.3f}s")
    # The operator's data (about 110 MB) fits in the 256 MB Infinity Cache, so back-to-back
    # launches read much of it from there.  Cold: reading a 1 GiB buffer between launches
    # evicts it and each launch is timed alone.  The flush reads rather than writes: a fill
    # would leave 1 GiB of dirty lines whose write-back lands inside the timed launch (a plain
    # 110 MiB read then takes 40 us instead of 21).
    flush = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device=dev)
    for _ in range(2):
        flush.sum()
    cold = []
    for it in range(12):
        flush.sum()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        mf.matvec(x, y, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        if it >= 2:
            cold.append(e0.elapsed_time(e1))
    del flush
    cold_ms = sorted(cold)[len(cold) // 2]
    log(f"MatVec FP64 after an Infinity Cache flush: {cold_ms*1e3:.1f} us ({alg_bytes / (cold_ms * 1e-3) / 1e9:.0f} GB/s "
        f"algorithmic, {100 * alg_bytes / (cold_ms * 1e-3) / 1e9 / HBM_PEAK_GBS:.1f}% of HBM peak)")

    if args.sweep_tiles:
        for te, blk in [(t_, b_) for t_ in (512, 1024, 2048, 4096) for b_ in (256, 512, 1024)]:
            try:
                mf.set_tile(te, blk)
            except da.DendroError as e:
                log(f"tile {te}: {e}")
                continue
            for _ in range(5):
                mf.matvec(x, y, stream=stream)
            ev0.record(stream)
            for _ in range(50):
                mf.matvec(x, y, stream=stream)
            ev1.record(stream)
            torch.cuda.synchronize()
            ti = mf.tile_info()
            log(f"tile {te} block {blk}: {ev0.elapsed_time(ev1) / 50 * 1e3:.1f} us  {ti}")
        mf.autotune(da.F64)

    if args.calib_copy:
        a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        c = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()

    # FP32 MatVec rate (V-cycle precision of the mixed solver)
    x32, y32 = x.float(), torch.empty(N, dtype=torch.float32, device=dev)
    for _ in range(args.warmup):
        mf.matvec(x32, y32, prec=da.F32, stream=stream)
    ev0.record(stream)
    for _ in range(args.steps):
        mf.matvec(x32, y32, prec=da.F32, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms32 = ev0.elapsed_time(ev1) / args.steps

    # Dirichlet-masked FP64 MatVec on the same mesh: boundary columns dropped, identity rows
    # (a mask pass over x, the operator, and the boundary rows)
    dirichlet = {}
    if not args.no_dirichlet:
        bdy = damg.boundary(damg.nlevels - 1)
        mf.set_boundary(bdy)
        for _ in range(args.warmup):
            mf.matvec_dirichlet(x, y, stream=stream)
        ev0.record(stream)
        for _ in range(args.steps):
            mf.matvec_dirichlet(x, y, stream=stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        msd = ev0.elapsed_time(ev1) / args.steps
        log(f"Dirichlet MatVec FP64: {msd*1e3:.1f} us/call ({int(bdy.sum())} boundary nodes), "
            f"{N / (msd * 1e-3) / 1e9:.2f} GDoF/s")
        dirichlet = {"matvec_dirichlet_f64_us": round(msd * 1e3, 2), "n_boundary_nodes": int(bdy.sum())}

    extra = {}
    if not args.no_pcg:
        b = torch.tensor(np.random.default_rng(22).uniform(-1, 1, N), dtype=torch.float64, device=dev)
        for name, prec in (("fp64", da.F64), ("mixed", da.F32)):
            p = da.default_params(omega=2.0 / 3.0, nu_pre=2, nu_post=2, max_it=500, rtol=1e-8,
                                  norm_type=da.NORM_UNPRECONDITIONED, null_space=1, vcycle_prec=prec,
                                  use_graph=1)
            s = da.Solver(meshes, trans, p)
            xs = torch.zeros(N, dtype=torch.float64, device=dev)
            s.solve(b, xs)                                   # warm-up solve
            st, _ = s.solve(b, xs)
            extra[f"pcg_{name}"] = {"seconds": round(st.seconds, 5), "iterations": st.iterations,
                                    "rel_true": st.rel_true, "converged": bool(st.converged)}
            log(f"PCG {name}: {st.iterations} its, {st.seconds*1e3:.1f} ms, rel_true {st.rel_true:.2e}")
            del s
        extra["pcg_speedup_mixed_vs_fp64"] = round(extra["pcg_fp64"]["seconds"] / extra["pcg_mixed"]["seconds"], 3)

    if not args.no_pcg and not args.no_learned:
        extra["pcg_mixed_learned"] = bench_learned(meshes, trans, N, dev, extra)

    if not args.no_multi:
        extra["multi_rhs"] = bench_multi(args, meshes, trans, N, dev, stream, extra)

    lane = "true" if mf.tile_info()["lane_stencils"] else "false"
    kernel = f"dgh::tile_op_kernel<double, double, false, 0, {tuned['block']}, false, {lane}>"
    traffic = measured_traffic(kernel, mf.tile_info()["tile_elems"], tuned["block"])
    out = {
        "metric": METRIC,
        "value": round(gdofs, 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Jacobian2 MatVec + 5-level GMG-PCG, octree from 1M log-normal points, "
                               "max depth 12, kappa=10^U(-1,1), FP64",
                   "n_elem": E, "n_node": N, "n_levels": len(levels), "n_hanging_elem": F.n_hanging},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "kernel": kernel,
                     "tile_elems": mf.tile_info()["tile_elems"], "block": tuned["block"],
                     "n_tiles": mf.tile_info()["n_tiles"], "n_records": mf.tile_info()["n_records"],
                     "cold_cache_us": round(cold_ms * 1e3, 2),
                     "cold_cache_frac": round(alg_bytes / (cold_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "matvec_fp32_ms": round(ms32, 5),
        "matvec_fp32_gdofs": round(N / (ms32 * 1e-3) / 1e9, 4),
        **dirichlet,
    }
    out.update(extra)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(F, coefs[-1][0])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
